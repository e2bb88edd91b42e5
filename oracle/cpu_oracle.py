"""ctypes front end for the CPU oracles (TEST INFRASTRUCTURE ONLY).

Loads oracle/_build/libhcm_oracle.so (the clean-room C restatement of
client/include/client/BlockMeshAlgo.hpp:494-521) and, when it was built in this
container, oracle/_ref/libhcref.so (the real reference mesher). Only tests/,
__graft_entry__.smoke() and bench.py's CPU-baseline leg may import this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libhcm_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libhcref.so")

MAX_MESHES = 64
VOX = 4096


class MeshOut(ctypes.Structure):
    _fields_ = [
        ("n_meshes", ctypes.c_uint32),
        ("vert_begin", ctypes.c_uint32 * MAX_MESHES),
        ("n_verts", ctypes.c_uint32 * MAX_MESHES),
        ("idx_begin", ctypes.c_uint32 * MAX_MESHES),
        ("n_idx", ctypes.c_uint32 * MAX_MESHES),
        ("aabb", (ctypes.c_uint32 * 6) * MAX_MESHES),
        ("transparent", ctypes.c_uint32 * MAX_MESHES),
        ("vert_cap", ctypes.c_uint32),
        ("idx_cap", ctypes.c_uint32),
        ("verts", ctypes.POINTER(ctypes.c_uint32)),
        ("idx", ctypes.POINTER(ctypes.c_uint16)),
    ]


class Registry(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_uint32),
        ("n_meshes", ctypes.c_uint32),
        ("n_tex", ctypes.c_uint32),
        ("row_of", ctypes.c_uint16 * 65536),
        ("rows", ctypes.c_void_p),
        ("meshes", ctypes.c_void_p),
        ("tex_flags", ctypes.c_uint8 * 8192),
        ("liquid_source_id", ctypes.c_uint8),
        ("liquid_flow_id", ctypes.c_uint8),
        ("liquid_flow_tex", ctypes.c_uint16),
    ]


class Mesh:
    """One BlockMesh: vertices (n, 2) u32, indices (m,) u16, aabb (6,), transparent flag."""

    __slots__ = ("vertices", "indices", "aabb", "transparent")

    def __init__(self, vertices, indices, aabb, transparent):
        self.vertices = vertices
        self.indices = indices
        self.aabb = aabb
        self.transparent = transparent


def _ptr_array(arrs, ctype):
    p = (ctypes.POINTER(ctype) * 27)()
    for i, a in enumerate(arrs):
        p[i] = a.ctypes.data_as(ctypes.POINTER(ctype))
    return p


class _Out:
    def __init__(self, vcap=1 << 20):
        self.verts = np.zeros(2 * vcap, np.uint32)
        self.idx = np.zeros(2 * vcap, np.uint16)
        self.s = MeshOut()
        self.s.vert_cap = vcap
        self.s.idx_cap = 2 * vcap
        self.s.verts = self.verts.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
        self.s.idx = self.idx.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))

    def meshes(self):
        s = self.s
        res = []
        for m in range(s.n_meshes):
            vb, nv, ib, ni = s.vert_begin[m], s.n_verts[m], s.idx_begin[m], s.n_idx[m]
            res.append(Mesh(self.verts[2 * vb:2 * (vb + nv)].reshape(-1, 2).copy(),
                            self.idx[ib:ib + ni].copy(),
                            tuple(int(v) for v in s.aabb[m]), int(s.transparent[m])))
        return res


class _Base:
    def _chunks(self, blocks27, light27):
        b = [np.ascontiguousarray(x, dtype=np.uint16).reshape(-1) for x in blocks27]
        l = [np.ascontiguousarray(x, dtype=np.uint8).reshape(-1) for x in light27]
        assert len(b) == 27 and len(l) == 27 and all(x.size == VOX for x in b + l)
        return b, l


class Oracle(_Base):
    """The clean-room C restatement."""

    def __init__(self, registry_path):
        self.lib = ctypes.CDLL(ORACLE_SO)
        self.lib.orc_registry_load.argtypes = [ctypes.c_char_p, ctypes.POINTER(Registry)]
        self.lib.orc_mesh_chunk.argtypes = [ctypes.POINTER(Registry), ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.POINTER(MeshOut)]
        self.lib.orc_mesh_many.argtypes = [ctypes.POINTER(Registry), ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_void_p]
        self.lib.orc_light_rebuild.argtypes = [ctypes.POINTER(Registry), ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        self.lib.orc_relight_many.argtypes = [ctypes.POINTER(Registry), ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32,
                                              ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        self.reg = Registry()
        r = self.lib.orc_registry_load(registry_path.encode(), ctypes.byref(self.reg))
        if r != 0:
            raise RuntimeError(f"orc_registry_load({registry_path}) failed: {r}")
        self.out = _Out()

    def mesh(self, blocks27, light27):
        b, l = self._chunks(blocks27, light27)
        r = self.lib.orc_mesh_chunk(ctypes.byref(self.reg), _ptr_array(b, ctypes.c_uint16),
                                    _ptr_array(l, ctypes.c_uint8), ctypes.byref(self.out.s))
        if r != 0:
            raise RuntimeError(f"orc_mesh_chunk failed: {r}")
        return self.out.meshes()

    def light_rebuild(self, light_level, blocks27, heights27):
        """The mesh task's light rebuild: -> (window (46, 46, 46) u8 indexed [y, z, x] over [-15, 31),
        halo (18, 18, 18) u8 indexed [y, z, x] over [-1, 17))."""
        b = [np.ascontiguousarray(x, dtype=np.uint16).reshape(-1) for x in blocks27]
        h = [np.ascontiguousarray(x, dtype=np.int8).reshape(-1) for x in heights27]
        assert len(b) == 27 and len(h) == 27 and all(x.size == VOX for x in b) and all(x.size == 256 for x in h)
        lv = np.ascontiguousarray(light_level, np.uint8)
        assert lv.size == 65536
        win = np.zeros(46 ** 3, np.uint8)
        halo = np.zeros(18 ** 3, np.uint8)
        r = self.lib.orc_light_rebuild(ctypes.byref(self.reg), lv.ctypes.data, _ptr_array(b, ctypes.c_uint16),
                                       _ptr_array(h, ctypes.c_int8), win.ctypes.data, halo.ctypes.data)
        if r != 0:
            raise RuntimeError(f"orc_light_rebuild failed: {r}")
        return win.reshape(46, 46, 46), halo.reshape(18, 18, 18)

    def relight_many(self, light_level, blocks, heights, nbr, n_threads, mesh=True, halo=False):
        """Light rebuild (and meshing) of every job: -> (halo (J, 18, 18, 18) u8 or None, quads (J,) or None)."""
        blocks = np.ascontiguousarray(blocks, np.uint16)
        heights = np.ascontiguousarray(heights, np.int8)
        nbr = np.ascontiguousarray(nbr, np.uint32)
        lv = np.ascontiguousarray(light_level, np.uint8)
        n = nbr.shape[0]
        h = np.zeros((n, 18, 18, 18), np.uint8) if halo else None
        q = np.zeros(n, np.uint64) if mesh else None
        r = self.lib.orc_relight_many(ctypes.byref(self.reg), lv.ctypes.data, blocks.ctypes.data, heights.ctypes.data,
                                      nbr.ctypes.data, n, n_threads, 1 if mesh else 0,
                                      h.ctypes.data if halo else None, q.ctypes.data if mesh else None)
        if r != 0:
            raise RuntimeError(f"orc_relight_many failed: {r}")
        return h, q

    def mesh_many(self, blocks, light, nbr, n_threads):
        """blocks (S, 4096) u16, light (S, 4096) u8, nbr (J, 27) u32 -> per-job quad counts."""
        blocks = np.ascontiguousarray(blocks, np.uint16)
        light = np.ascontiguousarray(light, np.uint8)
        nbr = np.ascontiguousarray(nbr, np.uint32)
        q = np.zeros(nbr.shape[0], np.uint64)
        r = self.lib.orc_mesh_many(ctypes.byref(self.reg), blocks.ctypes.data, light.ctypes.data, nbr.ctypes.data,
                                   nbr.shape[0], n_threads, q.ctypes.data)
        if r != 0:
            raise RuntimeError(f"orc_mesh_many failed: {r}")
        return q


class Reference(_Base):
    """The real reference mesher compiled from /root/reference (only where it was built)."""

    def __init__(self):
        self.lib = ctypes.CDLL(REF_SO)
        self.lib.hcref_mesh_chunk.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(MeshOut)]
        self.lib.hcref_mesh_many.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                             ctypes.c_uint32, ctypes.c_void_p]
        self.lib.hcref_export_registry.argtypes = [ctypes.c_char_p]
        self.lib.hcref_export_light_levels.argtypes = [ctypes.c_char_p]
        self.lib.hcref_light_rebuild.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        self.lib.hcref_relight_many.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint32,
                                                ctypes.c_uint32, ctypes.c_void_p]
        self.out = _Out()

    @staticmethod
    def available():
        return os.path.exists(REF_SO)

    def export_registry(self, path):
        r = self.lib.hcref_export_registry(path.encode())
        if r != 0:
            raise RuntimeError(f"hcref_export_registry failed: {r}")

    def light_levels(self):
        """Block::GetLightLevel for every Block::m_data value, u8[65536]."""
        import tempfile
        with tempfile.NamedTemporaryFile(suffix=".bin") as f:
            r = self.lib.hcref_export_light_levels(f.name.encode())
            if r != 0:
                raise RuntimeError(f"hcref_export_light_levels failed: {r}")
            return np.fromfile(f.name, np.uint8)

    def light_rebuild(self, blocks27, heights27):
        """The reference's light rebuild -> window (46, 46, 46) u8 indexed [y, z, x] over [-15, 31)."""
        b = [np.ascontiguousarray(x, dtype=np.uint16).reshape(-1) for x in blocks27]
        h = [np.ascontiguousarray(x, dtype=np.int8).reshape(-1) for x in heights27]
        win = np.zeros(46 ** 3, np.uint8)
        r = self.lib.hcref_light_rebuild(_ptr_array(b, ctypes.c_uint16), _ptr_array(h, ctypes.c_int8),
                                         win.ctypes.data)
        if r != 0:
            raise RuntimeError(f"hcref_light_rebuild failed: {r}")
        return win.reshape(46, 46, 46)

    def mesh(self, blocks27, light27):
        b, l = self._chunks(blocks27, light27)
        r = self.lib.hcref_mesh_chunk(_ptr_array(b, ctypes.c_uint16), _ptr_array(l, ctypes.c_uint8),
                                      ctypes.byref(self.out.s))
        if r != 0:
            raise RuntimeError(f"hcref_mesh_chunk failed: {r}")
        return self.out.meshes()

    def relight_many(self, light_level, blocks, heights, nbr, n_threads, mesh=True, halo=False):
        """The reference's whole mesh task per job (light rebuild, then meshing): -> (None, quads (J,)).
        Same call shape as Oracle.relight_many (the reference resolves light levels itself)."""
        assert mesh and not halo
        blocks = np.ascontiguousarray(blocks, np.uint16)
        heights = np.ascontiguousarray(heights, np.int8)
        nbr = np.ascontiguousarray(nbr, np.uint32)
        q = np.zeros(nbr.shape[0], np.uint64)
        self.lib.hcref_relight_many(blocks.ctypes.data, heights.ctypes.data, nbr.ctypes.data, nbr.shape[0],
                                    n_threads, q.ctypes.data)
        return None, q

    def mesh_many(self, blocks, light, nbr, n_threads):
        blocks = np.ascontiguousarray(blocks, np.uint16)
        light = np.ascontiguousarray(light, np.uint8)
        nbr = np.ascontiguousarray(nbr, np.uint32)
        q = np.zeros(nbr.shape[0], np.uint64)
        self.lib.hcref_mesh_many(blocks.ctypes.data, light.ctypes.data, nbr.ctypes.data, nbr.shape[0], n_threads,
                                 q.ctypes.data)
        return q


def meshes_equal(a, b):
    """Exact equality of two mesh lists (order, vertices, indices, aabb, transparent flag)."""
    if len(a) != len(b):
        return False
    for x, y in zip(a, b):
        if (x.transparent != y.transparent or x.aabb != y.aabb or not np.array_equal(x.vertices, y.vertices)
                or not np.array_equal(x.indices, y.indices)):
            return False
    return True
