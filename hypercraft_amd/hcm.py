"""ctypes bindings for libhcm.so (the C ABI in include/hcm.h).

Python is plumbing here: it drives the C ABI from the tests and the benchmark. The product
is libhcm.so, whose meshing runs only in the gfx950 kernels; if the library is missing or no
gfx950 device is visible, every entry point that meshes raises instead of falling back.
"""
import ctypes
import os

import numpy as np

from . import build as _build

# HCM_LIB selects another build of the library (for example an instrumented one)
LIB_PATH = os.environ.get("HCM_LIB", _build.LIB)

HCM_OK = 0
HCM_ERR_INVALID = -1
HCM_ERR_HIP = -2
HCM_ERR_POOL_FULL = -3
HCM_ERR_CAPACITY = -4
HCM_ERR_NO_DEVICE = -5
HCM_ERR_BUSY = -6

HALO_BLOCKS = 20 * 20 * 18  # blocks over x, z in [-2, 18), y in [-1, 17): index (x+2) + 20 ((z+2) + 20 (y+1))
HALO_LIGHT = 18 * 18 * 18   # light over [-1, 17)^3: index (x+1) + 18 ((z+1) + 18 (y+1))

NOALLOC_MAX_CHUNKS = 8  # HCM_NOALLOC_MAX_CHUNKS

RUN_PERSISTENT = 0
RUN_PER_CHUNK = 1
LIGHT_STORE = 0    # mesh with the store's light bytes
LIGHT_REBUILD = 1  # rebuild light from blocks and sunlight heights first (ChunkMeshTask.cpp:48-111)

HCM_GEN_PLANTS = 1


class MeshVertexC(ctypes.Structure):
    _fields_ = [("x", ctypes.c_uint8), ("y", ctypes.c_uint8), ("z", ctypes.c_uint8), ("ao", ctypes.c_uint8)]


class MeshFaceC(ctypes.Structure):
    _fields_ = [("axis", ctypes.c_uint8), ("light_face", ctypes.c_uint8), ("render_face", ctypes.c_uint8),
                ("pad", ctypes.c_uint8), ("texture", ctypes.c_uint16), ("pad2", ctypes.c_uint16),
                ("v", MeshVertexC * 4)]


class CustomMeshC(ctypes.Structure):
    _fields_ = [("face_count", ctypes.c_uint32), ("kind", ctypes.c_uint32), ("flow_height", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32), ("faces", MeshFaceC * 16)]


class BlockRowC(ctypes.Structure):
    _fields_ = [("tex", ctypes.c_uint16 * 6), ("transparency", ctypes.c_uint8), ("collision", ctypes.c_uint8),
                ("mesh", ctypes.c_uint16)]


class RegistryC(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_uint32), ("n_meshes", ctypes.c_uint32), ("n_tex", ctypes.c_uint32),
                ("row_of", ctypes.POINTER(ctypes.c_uint16)), ("rows", ctypes.POINTER(BlockRowC)),
                ("meshes", ctypes.POINTER(CustomMeshC)), ("tex_flags", ctypes.POINTER(ctypes.c_uint8)),
                ("liquid_source_id", ctypes.c_uint8), ("liquid_flow_id", ctypes.c_uint8),
                ("liquid_flow_tex", ctypes.c_uint16), ("light_level", ctypes.POINTER(ctypes.c_uint8))]


class MeshDescC(ctypes.Structure):
    _fields_ = [("vert_begin", ctypes.c_uint32), ("vert_count", ctypes.c_uint32), ("idx_begin", ctypes.c_uint32),
                ("idx_count", ctypes.c_uint32), ("aabb", ctypes.c_uint32 * 6), ("transparent", ctypes.c_uint32)]


class HostResultC(ctypes.Structure):
    _fields_ = [("vertices", ctypes.c_void_p), ("vertex_capacity", ctypes.c_uint32),
                ("indices", ctypes.c_void_p), ("index_capacity", ctypes.c_uint32),
                ("meshes", ctypes.POINTER(MeshDescC)), ("mesh_capacity", ctypes.c_uint32),
                ("vertex_count", ctypes.c_uint32), ("index_count", ctypes.c_uint32), ("mesh_count", ctypes.c_uint32)]


class JobResultC(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("first_mesh", ctypes.c_uint32), ("mesh_count", ctypes.c_uint32),
                ("quad_count", ctypes.c_uint32)]


class TuningC(ctypes.Structure):
    _fields_ = [("wg_per_cu", ctypes.c_uint32), ("threads", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 2)]


class ServerConfigC(ctypes.Structure):
    _fields_ = [("max_chunks", ctypes.c_uint32), ("max_batch", ctypes.c_uint32), ("max_inflight", ctypes.c_uint32),
                ("threads", ctypes.c_uint32), ("pool_quads", ctypes.c_uint64), ("keep_history", ctypes.c_uint32),
                ("light_mode", ctypes.c_uint32), ("reserved", ctypes.c_uint32 * 2)]


class CompletionC(ctypes.Structure):
    _fields_ = [("tag", ctypes.c_uint64), ("kind", ctypes.c_uint32), ("remeshed", ctypes.c_uint32),
                ("epoch", ctypes.c_uint64), ("latency_ms", ctypes.c_double)]


class ServerStatsC(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("requests", "completed", "batches", "chunks_meshed", "quads_meshed",
                                               "retries", "pool_used_quads", "pool_free_quads")] + \
               [("gpu_ms", ctypes.c_double)]


REQ_LOAD = 1
REQ_EDIT = 2

JOB_RESULT_DTYPE = np.dtype([("status", "<i4"), ("first_mesh", "<u4"), ("mesh_count", "<u4"), ("quad_count", "<u4")])
MESH_INFO_DTYPE = np.dtype([("index_count", "<u4"), ("first_index", "<u4"), ("vertex_offset", "<u4"),
                            ("aabb", "<f4", (6,)), ("base_position", "<i4", (3,)), ("transparent", "<u4")])
DRAW_CMD_DTYPE = np.dtype([("index_count", "<u4"), ("instance_count", "<u4"), ("first_index", "<u4"),
                           ("vertex_offset", "<u4"), ("first_instance", "<u4")])
MESH_REC_DTYPE = np.dtype([("aabb", "<u4", (6,)), ("first_quad", "<u4"), ("quad_count", "<u4"), ("job", "<u4"),
                           ("transparent", "<u4")])

# (name, restype, argtypes) for every entry point of include/hcm.h and include/hcm_workload.h
_VP = ctypes.c_void_p
SIGNATURES = [
    ("hcm_registry_default", ctypes.POINTER(RegistryC), []),
    ("hcm_create", ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.POINTER(_VP)]),
    ("hcm_destroy", None, [_VP]),
    ("hcm_last_hip_error", ctypes.c_int, []),
    ("hcm_device_ok", ctypes.c_int, []),
    ("hcm_mesh_chunk", ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(HostResultC)]),
    ("hcm_mesh_halo", ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(HostResultC)]),
    ("hcm_mesh_halos_noalloc", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_uint32, _VP, _VP, ctypes.c_uint32, _VP,
                                              ctypes.c_uint32, _VP, _VP]),
    ("hcm_mesh_chunk_relight", ctypes.c_int, [_VP, _VP, _VP, ctypes.POINTER(HostResultC), _VP]),
    ("hcm_store_create", ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.POINTER(_VP)]),
    ("hcm_store_destroy", None, [_VP]),
    ("hcm_store_upload", ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_uint32, _VP, _VP]),
    ("hcm_store_device_ptrs", ctypes.c_int, [_VP, ctypes.POINTER(_VP), ctypes.POINTER(_VP)]),
    ("hcm_store_upload_sunlight", ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.c_uint32, _VP]),
    ("hcm_batch_create", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_uint32, ctypes.POINTER(_VP)]),
    ("hcm_batch_destroy", None, [_VP]),
    ("hcm_batch_run", ctypes.c_int, [_VP, ctypes.c_uint32]),
    ("hcm_batch_run_repeat", ctypes.c_int, [_VP, ctypes.c_uint32]),
    ("hcm_batch_sync", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_float)]),
    ("hcm_batch_results", ctypes.c_int, [_VP, _VP]),
    ("hcm_batch_fetch", ctypes.c_int, [_VP, ctypes.c_uint32, ctypes.POINTER(HostResultC)]),
    ("hcm_batch_totals", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64)]),
    ("hcm_batch_download", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_uint32, _VP, _VP, ctypes.c_uint64]),
    ("hcm_batch_set_light", ctypes.c_int, [_VP, ctypes.c_uint32]),
    ("hcm_batch_mesh_infos", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    ("hcm_batch_cull", ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32), _VP,
                                      ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    ("hcm_batch_cull_occlusion", ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32),
                                                _VP, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint32)]),
    ("hcm_ctx_set_depth", ctypes.c_int, [_VP, _VP, ctypes.c_uint32, ctypes.c_uint32]),
    ("hcm_ctx_depth_level", ctypes.c_int, [_VP, ctypes.c_uint32, _VP, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint32),
                                           ctypes.POINTER(ctypes.c_uint32)]),
    ("hcm_batch_fetch_light", ctypes.c_int, [_VP, ctypes.c_uint32, _VP]),
    ("hcm_set_tuning", ctypes.c_int, [_VP, ctypes.POINTER(TuningC)]),
    ("hcm_get_tuning", ctypes.c_int, [_VP, ctypes.POINTER(TuningC)]),
    ("hcm_server_create", ctypes.c_int, [_VP, ctypes.POINTER(ServerConfigC), ctypes.POINTER(_VP)]),
    ("hcm_server_destroy", None, [_VP]),
    ("hcm_server_load", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _VP, _VP,
                                       ctypes.c_uint64]),
    ("hcm_server_load_sunlit", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _VP, _VP,
                                              ctypes.c_uint64]),
    ("hcm_server_edit", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint16,
                                       ctypes.c_uint8, ctypes.c_uint64]),
    ("hcm_server_poll", ctypes.c_int, [_VP, ctypes.POINTER(CompletionC), ctypes.c_uint32,
                                       ctypes.POINTER(ctypes.c_uint32)]),
    ("hcm_server_drain", ctypes.c_int, [_VP]),
    ("hcm_server_chunk_mesh", ctypes.c_int, [_VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.POINTER(HostResultC)]),
    ("hcm_server_get_stats", ctypes.c_int, [_VP, ctypes.POINTER(ServerStatsC)]),
    ("hcm_server_history_size", ctypes.c_int, [_VP, ctypes.POINTER(ctypes.c_uint64)]),
    ("hcm_server_history_get", ctypes.c_int, [_VP, ctypes.c_uint64, ctypes.POINTER(ctypes.c_int32),
                                              ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(HostResultC)]),
    ("hcm_gen_terrain_chunk", None, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_uint32, _VP, _VP]),
    ("hcm_gen_cave_chunk", None, [ctypes.c_uint64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                  ctypes.c_uint32, _VP, _VP]),
    ("hcm_gen_random_chunk", None, [ctypes.c_uint64, ctypes.c_uint64, _VP, _VP]),
    ("hcm_gen_opaque_values", ctypes.c_uint32, [_VP, ctypes.c_uint32]),
]

_lib = None


def lib():
    """Loads libhcm.so (built by hypercraft_amd/build.py). Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        l = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class HcmError(RuntimeError):
    def __init__(self, what, code):
        hip = lib().hcm_last_hip_error() if code == HCM_ERR_HIP else 0
        super().__init__(f"{what} failed: status {code}" + (f" (hip error {hip})" if hip else ""))
        self.code = code


def check(what, code):
    if code != HCM_OK:
        raise HcmError(what, code)


def device_ok():
    return bool(lib().hcm_device_ok())


def require_device():
    if not device_ok():
        raise RuntimeError("no gfx950 device visible: the HIP mesher has no CPU fallback")


class Mesh:
    """One BlockMesh: vertices (n, 2) u32, indices (m,) u16, aabb (6-tuple), transparent flag."""

    __slots__ = ("vertices", "indices", "aabb", "transparent")

    def __init__(self, vertices, indices, aabb, transparent):
        self.vertices = vertices
        self.indices = indices
        self.aabb = aabb
        self.transparent = transparent


def _ptrs27(arrs, dtype, size=4096):
    arrs = [np.ascontiguousarray(a, dtype=dtype).reshape(-1) for a in arrs]
    assert len(arrs) == 27 and all(a.size == size for a in arrs)
    p = (ctypes.c_void_p * 27)(*[a.ctypes.data for a in arrs])
    return p, arrs


class _HostOut:
    def __init__(self, vcap=1 << 19, mcap=64):
        self.v = np.zeros((vcap, 2), np.uint32)
        self.i = np.zeros(vcap * 2, np.uint16)
        self.m = (MeshDescC * mcap)()
        self.c = HostResultC(self.v.ctypes.data, vcap, self.i.ctypes.data, vcap * 2,
                             ctypes.cast(self.m, ctypes.POINTER(MeshDescC)), mcap, 0, 0, 0)

    def meshes(self):
        out = []
        for k in range(self.c.mesh_count):
            d = self.m[k]
            out.append(Mesh(self.v[d.vert_begin:d.vert_begin + d.vert_count].copy(),
                            self.i[d.idx_begin:d.idx_begin + d.idx_count].copy(),
                            tuple(int(a) for a in d.aabb), int(d.transparent)))
        return out


class Context:
    """An hcm_ctx: registry on the device plus a quad pool of pool_quads quads."""

    def __init__(self, pool_quads=1 << 22, registry=None):
        require_device()
        l = lib()
        self._ptr = ctypes.c_void_p()
        reg = registry if registry is not None else l.hcm_registry_default()
        check("hcm_create", l.hcm_create(ctypes.cast(reg, ctypes.c_void_p), pool_quads, ctypes.byref(self._ptr)))
        self._out = _HostOut()

    @property
    def ptr(self):
        return self._ptr

    def set_depth(self, depth):
        """The depth hierarchy for occlusion culling, from a (height, width) float32 depth image."""
        d = np.ascontiguousarray(depth, np.float32)
        check("hcm_ctx_set_depth", lib().hcm_ctx_set_depth(self._ptr, d.ctypes.data, d.shape[1], d.shape[0]))

    def depth_level(self, level):
        """Level `level` of the depth hierarchy as a (height, width) float32 array."""
        w, h = ctypes.c_uint32(), ctypes.c_uint32()
        check("hcm_ctx_depth_level", lib().hcm_ctx_depth_level(self._ptr, level, None, 0, ctypes.byref(w),
                                                               ctypes.byref(h)))
        out = np.zeros((h.value, w.value), np.float32)
        check("hcm_ctx_depth_level", lib().hcm_ctx_depth_level(self._ptr, level, out.ctypes.data, out.size,
                                                               ctypes.byref(w), ctypes.byref(h)))
        return out

    def close(self):
        if self._ptr:
            lib().hcm_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tuning(self, wg_per_cu=0, threads=256):
        t = TuningC(wg_per_cu, threads, (ctypes.c_uint32 * 2)(0, 0))
        check("hcm_set_tuning", lib().hcm_set_tuning(self._ptr, ctypes.byref(t)))

    def get_tuning(self):
        t = TuningC()
        check("hcm_get_tuning", lib().hcm_get_tuning(self._ptr, ctypes.byref(t)))
        return {"wg_per_cu": t.wg_per_cu, "threads": t.threads}

    def mesh_chunk(self, blocks27, light27):
        """Plain path: meshes slot 26 of 27 host chunks; returns the reference's list of BlockMesh."""
        pb, keep_b = _ptrs27(blocks27, np.uint16)
        pl, keep_l = _ptrs27(light27, np.uint8)
        check("hcm_mesh_chunk", lib().hcm_mesh_chunk(self._ptr, pb, pl, ctypes.byref(self._out.c)))
        return self._out.meshes()

    def mesh_halo(self, halo_blocks, halo_light):
        """Plain path from a packed halo (halo_of_chunks): the drop-in call's 20 KB upload."""
        b = np.ascontiguousarray(halo_blocks, np.uint16).reshape(HALO_BLOCKS)
        l = np.ascontiguousarray(halo_light, np.uint8).reshape(HALO_LIGHT)
        check("hcm_mesh_halo", lib().hcm_mesh_halo(self._ptr, b.ctypes.data, l.ctypes.data, ctypes.byref(self._out.c)))
        return self._out.meshes()

    def mesh_chunk_relight(self, blocks27, heights27):
        """The whole mesh task: light rebuilt from 27 host chunks' blocks and sunlight heights (256 int8
        each), then meshed. Returns (meshes, rebuilt light over [-1, 17)^3 as (18, 18, 18) u8 [y, z, x])."""
        pb, keep_b = _ptrs27(blocks27, np.uint16)
        ph, keep_h = _ptrs27(heights27, np.int8, size=256)
        light = np.zeros(18 ** 3, np.uint8)
        check("hcm_mesh_chunk_relight", lib().hcm_mesh_chunk_relight(self._ptr, pb, ph, ctypes.byref(self._out.c),
                                                                      light.ctypes.data))
        return self._out.meshes(), light.reshape(18, 18, 18)


class Store:
    """Device-resident chunk slots (hcm_store)."""

    def __init__(self, ctx, n_slots):
        self.ctx = ctx
        self._ptr = ctypes.c_void_p()
        self.n_slots = n_slots
        check("hcm_store_create", lib().hcm_store_create(ctx.ptr, n_slots, ctypes.byref(self._ptr)))

    @property
    def ptr(self):
        return self._ptr

    def upload(self, blocks, light, first=0):
        blocks = np.ascontiguousarray(blocks, np.uint16).reshape(-1, 4096)
        light = np.ascontiguousarray(light, np.uint8).reshape(-1, 4096)
        assert blocks.shape == light.shape
        check("hcm_store_upload", lib().hcm_store_upload(self._ptr, first, blocks.shape[0], blocks.ctypes.data,
                                                          light.ctypes.data))

    def upload_sunlight(self, heights, first=0):
        """Sunlight heights (n, 256) int8 per chunk (Chunk::m_sunlight_heights, index x + 16 z)."""
        heights = np.ascontiguousarray(heights, np.int8).reshape(-1, 256)
        check("hcm_store_upload_sunlight", lib().hcm_store_upload_sunlight(self._ptr, first, heights.shape[0],
                                                                            heights.ctypes.data))

    def close(self):
        if self._ptr:
            lib().hcm_store_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A set of chunk jobs (27 slot indices each) over a Store (hcm_batch)."""

    def __init__(self, ctx, store, nbr):
        self.ctx = ctx
        self.store = store
        nbr = np.ascontiguousarray(nbr, np.uint32).reshape(-1, 27)
        self.n_jobs = nbr.shape[0]
        self._ptr = ctypes.c_void_p()
        check("hcm_batch_create", lib().hcm_batch_create(ctx.ptr, store.ptr, nbr.ctypes.data, self.n_jobs,
                                                          ctypes.byref(self._ptr)))

    def run(self, mode=RUN_PERSISTENT):
        check("hcm_batch_run", lib().hcm_batch_run(self._ptr, mode))

    def run_repeat(self, times):
        """`times` persistent passes back to back; sync() then times all of them together."""
        check("hcm_batch_run_repeat", lib().hcm_batch_run_repeat(self._ptr, times))

    def set_light(self, mode):
        """LIGHT_STORE: mesh with the store's light bytes; LIGHT_REBUILD: rebuild light first."""
        check("hcm_batch_set_light", lib().hcm_batch_set_light(self._ptr, mode))

    def mesh_infos(self, chunk_pos):
        """The mesh info of every mesh record of the last run (chunk_pos: (n_jobs, 3) chunk coordinates)."""
        pos = np.ascontiguousarray(chunk_pos, np.int32).reshape(self.n_jobs, 3)
        n = ctypes.c_uint32()
        check("hcm_batch_mesh_infos", lib().hcm_batch_mesh_infos(self._ptr, pos.ctypes.data, None, 0, ctypes.byref(n)))
        out = np.zeros(max(n.value, 1), MESH_INFO_DTYPE)
        check("hcm_batch_mesh_infos", lib().hcm_batch_mesh_infos(self._ptr, pos.ctypes.data, out.ctypes.data,
                                                                  out.shape[0], ctypes.byref(n)))
        return out[:n.value]

    def cull(self, view_position, view_projection, cap=1 << 20, occlusion=False):
        """Frustum culling of the last mesh_infos, and with occlusion=True the depth-hierarchy test against
        the context's depth (Context.set_depth): (opaque, transparent) draw commands."""
        vp = np.ascontiguousarray(view_position, np.float32).reshape(3)
        m = np.ascontiguousarray(view_projection, np.float32).reshape(16)
        o = np.zeros(cap, DRAW_CMD_DTYPE)
        t = np.zeros(cap, DRAW_CMD_DTYPE)
        no, nt = ctypes.c_uint32(), ctypes.c_uint32()
        name = "hcm_batch_cull_occlusion" if occlusion else "hcm_batch_cull"
        check(name, getattr(lib(), name)(self._ptr, vp.ctypes.data, m.ctypes.data, o.ctypes.data, cap,
                                         ctypes.byref(no), t.ctypes.data, cap, ctypes.byref(nt)))
        return o[:no.value], t[:nt.value]

    def fetch_light(self, job):
        """One job's rebuilt light over [-1, 17)^3 after a LIGHT_REBUILD run: (18, 18, 18) u8 [y, z, x]."""
        out = np.zeros(18 ** 3, np.uint8)
        check("hcm_batch_fetch_light", lib().hcm_batch_fetch_light(self._ptr, job, out.ctypes.data))
        return out.reshape(18, 18, 18)

    def sync(self):
        ms = ctypes.c_float()
        check("hcm_batch_sync", lib().hcm_batch_sync(self._ptr, ctypes.byref(ms)))
        return ms.value

    def results(self):
        r = np.zeros(self.n_jobs, JOB_RESULT_DTYPE)
        check("hcm_batch_results", lib().hcm_batch_results(self._ptr, r.ctypes.data))
        return r

    def totals(self):
        q, m = ctypes.c_uint64(), ctypes.c_uint64()
        check("hcm_batch_totals", lib().hcm_batch_totals(self._ptr, ctypes.byref(q), ctypes.byref(m)))
        return q.value, m.value

    def fetch(self, job, out=None):
        out = out or _HostOut()
        check("hcm_batch_fetch", lib().hcm_batch_fetch(self._ptr, job, ctypes.byref(out.c)))
        return out.meshes()

    def download(self):
        """Everything of the last run: (results, mesh records, vertices (Q*4, 2) u32, indices (Q*6,) u16)."""
        quads, meshes = self.totals()
        res = np.zeros(self.n_jobs, JOB_RESULT_DTYPE)
        recs = np.zeros(max(meshes, 1), MESH_REC_DTYPE)
        verts = np.zeros((max(quads, 1) * 4, 2), np.uint32)
        idx = np.zeros(max(quads, 1) * 6, np.uint16)
        check("hcm_batch_download", lib().hcm_batch_download(self._ptr, res.ctypes.data, recs.ctypes.data,
                                                              recs.shape[0], verts.ctypes.data, idx.ctypes.data,
                                                              max(quads, 1)))
        return res, recs[:meshes], verts[:quads * 4], idx[:quads * 6]

    def close(self):
        if self._ptr:
            lib().hcm_batch_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous batching of chunk loads and voxel edits (hcm_server)."""

    def __init__(self, ctx, max_chunks, max_batch=4096, max_inflight=2, threads=0, pool_quads=1 << 24,
                 keep_history=False, light_mode=LIGHT_STORE):
        self.ctx = ctx
        cfg = ServerConfigC(max_chunks, max_batch, max_inflight, threads, pool_quads, int(keep_history),
                            light_mode, (ctypes.c_uint32 * 2)(0, 0))
        self._ptr = ctypes.c_void_p()
        check("hcm_server_create", lib().hcm_server_create(ctx.ptr, ctypes.byref(cfg), ctypes.byref(self._ptr)))
        self._out = _HostOut()
        self._done = (CompletionC * 4096)()

    def load(self, cx, cy, cz, blocks, light, tag=0):
        b = np.ascontiguousarray(blocks, np.uint16).reshape(4096)
        l = np.ascontiguousarray(light, np.uint8).reshape(4096)
        check("hcm_server_load", lib().hcm_server_load(self._ptr, cx, cy, cz, b.ctypes.data, l.ctypes.data, tag))

    def load_sunlit(self, cx, cy, cz, blocks, heights, tag=0):
        """A LIGHT_REBUILD server's load: blocks (4096 u16) and sunlight heights (256 int8)."""
        b = np.ascontiguousarray(blocks, np.uint16).reshape(4096)
        h = np.ascontiguousarray(heights, np.int8).reshape(256)
        check("hcm_server_load_sunlit", lib().hcm_server_load_sunlit(self._ptr, cx, cy, cz, b.ctypes.data,
                                                                      h.ctypes.data, tag))

    def edit(self, x, y, z, block, light, tag=0):
        check("hcm_server_edit", lib().hcm_server_edit(self._ptr, x, y, z, block, light, tag))

    def poll(self):
        """Completed requests: list of dicts (tag, kind, remeshed, epoch, latency_ms)."""
        out = []
        while True:
            n = ctypes.c_uint32()
            check("hcm_server_poll", lib().hcm_server_poll(self._ptr, self._done, len(self._done), ctypes.byref(n)))
            for k in range(n.value):
                c = self._done[k]
                out.append({"tag": c.tag, "kind": c.kind, "remeshed": c.remeshed, "epoch": c.epoch,
                            "latency_ms": c.latency_ms})
            if n.value < len(self._done):
                return out

    def drain(self):
        check("hcm_server_drain", lib().hcm_server_drain(self._ptr))

    def chunk_mesh(self, cx, cy, cz):
        check("hcm_server_chunk_mesh", lib().hcm_server_chunk_mesh(self._ptr, cx, cy, cz, ctypes.byref(self._out.c)))
        return self._out.meshes()

    def stats(self):
        st = ServerStatsC()
        check("hcm_server_get_stats", lib().hcm_server_get_stats(self._ptr, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in ServerStatsC._fields_}

    def history(self):
        """Every (chunk position, epoch, meshes) result recorded with keep_history."""
        n = ctypes.c_uint64()
        check("hcm_server_history_size", lib().hcm_server_history_size(self._ptr, ctypes.byref(n)))
        out = []
        pos = (ctypes.c_int32 * 3)()
        ep = ctypes.c_uint64()
        for i in range(n.value):
            check("hcm_server_history_get", lib().hcm_server_history_get(self._ptr, i, pos, ctypes.byref(ep),
                                                                          ctypes.byref(self._out.c)))
            out.append((tuple(pos), ep.value, self._out.meshes()))
        return out

    def close(self):
        if self._ptr:
            lib().hcm_server_destroy(self._ptr)
            self._ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def halo_of_chunks(blocks27, light27):
    """The packed halo (include/hcm.h, HCM_HALO_BLOCKS / HCM_HALO_LIGHT) of the chunk in slot 26 of 27
    chunks in neighbour-slot order: (blocks (18, 20, 20) u16 [y, z, x], light (18, 18, 18) u8 [y, z, x])."""
    lut = {-1: 1, 0: 2, 1: 0}
    b = np.zeros((18, 20, 20), np.uint16)
    l = np.zeros((18, 18, 18), np.uint8)
    for cy in (-1, 0, 1):
        for cz in (-1, 0, 1):
            for cx in (-1, 0, 1):
                slot = 9 * lut[cx] + 3 * lut[cy] + lut[cz]
                cb = np.asarray(blocks27[slot], np.uint16).reshape(16, 16, 16)
                cl = np.asarray(light27[slot], np.uint8).reshape(16, 16, 16)
                # world coordinate w of local c in this chunk: w = 16 * chunk + c
                def span(c, lo, hi):
                    w0, w1 = max(lo, 16 * c), min(hi, 16 * c + 16)
                    return (slice(w0 - lo, w1 - lo), slice(w0 - 16 * c, w1 - 16 * c)) if w0 < w1 else None
                sy, sz, sx = span(cy, -1, 17), span(cz, -2, 18), span(cx, -2, 18)
                if sy and sz and sx:
                    b[sy[0], sz[0], sx[0]] = cb[sy[1], sz[1], sx[1]]
                sz, sx = span(cz, -1, 17), span(cx, -1, 17)
                if sy and sz and sx:
                    l[sy[0], sz[0], sx[0]] = cl[sy[1], sz[1], sx[1]]
    return b, l


def job_meshes(job, res, recs, verts, idx):
    """The meshes of one job from a download(): list of Mesh in the reference's order."""
    r = res[job]
    if r["status"] != HCM_OK:
        raise HcmError(f"job {job}", int(r["status"]))
    out = []
    for m in recs[r["first_mesh"]:r["first_mesh"] + r["mesh_count"]]:
        q0, n = int(m["first_quad"]), int(m["quad_count"])
        out.append(Mesh(verts[4 * q0:4 * (q0 + n)], idx[6 * q0:6 * (q0 + n)],
                        tuple(int(a) for a in m["aabb"]), int(m["transparent"])))
    return out


# ---- workload generators (include/hcm_workload.h) ----
def gen_terrain_chunk(seed, cx, cy, cz, flags=0):
    b = np.zeros(4096, np.uint16)
    l = np.zeros(4096, np.uint8)
    lib().hcm_gen_terrain_chunk(seed, cx, cy, cz, flags, b.ctypes.data, l.ctypes.data)
    return b, l


def gen_cave_chunk(seed, cx, cy, cz, flags=0):
    b = np.zeros(4096, np.uint16)
    l = np.zeros(4096, np.uint8)
    lib().hcm_gen_cave_chunk(seed, cx, cy, cz, flags, b.ctypes.data, l.ctypes.data)
    return b, l


def gen_random_chunk(seed, index):
    b = np.zeros(4096, np.uint16)
    l = np.zeros(4096, np.uint8)
    lib().hcm_gen_random_chunk(seed, index, b.ctypes.data, l.ctypes.data)
    return b, l


def opaque_values():
    out = np.zeros(64, np.uint16)
    n = lib().hcm_gen_opaque_values(out.ctypes.data, 64)
    return out[:n]


def registry_arrays():
    """The built-in registry as numpy arrays (for checks against the exported reference table)."""
    r = lib().hcm_registry_default().contents
    row_of = np.ctypeslib.as_array(r.row_of, shape=(65536,)).copy()
    rows = np.frombuffer(ctypes.string_at(r.rows, 16 * r.n_rows), np.uint8).reshape(r.n_rows, 16).copy()
    meshes = np.frombuffer(ctypes.string_at(r.meshes, 400 * r.n_meshes), np.uint8).reshape(r.n_meshes, 400).copy()
    tex_flags = np.frombuffer(ctypes.string_at(r.tex_flags, r.n_tex), np.uint8).copy()
    light_level = np.ctypeslib.as_array(r.light_level, shape=(65536,)).copy()
    return {"row_of": row_of, "rows": rows, "meshes": meshes, "tex_flags": tex_flags,
            "liquid": (r.liquid_source_id, r.liquid_flow_id, r.liquid_flow_tex), "light_level": light_level}
