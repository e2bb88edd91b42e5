"""hcm_mesh_halos_noalloc, the entry point that allocates nothing (needs an MI355X): called through the C
ABI on device buffers this test allocates, twice in a row (eight halos, then three others), its meshes
equal hcm_mesh_halo's for the same halos and the reference's digests; more than HCM_NOALLOC_MAX_CHUNKS
halos are refused."""
import ctypes
import os
import sys

import numpy as np
import pytest

import meshdigest
from conftest import GOLDEN
from hypercraft_amd import hcm, workloads

pytestmark = pytest.mark.gpu

CASES_A = [("terrain_chunk", 0), ("terrain_chunk_plants", 0), ("lone_stone", 0), ("slab", 0), ("uniform_flow", 0),
           ("palette_sparse_flow", 0), ("palette_sparse_flow", 4), ("palette_sparse_custom", 0)]
CASES_B = [("palette_sparse_glassy", 8), ("leaf_checkerboard", 0), ("palette_sparse_logs", 12)]
QUAD_CAP, REC_CAP = 1 << 16, 64


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Hip:
    """The few HIP runtime calls the test needs for its own device buffers."""

    def __init__(self):
        # the HIP runtime the library uses (a debugging build on the CPU brings its own)
        self.h = hcm.lib() if os.environ.get("HCM_LIB", "").endswith("libhcm_cpu.so") else ctypes.CDLL("libamdhip64.so")
        self.bufs = []

    def check(self, what, r):
        assert r == 0, f"{what}: hip error {r}"

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check("hipMalloc", self.h.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(nbytes, 1))))
        self.bufs.append(p)
        return p

    def put(self, p, arr):
        a = np.ascontiguousarray(arr)
        self.check("hipMemcpy", self.h.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1))

    def get(self, p, arr):
        self.check("hipMemcpy", self.h.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(arr.nbytes), 2))
        return arr

    def free_all(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def halos(cases, sets):
    out = []
    for name, j in cases:
        blocks, light, nbr = sets[name]
        out.append(hcm.halo_of_chunks(*workloads.slice_jobs(blocks, light, nbr, j)))
    return out


def test_noalloc_equals_mesh_halo_twice(ctx):
    import test_oracle

    sets = test_oracle.sets()
    golden = np.load(os.path.join(GOLDEN, "golden.npz"))
    hip = Hip()
    try:
        stream = ctypes.c_void_p()
        hip.check("hipStreamCreate", hip.h.hipStreamCreate(ctypes.byref(stream)))
        n_max = hcm.NOALLOC_MAX_CHUNKS
        d_hb = hip.alloc(n_max * hcm.HALO_BLOCKS * 2)
        d_hl = hip.alloc(n_max * hcm.HALO_LIGHT)
        d_v = hip.alloc(QUAD_CAP * 4 * 8)
        d_i = hip.alloc(QUAD_CAP * 6 * 2)
        d_r = hip.alloc(REC_CAP * hcm.MESH_REC_DTYPE.itemsize)
        d_res = hip.alloc(n_max * hcm.JOB_RESULT_DTYPE.itemsize)
        for label, cases in (("eight halos", CASES_A), ("three other halos", CASES_B)):
            hs = halos(cases, sets)
            n = len(hs)
            hip.put(d_hb, np.stack([b.reshape(-1) for b, _ in hs]))
            hip.put(d_hl, np.stack([l.reshape(-1) for _, l in hs]))
            say(f"no-alloc: {label}")
            r = hcm.lib().hcm_mesh_halos_noalloc(ctx.ptr, d_hb, d_hl, n, d_v, d_i, QUAD_CAP, d_r, REC_CAP, d_res, stream)
            assert r == hcm.HCM_OK, r
            hip.check("hipStreamSynchronize", hip.h.hipStreamSynchronize(stream))
            res = hip.get(d_res, np.zeros(n, hcm.JOB_RESULT_DTYPE))
            assert (res["status"] == hcm.HCM_OK).all(), res
            recs = hip.get(d_r, np.zeros(REC_CAP, hcm.MESH_REC_DTYPE))
            quads = int(res["quad_count"].sum())
            verts = hip.get(d_v, np.zeros((max(quads, 1) * 4, 2), np.uint32))
            idx = hip.get(d_i, np.zeros(max(quads, 1) * 6, np.uint16))
            for j, ((name, job), (hb, hl)) in enumerate(zip(cases, hs)):
                got = meshdigest.digest(hcm.job_meshes(j, res, recs, verts, idx))
                assert got == meshdigest.digest(ctx.mesh_halo(hb, hl)), f"{label}: {name} job {job}"
                assert got == golden[name + "_digest"][job].tobytes(), f"{label}: {name} job {job}"
        r = hcm.lib().hcm_mesh_halos_noalloc(ctx.ptr, d_hb, d_hl, n_max + 1, d_v, d_i, QUAD_CAP, d_r, REC_CAP, d_res,
                                             stream)
        assert r == hcm.HCM_ERR_INVALID
        hip.h.hipStreamDestroy(stream)
    finally:
        hip.free_all()
