"""Seeded chunk workloads for the benchmark and the tests.

Each builder returns a store layout (blocks (S, 4096) u16, light (S, 4096) u8) and a job table
nbr (J, 27) u32 of store slots in the reference's neighbour-slot order
(common/include/common/Position.hpp:138-145). Voxels come from the integer generators of
include/hcm_workload.h, so every host produces identical inputs for a seed.
"""
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import hcm

# neighbour slot -> chunk offset (NeighbourIndex2CmpXYZ, Position.hpp:146-152: reverse of L = {1, -1, 0})
_REV = (1, -1, 0)
SLOT_OFFSETS = np.array([(_REV[s // 9], _REV[(s // 3) % 3], _REV[s % 3]) for s in range(27)], np.int32)

TERRAIN_SEED = 1
CAVE_SEED = 2
RANDOM_SEED = 3
MIXED_SEED = 4


def _gen_many(fn, seed, coords, flags=0, threads=8):
    coords = np.asarray(coords, np.int64).reshape(-1, 3)
    blocks = np.zeros((coords.shape[0], 4096), np.uint16)
    light = np.zeros((coords.shape[0], 4096), np.uint8)
    lib = hcm.lib()
    gen = getattr(lib, fn)

    def work(rng):
        for i in range(*rng):
            cx, cy, cz = (int(v) for v in coords[i])
            gen(seed, cx, cy, cz, flags, blocks[i].ctypes.data, light[i].ctypes.data)

    n = coords.shape[0]
    step = max(1, (n + threads - 1) // threads)
    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, [(i, min(n, i + step)) for i in range(0, n, step)]))
    return blocks, light


def region_jobs(dims, centers):
    """nbr table for jobs at chunk coords `centers` inside a grid of `dims` (dx, dy, dz) chunks whose
    slot index is x + dx * (z + dz * y) over grid coords starting at 0."""
    dx, dy, dz = dims
    c = np.asarray(centers, np.int64).reshape(-1, 1, 3) + SLOT_OFFSETS.reshape(1, 27, 3)
    assert c.min() >= 0 and (c[..., 0] < dx).all() and (c[..., 1] < dy).all() and (c[..., 2] < dz).all()
    return (c[..., 0] + dx * (c[..., 2] + dz * c[..., 1])).astype(np.uint32)


def _grid_coords(lo, dims):
    dx, dy, dz = dims
    ys, zs, xs = np.meshgrid(np.arange(dy), np.arange(dz), np.arange(dx), indexing="ij")
    g = np.stack([xs.ravel(), ys.ravel(), zs.ravel()], 1)  # slot order x + dx * (z + dz * y)
    return g + np.asarray(lo, np.int64)


def terrain_chunk(seed=TERRAIN_SEED, flags=0):
    """One 16^3 terrain chunk with its 26 neighbours: the 48^3 voxels x, z in [-16, 32), y in [0, 48),
    whose heightmap surface (y in [8, 40]) crosses the centre chunk (chunk (0, 1, 0))."""
    coords = _grid_coords((-1, 0, -1), (3, 3, 3))
    blocks, light = _gen_many("hcm_gen_terrain_chunk", seed, coords, flags)
    nbr = region_jobs((3, 3, 3), [(1, 1, 1)])
    return blocks, light, nbr


def terrain_region(n=(16, 16, 16), seed=TERRAIN_SEED, flags=0, y0=-5):
    """A region of terrain chunks (n chunks each way, lowest chunk y = y0) plus a one-chunk rim."""
    dims = (n[0] + 2, n[1] + 2, n[2] + 2)
    coords = _grid_coords((-1, y0 - 1, -1), dims)
    blocks, light = _gen_many("hcm_gen_terrain_chunk", seed, coords, flags)
    centers = _grid_coords((1, 1, 1), n)
    return blocks, light, region_jobs(dims, centers)


def cave_region(seed=CAVE_SEED, n=16, flags=0):
    """Cave terrain: a n^3-chunk region (chunks 0..n-1) plus a one-chunk rim, one job per region chunk."""
    dims = (n + 2, n + 2, n + 2)
    coords = _grid_coords((-1, -1, -1), dims)
    blocks, light = _gen_many("hcm_gen_cave_chunk", seed, coords, flags)
    centers = _grid_coords((1, 1, 1), (n, n, n))
    return blocks, light, region_jobs(dims, centers)


def random_dense(seed=RANDOM_SEED, n_jobs=1024):
    """Worst-case density: every job has its own 27 i.i.d. random chunks."""
    blocks = np.zeros((27 * n_jobs, 4096), np.uint16)
    light = np.zeros((27 * n_jobs, 4096), np.uint8)
    lib = hcm.lib()

    def work(rng):
        for i in range(*rng):
            lib.hcm_gen_random_chunk(seed, i, blocks[i].ctypes.data, light[i].ctypes.data)

    n = 27 * n_jobs
    step = (n + 7) // 8
    with ThreadPoolExecutor(8) as ex:
        list(ex.map(work, [(i, min(n, i + step)) for i in range(0, n, step)]))
    nbr = np.arange(27 * n_jobs, dtype=np.uint32).reshape(n_jobs, 27)
    return blocks, light, nbr


def slice_jobs(blocks, light, nbr, jobs):
    """The 27 host chunks of one job (for the plain path and the CPU oracle)."""
    slots = nbr[jobs]
    return [blocks[s] for s in slots], [light[s] for s in slots]
