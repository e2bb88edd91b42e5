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
TRACE_SEED = 5


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


def terrain_plants_region(seed=MIXED_SEED):
    """Heightmap terrain with grass plants (cross meshes) on 5% of grass tops: a 16 x 3 x 16-chunk
    region (chunk y 0..2 holds the surface band, y in [8, 40]) plus a one-chunk rim; 768 jobs."""
    return terrain_region(n=(16, 3, 16), seed=seed, flags=hcm.HCM_GEN_PLANTS, y0=0)


def mixed_pool():
    """One store holding the cave region and the plant-terrain region. Returns (blocks, light, nbr,
    n_cave): jobs [0, n_cave) are the cave region's, the rest the plant terrain's."""
    cb, cl, cn = cave_region()
    tb, tl, tn = terrain_plants_region()
    blocks = np.concatenate([cb, tb])
    light = np.concatenate([cl, tl])
    nbr = np.concatenate([cn, tn + np.uint32(cb.shape[0])])
    return blocks, light, nbr, cn.shape[0]


def mixed_batches(n_pool, seed=MIXED_SEED, n_batches=200, max_size=4096):
    """Uneven batches: sizes log-uniform in [1, max_size], jobs drawn without replacement from a pool."""
    rng = np.random.default_rng(seed)
    sizes = np.floor(np.exp(rng.uniform(0.0, np.log(max_size + 1), n_batches))).astype(np.int64)
    sizes = np.clip(sizes, 1, min(max_size, n_pool))
    return [np.sort(rng.choice(n_pool, size=int(k), replace=False)).astype(np.int64) for k in sizes]


def spiral_columns(radius):
    """Chunk columns (x, z) with x^2 + z^2 <= radius^2, nearest first, counter-clockwise within a ring."""
    r = int(radius)
    xs, zs = np.meshgrid(np.arange(-r, r + 1), np.arange(-r, r + 1), indexing="ij")
    xs, zs = xs.ravel(), zs.ravel()
    keep = xs * xs + zs * zs <= r * r
    xs, zs = xs[keep], zs[keep]
    order = np.lexsort((np.arctan2(zs, xs), xs * xs + zs * zs))
    return np.stack([xs[order], zs[order]], 1)


def spiral_trace(radius=24, height=16, y0=-7, load_rate=20000.0, edit_rate=2000.0, seed=TRACE_SEED):
    """A serving trace: terrain chunks loaded column by column in spiral order (Poisson arrivals at
    load_rate per second) while single-voxel edits arrive (Poisson, edit_rate per second) at random
    voxels of chunks whose 26 neighbours have already arrived.

    Returns dict(coords (N, 3), blocks (N, 4096), light (N, 4096), events: list of
    (time_s, kind, payload)) with kind "load" -> chunk index, "edit" -> (x, y, z, block, light)."""
    cols = spiral_columns(radius)
    ys = np.arange(y0, y0 + height)
    coords = np.array([(x, y, z) for x, z in cols for y in ys], np.int64)
    blocks, light = _gen_many("hcm_gen_terrain_chunk", seed, coords)
    n = coords.shape[0]
    rng = np.random.default_rng(seed)
    t_load = np.cumsum(rng.exponential(1.0 / load_rate, n))
    # arrival index of each chunk's last neighbour: it becomes meshable when that one has arrived
    index = {tuple(c): i for i, c in enumerate(coords.tolist())}
    ready = np.full(n, -1, np.int64)
    for i, c in enumerate(coords.tolist()):
        last = 0
        for d in SLOT_OFFSETS.tolist():
            j = index.get((c[0] + d[0], c[1] + d[1], c[2] + d[2]))
            if j is None:
                last = -1
                break
            last = max(last, j)
        ready[i] = last
    order = np.argsort(np.where(ready < 0, n + 1, ready), kind="stable")
    ready_sorted = np.where(ready < 0, n + 1, ready)[order]
    palette = np.array([0, 0, 0, 1, 1, 2, 3, 4, 9, 0x10, 0x210, 14, 5, 0x105, 19, 20, 23, 0x121], np.uint16)
    events = [(float(t), "load", i) for i, t in enumerate(t_load)]
    t, t_end = 0.0, float(t_load[-1])
    while True:
        t += rng.exponential(1.0 / edit_rate)
        if t >= t_end:
            break
        arrived = int(np.searchsorted(t_load, t))  # loads submitted before t
        k = int(np.searchsorted(ready_sorted, arrived - 1, side="right"))
        if k == 0:
            continue
        c = coords[order[rng.integers(k)]]
        v = rng.integers(0, 16, 3)
        events.append((t, "edit", (int(16 * c[0] + v[0]), int(16 * c[1] + v[1]), int(16 * c[2] + v[2]),
                                   int(palette[rng.integers(len(palette))]), int(rng.integers(256)))))
    events.sort(key=lambda e: e[0])
    return {"coords": coords, "blocks": blocks, "light": light, "events": events}


def slice_jobs(blocks, light, nbr, jobs):
    """The 27 host chunks of one job (for the plain path and the CPU oracle)."""
    slots = nbr[jobs]
    return [blocks[s] for s in slots], [light[s] for s in slots]


# ---- inputs of the light rebuild (sunlight heights per chunk, torchlight sources) ----
LIGHT_SEED = 6
GLOWSTONE = 0x0018  # the reference's one light-emitting block (block/block/glowstone.hpp: light level 15)


def sunlight_heights(blocks, dims):
    """Per chunk of a grid store (slot x + dx * (z + dz * y)) its 16 x 16 sunlight heights (InnerPos1, index
    x + 16 z, Chunk.hpp:65-76): one above the highest non-air voxel of the world column, in chunk-local
    units clamped to [0, 16] as the terrain generator clamps its light map (DefaultTerrain.cpp:148-152)."""
    dx, dy, dz = dims
    solid = (blocks.reshape(dy, dz, dx, 16, 16, 16) & 0xFF) != 0  # [cy, cz, cx, ly, lz, lx]
    col = solid.transpose(1, 2, 4, 5, 0, 3).reshape(dz, dx, 16, 16, dy * 16)  # [cz, cx, lz, lx, world y]
    top = np.where(col.any(-1), dy * 16 - np.argmax(col[..., ::-1], -1), 0)  # first sunlit world y
    cy = np.arange(dy).reshape(dy, 1, 1, 1, 1)
    h = np.clip(top[None] - 16 * cy, 0, 16).astype(np.int8)  # [cy, cz, cx, lz, lx]
    return np.ascontiguousarray(h.reshape(dy * dz * dx, 256))


def place_glowstone(blocks, seed, fraction):
    """Turns a seeded fraction of the air voxels into glowstone (torchlight level 15 sources)."""
    rng = np.random.default_rng(seed)
    out = blocks.copy()
    air = np.flatnonzero((out.reshape(-1) & 0xFF) == 0)
    pick = rng.choice(air, size=int(len(air) * fraction), replace=False) if len(air) else air
    out.reshape(-1)[pick] = GLOWSTONE
    return out


def lit_cave_region(seed=LIGHT_SEED, n=16, glow=0.0005):
    """The cave region with glowstone in a fraction of its air and sunlight heights from its columns:
    (blocks, heights, nbr) for the light rebuild before meshing (ChunkMeshTask.cpp:48-111)."""
    blocks, _, nbr = cave_region(n=n)
    blocks = place_glowstone(blocks, seed, glow)
    return blocks, sunlight_heights(blocks, (n + 2, n + 2, n + 2)), nbr


def lit_terrain_chunk(seed=LIGHT_SEED, glow=0.002):
    """The single terrain chunk neighbourhood with glowstone and sunlight heights (one job)."""
    blocks, _, nbr = terrain_chunk()
    blocks = place_glowstone(blocks, seed, glow)
    return blocks, sunlight_heights(blocks, (3, 3, 3)), nbr


def slice_heights(heights, nbr, job):
    """The 27 sunlight-height arrays of one job."""
    return [heights[s] for s in nbr[job]]


def region_job_positions(n=16):
    """Chunk coordinates (J, 3) int32 of the jobs of cave_region / lit_cave_region (region chunks 0..n-1)."""
    return (_grid_coords((1, 1, 1), (n, n, n)) - 1).astype(np.int32)


# ---- cameras for the draw-list step ----
def look_at_projection(eye, target, fov_y_deg=70.0, aspect=16 / 9, near=0.1, far=512.0):
    """A camera for the draw-list workloads: the view-projection matrix (column-major, float32) looking
    from eye to target, in the view-position relative form the culling shader uses (the view translation
    is applied by subtracting the view position), with depth in [0, 1] between near and far."""
    eye, target = np.asarray(eye, np.float64), np.asarray(target, np.float64)
    f = target - eye
    f /= np.linalg.norm(f)
    up = np.array([0.0, 1.0, 0.0]) if abs(f[1]) < 0.99 else np.array([1.0, 0.0, 0.0])
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.eye(4)
    view[0, :3], view[1, :3], view[2, :3] = s, u, -f
    t = 1.0 / np.tan(np.radians(fov_y_deg) / 2)
    proj = np.zeros((4, 4))
    proj[0, 0], proj[1, 1] = t / aspect, t
    proj[2, 2], proj[2, 3] = far / (near - far), far * near / (near - far)
    proj[3, 2] = -1.0
    return (proj @ view).T.astype(np.float32).reshape(16)  # column-major
