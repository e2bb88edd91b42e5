"""Parity of the gfx950 mesher with the reference, through the C ABI (needs an MI355X).

Bit-exact: every job's mesh list (vertices, indices, AABB, transparent flag, mesh order) must have
the digest the real reference mesher produced for it (tests/golden/golden.npz), and, for the
plain-path cases, equal the CPU oracle's output.
"""
import os
import sys

import numpy as np
import pytest

import cases
import meshdigest
from conftest import GOLDEN
from hypercraft_amd import hcm, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(GOLDEN, "golden.npz"))


_SETS = {}


def gset(name):
    if name not in _SETS:
        import test_oracle

        _SETS.update(test_oracle.sets())
    return _SETS[name]


def test_device_present():
    assert hcm.device_ok(), "no gfx950 device: the mesher has no CPU fallback"


@pytest.mark.parametrize("name", ["lone_stone", "lone_glass", "slab", "uniform_air", "uniform_stone",
                                  "uniform_water", "uniform_flow", "terrain_chunk", "terrain_chunk_plants",
                                  "leaf_checkerboard"])
def test_plain_path_single_chunks(ctx, oracle, golden, name):
    blocks, light, nbr = gset(name)
    b27, l27 = workloads.slice_jobs(blocks, light, nbr, 0)
    say(f"plain path: {name}")
    got = ctx.mesh_chunk(b27, l27)
    assert meshdigest.digest(got) == golden[name + "_digest"][0].tobytes()
    want = oracle.mesh(b27, l27)
    assert meshdigest.digest(got) == meshdigest.digest(want)


@pytest.mark.parametrize("name", ["lone_stone", "slab", "uniform_flow", "terrain_chunk", "terrain_chunk_plants",
                                  "leaf_checkerboard", "palette_sparse_flow", "palette_sparse_custom"])
def test_plain_path_packed_halo(ctx, golden, name):
    """hcm_mesh_halo, the drop-in call's 20 KB form: the packed halo gives the reference's meshes (the
    leaf checkerboard's 24,576 quads take the second round trip)."""
    blocks, light, nbr = gset(name)
    for j in range(0, nbr.shape[0], 4):
        say(f"packed halo: {name} job {j}")
        got = ctx.mesh_halo(*hcm.halo_of_chunks(*workloads.slice_jobs(blocks, light, nbr, j)))
        assert meshdigest.digest(got) == golden[name + "_digest"][j].tobytes(), f"{name} job {j}"


def test_plain_path_terrain_chunk_full_meshes(ctx, golden):
    blocks, light, nbr = gset("terrain_chunk")
    got = ctx.mesh_chunk(*workloads.slice_jobs(blocks, light, nbr, 0))
    for k, m in enumerate(got):
        assert np.array_equal(m.vertices, golden[f"terrain_chunk_mesh{k}_vertices"])
        assert np.array_equal(m.indices, golden[f"terrain_chunk_mesh{k}_indices"])
        assert list(m.aabb) + [m.transparent] == golden[f"terrain_chunk_mesh{k}_meta"].tolist()


@pytest.mark.parametrize("fam", list(cases.FAMILIES))
def test_plain_path_palettes(ctx, golden, fam):
    name = "palette_sparse_" + fam
    blocks, light, nbr = gset(name)
    for j in range(0, nbr.shape[0], 3):
        say(f"plain path: {name} job {j}")
        got = ctx.mesh_chunk(*workloads.slice_jobs(blocks, light, nbr, j))
        assert meshdigest.digest(got) == golden[name + "_digest"][j].tobytes(), f"{name} job {j}"


def run_batch(ctx, blocks, light, nbr, mode=hcm.RUN_PERSISTENT):
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, light)
    batch = hcm.Batch(ctx, store, nbr)
    batch.run(mode)
    batch.sync()
    out = batch.download()
    batch.close()
    store.close()
    return out


def digests_of(out):
    res, recs, verts, idx = out
    n = res.shape[0]
    d = np.zeros((n, 32), np.uint8)
    for j in range(n):
        d[j] = np.frombuffer(meshdigest.digest(hcm.job_meshes(j, res, recs, verts, idx)), np.uint8)
    return d


def assert_digests(name, got, want):
    bad = np.nonzero((got != want).any(1))[0]
    assert bad.size == 0, f"{name}: {bad.size} jobs differ, first {bad[:10].tolist()}"


@pytest.mark.parametrize("name", ["palette_" + f for f in cases.FAMILIES] +
                         ["random_dense", "cave_region", "terrain_plants_region"])
def test_batched_persistent(ctx, golden, name):
    blocks, light, nbr = gset(name)
    say(f"batched persistent: {name} ({nbr.shape[0]} jobs)")
    out = run_batch(ctx, blocks, light, nbr)
    assert (out[0]["status"] == 0).all()
    assert np.array_equal(out[0]["quad_count"], golden[name + "_quads"])
    assert_digests(name, digests_of(out), golden[name + "_digest"])


def test_job_orders_give_the_same_meshes(ctx, golden):
    """Every job order the autotuner sweeps (the cost orders with the reference's quad counts as costs)
    meshes the cave region to the reference's meshes: job k of the reordered batch is region job perm[k]."""
    from hypercraft_amd import tune

    blocks, light, nbr = gset("cave_region")
    pos = workloads.region_job_positions()
    cost = golden["cave_region_quads"].astype(np.float64)
    for order in tune.ORDERS:
        perm = tune.job_order(order, pos, cost)
        say(f"job order {order}: cave_region")
        out = run_batch(ctx, blocks, light, nbr[perm])
        assert (out[0]["status"] == 0).all(), order
        assert_digests(order, digests_of(out), golden["cave_region_digest"][perm])


def test_repeated_runs_back_to_back(ctx, golden):
    """hcm_batch_run_repeat: three passes enqueued at once on a fresh batch (its first pass launched on its
    own, then two graph replays) and then five more leave the reference's meshes, and the event time of
    the replays is positive."""
    name = "terrain_plants_region"
    blocks, light, nbr = gset(name)
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, light)
    batch = hcm.Batch(ctx, store, nbr)
    try:
        for times in (3, 5):
            say(f"repeated runs: {name}, {times} passes")
            batch.run_repeat(times)
            assert batch.sync() > 0
            out = batch.download()
            assert (out[0]["status"] == 0).all()
            assert_digests(name, digests_of(out), golden[name + "_digest"])
    finally:
        batch.close()
        store.close()


def test_faceless_chunks(ctx, oracle):
    """The early exits (opaque cubes only, no chunk voxel with a face, no visible face) and the single
    voxels that defeat them: batched and plain results equal the oracle's."""
    names, blocks, light, nbr, want_faces = cases.faceless()
    say(f"faceless chunks: {len(names)} jobs")
    out = run_batch(ctx, blocks, light, nbr)
    assert (out[0]["status"] == 0).all()
    got = digests_of(out)
    for j, name in enumerate(names):
        b27, l27 = workloads.slice_jobs(blocks, light, nbr, j)
        want = meshdigest.digest(oracle.mesh(b27, l27))
        assert got[j].tobytes() == want, f"batched: {name}"
        assert meshdigest.digest(ctx.mesh_chunk(b27, l27)) == want, f"plain: {name}"
        assert (out[0]["quad_count"][j] > 0) == bool(want_faces[j]), name


def test_uniform_halo_of_self_showing_block(tmp_path):
    """A registry in which two neighbouring blocks of one value show faces to each other (glass with the
    leaves' texture on +x): a halo of that block alone must not take the uniform early exit; plain and
    batched results equal the oracle's on that registry."""
    import cpu_oracle
    from conftest import REGISTRY

    reg = cases.self_showing_glass(REGISTRY)
    path = str(tmp_path / "registry_glass.bin")
    reg.write(path)
    orc = cpu_oracle.Oracle(path)
    ctx2 = hcm.Context(pool_quads=1 << 20, registry=reg.c_registry(hcm))
    try:
        for value in (0x0009, 0x0109, 0x0001):
            blocks, light, nbr = cases.uniform(value)
            b27, l27 = workloads.slice_jobs(blocks, light, nbr, 0)
            want = meshdigest.digest(orc.mesh(b27, l27))
            say(f"self-showing uniform halo 0x{value:04x}")
            assert meshdigest.digest(ctx2.mesh_chunk(b27, l27)) == want
            out = run_batch(ctx2, blocks, light, nbr)
            assert (out[0]["status"] == 0).all()
            assert digests_of(out)[0].tobytes() == want
            assert (out[0]["quad_count"][0] == 32) == ((value & 0xff) == 0x09)
    finally:
        ctx2.close()


def test_mixed_uneven_batches(ctx, golden):
    """All 200 uneven batches (log-uniform sizes in 1..4096) drawn from one store holding the cave region
    and the plant terrain: each batch's first run (launched directly) and its captured graph's replay give
    the reference's meshes; every tenth batch also through per-chunk launches."""
    blocks, light, nbr, n_cave = workloads.mixed_pool()
    want = np.concatenate([golden["cave_region_digest"], golden["terrain_plants_region_digest"]])
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, light)
    chunks = 0
    try:
        for k, jobs in enumerate(workloads.mixed_batches(nbr.shape[0])):
            if k % 20 == 0:
                say(f"mixed batch {k}: {jobs.size} jobs")
            batch = hcm.Batch(ctx, store, nbr[jobs])
            modes = (hcm.RUN_PERSISTENT, hcm.RUN_PERSISTENT) + ((hcm.RUN_PER_CHUNK,) if k % 10 == 0 else ())
            for mode in modes:
                batch.run(mode)
                batch.sync()
                assert_digests(f"mixed batch {k} mode {mode}", digests_of(batch.download()), want[jobs])
            batch.close()
            chunks += jobs.size
        assert chunks > 100000
    finally:
        store.close()


def test_per_chunk_launches_match(ctx, golden):
    blocks, light, nbr = gset("cave_region")
    sub = nbr[::37]
    say(f"per-chunk launches: cave_region subset ({sub.shape[0]} jobs)")
    out = run_batch(ctx, blocks, light, sub, mode=hcm.RUN_PER_CHUNK)
    assert_digests("cave_region subset", digests_of(out), golden["cave_region_digest"][::37])


def test_run_modes_in_any_order(ctx, golden):
    """A batch's first persistent run (launched directly), its graph replays (whose last workgroup clears
    the counters) and per-chunk launches (after which the counters are reset) give the same meshes and
    totals in any order."""
    blocks, light, nbr = gset("cave_region")
    sub = np.arange(0, nbr.shape[0], 23)
    want_d = golden["cave_region_digest"][sub]
    want_q = int(golden["cave_region_quads"][sub].sum())
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, light)
    batch = hcm.Batch(ctx, store, nbr[sub])
    try:
        for k, mode in enumerate((hcm.RUN_PERSISTENT, hcm.RUN_PERSISTENT, hcm.RUN_PER_CHUNK, hcm.RUN_PERSISTENT,
                                  hcm.RUN_PER_CHUNK, hcm.RUN_PER_CHUNK, hcm.RUN_PERSISTENT, hcm.RUN_PERSISTENT)):
            say(f"run {k}: mode {mode}")
            batch.run(mode)
            batch.sync()
            quads, meshes = batch.totals()
            out = batch.download()
            assert quads == want_q and meshes == out[1].shape[0], f"run {k}"
            assert_digests(f"run {k}", digests_of(out), want_d)
    finally:
        batch.close()
        store.close()


@pytest.mark.parametrize("threads", [128, 512])
def test_workgroup_sizes_agree(ctx, golden, threads):
    blocks, light, nbr = gset("palette_all_values")
    ctx.set_tuning(wg_per_cu=0, threads=threads)
    try:
        say(f"threads {threads}: palette_all_values")
        out = run_batch(ctx, blocks, light, nbr)
    finally:
        ctx.set_tuning(wg_per_cu=0, threads=256)
    assert_digests(f"threads {threads}", digests_of(out), golden["palette_all_values_digest"])


def test_graph_replays_are_identical(ctx, golden):
    blocks, light, nbr = gset("palette_custom")
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, light)
    batch = hcm.Batch(ctx, store, nbr)
    try:
        for rep in range(3):
            say(f"graph replay {rep}")
            batch.run()
            batch.sync()
            out = batch.download()
            np.asarray(blocks[:10]).sum()  # host work between replays
            assert_digests(f"replay {rep}", digests_of(out), golden["palette_custom_digest"])
    finally:
        batch.close()
        store.close()


def test_pool_full_is_reported():
    small = hcm.Context(pool_quads=1000)
    try:
        blocks, light, nbr = gset("palette_glassy")
        say("pool full: palette_glassy on a 1000-quad pool")
        out = run_batch(small, blocks, light, nbr[:8])
        st = out[0]["status"]
        assert (st == hcm.HCM_ERR_POOL_FULL).any()
        assert set(st.tolist()) <= {0, hcm.HCM_ERR_POOL_FULL}
    finally:
        small.close()


def test_cpp_front_end_generate(golden, tmp_path):
    """hcm::BlockMeshGPU::Generate(get_block, get_light) (include/hcm.hpp), called like the reference's
    BlockMeshAlgo::Generate with callables returning GetData() types, gives the reference's meshes."""
    import subprocess

    import test_library
    from hypercraft_amd import build as hb

    say("C++ front end: terrain chunk through BlockMeshGPU::Generate")
    out = tmp_path / "meshes.bin"
    r = subprocess.run([hb.WRAPPER_CHECK, str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = test_library.read_wrapper_output(out)
    assert meshdigest.digest(got) == golden["terrain_chunk_digest"][0].tobytes()
