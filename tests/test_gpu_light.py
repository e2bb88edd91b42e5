"""Parity of the gfx950 light rebuild before meshing with the reference, through the C ABI (needs an MI355X).

The reference's mesh task rebuilds light from the blocks' light levels and the chunks' sunlight heights
before meshing (client/src/ChunkMeshTask.cpp:48-111). Bit-exact checks:
  - every job's rebuilt light over the mesher's 18^3 halo has the SHA-256 the reference's rebuild gave
    (tests/golden/light.npz), for every light set including the full 4096-job lit cave region, and
    equals the CPU oracle's on the small sets;
  - the meshes over the rebuilt light have the reference's mesh digests, in persistent and per-chunk runs;
  - switching a batch back to the store's light gives the store-light meshes again.
"""
import hashlib
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
    return np.load(os.path.join(GOLDEN, "light.npz"))


def lset(name):
    import test_light_oracle

    return test_light_oracle.sets()[name]


def relit_batch(ctx, blocks, heights, nbr):
    store = hcm.Store(ctx, blocks.shape[0])
    # the store's own light bytes are never read in rebuild mode: fill them with a value no rebuild makes
    store.upload(blocks, np.full(blocks.shape, 0xEE, np.uint8))
    store.upload_sunlight(heights)
    batch = hcm.Batch(ctx, store, nbr)
    batch.set_light(hcm.LIGHT_REBUILD)
    return store, batch


def digests_of(out):
    res, recs, verts, idx = out
    d = np.zeros((res.shape[0], 32), np.uint8)
    for j in range(res.shape[0]):
        d[j] = np.frombuffer(meshdigest.digest(hcm.job_meshes(j, res, recs, verts, idx)), np.uint8)
    return d


def assert_digests(name, got, want):
    bad = np.nonzero((got != want).any(1))[0]
    assert bad.size == 0, f"{name}: {bad.size} jobs differ, first {bad[:10].tolist()}"


SMALL = ["lit_glowstone", "lit_terrain_chunk", "lit_palette_all_values", "lit_palette_glassy",
         "lit_palette_custom", "lit_palette_flow"]


@pytest.mark.parametrize("order", ["morton_cost", "morton_tailcost"])
def test_relit_region_in_tuned_job_orders(ctx, golden, order):
    """The relit batch in the job orders the tuner applies (the light kernel shares the mesher's job list):
    job k of the reordered batch rebuilds and meshes region job perm[k] as the reference does (any costs
    give a valid order; these are seeded)."""
    from hypercraft_amd import tune

    blocks, heights, nbr = lset("lit_cave_region")
    cost = np.random.default_rng(7).integers(0, 2000, nbr.shape[0]).astype(np.float64)
    perm = tune.job_order(order, workloads.region_job_positions(), cost)
    say(f"light rebuild in order {order}: lit_cave_region")
    store, batch = relit_batch(ctx, blocks, heights, nbr[perm])
    try:
        batch.run()
        batch.sync()
        hd = golden["lit_cave_region_halo_digest"][perm]
        bad = [k for k in range(0, nbr.shape[0], 7)
               if hashlib.sha256(batch.fetch_light(k).tobytes()).digest() != hd[k].tobytes()]
        assert not bad, f"{order}: rebuilt light differs in {len(bad)} sampled jobs, first {bad[:10]}"
        out = batch.download()
        assert (out[0]["status"] == 0).all()
        assert_digests(order, digests_of(out), golden["lit_cave_region_mesh_digest"][perm])
    finally:
        batch.close()
        store.close()


@pytest.mark.parametrize("name", SMALL + ["lit_cave_region"])
def test_rebuilt_light_and_meshes_match_reference(ctx, golden, name):
    blocks, heights, nbr = lset(name)
    say(f"light rebuild: {name} ({nbr.shape[0]} jobs)")
    store, batch = relit_batch(ctx, blocks, heights, nbr)
    try:
        batch.run()
        batch.sync()
        hd = golden[name + "_halo_digest"]
        bad = [j for j in range(nbr.shape[0])
               if hashlib.sha256(batch.fetch_light(j).tobytes()).digest() != hd[j].tobytes()]
        assert not bad, f"{name}: rebuilt light differs in {len(bad)} jobs, first {bad[:10]}"
        out = batch.download()
        assert (out[0]["status"] == 0).all()
        assert_digests(name, digests_of(out), golden[name + "_mesh_digest"])
    finally:
        batch.close()
        store.close()


@pytest.mark.parametrize("name", SMALL)
def test_rebuilt_light_equals_oracle(ctx, oracle, golden, name):
    blocks, heights, nbr = lset(name)
    halo, _ = oracle.relight_many(golden["light_levels"], blocks, heights, nbr, 8, mesh=False, halo=True)
    store, batch = relit_batch(ctx, blocks, heights, nbr)
    try:
        batch.run()
        batch.sync()
        for j in range(nbr.shape[0]):
            assert np.array_equal(batch.fetch_light(j), halo[j]), f"{name} job {j}"
    finally:
        batch.close()
        store.close()


def test_lone_glowstone_closed_form(ctx):
    blocks, heights, nbr = cases.lone_glowstone((3, 12, 14))
    store, batch = relit_batch(ctx, blocks, heights, nbr)
    try:
        batch.run()
        batch.sync()
        got = batch.fetch_light(0)
    finally:
        batch.close()
        store.close()
    y, z, x = np.meshgrid(np.arange(-1, 17), np.arange(-1, 17), np.arange(-1, 17), indexing="ij")
    d = np.abs(x - 3) + np.abs(y - 12) + np.abs(z - 14)
    assert np.array_equal(got >> 4, np.maximum(15 - d, 0).astype(np.uint8))
    assert not (got & 15).any()


def test_per_chunk_and_replays_and_store_light(ctx, golden):
    """A subset of the lit cave region: per-chunk launches and graph replays give the reference's
    meshes over the rebuilt light; switching back to the store's light gives the store-light meshes."""
    blocks, heights, nbr = lset("lit_cave_region")
    sub = np.arange(0, nbr.shape[0], 29)
    want = golden["lit_cave_region_mesh_digest"][sub]
    store = hcm.Store(ctx, blocks.shape[0])
    light = np.zeros(blocks.shape, np.uint8)
    store.upload(blocks, light)
    store.upload_sunlight(heights)
    batch = hcm.Batch(ctx, store, nbr[sub])
    try:
        say(f"store light first ({sub.size} jobs)")
        batch.run()
        batch.sync()
        plain = digests_of(batch.download())
        batch.set_light(hcm.LIGHT_REBUILD)
        for mode in (hcm.RUN_PERSISTENT, hcm.RUN_PERSISTENT, hcm.RUN_PER_CHUNK):
            say(f"rebuild, mode {mode}")
            batch.run(mode)
            batch.sync()
            assert_digests(f"lit cave subset mode {mode}", digests_of(batch.download()), want)
        batch.set_light(hcm.LIGHT_STORE)
        batch.run()
        batch.sync()
        assert_digests("store light again", digests_of(batch.download()), plain)
    finally:
        batch.close()
        store.close()


def test_torch_levels_below_15_equal_oracle(oracle, golden):
    """A registry in which stone, dirt and glass emit torchlight 7, 11 and 3: sources of every level
    enter the propagation at their own level."""
    import ctypes

    levels = golden["light_levels"].copy()
    for block, lv in ((1, 7), (3, 11), (9, 3)):
        levels[block | np.arange(256) << 8] = lv
    base = hcm.lib().hcm_registry_default().contents
    reg = hcm.RegistryC()
    for field, _ in hcm.RegistryC._fields_:
        setattr(reg, field, getattr(base, field))
    reg.light_level = levels.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
    ctx2 = hcm.Context(pool_quads=1 << 20, registry=ctypes.pointer(reg))
    blocks, _, nbr = cases.palette("all_values", 77, 6, density=0.3)
    heights = cases.random_heights(78, blocks.shape[0])
    halo, _ = oracle.relight_many(levels, blocks, heights, nbr, 8, mesh=False, halo=True)
    say("light rebuild with torchlight levels 3, 7, 11")
    store, batch = relit_batch(ctx2, blocks, heights, nbr)
    try:
        batch.run()
        batch.sync()
        for j in range(nbr.shape[0]):
            assert np.array_equal(batch.fetch_light(j), halo[j]), f"job {j}"
        assert ((halo >> 4) % 15 != 0).any()  # levels other than 0 and 15 occur
    finally:
        batch.close()
        store.close()
        ctx2.close()


@pytest.mark.parametrize("name", SMALL)
def test_plain_path_whole_mesh_task(ctx, golden, name):
    """hcm_mesh_chunk_relight: light rebuild + meshing of one chunk from host arrays."""
    blocks, heights, nbr = lset(name)
    for j in range(0, nbr.shape[0], 3):
        say(f"plain whole mesh task: {name} job {j}")
        b27 = [blocks[s] for s in nbr[j]]
        meshes, light = ctx.mesh_chunk_relight(b27, workloads.slice_heights(heights, nbr, j))
        assert hashlib.sha256(light.tobytes()).digest() == golden[name + "_halo_digest"][j].tobytes()
        assert meshdigest.digest(meshes) == golden[name + "_mesh_digest"][j].tobytes()
