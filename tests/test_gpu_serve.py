"""Serving (hcm_server): continuous batching of chunk loads and voxel edits (needs an MI355X).

The claim under test is that every mesh the server produces in the mix of concurrent batches is the
mesh of that chunk meshed alone, by the CPU oracle, on the world as it stood at the mesh's epoch
(the number of requests applied before its batch) -- bit-exact -- and that every request completes
once, after every chunk it made stale holds a mesh at or after its epoch.
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

SEED = 11
DIMS = (4, 3, 4)  # chunk grid x, y, z: the inner 2 x 1 x 2 chunks have all 26 neighbours
EDIT_PALETTE = np.concatenate([cases.CUSTOM, cases.GLASSY, cases.FLOW, cases.LOGS, [0, 0, 0, 1, 2, 3]]).astype(np.uint16)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class World:
    """Host copy of the world: chunk position -> (blocks, light), edited like the server's."""

    def __init__(self):
        self.chunks = {}

    def load(self, pos, blocks, light):
        self.chunks[pos] = (blocks.copy(), light.copy())

    def edit(self, x, y, z, block, light):
        c = (x // 16, y // 16, z // 16)
        i = (x % 16) + 16 * ((z % 16) + 16 * (y % 16))
        b, l = self.chunks[c]
        b[i], l[i] = block, light

    def meshable(self, pos):
        return all(tuple(int(v) for v in np.add(pos, d)) in self.chunks for d in workloads.SLOT_OFFSETS)

    def mesh_alone(self, oracle, pos):
        nb = [self.chunks[tuple(int(v) for v in np.add(pos, d))] for d in workloads.SLOT_OFFSETS]
        return oracle.mesh([b for b, _ in nb], [l for _, l in nb])


def terrain(pos):
    return hcm.gen_terrain_chunk(SEED, *pos)


def grid():
    return [(x, y, z) for y in range(DIMS[1]) for z in range(DIMS[2]) for x in range(DIMS[0])]


def inner():
    return [p for p in grid() if all(0 < p[a] < DIMS[a] - 1 for a in range(3))]


def check_history(srv, oracle, log, world=None):
    """Replays the request log and compares every recorded mesh with the oracle alone at its epoch."""
    hist = sorted(srv.history(), key=lambda h: h[1])
    world = world if world is not None else World()
    applied = 0
    for pos, epoch, meshes in hist:
        while applied < epoch:
            op = log[applied]
            if op[0] == "load":
                world.load(*op[1:])
            else:
                world.edit(*op[1:])
            applied += 1
        assert world.meshable(pos), (pos, epoch)
        want = world.mesh_alone(oracle, pos)
        assert meshdigest.digest(meshes) == meshdigest.digest(want), f"chunk {pos} at epoch {epoch}"
    return len(hist)


def check_final(srv, oracle, world):
    for pos in world.chunks:
        if world.meshable(pos):
            got = srv.chunk_mesh(*pos)
            assert meshdigest.digest(got) == meshdigest.digest(world.mesh_alone(oracle, pos)), pos


def check_completions(done, n_requests, log):
    tags = sorted(c["tag"] for c in done)
    assert tags == list(range(n_requests)), "every request completes exactly once"
    for c in done:
        assert c["epoch"] == c["tag"] + 1  # one submitting thread: request k has epoch k + 1
        assert c["kind"] == (hcm.REQ_LOAD if log[c["tag"]][0] == "load" else hcm.REQ_EDIT)
        assert c["latency_ms"] >= 0


def load_all(srv, world, log, order):
    for pos in order:
        b, l = terrain(pos)
        srv.load(*pos, b, l, tag=len(log))
        world.load(pos, b, l)
        log.append(("load", pos, b, l))


def test_server_loads_match_meshing_alone(ctx, oracle):
    rng = np.random.default_rng(1)
    order = [grid()[i] for i in rng.permutation(len(grid()))]
    srv = hcm.Server(ctx, max_chunks=64, max_batch=3, max_inflight=3, pool_quads=1 << 22, keep_history=True)
    try:
        world, log = World(), []
        say("serve: loading 48 chunks in random order")
        load_all(srv, world, log, order)
        srv.drain()
        done = srv.poll()
        check_completions(done, len(log), log)
        n = check_history(srv, oracle, log)
        assert n >= len(inner())
        check_final(srv, oracle, world)
        st = srv.stats()
        assert st["completed"] == st["requests"] == len(log)
        assert st["chunks_meshed"] == n
    finally:
        srv.close()


def halo_chunks(world, x, y, z):
    """Meshable chunks whose 18^3 halo holds world voxel (x, y, z)."""
    out = set()
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                c = ((x + dx) // 16, (y + dy) // 16, (z + dz) // 16)
                if world.meshable(c):
                    out.add(c)
    return out


def _edits(rng, n):
    """Random voxel edits in the inner chunks, half of them on chunk borders (halo of the neighbours)."""
    out = []
    for _ in range(n):
        c = inner()[rng.integers(len(inner()))]
        v = rng.integers(0, 16, 3)
        if rng.random() < 0.5:
            a = rng.integers(3)
            v[a] = 0 if rng.random() < 0.5 else 15
        if rng.random() < 0.2:
            v[:] = [0 if rng.random() < 0.5 else 15 for _ in range(3)]  # corner: seven neighbours
        x, y, z = (int(16 * c[a] + v[a]) for a in range(3))
        out.append((x, y, z, int(EDIT_PALETTE[rng.integers(len(EDIT_PALETTE))]), int(rng.integers(256))))
    return out


@pytest.mark.parametrize("max_batch,max_inflight", [(1, 1), (4, 3), (64, 2)])
def test_server_edits_match_meshing_alone(ctx, oracle, max_batch, max_inflight):
    rng = np.random.default_rng(100 + max_batch)
    srv = hcm.Server(ctx, max_chunks=64, max_batch=max_batch, max_inflight=max_inflight, pool_quads=1 << 22,
                     keep_history=True)
    try:
        world, log = World(), []
        load_all(srv, world, log, grid())
        say(f"serve: 300 edits, max_batch {max_batch}, max_inflight {max_inflight}")
        for k, (x, y, z, blk, lt) in enumerate(_edits(rng, 300)):
            srv.edit(x, y, z, blk, lt, tag=len(log))
            world.edit(x, y, z, blk, lt)
            log.append(("edit", x, y, z, blk, lt))
            if k % 50 == 49:
                srv.drain()
        srv.drain()
        done = srv.poll()
        check_completions(done, len(log), log)
        for c in done:
            if c["kind"] == hcm.REQ_EDIT:
                _, x, y, z = log[c["tag"]][:4]
                assert c["remeshed"] == len(halo_chunks(world, x, y, z))
        n = check_history(srv, oracle, log)
        say(f"serve: {n} recorded meshes equal the oracle alone at their epochs")
        check_final(srv, oracle, world)
    finally:
        srv.close()


def test_server_reload_replaces_chunk(ctx, oracle):
    srv = hcm.Server(ctx, max_chunks=64, max_batch=8, max_inflight=2, pool_quads=1 << 22, keep_history=True)
    try:
        world, log = World(), []
        load_all(srv, world, log, grid())
        say("serve: reload the inner chunks with cave chunks, edits in between")
        rng = np.random.default_rng(5)
        for pos in inner():
            b, l = hcm.gen_cave_chunk(SEED, *pos)
            srv.load(*pos, b, l, tag=len(log))
            world.load(pos, b, l)
            log.append(("load", pos, b, l))
            for x, y, z, blk, lt in _edits(rng, 5):
                srv.edit(x, y, z, blk, lt, tag=len(log))
                world.edit(x, y, z, blk, lt)
                log.append(("edit", x, y, z, blk, lt))
        srv.drain()
        check_completions(srv.poll(), len(log), log)
        check_history(srv, oracle, log)
        check_final(srv, oracle, world)
    finally:
        srv.close()


def test_server_pool_pressure(ctx, oracle):
    """A pool with room for two arenas: batches wait for memory and still produce the same meshes."""
    srv = hcm.Server(ctx, max_chunks=64, max_batch=2, max_inflight=4, pool_quads=3 << 16, keep_history=True)
    try:
        world, log = World(), []
        load_all(srv, world, log, grid())
        rng = np.random.default_rng(9)
        say("serve: edits under pool pressure")
        for x, y, z, blk, lt in _edits(rng, 100):
            srv.edit(x, y, z, blk, lt, tag=len(log))
            world.edit(x, y, z, blk, lt)
            log.append(("edit", x, y, z, blk, lt))
        srv.drain()
        check_completions(srv.poll(), len(log), log)
        check_history(srv, oracle, log)
        check_final(srv, oracle, world)
        st = srv.stats()
        assert st["pool_used_quads"] + st["pool_free_quads"] == 3 << 16
        live = sum(sum(len(m.vertices) // 4 for m in srv.chunk_mesh(*p)) for p in inner())
        assert st["pool_used_quads"] == live  # superseded meshes and unused arena tails went back to the pool
    finally:
        srv.close()


class LitWorld(World):
    """Host world of a LIGHT_REBUILD server: chunk position -> (blocks, sunlight heights); a chunk is
    meshed alone by the oracle's light rebuild followed by the oracle's mesher (ChunkMeshTask.cpp:41-122)."""

    def __init__(self, levels):
        super().__init__()
        self.levels = levels
        from test_registry import load_fixture

        reg = load_fixture()
        self.vertical = reg["rows"][reg["row_of"], 12] == 2  # Block::GetVerticalLightPass by Block value

    def edit(self, x, y, z, block, light):
        """Sets the block, then floods sunlight down the edited column when the chunk above is loaded
        (ChunkFloodSunlightTask.cpp:38-54). Returns whether the column's sunlight height changed, or None
        when the voxel already holds the block: such an edit changes nothing (ChunkSetBlockTask.cpp:89)."""
        c = (x // 16, y // 16, z // 16)
        blocks, heights = self.chunks[c]
        i = (x % 16) + 16 * ((z % 16) + 16 * (y % 16))
        if blocks[i] == block:
            return None
        blocks[i] = block
        up = (c[0], c[1] + 1, c[2])
        if up not in self.chunks:
            return False
        xz = (x % 16) + 16 * (z % 16)
        if self.chunks[up][1][xz] > 0:
            h = 16
        else:
            h = 15
            while h >= 0 and self.vertical[blocks[xz + 256 * h]]:
                h -= 1
            h += 1
        if h == heights[xz]:
            return False
        heights[xz] = h
        return True

    def mesh_alone(self, oracle, pos):
        nb = [self.chunks[tuple(int(v) for v in np.add(pos, d))] for d in workloads.SLOT_OFFSETS]
        b27 = [b for b, _ in nb]
        win, _ = oracle.light_rebuild(self.levels, b27, [h for _, h in nb])
        return oracle.mesh(b27, cases.light27_from_window(win))


def light_chunks(world, x, y, z):
    """Meshable chunks within light distance 14 of world voxel (x, y, z) (ChunkMeshTask.inl:9-15)."""
    def dist(v):
        return -1 - v if v < -1 else (v - 17 if v >= 17 else 0)

    c = (x // 16, y // 16, z // 16)
    out = set()
    for d in workloads.SLOT_OFFSETS:
        n = tuple(int(c[a] + d[a]) for a in range(3))
        if world.meshable(n) and sum(dist((x, y, z)[a] - 16 * n[a]) for a in range(3)) < 15:
            out.add(n)
    return out


@pytest.mark.parametrize("max_batch,max_inflight", [(1, 1), (8, 3)])
def test_server_relight_matches_mesh_task_alone(ctx, oracle, max_batch, max_inflight):
    """A LIGHT_REBUILD server: loads carry sunlight heights, every batch rebuilds light before meshing,
    edits remesh the chunks within light distance 14; every mesh equals the whole mesh task run alone."""
    levels = np.load(os.path.join(GOLDEN, "light.npz"))["light_levels"]
    rng = np.random.default_rng(300 + max_batch)
    srv = hcm.Server(ctx, max_chunks=64, max_batch=max_batch, max_inflight=max_inflight, pool_quads=1 << 22,
                     keep_history=True, light_mode=hcm.LIGHT_REBUILD)
    try:
        world, log = LitWorld(levels), []
        say(f"serve relight: loads, max_batch {max_batch}, max_inflight {max_inflight}")
        for i, pos in enumerate(grid()):
            b, _ = terrain(pos)
            b = workloads.place_glowstone(b.reshape(1, -1), 400 + i, 0.003).reshape(-1)
            h = cases.random_heights(500 + i, 1)[0]
            srv.load_sunlit(*pos, b, h, tag=len(log))
            world.chunks[pos] = (b.copy(), h.copy())
            log.append(("load", pos, b, h))
        say("serve relight: 120 edits")
        palette = np.array([0, 0, 1, 3, 9, 0x18, 0x10, 0x0E], np.uint16)  # air, stone, dirt, glass, glowstone, leaves, water
        flooded, noop = set(), set()
        for k, (x, y, z, _, _) in enumerate(_edits(rng, 120)):
            blk = int(palette[rng.integers(len(palette))])
            if k % 10 == 9:  # every tenth edit sets a voxel to the block it holds
                c = (x // 16, y // 16, z // 16)
                blk = int(world.chunks[c][0][(x % 16) + 16 * ((z % 16) + 16 * (y % 16))])
            srv.edit(x, y, z, blk, 0, tag=len(log))
            changed = world.edit(x, y, z, blk, 0)
            if changed:
                flooded.add(len(log))
            elif changed is None:
                noop.add(len(log))
            log.append(("edit", x, y, z, blk, 0))
            if k % 40 == 39:
                srv.drain()
        assert flooded, "some edits move a sunlight height"
        assert len(noop) >= 12, "some edits change nothing"
        srv.drain()
        done = srv.poll()
        check_completions(done, len(log), log)
        for c in done:
            if c["kind"] == hcm.REQ_EDIT:
                _, x, y, z = log[c["tag"]][:4]
                want = set() if c["tag"] in noop else light_chunks(world, x, y, z)
                if c["tag"] in flooded:  # a moved sunlight height remeshes the chunk and its 26 neighbours
                    want |= {n for n in (tuple(int(v) for v in np.add((x // 16, y // 16, z // 16), d))
                                         for d in workloads.SLOT_OFFSETS) if world.meshable(n)}
                assert c["remeshed"] == len(want)
        n = check_history(srv, oracle, log, LitWorld(levels))
        say(f"serve relight: {n} recorded meshes equal the oracle's mesh task alone at their epochs")
        check_final(srv, oracle, world)
        with pytest.raises(hcm.HcmError):
            srv.load(0, 0, 0, *terrain((0, 0, 0)))  # light bytes are not what this server takes
    finally:
        srv.close()


def test_server_edit_remesh_goes_first(ctx, oracle):
    """Remeshes an edit causes are taken before those of earlier loads (ChunkSetBlockTask.cpp:119 pushes
    them at high priority): behind a backlog of 108 reloads (16 chunks to remesh, one per batch), the first
    mesh made at or after the edit's epoch is the edited chunk's, which the reloads queued last."""
    dims = (6, 3, 6)
    chunks = [(x, y, z) for y in range(dims[1]) for z in range(dims[2]) for x in range(dims[0])]
    srv = hcm.Server(ctx, max_chunks=128, max_batch=1, max_inflight=1, pool_quads=1 << 22, keep_history=True)
    try:
        world, log = World(), []
        load_all(srv, world, log, chunks)
        srv.drain()
        say("serve: 108 reloads, then an edit in the chunk they remesh last")
        for pos in chunks:
            b, l = hcm.gen_cave_chunk(SEED, *pos)
            srv.load(*pos, b, l, tag=len(log))
            world.load(pos, b, l)
            log.append(("load", pos, b, l))
        x, y, z = 16 * 4 + 7, 16 * 1 + 8, 16 * 4 + 9  # inside chunk (4, 1, 4), away from its borders
        srv.edit(x, y, z, 9, 0x3A, tag=len(log))
        world.edit(x, y, z, 9, 0x3A)
        log.append(("edit", x, y, z, 9, 0x3A))
        edit_epoch = len(log)
        srv.drain()
        done = srv.poll()
        check_completions(done, len(log), log)
        after = [h for h in srv.history() if h[1] >= edit_epoch]
        assert after and tuple(after[0][0]) == (4, 1, 4), [tuple(h[0]) for h in after[:4]]
        assert sum(1 for h in after if tuple(h[0]) == (4, 1, 4)) == 1
        edit_done = next(c for c in done if c["tag"] == edit_epoch - 1)
        assert edit_done["remeshed"] == 1
        check_history(srv, oracle, log)
        check_final(srv, oracle, world)
    finally:
        srv.close()


def test_server_spiral_trace_final_meshes(ctx):
    """The whole serving trace of the benchmark (spiral loads of 28,688 terrain chunks, radius 24 x 16
    chunks, and 2,911 single-voxel edits), submitted as fast as the server takes it: every request
    completes, and each of the 22,470 meshable chunks then holds the mesh the reference mesher makes of
    it on the trace's final world (tests/golden/trace.npz, written by make_golden.py trace)."""
    gold = np.load(os.path.join(GOLDEN, "trace.npz"))
    tr = workloads.spiral_trace()
    coords, blocks, light, events = tr["coords"], tr["blocks"], tr["light"], tr["events"]
    assert gold["counts"].tolist() == [coords.shape[0], sum(1 for e in events if e[1] == "edit")]
    srv = hcm.Server(ctx, max_chunks=coords.shape[0], max_batch=4096, max_inflight=2, pool_quads=1 << 24)
    try:
        say(f"serve: spiral trace, {len(events)} requests")
        for tag, (_, kind, payload) in enumerate(events):
            if kind == "load":
                c = coords[payload]
                srv.load(int(c[0]), int(c[1]), int(c[2]), blocks[payload], light[payload], tag)
            else:
                srv.edit(*payload, tag)
        srv.drain()
        done = srv.poll()
        assert sorted(c["tag"] for c in done) == list(range(len(events)))
        say(f"serve: comparing {gold['positions'].shape[0]} final meshes with the reference's")
        bad = [tuple(p) for p, d in zip(gold["positions"].tolist(), gold["digest"])
               if meshdigest.digest(srv.chunk_mesh(*p)) != d.tobytes()]
        assert not bad, f"{len(bad)} chunks differ, first {bad[:5]}"
    finally:
        srv.close()


def test_server_pool_full_is_reported(ctx):
    srv = hcm.Server(ctx, max_chunks=64, max_batch=4, max_inflight=2, pool_quads=64)
    try:
        world, log = World(), []
        say("serve: pool of 64 quads")
        with pytest.raises(hcm.HcmError) as e:
            load_all(srv, world, log, grid())
            srv.drain()
        assert e.value.code == hcm.HCM_ERR_POOL_FULL
    finally:
        srv.close()


def test_server_rejects_bad_requests(ctx):
    srv = hcm.Server(ctx, max_chunks=2, max_batch=4, max_inflight=1, pool_quads=1 << 20)
    try:
        b, l = terrain((0, 0, 0))
        with pytest.raises(hcm.HcmError) as e:
            srv.edit(1, 2, 3, 1, 0)
        assert e.value.code == hcm.HCM_ERR_INVALID  # chunk not loaded
        srv.load(0, 0, 0, b, l)
        srv.load(-1, 0, 0, b, l)
        srv.load(0, 0, 0, b, l)  # a reload takes no new slot
        with pytest.raises(hcm.HcmError) as e:
            srv.load(5, 0, 0, b, l)
        assert e.value.code == hcm.HCM_ERR_BUSY
        srv.edit(-1, 0, 0, 1, 0)  # voxel (-1, 0, 0) lies in chunk (-1, 0, 0)
        srv.drain()
        assert len(srv.poll()) == 4
        with pytest.raises(hcm.HcmError) as e:
            srv.chunk_mesh(0, 0, 0)  # not meshable: neighbours missing
        assert e.value.code == hcm.HCM_ERR_INVALID
    finally:
        srv.close()
