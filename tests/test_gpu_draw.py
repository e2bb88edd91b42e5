"""Draw lists on the GPU against the oracle (needs an MI355X): the mesh info of every mesh of the cave
batch (WorldRenderer::PushChunkMesh, WorldRenderer.hpp:69-87) bit for bit, and the frustum-culled draw
commands (chunk_culling.comp) for several cameras as sets (the shader's atomics leave their order open).
"""
import sys

import numpy as np
import pytest

import draw_oracle as do
from hypercraft_amd import hcm, workloads

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_mesh_infos_and_culling_match_oracle(ctx):
    blocks, light, nbr = workloads.cave_region()
    pos = workloads.region_job_positions()
    store = hcm.Store(ctx, blocks.shape[0])
    store.upload(blocks, light)
    batch = hcm.Batch(ctx, store, nbr)
    try:
        batch.run()
        batch.sync()
        _, recs, _, _ = batch.download()
        say(f"mesh infos: {recs.shape[0]} meshes")
        got = batch.mesh_infos(pos)
        want = do.mesh_infos(recs, pos)
        for f in ("index_count", "first_index", "vertex_offset", "base_position", "transparent"):
            assert np.array_equal(got[f], want[f]), f
        assert np.array_equal(got["aabb"].view(np.uint32), want["aabb"].view(np.uint32))
        rng = np.random.default_rng(12)
        for k in range(6):
            eye = rng.uniform(-40, 300, 3).astype(np.float32)
            vp = do.look_at_projection(eye, rng.uniform(0, 256, 3), fov_y_deg=float(rng.uniform(40, 100)))
            say(f"camera {k}: eye {eye.tolist()}")
            o, t = batch.cull(eye, vp)
            wo, wt = do.cull(want, eye, vp)
            for g, w in ((o, wo), (t, wt)):
                g = np.stack([g[f] for f in hcm.DRAW_CMD_DTYPE.names], 1)
                assert np.array_equal(g[np.argsort(g[:, 4])], w), "draw commands"
            assert 0 < len(o) + len(t) < recs.shape[0]
    finally:
        batch.close()
        store.close()
