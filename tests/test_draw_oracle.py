"""Pins the draw-list oracle (oracle/draw_oracle.py) before it checks the GPU culling.

Closed forms of WorldRenderer::PushChunkMesh's mesh info (WorldRenderer.hpp:69-87) and of the frustum
test of chunk_culling.comp:37-88 (boxes in front of, behind, beside and across the view), plus agreement
of the float32 decisions with a float64 evaluation for every mesh not within rounding of a clip plane.
"""
import numpy as np

import draw_oracle as do
from hypercraft_amd import hcm, workloads


def rec(aabb, first_quad=0, quad_count=1, job=0, transparent=0):
    r = np.zeros(1, hcm.MESH_REC_DTYPE)
    r["aabb"][0] = aabb
    r["first_quad"], r["quad_count"], r["job"], r["transparent"] = first_quad, quad_count, job, transparent
    return r


def test_mesh_info_closed_form():
    r = rec((0, 8, 16, 256, 128, 200), first_quad=1000, quad_count=7, job=1, transparent=1)
    info = do.mesh_infos(r, np.array([[0, 0, 0], [2, -1, 3]], np.int32))
    assert info["aabb"][0].tolist() == [32.0, -15.5, 49.0, 48.0, -8.0, 60.5]
    assert info["base_position"][0].tolist() == [32, -16, 48]
    assert (info["index_count"][0], info["first_index"][0], info["vertex_offset"][0]) == (42, 6000, 4000)
    assert info["transparent"][0] == 1


def boxes(*aabbs):
    return {"aabb": np.array(aabbs, np.float32), "index_count": np.arange(len(aabbs), dtype=np.uint32) * 6 + 6,
            "first_index": np.zeros(len(aabbs), np.uint32), "vertex_offset": np.zeros(len(aabbs), np.uint32),
            "transparent": np.array([i % 2 for i in range(len(aabbs))], np.uint32)}


def test_frustum_closed_forms():
    eye = np.array([0.0, 0.0, 0.0], np.float32)
    vp = workloads.look_at_projection(eye, (0.0, 0.0, -1.0))  # looking down -z
    infos = boxes((-1, -1, -12, 1, 1, -10),      # in front: kept
                  (-1, -1, 10, 1, 1, 12),        # behind the eye: every w <= 0
                  (-200, -1, -12, -150, 1, -10), # far to the left of the view
                  (-1, -1, -600, 1, 1, -550),    # beyond the far plane
                  (-50, -1, -12, 50, 1, -10),    # wider than the view: corners on both sides
                  (-1, -1, -1, 1, 1, 1))         # around the eye: corners in front and behind
    assert do.keep_mask(infos["aabb"], eye, vp).tolist() == [True, False, False, False, True, True]
    o, t = do.cull(infos, eye, vp)
    assert sorted(o[:, 4].tolist() + t[:, 4].tolist()) == [0, 4, 5]
    assert all(c[1] == 1 for c in np.concatenate([o, t]))


def test_float32_decisions_agree_with_float64():
    rng = np.random.default_rng(4)
    lo = rng.uniform(-256, 256, (4000, 3))
    aabb = np.concatenate([lo, lo + rng.uniform(0.0625, 16, (4000, 3))], 1).astype(np.float32)
    for _ in range(10):
        eye = rng.uniform(-100, 100, 3).astype(np.float32)
        vp = workloads.look_at_projection(eye, eye + rng.normal(size=3))
        keep32 = do.keep_mask(aabb, eye, vp)
        # float64: the same test, with a margin: corners within 1e-3 of a plane are left out
        m = vp.astype(np.float64).reshape(4, 4).T
        rel = aabb.astype(np.float64) - np.concatenate([eye, eye]).astype(np.float64)
        corners = np.stack([np.stack([rel[:, 3 if k & 4 else 0], rel[:, 4 if k & 2 else 1], rel[:, 5 if k & 1 else 2],
                                      np.ones(len(rel))], 1) for k in range(8)], 1)  # (M, 8, 4)
        c = corners @ m.T
        w = c[..., 3]
        margin = np.min(np.abs(np.concatenate([c[..., :3] - w[..., None], c[..., :3] + w[..., None], w[..., None]],
                                              -1)), axis=(1, 2))
        out_all = [(c[..., a] < -w).all(1) | (c[..., a] > w).all(1) for a in range(3)]
        keep64 = ~(out_all[0] | out_all[1] | out_all[2] | (w <= 0).all(1))
        clear = margin > 1e-3
        assert clear.mean() > 0.99
        assert np.array_equal(keep32[clear], keep64[clear])
        assert 0 < keep32.sum() < len(keep32)


# ---- depth hierarchy and occlusion (depth_hierarchy.comp, chunk_culling.comp:84-90) ----
def test_depth_pyramid_closed_forms():
    rng = np.random.default_rng(21)
    img = rng.random((16, 16), dtype=np.float32)
    p = do.depth_pyramid(img)
    assert [lv.shape for lv in p] == [(16, 16), (8, 8), (4, 4), (2, 2), (1, 1)]
    for L in range(5):  # power-of-two sizes: the max over each 2^L x 2^L block
        s = 1 << L
        assert np.array_equal(p[L], img.reshape(16 // s, s, 16 // s, s).max(axis=(1, 3)))
    # odd width and height: every texel also takes the third column and row (3 x 3 footprints)
    img = rng.random((5, 7), dtype=np.float32)
    p = do.depth_pyramid(img)
    assert [lv.shape for lv in p] == [(5, 7), (2, 3), (1, 1)]
    for y in range(2):
        for x in range(3):
            assert p[1][y, x] == img[2 * y:2 * y + 3, 2 * x:2 * x + 3].max()
    assert p[2][0, 0] == p[1].max()
    # one row: the texels past the edge read 0; level count is the bit length of width | height
    assert [lv.tolist() for lv in do.depth_pyramid(np.array([[0.25, 0.75, 0.5]], np.float32))] == \
        [[[0.25, 0.75, 0.5]], [[0.75]]]
    for h, w in ((1080, 1920), (777, 333), (3, 3), (2, 1), (1, 2), (9, 1000)):
        img = rng.random((h, w), dtype=np.float32)
        p = do.depth_pyramid(img)
        assert len(p) == (w | h).bit_length() and p[-1].shape == (1, 1)
        for a, b in zip(p, p[1:]):
            assert b.shape == (max(a.shape[0] >> 1, 1), max(a.shape[1] >> 1, 1))
            assert b.max() == a.max()  # every texel is under some texel of the next level


def test_ceil_log2_closed_forms():
    m = np.array([1, 2, 3, 4, 0.5, 0.75, 1.0000001, 1023, 1024, 1025, 2.0 ** -20, 3e-30], np.float32)
    assert do.ceil_log2(m).tolist() == [0, 1, 2, 2, -1, 0, 1, 10, 10, 11, -20, -98]
    import math
    rng = np.random.default_rng(3)
    r = (rng.random(10000) * 10.0 ** rng.uniform(-6, 6, 10000)).astype(np.float32)
    assert do.ceil_log2(r).tolist() == [math.ceil(math.log2(float(v))) for v in r]


def wall_scene():
    eye = np.array([0.0, 0.0, 0.0], np.float32)
    vp = workloads.look_at_projection(eye, (0.0, 0.0, -1.0))
    zc = lambda d: float(do.project(vp, np.float32(0), np.float32(0), np.float32(-d))[2] /
                         do.project(vp, np.float32(0), np.float32(0), np.float32(-d))[3])
    # a 128 x 72 depth image: a wall at distance 20 over the middle third, the far plane elsewhere
    img = np.ones((72, 128), np.float32)
    img[24:48, 43:85] = zc(20.0)
    return eye, vp, img, zc


def test_occlusion_closed_forms():
    eye, vp, img, zc = wall_scene()
    infos = boxes((-1, -1, -31, 1, 1, -30),     # behind the wall, inside its screen rectangle: occluded
                  (-1, -1, -11, 1, 1, -10),     # in front of the wall: kept
                  (-25, -1, -31, -23, 1, -30),  # behind the wall's depth but left of it on screen: kept
                  (-30, -1, -31, 30, 1, -30),   # wider than the wall: its coarse texel sees the far plane
                  (-1, -1, -1, 1, 1, 1))        # around the eye (some w <= 0): never tested
    p = do.depth_pyramid(img)
    assert do.keep_mask(infos["aabb"], eye, vp, p).tolist() == [False, True, True, True, True]
    assert do.keep_mask(infos["aabb"], eye, vp).tolist() == [True] * 5
    far = do.depth_pyramid(np.ones((72, 128), np.float32))  # everything at the far plane: nothing hidden
    assert do.keep_mask(infos["aabb"], eye, vp, far).all()
    near = do.depth_pyramid(np.zeros((72, 128), np.float32))  # everything at the near plane
    assert do.keep_mask(infos["aabb"], eye, vp, near).tolist() == [False, False, False, False, True]
    o, t = do.cull(infos, eye, vp, p)
    assert sorted(o[:, 4].tolist() + t[:, 4].tolist()) == [1, 2, 3, 4]


def test_occlusion_float32_agrees_with_float64():
    """The occlusion decision in float64 with math.log2, for boxes whose decision is not within rounding
    of a mip boundary, a texel edge or the depth comparison."""
    import math
    rng = np.random.default_rng(8)
    n = 3000
    lo = rng.uniform(-64, 64, (n, 3))
    aabb = np.concatenate([lo, lo + rng.uniform(0.5, 16, (n, 3))], 1).astype(np.float32)
    yy, xx = np.mgrid[0:90, 0:160]
    checked = 0
    for k in range(6):
        eye = (rng.uniform(-20, 20, 3) + np.array([0, 0, 90])).astype(np.float32)
        vp = workloads.look_at_projection(eye, (0.0, 0.0, 0.0))
        img = (0.985 + 0.015 * np.sin(xx / (7 + k)) * np.cos(yy / 5)).astype(np.float32)
        p = do.depth_pyramid(img)
        bits, cmin, cmax = do.clip_info(aabb, eye, vp)
        test = ((bits & 127) == 0) & ((bits & 128) != 0)
        got = do.occluded(p, cmin[test], cmax[test])
        # float64 restatement
        m = vp.astype(np.float64).reshape(4, 4).T
        rel = aabb[test].astype(np.float64) - np.concatenate([eye, eye]).astype(np.float64)
        cs = np.stack([np.stack([rel[:, 3 if c & 4 else 0], rel[:, 4 if c & 2 else 1], rel[:, 5 if c & 1 else 2],
                                 np.ones(len(rel))], 1) for c in range(8)], 1) @ m.T
        clip = cs[..., :3] / cs[..., 3:]
        mn, mx = clip.min(1), clip.max(1)
        u0, v0, u1, v1 = mn[:, 0] * .5 + .5, mn[:, 1] * .5 + .5, mx[:, 0] * .5 + .5, mx[:, 1] * .5 + .5
        mh, mw = p[1].shape
        mm = np.maximum(u1 - u0, v1 - v0) * max(mw, mh)
        want = np.zeros(len(mm), bool)
        clear = np.ones(len(mm), bool)
        for i in range(len(mm)):
            lg = math.log2(mm[i]) if mm[i] > 0 else -math.inf
            mip = int(min(max(math.ceil(lg), 0), len(p) - 2)) if lg > -math.inf else 0
            if lg > -math.inf and abs(lg - round(lg)) < 1e-4:
                clear[i] = False
            lv = p[mip + 1]
            fx, fy = (u0[i] + u1[i]) * .5 * lv.shape[1], (v0[i] + v1[i]) * .5 * lv.shape[0]
            if abs(fx - round(fx)) < 1e-4 or abs(fy - round(fy)) < 1e-4:
                clear[i] = False
            x = min(max(int(math.floor(fx)), 0), lv.shape[1] - 1)
            y = min(max(int(math.floor(fy)), 0), lv.shape[0] - 1)
            if abs(float(lv[y, x]) - mn[i, 2]) < 1e-5:
                clear[i] = False
            want[i] = lv[y, x] < mn[i, 2]
        assert clear.mean() > 0.95
        assert np.array_equal(got[clear], want[clear])
        checked += clear.sum()
        assert 0 < got.sum() < len(got)
    assert checked > 1000
