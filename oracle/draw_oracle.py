"""CPU oracle of the draw-list step after meshing (TEST INFRASTRUCTURE ONLY).

Restates, in numpy float32 with the same operation order as the gfx950 kernels:
  - the mesh info of each BlockMesh, as WorldRenderer::PushChunkMesh builds it
    (client/include/client/WorldRenderer.hpp:69-87): world-space float AABB = AABB<uint32_t> / 16 +
    chunk position * 16, the base position, and the draw ranges in the quad pool;
  - the frustum test of client/shader/src/chunk_culling.comp:37-88: the eight corners of the AABB,
    relative to the view position, through the column-major view-projection matrix (each product row
    summed left to right without fused multiply-adds); culled when every corner lies outside one clip
    plane or behind the eye;
  - the depth hierarchy of client/shader/src/depth_hierarchy.comp as DepthHierarchyPass.hpp:166-196 chains
    it, and the occlusion test of chunk_culling.comp:84-90: clip-space box of the corners (IEEE division),
    mip ceil(log2(...)) from the exact binary exponent, nearest texel clamped to the edge.
Only tests/ may import this module.
"""
import numpy as np

f32 = np.float32


def mesh_infos(recs, chunk_pos):
    """recs: hcm_mesh_rec structured array (aabb, first_quad, quad_count, job, transparent);
    chunk_pos: (J, 3) int32. -> dict of numpy arrays, one entry per record."""
    base = (np.asarray(chunk_pos, np.int32)[recs["job"]] * 16).astype(np.int32)
    aabb = recs["aabb"].astype(np.float32) / f32(16.0)
    aabb = aabb + np.concatenate([base, base], 1).astype(np.float32)
    return {"index_count": 6 * recs["quad_count"].astype(np.uint32), "first_index": 6 * recs["first_quad"].astype(np.uint32),
            "vertex_offset": 4 * recs["first_quad"].astype(np.uint32), "aabb": aabb.astype(np.float32),
            "base_position": base, "transparent": recs["transparent"].astype(np.uint32)}


def cull_bits(vp, x, y, z):
    """GetCullBits (chunk_culling.comp:37-47) of vp * (x, y, z, 1) for arrays of points (float32)."""
    vp = np.asarray(vp, np.float32).reshape(16)
    c = [((vp[r] * x + vp[4 + r] * y) + vp[8 + r] * z) + vp[12 + r] for r in range(4)]
    w = c[3]
    bits = ((c[0] < -w) * 1 | (c[0] > w) * 2 | (c[1] < -w) * 4 | (c[1] > w) * 8 | (c[2] < -w) * 16 |
            (c[2] > w) * 32 | (w <= 0) * 64 | (w > 0) * 128)
    return bits.astype(np.uint32)


def project(vp, x, y, z):
    """view_projection * (x, y, z, 1), each row summed left to right in float32: [x, y, z, w] arrays."""
    vp = np.asarray(vp, np.float32).reshape(16)
    return [((vp[r] * x + vp[4 + r] * y) + vp[8 + r] * z) + vp[12 + r] for r in range(4)]


def gmin(a, b):
    return np.where(b < a, b, a)  # GLSL min(a, b)


def gmax(a, b):
    return np.where(a < b, b, a)  # GLSL max(a, b)


def clip_info(aabb, view_pos, view_proj):
    """GetClipInfo (:49-80) per mesh: (cull bits ANDed over the corners, clip min (M, 3), clip max (M, 3)),
    corners in the shader's order."""
    vpos = np.asarray(view_pos, np.float32)
    lo = (aabb[:, :3] - vpos).astype(np.float32)
    hi = (aabb[:, 3:] - vpos).astype(np.float32)
    bits = np.full(aabb.shape[0], 0xFF, np.uint32)
    cmin = cmax = None
    for k in range(8):
        x = hi[:, 0] if k & 4 else lo[:, 0]
        y = hi[:, 1] if k & 2 else lo[:, 1]
        z = hi[:, 2] if k & 1 else lo[:, 2]
        bits &= cull_bits(view_proj, x, y, z)
        c = project(view_proj, x, y, z)
        with np.errstate(divide="ignore", invalid="ignore"):
            clip = np.stack([c[a] / c[3] for a in range(3)], 1).astype(np.float32)
        cmin = clip if cmin is None else gmin(cmin, clip)
        cmax = clip if cmax is None else gmax(cmax, clip)
    return bits, cmin, cmax


def depth_pyramid(depth):
    """The depth hierarchy of a (height, width) float32 depth image: [level 0 = the image, level 1, ...],
    myvk QueryMipLevel levels (bit length of width | height), each texel the largest of the 2 x 2 texels
    under it plus the third column / row where the level above is odd-sized; texels past the edge read 0."""
    levels = [np.ascontiguousarray(depth, np.float32)]
    h0, w0 = levels[0].shape
    n = int(w0 | h0).bit_length()
    for _ in range(1, n):
        prev = levels[-1]
        ph, pw = prev.shape
        ch, cw = max(ph >> 1, 1), max(pw >> 1, 1)
        pad = np.zeros((2 * ch + 3, 2 * cw + 3), np.float32)
        pad[:ph, :pw] = prev
        at = lambda dy, dx: pad[dy:dy + 2 * ch:2, dx:dx + 2 * cw:2]
        d = gmax(gmax(at(0, 0), at(0, 1)), gmax(at(1, 1), at(1, 0)))
        if pw & 1:
            d = gmax(d, gmax(at(0, 2), at(1, 2)))
            if ph & 1:
                d = gmax(d, at(2, 2))
        if ph & 1:
            d = gmax(d, gmax(at(2, 0), at(2, 1)))
        levels.append(d.astype(np.float32))
    return levels


def ceil_log2(m):
    """ceil(log2(m)) for m > 0 finite, exactly, from the binary exponent: m = f * 2^e, f in [0.5, 1)."""
    f, e = np.frexp(np.asarray(m, np.float32))
    return np.where(f == np.float32(0.5), e - 1, e)


def occluded(pyramid, cmin, cmax):
    """OcclusionCull (:84-90) for boxes whose corners all have w > 0: the nearest texel at the box centre,
    at mip ceil(log2(max uv extent * max(mip 0 width, height))) of the texture holding levels 1 .. n - 1."""
    h = np.float32(0.5)
    u0, v0 = cmin[:, 0] * h + h, cmin[:, 1] * h + h
    u1, v1 = cmax[:, 0] * h + h, cmax[:, 1] * h + h
    mh, mw = pyramid[1].shape
    m = gmax(u1 - u0, v1 - v0) * np.float32(max(mw, mh))
    q = len(pyramid) - 2
    with np.errstate(invalid="ignore"):
        pos = m > 0
        fin = np.isfinite(m)
    n = np.zeros(m.shape, np.int64)
    ok = pos & fin
    n[ok] = ceil_log2(m[ok])
    n[pos & ~fin] = q
    mip = np.clip(n, 0, q)
    out = np.zeros(m.shape, bool)
    for L in np.unique(mip) + 1:
        sel = mip + 1 == L
        lv = pyramid[L]
        lh, lw = lv.shape
        fx = np.floor(((u0[sel] + u1[sel]) * h) * np.float32(lw))
        fy = np.floor(((v0[sel] + v1[sel]) * h) * np.float32(lh))
        x = np.where(~(fx > 0), 0, np.where(fx >= lw, lw - 1, np.nan_to_num(fx))).astype(np.int64)
        y = np.where(~(fy > 0), 0, np.where(fy >= lh, lh - 1, np.nan_to_num(fy))).astype(np.int64)
        out[sel] = lv[y, x] < cmin[sel, 2]
    return out


def keep_mask(aabb, view_pos, view_proj, pyramid=None):
    """Neither FrustumCull (:82) nor, with a depth pyramid, OcclusionCull (:84-90), per mesh: aabb (M, 6)
    float32 world-space min xyz, max xyz."""
    bits, cmin, cmax = clip_info(aabb, view_pos, view_proj)
    keep = (bits & 127) == 0
    if pyramid is not None:
        test = keep & ((bits & 128) != 0)
        if test.any():
            occ = np.zeros(keep.shape, bool)
            occ[test] = occluded(pyramid, cmin[test], cmax[test])
            keep &= ~occ
    return keep


def cull(infos, view_pos, view_proj, pyramid=None):
    """-> (opaque, transparent): (K, 5) uint32 draw commands of the kept meshes, by info index."""
    keep = keep_mask(infos["aabb"], view_pos, view_proj, pyramid)
    idx = np.nonzero(keep)[0].astype(np.uint32)
    cmds = np.stack([infos["index_count"][idx], np.ones_like(idx), infos["first_index"][idx],
                     infos["vertex_offset"][idx], idx], 1).astype(np.uint32)
    tr = infos["transparent"][idx] != 0
    return cmds[~tr], cmds[tr]
