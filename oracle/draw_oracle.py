"""CPU oracle of the draw-list step after meshing (TEST INFRASTRUCTURE ONLY).

Restates, in numpy float32 with the same operation order as the gfx950 kernels:
  - the mesh info of each BlockMesh, as WorldRenderer::PushChunkMesh builds it
    (client/include/client/WorldRenderer.hpp:69-87): world-space float AABB = AABB<uint32_t> / 16 +
    chunk position * 16, the base position, and the draw ranges in the quad pool;
  - the frustum test of client/shader/src/chunk_culling.comp:37-88: the eight corners of the AABB,
    relative to the view position, through the column-major view-projection matrix (each product row
    summed left to right without fused multiply-adds); culled when every corner lies outside one clip
    plane or behind the eye.
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


def keep_mask(aabb, view_pos, view_proj):
    """FrustumCull (:84) negated, per mesh: aabb (M, 6) float32 world-space min xyz, max xyz."""
    vpos = np.asarray(view_pos, np.float32)
    lo = (aabb[:, :3] - vpos).astype(np.float32)
    hi = (aabb[:, 3:] - vpos).astype(np.float32)
    bits = np.full(aabb.shape[0], 0xFF, np.uint32)
    for k in range(8):
        x = hi[:, 0] if k & 4 else lo[:, 0]
        y = hi[:, 1] if k & 2 else lo[:, 1]
        z = hi[:, 2] if k & 1 else lo[:, 2]
        bits &= cull_bits(view_proj, x, y, z)
    return (bits & 127) == 0


def cull(infos, view_pos, view_proj):
    """-> (opaque, transparent): (K, 5) uint32 draw commands of the kept meshes, by info index."""
    keep = keep_mask(infos["aabb"], view_pos, view_proj)
    idx = np.nonzero(keep)[0].astype(np.uint32)
    cmds = np.stack([infos["index_count"][idx], np.ones_like(idx), infos["first_index"][idx],
                     infos["vertex_offset"][idx], idx], 1).astype(np.uint32)
    tr = infos["transparent"][idx] != 0
    return cmds[~tr], cmds[tr]


def look_at_projection(eye, target, fov_y_deg=70.0, aspect=16 / 9, near=0.1, far=512.0):
    """A view-projection matrix (column-major, float32) looking from eye to target, in the view-position
    relative form the shader uses (the view translation is applied by subtracting the view position)."""
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
