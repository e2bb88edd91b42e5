// hcm_draw.hip -- gfx950 kernels that turn a batch's meshes into a renderer's draw list.
//
//   mesh_info_kernel  one thread per mesh record: the culling pass's MeshInfo (index count, first index,
//                     vertex offset, world-space float AABB, base position, transparent flag), as
//                     WorldRenderer::PushChunkMesh builds it (client/include/client/WorldRenderer.hpp:69-87):
//                     AABB<uint32_t> / 16 + chunk position * 16, in float.
//   cull_kernel       one thread per mesh: the frustum test of client/shader/src/chunk_culling.comp:37-88
//                     (the eight AABB corners through the view-projection matrix; culled when every corner
//                     lies outside one clip plane or behind the eye), then wave-ballot compaction of the
//                     survivors into opaque and transparent DrawIndexedIndirect commands (:102-131). The
//                     matrix products are evaluated without fused multiply-adds, left to right, so the cull
//                     decisions are reproducible on any host. With a depth hierarchy it also runs the
//                     shader's occlusion test (:84-90, :104).
//   depth_level_kernel one thread per texel of a depth-hierarchy level: the largest of the 2x2 texels under
//                     it in the level above, plus the third column / row where that level's width / height
//                     is odd (client/shader/src/depth_hierarchy.comp, DepthHierarchyPass.hpp:166-196).
#include <hip/hip_runtime.h>

#include "hcm_internal.h"

namespace hcm {

namespace {

// One IEEE single-precision operation each, rounded to nearest even and never fused with another: the
// device library's rounded multiply, add and subtract (HIP's __fmul_rn and __fadd_rn are plain * and +,
// which hipcc's default -ffp-contract=fast fuses into FMAs), and the correctly rounded division hipcc
// emits for / (its default -fhip-fp32-correctly-rounded-divide-sqrt; a division is never fused).
__device__ __forceinline__ float f_mul(float a, float b) { return __ocml_mul_rte_f32(a, b); }
__device__ __forceinline__ float f_add(float a, float b) { return __ocml_add_rte_f32(a, b); }
__device__ __forceinline__ float f_sub(float a, float b) { return __ocml_sub_rte_f32(a, b); }
__device__ __forceinline__ float f_div(float a, float b) { return a / b; }

} // namespace

__global__ void mesh_info_kernel(const hcm_mesh_rec *recs, uint32_t n, const int32_t *chunk_pos, hcm_mesh_info *out) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const hcm_mesh_rec r = recs[i];
	hcm_mesh_info m;
	m.index_count = 6u * r.quad_count;
	m.first_index = 6u * r.first_quad;
	m.vertex_offset = 4u * r.first_quad;
	for (int a = 0; a < 3; ++a) {
		const int32_t base = chunk_pos[3 * r.job + a] * 16;
		m.base_position[a] = base;
		m.aabb[a] = f_add(f_div(float(r.aabb[a]), 16.0f), float(base));
		m.aabb[3 + a] = f_add(f_div(float(r.aabb[3 + a]), 16.0f), float(base));
	}
	m.transparent = r.transparent;
	out[i] = m;
}

namespace {

// view_projection * (x, y, z, 1), column-major, each row summed left to right.
__device__ __forceinline__ void project(const float *vp, float x, float y, float z, float c[4]) {
#pragma unroll
	for (int r = 0; r < 4; ++r)
		c[r] = f_add(f_add(f_add(f_mul(vp[r], x), f_mul(vp[4 + r], y)), f_mul(vp[8 + r], z)),
		                 vp[12 + r]);
}

// GetCullBits (chunk_culling.comp:37-47).
__device__ __forceinline__ uint32_t cull_bits(const float c[4]) {
	const float w = c[3];
	return (c[0] < -w ? 1u : 0u) | (c[0] > w ? 2u : 0u) | (c[1] < -w ? 4u : 0u) | (c[1] > w ? 8u : 0u) |
	       (c[2] < -w ? 16u : 0u) | (c[2] > w ? 32u : 0u) | (w <= 0.0f ? 64u : 0u) | (w > 0.0f ? 128u : 0u);
}

// GLSL min / max (y < x ? y : x, x < y ? y : x).
__device__ __forceinline__ float gmin(float x, float y) { return y < x ? y : x; }
__device__ __forceinline__ float gmax(float x, float y) { return x < y ? y : x; }

// OcclusionCull (chunk_culling.comp:84-90) for a box whose corners all have w > 0. The texture level is
// ceil(log2(m)) taken exactly from m's binary exponent, clamped to the texture's mips (the sampler's
// minLod 0 and unbounded maxLod; the level's fraction is 0, so the linear mip blend reads one level);
// the texel is the nearest one, clamped to the edge.
__device__ __forceinline__ bool occluded(const DepthPyramid &h, const float cmin[3], const float cmax[3]) {
	const float u0 = f_add(f_mul(cmin[0], 0.5f), 0.5f), v0 = f_add(f_mul(cmin[1], 0.5f), 0.5f);
	const float u1 = f_add(f_mul(cmax[0], 0.5f), 0.5f), v1 = f_add(f_mul(cmax[1], 0.5f), 0.5f);
	const float m = f_mul(gmax(f_sub(u1, u0), f_sub(v1, v0)), float(max(h.width[1], h.height[1])));
	const int q = int(h.n_levels) - 2; // the texture's last mip
	int mip = 0;
	if (m > 0.0f) {
		if (isinf(m)) {
			mip = q;
		} else {
			int e;
			const float f = frexpf(m, &e); // m = f * 2^e, f in [0.5, 1)
			const int n = f == 0.5f ? e - 1 : e;
			mip = n < 0 ? 0 : (n > q ? q : n);
		}
	}
	const int L = mip + 1;
	const uint32_t w = h.width[L], hh = h.height[L];
	const float fx = floorf(f_mul(f_mul(f_add(u0, u1), 0.5f), float(w)));
	const float fy = floorf(f_mul(f_mul(f_add(v0, v1), 0.5f), float(hh)));
	const uint32_t x = !(fx > 0.0f) ? 0u : (fx >= float(w) ? w - 1u : uint32_t(fx));
	const uint32_t y = !(fy > 0.0f) ? 0u : (fy >= float(hh) ? hh - 1u : uint32_t(fy));
	return h.data[h.offset[L] + y * w + x] < cmin[2];
}

} // namespace

__global__ void depth_level_kernel(const float *prev, uint32_t pw, uint32_t ph, float *cur, uint32_t cw, uint32_t ch) {
	const uint32_t x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
	if (x >= cw || y >= ch)
		return;
	// texels outside the level above read 0 (robust image access), which never raises a depth in [0, 1]
	auto at = [&](uint32_t px, uint32_t py) { return px < pw && py < ph ? prev[py * pw + px] : 0.0f; };
	const uint32_t bx = 2 * x, by = 2 * y;
	float d = gmax(gmax(at(bx, by), at(bx + 1, by)), gmax(at(bx + 1, by + 1), at(bx, by + 1)));
	if (pw & 1u) {
		d = gmax(d, gmax(at(bx + 2, by), at(bx + 2, by + 1)));
		if (ph & 1u)
			d = gmax(d, at(bx + 2, by + 2));
	}
	if (ph & 1u)
		d = gmax(d, gmax(at(bx, by + 2), at(bx + 1, by + 2)));
	cur[y * cw + x] = d;
}

__global__ void cull_kernel(const hcm_mesh_info *infos, uint32_t n, CullParams cp, hcm_draw_cmd *opaque,
                            hcm_draw_cmd *transparent, uint32_t *counts) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	bool keep = false;
	hcm_mesh_info m;
	if (i < n) {
		m = infos[i];
		// GetClipInfo (:49-82): corners relative to the view position, their clip-space bounds
		const float lo[3] = {f_sub(m.aabb[0], cp.view_pos[0]), f_sub(m.aabb[1], cp.view_pos[1]),
		                     f_sub(m.aabb[2], cp.view_pos[2])};
		const float hi[3] = {f_sub(m.aabb[3], cp.view_pos[0]), f_sub(m.aabb[4], cp.view_pos[1]),
		                     f_sub(m.aabb[5], cp.view_pos[2])};
		uint32_t bits = 0xffu;
		float cmin[3], cmax[3];
#pragma unroll
		for (int k = 0; k < 8; ++k) {
			float c[4];
			project(cp.view_proj, (k & 4) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 1) ? hi[2] : lo[2], c);
			bits &= cull_bits(c);
#pragma unroll
			for (int a = 0; a < 3; ++a) {
				const float v = f_div(c[a], c[3]);
				cmin[a] = k ? gmin(cmin[a], v) : v;
				cmax[a] = k ? gmax(cmax[a], v) : v;
			}
		}
		keep = (bits & 127u) == 0u; // FrustumCull (:82)
		if (keep && cp.hiz.data && (bits & 128u)) // OcclusionCull (:84-90): every corner has w > 0
			keep = !occluded(cp.hiz, cmin, cmax);
	}
	// one atomic per wavefront and pass, slots in lane order (:102-131)
	const uint32_t lane = threadIdx.x & 63u;
	for (int pass = 0; pass < 2; ++pass) {
		const bool mine = keep && (m.transparent != 0) == (pass == 1);
		const uint64_t ballot = __ballot(mine);
		if (!ballot)
			continue;
		uint32_t base = 0;
		if (lane == __ffsll((unsigned long long)ballot) - 1)
			base = atomicAdd(&counts[pass], (uint32_t)__popcll(ballot));
		base = (uint32_t)__shfl((int)base, __ffsll((unsigned long long)ballot) - 1);
		if (mine) {
			const uint32_t slot = base + (uint32_t)__popcll(ballot & ((1ull << lane) - 1ull));
			(pass ? transparent : opaque)[slot] = hcm_draw_cmd{m.index_count, 1u, m.first_index, m.vertex_offset, i};
		}
	}
}

hipError_t launch_mesh_infos(const hcm_mesh_rec *recs, uint32_t n, const int32_t *chunk_pos, hcm_mesh_info *out,
                             hipStream_t stream) {
	if (n == 0)
		return hipSuccess;
	hipLaunchKernelGGL(mesh_info_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, recs, n, chunk_pos, out);
	return hipGetLastError();
}

hipError_t launch_depth_level(const float *prev, uint32_t pw, uint32_t ph, float *cur, uint32_t cw, uint32_t ch,
                              hipStream_t stream) {
	const dim3 block(16, 16), grid((cw + 15) / 16, (ch + 15) / 16);
	hipLaunchKernelGGL(depth_level_kernel, grid, block, 0, stream, prev, pw, ph, cur, cw, ch);
	return hipGetLastError();
}

hipError_t launch_cull(const hcm_mesh_info *infos, uint32_t n, const CullParams &cp, hcm_draw_cmd *opaque,
                       hcm_draw_cmd *transparent, uint32_t *counts, hipStream_t stream) {
	if (n == 0)
		return hipSuccess;
	hipLaunchKernelGGL(cull_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, infos, n, cp, opaque, transparent,
	                   counts);
	return hipGetLastError();
}

} // namespace hcm
