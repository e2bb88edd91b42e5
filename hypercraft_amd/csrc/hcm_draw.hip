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
//                     decisions are reproducible on any host.
#include <hip/hip_runtime.h>

#include "hcm_internal.h"

namespace hcm {

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
		m.aabb[a] = __fadd_rn(__fdiv_rn(float(r.aabb[a]), 16.0f), float(base));
		m.aabb[3 + a] = __fadd_rn(__fdiv_rn(float(r.aabb[3 + a]), 16.0f), float(base));
	}
	m.transparent = r.transparent;
	out[i] = m;
}

namespace {

// GetCullBits (chunk_culling.comp:37-47) of view_projection * (x, y, z, 1), column-major.
__device__ __forceinline__ uint32_t cull_bits(const float *vp, float x, float y, float z) {
	float c[4];
#pragma unroll
	for (int r = 0; r < 4; ++r)
		c[r] = __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(vp[r], x), __fmul_rn(vp[4 + r], y)), __fmul_rn(vp[8 + r], z)),
		                 vp[12 + r]);
	const float w = c[3];
	return (c[0] < -w ? 1u : 0u) | (c[0] > w ? 2u : 0u) | (c[1] < -w ? 4u : 0u) | (c[1] > w ? 8u : 0u) |
	       (c[2] < -w ? 16u : 0u) | (c[2] > w ? 32u : 0u) | (w <= 0.0f ? 64u : 0u) | (w > 0.0f ? 128u : 0u);
}

} // namespace

__global__ void cull_kernel(const hcm_mesh_info *infos, uint32_t n, CullParams cp, hcm_draw_cmd *opaque,
                            hcm_draw_cmd *transparent, uint32_t *counts) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	bool keep = false;
	hcm_mesh_info m;
	if (i < n) {
		m = infos[i];
		// GetClipInfo (:49-82): corners relative to the view position
		const float lo[3] = {__fsub_rn(m.aabb[0], cp.view_pos[0]), __fsub_rn(m.aabb[1], cp.view_pos[1]),
		                     __fsub_rn(m.aabb[2], cp.view_pos[2])};
		const float hi[3] = {__fsub_rn(m.aabb[3], cp.view_pos[0]), __fsub_rn(m.aabb[4], cp.view_pos[1]),
		                     __fsub_rn(m.aabb[5], cp.view_pos[2])};
		uint32_t bits = 0xffu;
#pragma unroll
		for (int k = 0; k < 8; ++k)
			bits &= cull_bits(cp.view_proj, (k & 4) ? hi[0] : lo[0], (k & 2) ? hi[1] : lo[1], (k & 1) ? hi[2] : lo[2]);
		keep = (bits & 127u) == 0u; // FrustumCull (:84)
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

hipError_t launch_cull(const hcm_mesh_info *infos, uint32_t n, const CullParams &cp, hcm_draw_cmd *opaque,
                       hcm_draw_cmd *transparent, uint32_t *counts, hipStream_t stream) {
	if (n == 0)
		return hipSuccess;
	hipLaunchKernelGGL(cull_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, infos, n, cp, opaque, transparent,
	                   counts);
	return hipGetLastError();
}

} // namespace hcm
