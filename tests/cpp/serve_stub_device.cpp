// The device side of the serving loop replaced by CPU code, for the sanitizer builds of hcm_serve.cpp
// (tests/cpp/serve_check.cpp). The mesher here is a stand-in with the kernel's bookkeeping -- one
// contiguous quad range per job from the pool cursor, mesh records, per-job status, POOL_FULL past the
// arena, the last workgroup's totals and counter reset -- and a toy mesh: (1 + block sum % 3) quads per
// job whose words depend on the center chunk's blocks, so stale data would show in the final meshes.
#include <hip/hip_runtime.h>

#include <vector>

#include "hcm_internal.h"

namespace hcm {

int ctx_info(hcm_ctx *, uint32_t threads, CtxInfo *out) {
	static std::vector<uint8_t> light_info = [] {
		std::vector<uint8_t> t(65536, 0);
		for (uint32_t d = 0; d < 65536; ++d)
			if ((d & 0xffu) == 0 || (d & 0xffu) == 9) // air and glass pass light; air passes sunlight straight down
				t[d] = kLightInfoPass | ((d & 0xffu) == 0 ? kLightInfoVertical : 0);
		return t;
	}();
	*out = CtxInfo{};
	out->n_cu = 2;
	out->threads = threads ? threads : 256;
	out->grid = 4;
	out->light_info = light_info.data();
	out->light_info_host = light_info.data();
	return HCM_OK;
}

uint32_t toy_quads(const uint16_t *center) {
	uint32_t sum = 0;
	for (int i = 0; i < HCM_CHUNK_VOXELS; ++i)
		sum += center[i];
	return 1 + sum % 3;
}

hipError_t launch_mesh(const MeshParams &p, int, int, hipStream_t) {
	for (uint32_t j = 0; j < p.n_jobs; ++j) {
		const uint32_t center = p.nbr[27 * (size_t)j + 26];
		const uint16_t *b = p.blocks + (size_t)center * HCM_CHUNK_VOXELS;
		const uint32_t q = toy_quads(b);
		const uint32_t base = p.ctr->pool_cursor, rec = p.ctr->mesh_cursor;
		p.ctr->pool_cursor += q;
		p.ctr->mesh_cursor += 1;
		if ((uint64_t)base + q > p.pool_cap || rec + 1 > p.rec_cap) {
			p.results[j] = hcm_job_result{HCM_ERR_POOL_FULL, 0u, 0u, q};
			continue;
		}
		for (uint32_t k = 0; k < 4 * q; ++k)
			p.pool_verts[4 * (size_t)base + k] = make_uint2(b[k % HCM_CHUNK_VOXELS] | k << 16, center);
		for (uint32_t k = 0; k < 6 * q; ++k)
			p.pool_idx[6 * (size_t)base + k] = uint16_t(k);
		hcm_mesh_rec r{};
		r.first_quad = base;
		r.quad_count = q;
		r.job = j;
		p.recs[rec] = r;
		p.results[j] = hcm_job_result{HCM_OK, rec, 1u, q};
	}
	if (p.self_reset) {
		p.ctr->total_quads = p.ctr->pool_cursor;
		p.ctr->total_meshes = p.ctr->mesh_cursor;
		p.ctr->pool_cursor = p.ctr->mesh_cursor = p.ctr->done = 0;
		for (auto &q : p.ctr->queue)
			q.v = 0;
	}
	return hipSuccess;
}

hipError_t launch_edits(uint16_t *blocks, uint8_t *light, int8_t *heights, const VoxelEdit *edits, uint32_t n,
                        hipStream_t) {
	for (uint32_t i = 0; i < n; ++i) {
		const VoxelEdit &e = edits[i];
		if (e.local >= HCM_CHUNK_VOXELS) {
			heights[(size_t)e.slot * 256 + (e.local - HCM_CHUNK_VOXELS)] = int8_t(uint8_t(e.block));
			continue;
		}
		blocks[(size_t)e.slot * HCM_CHUNK_VOXELS + e.local] = uint16_t(e.block);
		light[(size_t)e.slot * HCM_CHUNK_VOXELS + e.local] = uint8_t(e.light);
	}
	return hipSuccess;
}

hipError_t launch_upload_scatter(const uint16_t *sblocks, const uint8_t *slight, const int8_t *sheights,
                                 const uint2 *moves, uint32_t n, uint16_t *blocks, uint8_t *light, int8_t *heights,
                                 hipStream_t) {
	for (uint32_t i = 0; i < n; ++i) {
		const size_t k = moves[i].x, slot = moves[i].y;
		std::memcpy(blocks + slot * HCM_CHUNK_VOXELS, sblocks + k * HCM_CHUNK_VOXELS, HCM_CHUNK_VOXELS * 2);
		if (slight)
			std::memcpy(light + slot * HCM_CHUNK_VOXELS, slight + k * HCM_CHUNK_VOXELS, HCM_CHUNK_VOXELS);
		if (sheights)
			std::memcpy(heights + slot * 256, sheights + k * 256, 256);
	}
	return hipSuccess;
}

hipError_t launch_light_prep(const LightParams &, hipStream_t) { return hipSuccess; }
hipError_t launch_light(const LightParams &p, int, hipStream_t) {
	std::memset(p.halo_light, 0, (size_t)p.n_jobs * kLightJobBytes);
	return hipSuccess;
}
hipError_t launch_reset(Counters *ctr, hipStream_t) {
	*ctr = Counters{};
	return hipSuccess;
}

} // namespace hcm
