// hcm.hpp -- C++ front end of the gfx950 chunk mesher, in the shape of the reference's mesher.
//
// hcm::BlockMeshGPU::Generate(get_block, get_light) takes the same two callables as
// hc::client::BlockMeshAlgo<Config>::Generate (client/include/client/BlockMeshAlgo.hpp:494-521) for a
// 16^3 chunk and returns the same std::vector<BlockMesh> (client/include/client/BlockVertex.hpp:8-25):
// packed vertices, u16 indices, AABB<uint32_t> in 1/16 block units, and the transparent flag, in the
// reference's mesh order. The callables are evaluated only where the mesher reads: blocks over
// x, z in [-2, 18), y in [-1, 17) (the 18^3 halo plus the x/z cells flowing water reads around halo
// cells, BlockMeshAlgo.hpp:202-236), light over the 18^3 halo [-1, 17)^3 -- the packed halo of
// hcm_mesh_halo; the meshing itself runs on the GPU through include/hcm.h.
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "hcm.h"

namespace hcm {

struct BlockVertex { // the reference's BlockVertex words (BlockVertex.hpp:8-18)
	uint32_t x10_y10_z10_axis2, tex10_trans3_face3_ao2_sl6_tl6;
};
static_assert(sizeof(BlockVertex) == 8, "two packed words");

struct AABB { // AABB<uint32_t> (dep/AABB.hpp:9-60)
	uint32_t min[3], max[3];
};

struct BlockMesh { // BlockMesh (BlockVertex.hpp:20-25)
	std::vector<BlockVertex> vertices;
	std::vector<uint16_t> indices;
	AABB aabb{};
	bool transparent{};
};

class Error : public std::runtime_error {
public:
	Error(const char *what, int code) : std::runtime_error(std::string(what) + " failed: " + std::to_string(code)), code(code) {}
	int code;
};

class BlockMeshGPU {
public:
	explicit BlockMeshGPU(const hcm_registry *registry = hcm_registry_default(), uint64_t pool_quads = 0) {
		int r = hcm_create(registry, pool_quads, &m_ctx);
		if (r != HCM_OK)
			throw Error("hcm_create", r);
	}
	~BlockMeshGPU() { hcm_destroy(m_ctx); }
	BlockMeshGPU(const BlockMeshGPU &) = delete;
	BlockMeshGPU &operator=(const BlockMeshGPU &) = delete;

	// Same call shape as BlockMeshAlgo::Generate: get_block(x, y, z) returns something convertible to the
	// 16-bit Block::m_data (e.g. via .GetData()) and get_light(x, y, z) the 8-bit Light::m_data.
	// The callables fill the chunk's packed halo (include/hcm.h), the only data sent to the device.
	template <typename GetBlockFunc, typename GetLightFunc>
	std::vector<BlockMesh> Generate(GetBlockFunc &&get_block_func, GetLightFunc &&get_light_func) {
		for (int y = -1; y < 17; ++y)
			for (int z = -2; z < 18; ++z)
				for (int x = -2; x < 18; ++x) {
					m_halo_blocks[(x + 2) + 20 * ((z + 2) + 20 * (y + 1))] = (uint16_t)data_of(get_block_func(x, y, z));
					if (x >= -1 && x < 17 && z >= -1 && z < 17)
						m_halo_light[(x + 1) + 18 * ((z + 1) + 18 * (y + 1))] = (uint8_t)data_of(get_light_func(x, y, z));
				}
		return Collect("hcm_mesh_halo", [&](hcm_host_result *out) {
			return hcm_mesh_halo(m_ctx, m_halo_blocks.data(), m_halo_light.data(), out);
		});
	}

	// The 27 chunks directly, in neighbour-slot order (common/include/common/Position.hpp:138-145);
	// e.g. neighbour_chunks[i]->GetBlockData() at client/src/ChunkMeshTask.cpp:117-119. (A distinct name:
	// arrays of pointers would otherwise bind to the callable overload above.)
	std::vector<BlockMesh> GenerateChunks(const uint16_t *const blocks[27], const uint8_t *const light[27]) {
		return Collect("hcm_mesh_chunk",
		               [&](hcm_host_result *out) { return hcm_mesh_chunk(m_ctx, blocks, light, out); });
	}

	// The whole of ChunkTaskRunner<kMesh>::Run (client/src/ChunkMeshTask.cpp:41-122): light rebuilt from
	// the 27 chunks' blocks and sunlight heights (Chunk::m_sunlight_heights, 16 x 16 int8 each, index
	// x + 16 z), then meshed. heights[i] belongs to the chunk of blocks[i].
	std::vector<BlockMesh> RunMeshTask(const uint16_t *const blocks[27], const int8_t *const heights[27]) {
		return Collect("hcm_mesh_chunk_relight", [&](hcm_host_result *out) {
			return hcm_mesh_chunk_relight(m_ctx, blocks, heights, out, nullptr);
		});
	}

	hcm_ctx *context() const { return m_ctx; }

private:
	template <typename Call> std::vector<BlockMesh> Collect(const char *what, Call &&call) {
		hcm_host_result out{};
		for (;;) {
			out.vertices = reinterpret_cast<uint32_t(*)[2]>(m_verts.data());
			out.vertex_capacity = (uint32_t)m_verts.size();
			out.indices = m_idx.data();
			out.index_capacity = (uint32_t)m_idx.size();
			out.meshes = m_desc.data();
			out.mesh_capacity = (uint32_t)m_desc.size();
			int r = call(&out);
			if (r == HCM_ERR_CAPACITY) {
				m_verts.resize(out.vertex_count);
				m_idx.resize(out.index_count);
				m_desc.resize(out.mesh_count);
				continue;
			}
			if (r != HCM_OK)
				throw Error(what, r);
			break;
		}
		std::vector<BlockMesh> meshes(out.mesh_count);
		for (uint32_t m = 0; m < out.mesh_count; ++m) {
			const hcm_mesh_desc &d = m_desc[m];
			meshes[m].vertices.assign(m_verts.begin() + d.vert_begin, m_verts.begin() + d.vert_begin + d.vert_count);
			meshes[m].indices.assign(m_idx.begin() + d.idx_begin, m_idx.begin() + d.idx_begin + d.idx_count);
			for (int a = 0; a < 3; ++a)
				meshes[m].aabb.min[a] = d.aabb[a], meshes[m].aabb.max[a] = d.aabb[3 + a];
			meshes[m].transparent = d.transparent != 0;
		}
		return meshes;
	}

	template <typename T> static auto data_of(const T &v) -> decltype(v.GetData()) { return v.GetData(); }
	static uint32_t data_of(uint32_t v) { return v; }

	hcm_ctx *m_ctx = nullptr;
	std::array<uint16_t, HCM_HALO_BLOCKS> m_halo_blocks{};
	std::array<uint8_t, HCM_HALO_LIGHT> m_halo_light{};
	std::vector<BlockVertex> m_verts = std::vector<BlockVertex>(1 << 16);
	std::vector<uint16_t> m_idx = std::vector<uint16_t>(3 << 15);
	std::vector<hcm_mesh_desc> m_desc = std::vector<hcm_mesh_desc>(8);
};

} // namespace hcm
