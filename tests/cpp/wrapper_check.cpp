// Meshes the 16^3 terrain chunk (chunk (0, 1, 0) of the heightmap generator, seed 1) through
// hcm::BlockMeshGPU::Generate(get_block, get_light) -- the reference's call shape
// (client/include/client/BlockMeshAlgo.hpp:494-521) with callables returning types that expose
// GetData(), as block::Block and block::Light do -- and writes the mesh list to a file:
//   u32 mesh_count, then per mesh: u32 n_vertices, u32 n_indices, u32 aabb[6], u32 transparent,
//   the vertex words (2 x u32 each) and the u16 indices.
// Exit status: 0 meshed, 3 no gfx950 device, 1 other failure.
#include <cstdio>
#include <cstring>
#include <vector>

#include "hcm.hpp"
#include "hcm_workload.h"

namespace {
struct Block {
	uint16_t data;
	uint16_t GetData() const { return data; }
};
struct Light {
	uint8_t data;
	uint8_t GetData() const { return data; }
};
} // namespace

int main(int argc, char **argv) {
	if (argc != 2) {
		std::fprintf(stderr, "usage: %s OUT_FILE\n", argv[0]);
		return 1;
	}
	// 27 chunks around chunk (0, 1, 0): blocks[dx + 1][dy + 1][dz + 1]
	std::vector<uint16_t> blocks(27 * 4096);
	std::vector<uint8_t> light(27 * 4096);
	for (int dx = -1; dx <= 1; ++dx)
		for (int dy = -1; dy <= 1; ++dy)
			for (int dz = -1; dz <= 1; ++dz) {
				const int i = 9 * (dx + 1) + 3 * (dy + 1) + (dz + 1);
				hcm_gen_terrain_chunk(1, dx, 1 + dy, dz, 0, &blocks[4096 * i], &light[4096 * i]);
			}
	auto at = [&](int x, int y, int z) {
		const int cx = x < 0 ? -1 : (x >= 16 ? 1 : 0), cy = y < 0 ? -1 : (y >= 16 ? 1 : 0);
		const int cz = z < 0 ? -1 : (z >= 16 ? 1 : 0);
		const int i = 9 * (cx + 1) + 3 * (cy + 1) + (cz + 1);
		return 4096 * i + (x & 15) + 16 * ((z & 15) + 16 * (y & 15));
	};
	std::vector<hcm::BlockMesh> meshes;
	try {
		hcm::BlockMeshGPU mesher;
		meshes = mesher.Generate([&](int x, int y, int z) { return Block{blocks[at(x, y, z)]}; },
		                         [&](int x, int y, int z) { return Light{light[at(x, y, z)]}; });
	} catch (const hcm::Error &e) {
		std::fprintf(stderr, "%s\n", e.what());
		return e.code == HCM_ERR_NO_DEVICE ? 3 : 1;
	}
	FILE *f = std::fopen(argv[1], "wb");
	if (!f)
		return 1;
	const uint32_t n = (uint32_t)meshes.size();
	std::fwrite(&n, 4, 1, f);
	for (const auto &m : meshes) {
		const uint32_t hdr[9] = {(uint32_t)m.vertices.size(), (uint32_t)m.indices.size(), m.aabb.min[0], m.aabb.min[1],
		                         m.aabb.min[2], m.aabb.max[0], m.aabb.max[1], m.aabb.max[2], m.transparent ? 1u : 0u};
		std::fwrite(hdr, 4, 9, f);
		std::fwrite(m.vertices.data(), 8, m.vertices.size(), f);
		std::fwrite(m.indices.data(), 2, m.indices.size(), f);
	}
	std::fclose(f);
	return 0;
}
