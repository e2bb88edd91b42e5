// ref_driver.cpp -- wraps the REAL reference mesher for use as a test oracle (TEST INFRASTRUCTURE ONLY).
//
// Compiled by oracle/Makefile against the reference sources where they lie (REF=/root/reference);
// the result goes to oracle/_ref/libhcref.so and is never part of the product. It exposes:
//   hcref_mesh_chunk      -- BlockMeshAlgo<Config>::Generate (client/include/client/BlockMeshAlgo.hpp:494-521)
//                            with Config = BlockAlgoConfig<int8_t, {0,0,0,16,16,16}, YZX>, fed by the same
//                            27-chunk lambdas as client/src/ChunkMeshTask.cpp:113-122 (light is an input
//                            array instead of the rebuilt light buffer);
//   hcref_mesh_many       -- the same over many jobs on a std::thread pool (CPU baseline, one chunk per task
//                            like the reference's worker threads, client/src/WorldWorker.cpp:18-27);
//   hcref_light_rebuild   -- the mesh task's light rebuild (client/src/ChunkMeshTask.cpp:48-111): the two
//                            initialisation loops are restated here around the reference's own
//                            BlockLightAlgo<Config, 14> (client/include/client/BlockLightAlgo.hpp,
//                            ChunkMeshTask.inl:53-56), whose seed test and PropagateLight do the work;
//   hcref_relight_many    -- the whole mesh task (light rebuild, then Generate over the rebuilt light)
//                            over many jobs on a std::thread pool (CPU baseline of the relit batch);
//   hcref_export_light_levels -- Block::GetLightLevel for all 65,536 Block::m_data values;
//   hcref_export_registry -- writes the reference's 65,536-entry block property table
//                            (block/generator/Block.cpp.in:30-32) and texture flags
//                            (texture/internal/generated/block_texture_transparency.inl:1-3) in the
//                            flattened format read by orc_registry_load.
#include <array>
#include <atomic>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include <client/BlockLightAlgo.hpp>
#include <client/BlockMeshAlgo.hpp>

#include "hcm_oracle.h"

namespace {

using hc::block::Block;
using hc::block::Light;
using Config = hc::client::BlockAlgoConfig<int8_t, hc::client::BlockAlgoBound<int8_t>{0, 0, 0, 16, 16, 16},
                                           hc::client::kBlockAlgoSwizzleYZX>;
using Algo = hc::client::BlockMeshAlgo<Config>;

// Neighbour slot of a chunk-relative position, Position.hpp:138-145 restated for a 16^3 chunk.
inline uint32_t slot_of(int x, int y, int z) {
	constexpr uint32_t kL[3] = {1, 2, 0};
	auto c = [](int v) { return v < 0 ? 0 : (v >= 16 ? 2 : 1); };
	return kL[c(x)] * 9u + kL[c(y)] * 3u + kL[c(z)];
}
inline uint32_t local_of(int x, int y, int z) {
	return uint32_t((x + 16) % 16) + (uint32_t((y + 16) % 16) * 16u + uint32_t((z + 16) % 16)) * 16u;
}

int run_one(Algo &algo, const uint16_t *const blocks[27], const uint8_t *const light[27], orc_mesh_out *out) {
	auto meshes = algo.Generate(
	    [&](auto x, auto y, auto z) -> Block {
		    Block b;
		    b.SetData(blocks[slot_of(x, y, z)][local_of(x, y, z)]);
		    return b;
	    },
	    [&](auto x, auto y, auto z) -> Light {
		    Light l;
		    l.SetData(light[slot_of(x, y, z)][local_of(x, y, z)]);
		    return l;
	    });
	if (!out)
		return 0;
	if (meshes.size() > ORC_MAX_MESHES)
		return ORC_ERR_CAPACITY;
	uint32_t nv = 0, ni = 0;
	for (auto &m : meshes) {
		nv += (uint32_t)m.vertices.size();
		ni += (uint32_t)m.indices.size();
	}
	out->n_meshes = (uint32_t)meshes.size();
	if (nv > out->vert_cap || ni > out->idx_cap)
		return ORC_ERR_CAPACITY;
	uint32_t vb = 0, ib = 0;
	for (size_t i = 0; i < meshes.size(); ++i) {
		auto &m = meshes[i];
		out->vert_begin[i] = vb;
		out->n_verts[i] = (uint32_t)m.vertices.size();
		out->idx_begin[i] = ib;
		out->n_idx[i] = (uint32_t)m.indices.size();
		auto mn = m.aabb.GetMin(), mx = m.aabb.GetMax();
		out->aabb[i][0] = mn.x, out->aabb[i][1] = mn.y, out->aabb[i][2] = mn.z;
		out->aabb[i][3] = mx.x, out->aabb[i][4] = mx.y, out->aabb[i][5] = mx.z;
		out->transparent[i] = m.transparent ? 1u : 0u;
		for (auto &v : m.vertices) {
			out->verts[2 * vb] = v.x10_y10_z10_axis2;
			out->verts[2 * vb + 1] = v.tex10_trans3_face3_ao2_sl6_tl6;
			++vb;
		}
		for (auto ix : m.indices)
			out->idx[ib++] = ix;
	}
	return (int)nv / 4;
}

} // namespace

extern "C" int hcref_mesh_chunk(const uint16_t *const blocks[27], const uint8_t *const light[27],
                                orc_mesh_out *out) {
	thread_local Algo *algo = new Algo{};
	int r = run_one(*algo, blocks, light, out);
	return r < 0 ? r : 0;
}

extern "C" int hcref_mesh_many(const uint16_t *blocks, const uint8_t *light, const uint32_t *nbr,
                               uint32_t n_jobs, uint32_t n_threads, uint64_t *quads_out) {
	std::atomic<uint32_t> next{0};
	auto worker = [&]() {
		Algo *algo = new Algo{};
		for (;;) {
			uint32_t j = next.fetch_add(1);
			if (j >= n_jobs)
				break;
			const uint16_t *b[27];
			const uint8_t *l[27];
			for (int s = 0; s < 27; ++s) {
				b[s] = blocks + (size_t)nbr[27 * (size_t)j + s] * ORC_CHUNK_VOXELS;
				l[s] = light + (size_t)nbr[27 * (size_t)j + s] * ORC_CHUNK_VOXELS;
			}
			auto meshes = algo->Generate(
			    [&](auto x, auto y, auto z) -> Block {
				    Block bk;
				    bk.SetData(b[slot_of(x, y, z)][local_of(x, y, z)]);
				    return bk;
			    },
			    [&](auto x, auto y, auto z) -> Light {
				    Light lt;
				    lt.SetData(l[slot_of(x, y, z)][local_of(x, y, z)]);
				    return lt;
			    });
			uint64_t q = 0;
			for (auto &m : meshes)
				q += m.vertices.size() / 4;
			if (quads_out)
				quads_out[j] = q;
		}
		delete algo;
	};
	if (n_threads <= 1) {
		worker();
		return 0;
	}
	std::vector<std::thread> th;
	for (uint32_t t = 0; t < n_threads; ++t)
		th.emplace_back(worker);
	for (auto &t : th)
		t.join();
	return 0;
}

extern "C" int hcref_export_registry(const char *path) {
	using hc::texture::BlockTexture;
	struct RowKey {
		uint16_t tex[6];
		uint8_t transparency, collision;
		uint16_t mesh;
		bool operator<(const RowKey &r) const { return std::memcmp(this, &r, sizeof(RowKey)) < 0; }
	};
	static_assert(sizeof(RowKey) == 16);
	std::map<const hc::block::BlockMesh *, uint16_t> mesh_ids;
	std::vector<orc_mesh> meshes;
	std::map<RowKey, uint16_t> row_ids;
	std::vector<RowKey> rows;
	static uint16_t row_of[65536];

	for (uint32_t data = 0; data < 65536; ++data) {
		Block b;
		b.SetData((uint16_t)data);
		RowKey k;
		std::memset(&k, 0, sizeof(k));
		for (uint32_t f = 0; f < 6; ++f)
			k.tex[f] = b.GetTexture((hc::block::BlockFace)f).GetData();
		k.transparency = (uint8_t)b.GetTransparency();
		k.collision = (uint8_t)b.GetCollision();
		const hc::block::BlockMesh *cm = b.GetCustomMesh();
		k.mesh = ORC_NO_MESH;
		if (cm) {
			auto it = mesh_ids.find(cm);
			if (it == mesh_ids.end()) {
				orc_mesh m;
				std::memset(&m, 0, sizeof(m));
				m.face_count = cm->face_count;
				if (cm->p_dynamic_mesh_func) {
					// The reference has exactly one dynamic mesh family: the liquid flow mesh
					// (block/block/public/Liquid.hpp:284-302). Its template Variant is observable
					// through the top face height when all 8 horizontal neighbours are obstacles.
					static const int8_t kFetch[8][3] = {{-1, 0, -1}, {0, 0, -1}, {1, 0, -1}, {-1, 0, 0},
					                                    {1, 0, 0},   {-1, 0, 1}, {0, 0, 1},  {1, 0, 1}};
					if (cm->fetch_neighbour_count != 8)
						return ORC_ERR_FORMAT;
					for (int i = 0; i < 8; ++i)
						if (cm->fetch_neighbours[i].x != kFetch[i][0] || cm->fetch_neighbours[i].y != kFetch[i][1] ||
						    cm->fetch_neighbours[i].z != kFetch[i][2])
							return ORC_ERR_FORMAT;
					Block nb[hc::block::kBlockDynamicMeshMaxNeighbours];
					for (auto &x : nb)
						x = Block{hc::block::Blocks::kStone};
					hc::block::BlockMeshFace faces[hc::block::kBlockMeshMaxFaceCount];
					uint32_t fc = 0;
					BlockTexture tex6[6];
					bool dyn = false;
					cm->p_dynamic_mesh_func(nb, faces, &fc, tex6, &dyn);
					m.kind = ORC_MESH_LIQUID_FLOW;
					m.flow_height = fc ? faces[0].vertices[0].y : 16u;
				} else {
					m.kind = ORC_MESH_STATIC;
					for (uint32_t f = 0; f < cm->face_count && f < ORC_MAX_FACES; ++f) {
						const auto &src = cm->faces[f];
						auto &dst = m.faces[f];
						dst.axis = src.axis;
						dst.light_face = src.light_face;
						dst.render_face = src.render_face;
						dst.texture = src.texture.GetData();
						for (int v = 0; v < 4; ++v) {
							dst.v[v].x = src.vertices[v].x;
							dst.v[v].y = src.vertices[v].y;
							dst.v[v].z = src.vertices[v].z;
							dst.v[v].ao = src.vertices[v].ao;
						}
					}
				}
				uint16_t id = (uint16_t)meshes.size();
				meshes.push_back(m);
				it = mesh_ids.emplace(cm, id).first;
			}
			k.mesh = it->second;
		}
		auto rit = row_ids.find(k);
		if (rit == row_ids.end()) {
			rit = row_ids.emplace(k, (uint16_t)rows.size()).first;
			rows.push_back(k);
		}
		row_of[data] = rit->second;
	}

	// Texture flags for every id the texture enum defines (texture/register/block_textures:1-63).
	uint32_t n_tex = hc::texture::BlockTextures::kGlowstone + 1;
	std::vector<uint8_t> flags(n_tex);
	for (uint32_t t = 0; t < n_tex; ++t) {
		BlockTexture tx{(hc::texture::BlockTexID)t};
		flags[t] = (tx.IsTransparent() ? ORC_TEX_TRANSPARENT : 0u) | (tx.UseTransparentPass() ? ORC_TEX_TRANS_PASS : 0u) |
		           (tx.IsLiquid() ? ORC_TEX_LIQUID : 0u);
	}

	FILE *fp = std::fopen(path, "wb");
	if (!fp)
		return ORC_ERR_IO;
	char magic[8] = {'H', 'C', 'R', 'E', 'G', '0', '1', 0};
	uint32_t hdr[4] = {(uint32_t)rows.size(), (uint32_t)meshes.size(), n_tex, 0};
	std::fwrite(magic, 1, 8, fp);
	std::fwrite(hdr, 4, 4, fp);
	std::fwrite(row_of, 2, 65536, fp);
	std::fwrite(rows.data(), sizeof(RowKey), rows.size(), fp);
	std::fwrite(meshes.data(), sizeof(orc_mesh), meshes.size(), fp);
	std::fwrite(flags.data(), 1, flags.size(), fp);
	uint8_t liq[2] = {hc::block::Blocks::kWater, hc::block::Blocks::kFlowingWater};
	uint16_t liq_tex = hc::texture::BlockTextures::kWater;
	std::fwrite(liq, 1, 2, fp);
	std::fwrite(&liq_tex, 2, 1, fp);
	std::fclose(fp);
	return 0;
}

namespace {
using LightConfig = hc::client::BlockAlgoConfig<int8_t, hc::client::BlockAlgoBound<int8_t>{-1, -1, -1, 17, 17, 17}>;
using LightAlgo = hc::client::BlockLightAlgo<LightConfig, 14>;
inline uint32_t ext_index(int x, int y, int z) { return uint32_t(x + 15) + (uint32_t(z + 15) + uint32_t(y + 15) * 46u) * 46u; }
} // namespace

extern "C" int hcref_light_rebuild(const uint16_t *const blocks[27], const int8_t *const heights[27], uint8_t *win) {
	auto block_at = [&](auto x, auto y, auto z) -> Block {
		Block b;
		b.SetData(blocks[slot_of(x, y, z)][local_of(x, y, z)]);
		return b;
	};
	auto sun_at = [&](int x, int y, int z) -> bool { // Chunk::GetSunlightFromNeighbour (Chunk.hpp:73-79)
		return ((y + 16) % 16) >= heights[slot_of(x, y, z)][(x + 16) % 16 + ((z + 16) % 16) * 16];
	};
	std::vector<Light> buf(46 * 46 * 46);
	LightAlgo::Queue sun_q, torch_q;
	for (int8_t y = -15; y < 31; ++y)
		for (int8_t z = -15; z < 31; ++z)
			for (int8_t x = -15; x < 31; ++x) {
				Light light = {};
				light.SetTorchlight(block_at(x, y, z).GetLightLevel());
				light.SetSunlight(sun_at(x, y, z) ? 15 : 0);
				buf[ext_index(x, y, z)] = light;
				if (LightAlgo::IsBorderLightInterfere(x, y, z, light.GetTorchlight()))
					torch_q.push({{x, y, z}, light.GetTorchlight()});
			}
	auto sun = [&](int x, int y, int z) { return buf[ext_index(x, y, z)].GetSunlight(); };
	for (int8_t y = -15; y < 31; ++y)
		for (int8_t z = -15; z < 31; ++z)
			for (int8_t x = -15; x < 31; ++x) {
				auto s = sun(x, y, z);
				if (LightAlgo::IsBorderLightInterfere(x, y, z, s)) {
					if ((x == -15 || sun(x - 1, y, z) == 15) && (x == 30 || sun(x + 1, y, z) == 15) &&
					    (z == -15 || sun(x, y, z - 1) == 15) && (z == 30 || sun(x, y, z + 1) == 15) &&
					    (y == -15 || sun(x, y - 1, z) == 15) && (y == 30 || sun(x, y + 1, z) == 15))
						continue;
					sun_q.push({{x, y, z}, s});
				}
			}
	LightAlgo algo{};
	algo.PropagateLight(
	    &sun_q, block_at, [&](auto x, auto y, auto z) { return buf[ext_index(x, y, z)].GetSunlight(); },
	    [&](auto x, auto y, auto z, hc::block::LightLvl l) { buf[ext_index(x, y, z)].SetSunlight(l); });
	algo.PropagateLight(
	    &torch_q, block_at, [&](auto x, auto y, auto z) { return buf[ext_index(x, y, z)].GetTorchlight(); },
	    [&](auto x, auto y, auto z, hc::block::LightLvl l) { buf[ext_index(x, y, z)].SetTorchlight(l); });
	for (size_t i = 0; i < buf.size(); ++i)
		win[i] = buf[i].GetData();
	return 0;
}

extern "C" int hcref_export_light_levels(const char *path) {
	static uint8_t lv[65536];
	for (uint32_t data = 0; data < 65536; ++data) {
		Block b;
		b.SetData((uint16_t)data);
		lv[data] = (uint8_t)b.GetLightLevel();
	}
	FILE *fp = std::fopen(path, "wb");
	if (!fp)
		return ORC_ERR_IO;
	std::fwrite(lv, 1, sizeof(lv), fp);
	std::fclose(fp);
	return 0;
}

// The whole mesh task per job (ChunkMeshTask.cpp:41-127): light rebuild, then Generate reading light
// from the rebuilt buffer as the task's lambda does (:120-121). CPU baseline of the relit batch.
extern "C" int hcref_relight_many(const uint16_t *blocks, const int8_t *heights, const uint32_t *nbr,
                                  uint32_t n_jobs, uint32_t n_threads, uint64_t *quads_out) {
	std::atomic<uint32_t> next{0};
	auto worker = [&]() {
		Algo *algo = new Algo{};
		std::vector<uint8_t> win(46 * 46 * 46);
		for (;;) {
			uint32_t j = next.fetch_add(1);
			if (j >= n_jobs)
				break;
			const uint16_t *b[27];
			const int8_t *h[27];
			for (int s = 0; s < 27; ++s) {
				b[s] = blocks + (size_t)nbr[27 * (size_t)j + s] * ORC_CHUNK_VOXELS;
				h[s] = heights + (size_t)nbr[27 * (size_t)j + s] * 256;
			}
			hcref_light_rebuild(b, h, win.data());
			auto meshes = algo->Generate(
			    [&](auto x, auto y, auto z) -> Block {
				    Block bk;
				    bk.SetData(b[slot_of(x, y, z)][local_of(x, y, z)]);
				    return bk;
			    },
			    [&](auto x, auto y, auto z) -> Light {
				    Light lt;
				    lt.SetData(win[ext_index(x, y, z)]);
				    return lt;
			    });
			uint64_t q = 0;
			for (auto &m : meshes)
				q += m.vertices.size() / 4;
			if (quads_out)
				quads_out[j] = q;
		}
		delete algo;
	};
	if (n_threads <= 1) {
		worker();
		return 0;
	}
	std::vector<std::thread> th;
	for (uint32_t t = 0; t < n_threads; ++t)
		th.emplace_back(worker);
	for (auto &t : th)
		t.join();
	return 0;
}
