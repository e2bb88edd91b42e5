// Host-code checks built with g++ under sanitizers (no GPU): -fsanitize=address,undefined and, in a
// second build, -fsanitize=thread. Runs
//   1. the vertex-pool range allocator against a bitmap model under random alloc/release;
//   2. the serving loop of hcm_serve.cpp (scheduler thread, request queues, world model, stale and
//      priority queues, epochs, completions) over the CPU device of serve_stub_device.cpp, with four
//      threads submitting loads and edits, one polling completions and one reading stats and history,
//      in both light modes, then checks that every request completed exactly once and that every
//      meshable chunk's final mesh is the stand-in mesher's output for its final blocks;
//   3. the registry table and the workload generators (hcm_registry.cpp, hcm_workload.cpp).
// Exit status 0 when everything holds; the sanitizers abort on the first error they find.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <random>
#include <set>
#include <thread>
#include <tuple>
#include <vector>

#include "hcm.h"
#include "hcm_range_alloc.h"
#include "hcm_workload.h"

namespace hcm {
uint32_t toy_quads(const uint16_t *center);
}

static int g_fail = 0;
#define CHECK(c)                                                                                                   \
	do {                                                                                                           \
		if (!(c)) {                                                                                                \
			std::fprintf(stderr, "check failed at %s:%d: %s\n", __FILE__, __LINE__, #c);                          \
			++g_fail;                                                                                              \
		}                                                                                                          \
	} while (0)

static void check_range_allocator() {
	std::mt19937_64 rng(7);
	const uint64_t cap = 4096;
	hcm::RangeAllocator a(cap);
	std::vector<char> used(cap, 0);
	std::vector<std::pair<uint64_t, uint64_t>> live;
	for (int it = 0; it < 20000; ++it) {
		if (live.empty() || rng() % 3 != 0) {
			const uint64_t n = 1 + rng() % 200;
			uint64_t first = 0;
			// the model's first fit: lowest start of a free run of length >= n
			uint64_t want = cap, run = 0;
			for (uint64_t i = 0; i < cap; ++i) {
				run = used[i] ? 0 : run + 1;
				if (run == n) {
					want = i + 1 - n;
					break;
				}
			}
			// first fit over the allocator's free ranges equals the model's only when no free range is
			// split across a later merge; check the guarantee that matters: disjoint, in bounds, and
			// success whenever some free range is long enough
			uint64_t largest = 0;
			run = 0;
			for (uint64_t i = 0; i < cap; ++i) {
				run = used[i] ? 0 : run + 1;
				largest = std::max(largest, run);
			}
			CHECK(a.largest() == largest);
			const bool ok = a.alloc(n, &first);
			CHECK(ok == (want != cap));
			if (ok) {
				CHECK(first + n <= cap);
				for (uint64_t i = first; i < first + n; ++i) {
					CHECK(!used[i]);
					used[i] = 1;
				}
				live.emplace_back(first, n);
			}
		} else {
			const size_t k = rng() % live.size();
			a.release(live[k].first, live[k].second);
			for (uint64_t i = live[k].first; i < live[k].first + live[k].second; ++i)
				used[i] = 0;
			live[k] = live.back();
			live.pop_back();
		}
		uint64_t free_cells = 0;
		for (char u : used)
			free_cells += !u;
		CHECK(a.free_total() == free_cells);
	}
}

struct World {
	std::map<std::tuple<int, int, int>, std::vector<uint16_t>> blocks;
};

static std::vector<uint16_t> chunk_blocks(uint64_t seed, int x, int y, int z) {
	std::vector<uint16_t> b(HCM_CHUNK_VOXELS);
	std::vector<uint8_t> l(HCM_CHUNK_VOXELS);
	hcm_gen_cave_chunk(seed, x, y, z, 0, b.data(), l.data());
	return b;
}

static void check_server(uint32_t light_mode, uint32_t max_batch, uint32_t max_inflight) {
	hcm_server_config cfg{};
	cfg.max_chunks = 128;
	cfg.max_batch = max_batch;
	cfg.max_inflight = max_inflight;
	cfg.pool_quads = 1u << 12;
	cfg.keep_history = 1;
	cfg.light_mode = light_mode;
	int dummy_ctx = 0;
	hcm_server *srv = nullptr;
	CHECK(hcm_server_create(reinterpret_cast<hcm_ctx *>(&dummy_ctx), &cfg, &srv) == HCM_OK);
	if (!srv)
		return;
	const int dims[3] = {6, 3, 6};
	std::mutex wmu;
	World world;
	std::atomic<uint64_t> next_tag{0};
	std::atomic<bool> done_submitting{false};
	std::vector<hcm_completion> completions;
	std::mutex cmu;
	// a poller and a stats reader run while requests arrive
	std::thread poller([&] {
		std::vector<hcm_completion> buf(256);
		while (!done_submitting.load()) {
			uint32_t n = 0;
			CHECK(hcm_server_poll(srv, buf.data(), (uint32_t)buf.size(), &n) == HCM_OK);
			std::lock_guard<std::mutex> lk(cmu);
			completions.insert(completions.end(), buf.begin(), buf.begin() + n);
		}
	});
	std::thread stats([&] {
		while (!done_submitting.load()) {
			hcm_server_stats st;
			uint64_t h = 0;
			CHECK(hcm_server_get_stats(srv, &st) == HCM_OK);
			CHECK(hcm_server_history_size(srv, &h) == HCM_OK);
			CHECK(st.pool_used_quads + st.pool_free_quads == cfg.pool_quads);
		}
	});
	std::vector<std::thread> submitters;
	for (int t = 0; t < 4; ++t)
		submitters.emplace_back([&, t] {
			std::mt19937 rng(100 + t);
			std::vector<int8_t> heights(256);
			// thread t loads the columns x = t, t + 4, ... (each chunk by exactly one thread), reloads some
			for (int pass = 0; pass < 2; ++pass)
				for (int x = t; x < dims[0]; x += 4)
					for (int y = 0; y < dims[1]; ++y)
						for (int z = 0; z < dims[2]; ++z) {
							if (pass == 1 && rng() % 3)
								continue;
							std::vector<uint16_t> b = chunk_blocks(pass * 1000 + 5, x, y, z);
							std::vector<uint8_t> l(HCM_CHUNK_VOXELS, 0x2F);
							for (auto &h : heights)
								h = int8_t(rng() % 17);
							std::lock_guard<std::mutex> lk(wmu); // request order = world order
							const uint64_t tag = next_tag++;
							const int r = light_mode == HCM_LIGHT_REBUILD
							                  ? hcm_server_load_sunlit(srv, x, y, z, b.data(), heights.data(), tag)
							                  : hcm_server_load(srv, x, y, z, b.data(), l.data(), tag);
							CHECK(r == HCM_OK);
							world.blocks[{x, y, z}] = b;
						}
			for (int k = 0; k < 150; ++k) {
				const int x = int(rng() % (16 * dims[0])), y = int(rng() % (16 * dims[1])), z = int(rng() % (16 * dims[2]));
				const uint16_t blk = uint16_t(rng() % 4 == 0 ? 9 : rng() % 3);
				std::lock_guard<std::mutex> lk(wmu);
				auto it = world.blocks.find({x / 16, y / 16, z / 16});
				if (it == world.blocks.end())
					continue;
				const uint64_t tag = next_tag++;
				CHECK(hcm_server_edit(srv, x, y, z, blk, uint8_t(rng()), tag) == HCM_OK);
				it->second[(x % 16) + 16 * ((z % 16) + 16 * (y % 16))] = blk;
			}
		});
	for (auto &th : submitters)
		th.join();
	CHECK(hcm_server_drain(srv) == HCM_OK);
	done_submitting = true;
	poller.join();
	stats.join();
	{
		std::vector<hcm_completion> buf(4096);
		uint32_t n = 0;
		do {
			CHECK(hcm_server_poll(srv, buf.data(), (uint32_t)buf.size(), &n) == HCM_OK);
			completions.insert(completions.end(), buf.begin(), buf.begin() + n);
		} while (n);
	}
	std::set<uint64_t> tags;
	for (auto &c : completions)
		CHECK(tags.insert(c.tag).second);
	CHECK(tags.size() == next_tag.load());
	// final meshes: every chunk with all 26 neighbours holds the stand-in mesh of its final blocks
	std::vector<uint32_t> verts(2 << 16);
	std::vector<uint16_t> idx(1 << 16);
	std::vector<hcm_mesh_desc> meshes(8);
	int meshable = 0;
	for (auto &kv : world.blocks) {
		const int x = std::get<0>(kv.first), y = std::get<1>(kv.first), z = std::get<2>(kv.first);
		if (x < 1 || y < 1 || z < 1 || x >= dims[0] - 1 || y >= dims[1] - 1 || z >= dims[2] - 1)
			continue;
		++meshable;
		hcm_host_result out{};
		out.vertices = reinterpret_cast<uint32_t(*)[2]>(verts.data()), out.vertex_capacity = (uint32_t)verts.size() / 2;
		out.indices = idx.data(), out.index_capacity = (uint32_t)idx.size();
		out.meshes = meshes.data(), out.mesh_capacity = (uint32_t)meshes.size();
		CHECK(hcm_server_chunk_mesh(srv, x, y, z, &out) == HCM_OK);
		CHECK(out.mesh_count == 1 && out.vertex_count == 4 * hcm::toy_quads(kv.second.data()));
		for (uint32_t k = 0; k < out.vertex_count; ++k)
			CHECK((verts[2 * k] & 0xffffu) == kv.second[k % HCM_CHUNK_VOXELS]);
	}
	CHECK(meshable == 16);
	hcm_server_destroy(srv);
}

static void check_registry_and_generators() {
	const hcm_registry *r = hcm_registry_default();
	CHECK(r && r->n_rows > 0 && r->row_of && r->rows && r->light_level);
	for (uint32_t d = 0; d < 65536; ++d)
		CHECK(r->row_of[d] < r->n_rows);
	std::vector<uint16_t> b(HCM_CHUNK_VOXELS);
	std::vector<uint8_t> l(HCM_CHUNK_VOXELS);
	uint64_t sum = 0;
	for (int i = 0; i < 8; ++i) {
		hcm_gen_cave_chunk(3, i, i - 2, -i, 0, b.data(), l.data());
		for (int k = 0; k < HCM_CHUNK_VOXELS; ++k)
			sum += b[k] + l[k];
		hcm_gen_terrain_chunk(3, i, 0, -i, HCM_GEN_PLANTS, b.data(), l.data());
		hcm_gen_random_chunk(3, (uint32_t)i, b.data(), l.data());
	}
	CHECK(sum > 0);
}

int main(int argc, char **argv) {
	// argument: how many of the six server configurations to run (the thread-sanitizer build runs fewer)
	const int n_cfg = argc > 1 ? std::atoi(argv[1]) : 6;
	check_range_allocator();
	check_registry_and_generators();
	int k = 0;
	for (uint32_t mode : {HCM_LIGHT_STORE, HCM_LIGHT_REBUILD})
		for (auto cfg : {std::pair<uint32_t, uint32_t>{1, 1}, {4, 3}, {64, 2}})
			if (k++ < n_cfg) {
				std::fprintf(stderr, "server: light mode %u, max_batch %u, max_inflight %u\n", mode, cfg.first, cfg.second);
				check_server(mode, cfg.first, cfg.second);
			}
	if (g_fail) {
		std::fprintf(stderr, "%d checks failed\n", g_fail);
		return 1;
	}
	std::printf("serve_check: all checks passed\n");
	return 0;
}
