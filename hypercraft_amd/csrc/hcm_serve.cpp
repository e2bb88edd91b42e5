// hcm_serve.cpp -- continuous batching of chunk-load and block-edit requests (include/hcm.h, "Serving").
//
// The world's chunks live in device slots. Callers submit requests from any thread; one scheduler thread
// applies them to the world model in arrival order (request k gets epoch k), keeps the set of chunks whose
// mesh is stale, and, whenever fewer than max_inflight batches are on the device, forms a batch from what
// is pending:
//   copy stream:    first-time chunk loads -> their fresh slots (no running batch reads a fresh slot);
//   compute stream: wait for those copies; reloads of resident chunks (in stream order, after every
//                   earlier batch that may read them); the batch's voxel edits; then up to max_batch stale
//                   chunks meshed by the persistent kernel into an arena taken from the vertex pool, and
//                   the per-chunk results copied back.
// Uploads of batch i+1 overlap the meshing of batch i. A batch meshes chunks only when every pending data
// change has been flushed into it, so its meshes reflect all requests up to the epoch at its launch; a
// batch that cannot flush everything (upload or edit limits) carries data only, and its "applied epoch"
// is the last epoch before the first change it left pending. A request completes when a finished batch
// has applied it and every chunk it made stale holds a mesh from a batch at or after its epoch.
// Remesh rules follow the reference: a chunk is meshed only when all 26 neighbours are present
// (client/src/ChunkMeshTask.cpp:14-39), and an edit invalidates every chunk whose 18^3 halo holds the
// voxel -- the edited chunk plus the face/edge/corner neighbours the voxel touches. Remeshes an edit
// causes go ahead of those of loads when batches are formed, as the reference pushes them at high
// priority (client/src/ChunkSetBlockTask.cpp:119, served first by client/src/ChunkTaskPool.cpp:97-105).
#include <hip/hip_runtime.h>
#include <sys/prctl.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hcm_internal.h"
#include "hcm_range_alloc.h"

using namespace hcm;

#define HCM_CHECK(call)                                                                                            \
	do {                                                                                                           \
		if ((call) != hipSuccess)                                                                                  \
			return HCM_ERR_HIP;                                                                                    \
	} while (0)

namespace {

using Clock = std::chrono::steady_clock;

inline uint64_t chunk_key(int32_t cx, int32_t cy, int32_t cz) {
	return (uint64_t(uint32_t(cx) & 0x1fffffu) << 42) | (uint64_t(uint32_t(cy) & 0x1fffffu) << 21) |
	       uint64_t(uint32_t(cz) & 0x1fffffu);
}
inline int32_t sext21(uint64_t v) { return int32_t(uint32_t(v << 11)) >> 11; }
inline void key_pos(uint64_t k, int32_t p[3]) {
	p[0] = sext21(k >> 42 & 0x1fffffu), p[1] = sext21(k >> 21 & 0x1fffffu), p[2] = sext21(k & 0x1fffffu);
}
inline int32_t floor_div16(int32_t v) { return v >= 0 ? v / 16 : -((-v + 15) / 16); }

// neighbour slot -> chunk offset (common/include/common/Position.hpp:146-152)
constexpr int kRev[3] = {1, -1, 0};
inline void slot_offset(int s, int d[3]) { d[0] = kRev[s / 9], d[1] = kRev[(s / 3) % 3], d[2] = kRev[s % 3]; }

constexpr uint32_t kMaxUploads = 256;   // chunk uploads staged per batch (3 MB: a burst of loads goes up in
                                        // several batches rather than one long copy ahead of every edit)
constexpr uint32_t kMaxEdits = 65536;   // voxel edits applied per batch
constexpr uint64_t kMinArena = 1 << 16; // quads; more than any one chunk's meshes

struct Chunk {
	uint32_t slot = 0;
	uint32_t present = 0;           // chunks of its 3x3x3 neighbourhood in the world, itself included
	uint32_t nbr[27] = {};          // their slots in neighbour-slot order, once all 27 are present
	bool stale = false;             // mesh out of date and not in a launched batch
	bool urgent = false;            // stale because of an edit: queued with high priority
	bool has_mesh = false;
	uint64_t meshed_epoch = 0;      // epoch of the batch that produced the current meshes
	Range mesh;                     // their pool range
	std::vector<hcm_mesh_rec> recs; // their records (absolute first_quad)
	// HCM_LIGHT_REBUILD: host copies of the blocks and sunlight heights, for the sunlight flood of edits
	std::vector<uint16_t> blocks;
	std::vector<int8_t> heights;
};

struct Request {
	uint64_t tag = 0;
	uint32_t kind = 0;
	Clock::time_point t_submit;
	int32_t pos[3] = {0, 0, 0}; // chunk (load) or world voxel (edit)
	std::vector<uint16_t> blocks;
	std::vector<uint8_t> light;    // HCM_LIGHT_STORE loads
	std::vector<int8_t> heights;   // HCM_LIGHT_REBUILD loads: sunlight heights
	uint16_t block = 0;
	uint8_t light1 = 0;
	uint32_t slot = 0;             // load: the chunk's device slot
	uint64_t epoch = 0;            // set when the scheduler applies it
	std::vector<uint64_t> targets; // chunks it made stale
	size_t targets_done = 0;       // leading targets already meshed at or after its epoch
};

struct Upload {
	uint32_t slot;
	bool reload; // the slot already holds the chunk: copy in compute-stream order
	Request *req;
};

struct PendingEdit {
	VoxelEdit e;
	uint64_t first_epoch; // earliest request still pending for this voxel
};

struct BatchSlot { // device and pinned buffers of one in-flight batch
	uint32_t *d_nbr = nullptr;
	VoxelEdit *d_edits = nullptr;
	Counters *d_ctr = nullptr;
	hcm_job_result *d_res = nullptr;
	hcm_mesh_rec *d_recs = nullptr;
	uint32_t *h_nbr = nullptr;
	VoxelEdit *h_edits = nullptr;
	uint16_t *h_blocks = nullptr;
	uint8_t *h_light = nullptr;
	int8_t *h_heights = nullptr;
	hcm_job_result *h_res = nullptr;
	// the batch's chunk uploads: staged on the host, copied once, then scattered to their slots on the device
	uint16_t *d_up_blocks = nullptr;
	uint8_t *d_up_light = nullptr;
	int8_t *d_up_heights = nullptr;
	uint2 *d_moves = nullptr, *h_moves = nullptr;
	// light rebuild (HCM_LIGHT_REBUILD): the slots the jobs read, queues, rebuilt halo light
	uint32_t *d_rslots = nullptr, *h_rslots = nullptr;
	Counters *d_lctr = nullptr;
	uint8_t *d_halo = nullptr;
	hcm_mesh_rec *h_recs = nullptr;
	Counters *h_ctr = nullptr;
	hipEvent_t ev_up = nullptr, ev_k0 = nullptr, ev_k1 = nullptr, ev_done = nullptr;
	bool busy = false;
	bool arena_was_largest = false; // the arena took the largest free range there was
	uint64_t id = 0, epoch = 0;     // epoch: every request up to it is applied on the device
	uint32_t recs_copied = 0;       // mesh records copied back with the results (the usual count)
	std::vector<uint64_t> jobs;     // the first n_urgent were taken from the high-priority queue
	uint32_t n_urgent = 0;
	Range arena;
};

struct HistoryEntry {
	int32_t pos[3];
	uint64_t epoch;
	std::vector<hcm_mesh_rec> recs; // first_quad relative to verts/idx below
	std::vector<uint2> verts;
	std::vector<uint16_t> idx;
};

} // namespace

struct hcm_server {
	hcm_ctx *ctx = nullptr;
	hcm_server_config cfg{};
	CtxInfo info{};
	hipStream_t s_copy = nullptr, s_comp = nullptr;
	uint16_t *d_blocks = nullptr;
	uint8_t *d_light = nullptr;
	int8_t *d_heights = nullptr;  // HCM_LIGHT_REBUILD: [slot][256] sunlight heights
	uint16_t *d_packed = nullptr; // HCM_LIGHT_REBUILD: [slot][kLightPlanes][256]
	uint32_t *d_levels = nullptr;
	uint2 *d_verts = nullptr;
	uint16_t *d_idx = nullptr;
	std::vector<BatchSlot> slots;
	std::atomic<int> error{HCM_OK};

	// submission side (in_mu)
	std::mutex in_mu;
	std::condition_variable in_cv;
	std::deque<std::unique_ptr<Request>> incoming;
	std::unordered_map<uint64_t, uint32_t> accepted; // chunk -> slot, fixed at the first accepted load
	uint64_t submitted = 0;
	bool stop = false;

	// scheduler state: changed by the scheduler thread under world_mu
	std::mutex world_mu;
	std::unordered_map<uint64_t, Chunk> world;
	std::deque<uint64_t> stale_fifo; // stale chunks by time (loads, sunlight changes)
	std::deque<uint64_t> stale_hi;   // stale chunks an edit remeshes: taken first
	bool pool_blocked = false;       // an arena did not fit while batches ran: wait for one to finish
	std::deque<std::unique_ptr<Request>> waiting; // applied requests not yet complete, in epoch order
	bool completions_due = false;                 // something happened that may complete a waiting request
	std::deque<Upload> uploads;
	std::map<uint64_t, PendingEdit> edits; // (slot << 13 | local) -> last edit of that voxel or height
	uint64_t epoch = 0, batches = 0, applied_done = 0;
	RangeAllocator *pool = nullptr;
	double quads_per_chunk = 512.0; // running estimate that sizes arenas

	// completions (out_mu)
	std::mutex out_mu;
	std::condition_variable out_cv;
	std::deque<hcm_completion> done;
	uint64_t completed = 0;

	// stats and history (st_mu)
	std::mutex st_mu;
	hcm_server_stats stats{};
	std::vector<HistoryEntry> history;
	std::thread thread;

	bool relight() const { return cfg.light_mode == HCM_LIGHT_REBUILD; }

	void fail(int code) {
		error = code;
		std::lock_guard<std::mutex> lk(out_mu);
		out_cv.notify_all();
	}

	// A new chunk counts its present neighbours and is counted by them; a chunk whose count reaches 27
	// takes its neighbours' slots (fixed from their first load) once, for every batch that meshes it.
	void add_chunk(uint64_t key, Chunk &c) {
		int32_t p[3];
		key_pos(key, p);
		auto complete = [&](uint64_t k, Chunk &n) {
			if (++n.present < 27)
				return;
			int32_t q[3];
			key_pos(k, q);
			for (int s = 0; s < 27; ++s) {
				int d[3];
				slot_offset(s, d);
				n.nbr[s] = world.at(chunk_key(q[0] + d[0], q[1] + d[1], q[2] + d[2])).slot;
			}
		};
		for (int s = 0; s < 27; ++s) {
			int d[3];
			slot_offset(s, d);
			if (d[0] == 0 && d[1] == 0 && d[2] == 0)
				continue;
			const uint64_t nk = chunk_key(p[0] + d[0], p[1] + d[1], p[2] + d[2]);
			auto it = world.find(nk);
			if (it != world.end()) {
				complete(nk, it->second);
				++c.present;
			}
		}
		complete(key, c);
	}
	// A chunk made stale is queued once per priority: a chunk already queued by time moves ahead when an
	// edit marks it (ChunkMeshTask.inl:40 keeps the high priority flag until the task runs); the `stale`
	// flag makes the later of two queue entries a no-op.
	void queue_stale(uint64_t key, Chunk &c, bool urgent) {
		if (!c.stale) {
			c.stale = true;
			c.urgent = urgent;
			(urgent ? stale_hi : stale_fifo).push_back(key);
		} else if (urgent && !c.urgent) {
			c.urgent = true;
			stale_hi.push_back(key);
		}
	}
	void mark_stale(int32_t cx, int32_t cy, int32_t cz, std::vector<uint64_t> &targets, bool urgent) {
		const uint64_t key = chunk_key(cx, cy, cz);
		auto it = world.find(key);
		if (it == world.end() || it->second.present < 27)
			return;
		queue_stale(key, it->second, urgent);
		targets.push_back(key);
	}

	void apply(std::unique_ptr<Request> r) {
		r->epoch = ++epoch;
		if (r->kind == HCM_REQ_LOAD) {
			const uint64_t key = chunk_key(r->pos[0], r->pos[1], r->pos[2]);
			auto it = world.find(key);
			const bool reload = it != world.end();
			const uint32_t slot = r->slot;
			if (!reload) {
				Chunk c;
				c.slot = slot;
				it = world.emplace(key, c).first;
				add_chunk(key, it->second);
			}
			if (relight()) {
				it->second.blocks = r->blocks;
				it->second.heights = r->heights;
			}
			// pending edits of the chunk are superseded by the new data
			auto e = edits.lower_bound(uint64_t(slot) << 13);
			while (e != edits.end() && (e->first >> 13) == slot)
				e = edits.erase(e);
			uploads.push_back(Upload{slot, reload, r.get()});
			for (int s = 0; s < 27; ++s) {
				int d[3];
				slot_offset(s, d);
				mark_stale(r->pos[0] + d[0], r->pos[1] + d[1], r->pos[2] + d[2], r->targets, false);
			}
		} else {
			const int32_t cx = floor_div16(r->pos[0]), cy = floor_div16(r->pos[1]), cz = floor_div16(r->pos[2]);
			const int32_t lx = r->pos[0] - 16 * cx, ly = r->pos[1] - 16 * cy, lz = r->pos[2] - 16 * cz;
			Chunk &c = world.at(chunk_key(cx, cy, cz));
			const uint32_t local = uint32_t(lx) + 16u * (uint32_t(lz) + 16u * uint32_t(ly));
			if (relight() && c.blocks[local] == r->block) {
				// the block does not change: nothing to apply or remesh (ChunkSetBlockTask.cpp:89); the light
				// byte of an edit is not used when light is rebuilt. (A light-store server keeps no host copy
				// of the voxels and applies every edit.)
				waiting.push_back(std::move(r));
				completions_due = true;
				return;
			}
			auto pend = [&](const VoxelEdit &ve) {
				const uint64_t vkey = (uint64_t(ve.slot) << 13) | ve.local;
				auto e = edits.find(vkey);
				if (e == edits.end())
					edits.emplace(vkey, PendingEdit{ve, r->epoch});
				else
					e->second.e = ve;
			};
			pend(VoxelEdit{c.slot, local, r->block, r->light1});
			if (relight()) {
				// the sunlight flood of the edited column (ChunkSetBlockTask.cpp:94, ChunkFloodSunlightTask.cpp:38-54,
				// once the chunk above is loaded): below a chunk column that is not sunlit to its bottom nothing is
				// sunlit; else sunlight reaches down to the lowest run of blocks that pass it vertically
				c.blocks[local] = r->block;
				auto up = world.find(chunk_key(cx, cy + 1, cz));
				if (up != world.end()) {
					const uint32_t xz = uint32_t(lx) + 16u * uint32_t(lz);
					int h = 16;
					if (up->second.heights[xz] <= 0) {
						h = 15;
						while (h >= 0 && (info.light_info_host[c.blocks[xz + 256u * uint32_t(h)]] & kLightInfoVertical))
							--h;
						++h;
					}
					if (h != c.heights[xz]) { // ChunkSetSunlightTask.cpp:46-72: the chunk and all 26 neighbours
						c.heights[xz] = int8_t(h);
						pend(VoxelEdit{c.slot, HCM_CHUNK_VOXELS + xz, uint32_t(uint8_t(int8_t(h))), 0});
						for (int s = 0; s < 27; ++s) {
							int d[3];
							slot_offset(s, d);
							mark_stale(cx + d[0], cy + d[1], cz + d[2], r->targets, false); // ChunkSetSunlightTask.cpp:84
						}
					}
				}
			}
			if (relight()) {
				// the neighbours within light distance 14: ChunkShouldRemesh (ChunkMeshTask.inl:9-15), the
				// L1 distance of the voxel to the neighbour's bound [-1, 17)^3 below 15
				auto dist = [](int v) { return v < -1 ? -1 - v : (v >= 17 ? v - 17 : 0); };
				for (int s = 0; s < 27; ++s) {
					int d[3];
					slot_offset(s, d);
					if (dist(lx - 16 * d[0]) + dist(ly - 16 * d[1]) + dist(lz - 16 * d[2]) < 15)
						mark_stale(cx + d[0], cy + d[1], cz + d[2], r->targets, true);
				}
			} else {
				// a voxel on a chunk border lies in the halo of the neighbours across that border
				const int ox[2] = {0, lx == 0 ? -1 : (lx == 15 ? 1 : 0)};
				const int oy[2] = {0, ly == 0 ? -1 : (ly == 15 ? 1 : 0)};
				const int oz[2] = {0, lz == 0 ? -1 : (lz == 15 ? 1 : 0)};
				for (int a = 0; a < (ox[1] ? 2 : 1); ++a)
					for (int b = 0; b < (oy[1] ? 2 : 1); ++b)
						for (int d = 0; d < (oz[1] ? 2 : 1); ++d)
							mark_stale(cx + ox[a], cy + oy[b], cz + oz[d], r->targets, true);
			}
			std::sort(r->targets.begin(), r->targets.end()); // a chunk marked by two rules counts once
			r->targets.erase(std::unique(r->targets.begin(), r->targets.end()), r->targets.end());
		}
		waiting.push_back(std::move(r));
		completions_due = true;
	}

	bool any_busy() const {
		for (auto &b : slots)
			if (b.busy)
				return true;
		return false;
	}
	bool work_pending() const {
		return !stale_hi.empty() || !stale_fifo.empty() || !uploads.empty() || !edits.empty() || any_busy();
	}

	// Returns 0 when launched, 1 when there is nothing to launch now, or an error.
	int launch_batch(BatchSlot &b) {
		const uint32_t n_up = (uint32_t)std::min<size_t>(uploads.size(), kMaxUploads);
		const bool up_rest = uploads.size() > n_up;
		const uint32_t n_ed = up_rest ? 0u : (uint32_t)std::min<size_t>(edits.size(), kMaxEdits);
		const bool ed_rest = edits.size() > n_ed;
		// meshing only once every pending change is in this batch, and not while an arena that did not
		// fit waits for a running batch to return memory; edit remeshes first
		b.jobs.clear();
		b.n_urgent = 0;
		if (!up_rest && !ed_rest && !pool_blocked)
			for (std::deque<uint64_t> *q : {&stale_hi, &stale_fifo}) {
				while (!q->empty() && b.jobs.size() < cfg.max_batch) {
					const uint64_t key = q->front();
					q->pop_front();
					auto it = world.find(key);
					if (it != world.end() && it->second.stale) { // a key queued twice is taken once
						it->second.stale = false;
						it->second.urgent = false;
						b.jobs.push_back(key);
					}
				}
				if (q == &stale_hi)
					b.n_urgent = (uint32_t)b.jobs.size();
			}
		// arena admission by vertex-pool memory
		b.arena = Range{};
		b.arena_was_largest = false;
		if (!b.jobs.empty()) {
			const uint64_t want = std::min<uint64_t>(
			    pool->capacity(), std::max<uint64_t>(kMinArena, uint64_t(quads_per_chunk * 1.25 * double(b.jobs.size()))));
			const uint64_t largest = pool->largest();
			uint64_t first = 0, got = want;
			bool ok = pool->alloc(want, &first);
			if (!ok && !any_busy()) { // nothing will free memory soon: take the largest free range
				got = largest;
				if (got == 0)
					return HCM_ERR_POOL_FULL;
				ok = pool->alloc(got, &first);
			}
			b.arena_was_largest = ok && got >= largest;
			if (!ok) { // wait for running batches to return memory, the jobs back at the front of their queues
				for (size_t j = b.jobs.size(); j-- > 0;) {
					Chunk &c = world.at(b.jobs[j]);
					c.stale = true;
					c.urgent = j < b.n_urgent;
					(c.urgent ? stale_hi : stale_fifo).push_front(b.jobs[j]);
				}
				b.jobs.clear();
				b.n_urgent = 0;
				pool_blocked = true;
			} else
				b.arena = Range{first, got};
		}
		if (b.jobs.empty() && n_up == 0 && n_ed == 0)
			return 1;
		uint64_t applied = epoch;
		if (up_rest)
			applied = std::min(applied, uploads[n_up].req->epoch - 1);
		b.id = ++batches;
		b.busy = true;

		const size_t V = HCM_CHUNK_VOXELS;
		for (uint32_t i = 0; i < n_up; ++i) {
			Request *r = uploads[i].req;
			std::memcpy(b.h_blocks + i * V, r->blocks.data(), V * 2);
			if (relight())
				std::memcpy(b.h_heights + i * 256, r->heights.data(), 256);
			else
				std::memcpy(b.h_light + i * V, r->light.data(), V);
			std::vector<uint16_t>().swap(r->blocks);
			std::vector<uint8_t>().swap(r->light);
			std::vector<int8_t>().swap(r->heights);
		}
		// a slot's data, blocks and light bytes or sunlight heights, goes up with the batch's others in one
		// copy per array; first-time loads are scattered to their slots on the copy stream, reloads on the
		// compute stream after it, in order with the earlier batches that may read their slots
		uint32_t n_fresh = 0;
		for (uint32_t i = 0; i < n_up; ++i)
			if (!uploads[i].reload)
				b.h_moves[n_fresh++] = make_uint2(i, uploads[i].slot);
		for (uint32_t i = 0, k = n_fresh; i < n_up; ++i)
			if (uploads[i].reload)
				b.h_moves[k++] = make_uint2(i, uploads[i].slot);
		if (n_up) {
			HCM_CHECK(hipMemcpyAsync(b.d_up_blocks, b.h_blocks, n_up * V * 2, hipMemcpyHostToDevice, s_copy));
			if (relight())
				HCM_CHECK(hipMemcpyAsync(b.d_up_heights, b.h_heights, n_up * 256, hipMemcpyHostToDevice, s_copy));
			else
				HCM_CHECK(hipMemcpyAsync(b.d_up_light, b.h_light, n_up * V, hipMemcpyHostToDevice, s_copy));
			HCM_CHECK(hipMemcpyAsync(b.d_moves, b.h_moves, n_up * sizeof(uint2), hipMemcpyHostToDevice, s_copy));
		}
		HCM_CHECK(launch_upload_scatter(b.d_up_blocks, b.d_up_light, b.d_up_heights, b.d_moves, n_fresh, d_blocks,
		                                d_light, d_heights, s_copy));
		HCM_CHECK(hipEventRecord(b.ev_up, s_copy));
		HCM_CHECK(hipStreamWaitEvent(s_comp, b.ev_up, 0));
		HCM_CHECK(launch_upload_scatter(b.d_up_blocks, b.d_up_light, b.d_up_heights, b.d_moves + n_fresh,
		                                n_up - n_fresh, d_blocks, d_light, d_heights, s_comp));
		uploads.erase(uploads.begin(), uploads.begin() + n_up);

		uint32_t k = 0;
		for (auto it = edits.begin(); it != edits.end() && k < n_ed;) {
			b.h_edits[k++] = it->second.e;
			it = edits.erase(it);
		}
		for (auto &e : edits)
			applied = std::min(applied, e.second.first_epoch - 1);
		b.epoch = applied;
		if (n_ed) {
			HCM_CHECK(hipMemcpyAsync(b.d_edits, b.h_edits, n_ed * sizeof(VoxelEdit), hipMemcpyHostToDevice, s_comp));
			HCM_CHECK(launch_edits(d_blocks, d_light, d_heights, b.d_edits, n_ed, s_comp));
		}

		if (!b.jobs.empty()) {
			for (size_t j = 0; j < b.jobs.size(); ++j)
				std::memcpy(b.h_nbr + 27 * j, world.at(b.jobs[j]).nbr, 27 * sizeof(uint32_t));
			HCM_CHECK(hipMemcpyAsync(b.d_nbr, b.h_nbr, b.jobs.size() * 27 * 4, hipMemcpyHostToDevice, s_comp));
			HCM_CHECK(hipEventRecord(b.ev_k0, s_comp));
			if (relight()) { // rebuild the jobs' light from the slots they read (ChunkMeshTask.cpp:48-111)
				std::vector<uint32_t> rs(b.h_nbr, b.h_nbr + 27 * b.jobs.size());
				std::sort(rs.begin(), rs.end());
				rs.erase(std::unique(rs.begin(), rs.end()), rs.end());
				std::memcpy(b.h_rslots, rs.data(), rs.size() * 4);
				HCM_CHECK(hipMemcpyAsync(b.d_rslots, b.h_rslots, rs.size() * 4, hipMemcpyHostToDevice, s_comp));
				LightParams lp;
				std::memset(&lp, 0, sizeof(lp));
				lp.blocks = d_blocks;
				lp.heights = d_heights;
				lp.light_info = info.light_info;
				lp.slots = b.d_rslots;
				lp.n_slots = (uint32_t)rs.size();
				lp.n_jobs = (uint32_t)b.jobs.size();
				lp.packed = d_packed;
				lp.levels = d_levels;
				lp.nbr = b.d_nbr;
				lp.ctr = b.d_lctr;
				lp.halo_light = b.d_halo;
				lp.self_reset = 1; // the slot's light queues start at zero; the run's last workgroup clears them
				HCM_CHECK(launch_light_prep(lp, s_comp));
				HCM_CHECK(launch_light(lp, std::max(1, std::min(info.n_cu * 3, (int)b.jobs.size())), s_comp));
			}
			MeshParams p;
			std::memset(&p, 0, sizeof(p));
			p.reg = info.reg;
			p.blocks = d_blocks;
			p.light = d_light;
			p.nbr = b.d_nbr;
			p.n_jobs = (uint32_t)b.jobs.size();
			p.persistent = 1;
			p.pool_cap = (uint32_t)b.arena.count;
			p.ctr = b.d_ctr;
			p.pool_verts = d_verts + 4 * b.arena.first;
			p.pool_idx = d_idx + 6 * b.arena.first;
			p.recs = b.d_recs;
			p.rec_cap = rec_cap();
			p.results = b.d_res;
			p.halo_light = relight() ? b.d_halo : nullptr;
			p.self_reset = 1; // the slot's counters start at zero; the run's last workgroup clears them
			const int grid = (int)std::max<uint32_t>(1, std::min<uint32_t>(info.grid, (uint32_t)b.jobs.size()));
			HCM_CHECK(launch_mesh(p, (int)info.threads, grid, s_comp));
			HCM_CHECK(hipEventRecord(b.ev_k1, s_comp));
			HCM_CHECK(hipMemcpyAsync(b.h_res, b.d_res, b.jobs.size() * sizeof(hcm_job_result), hipMemcpyDeviceToHost,
			                         s_comp));
			HCM_CHECK(hipMemcpyAsync(b.h_ctr, b.d_ctr, sizeof(Counters), hipMemcpyDeviceToHost, s_comp));
			b.recs_copied = std::min<uint32_t>(rec_cap(), 3 * (uint32_t)b.jobs.size() + 64);
			HCM_CHECK(hipMemcpyAsync(b.h_recs, b.d_recs, b.recs_copied * sizeof(hcm_mesh_rec), hipMemcpyDeviceToHost,
			                         s_comp));
		}
		HCM_CHECK(hipEventRecord(b.ev_done, s_comp));
		return 0;
	}

	uint32_t rec_cap() const { return 3 * cfg.max_batch + 1024; }

	int finish_batch(BatchSlot &b) {
		if (!b.jobs.empty()) {
			float ms = 0;
			HCM_CHECK(hipEventElapsedTime(&ms, b.ev_k0, b.ev_k1));
			const uint32_t n_recs = std::min(b.h_ctr->total_meshes, rec_cap());
			if (n_recs > b.recs_copied) // more overflow meshes than usual: fetch the rest
				HCM_CHECK(hipMemcpy(b.h_recs + b.recs_copied, b.d_recs + b.recs_copied,
				                    (n_recs - b.recs_copied) * sizeof(hcm_mesh_rec), hipMemcpyDeviceToHost));
			std::vector<uint2> hv;
			std::vector<uint16_t> hi;
			const uint64_t used = std::min<uint64_t>(b.h_ctr->total_quads, b.arena.count);
			if (cfg.keep_history && used) {
				hv.resize(used * 4);
				hi.resize(used * 6);
				HCM_CHECK(hipMemcpy(hv.data(), d_verts + 4 * b.arena.first, used * 4 * sizeof(uint2), hipMemcpyDeviceToHost));
				HCM_CHECK(hipMemcpy(hi.data(), d_idx + 6 * b.arena.first, used * 12, hipMemcpyDeviceToHost));
			}
			uint64_t meshed = 0, retries = 0, quads = 0;
			std::vector<Range> kept; // arena ranges now holding live meshes
			for (size_t j = 0; j < b.jobs.size(); ++j) {
				const uint64_t key = b.jobs[j];
				Chunk &c = world.at(key);
				const hcm_job_result &res = b.h_res[j];
				if (res.status != HCM_OK) { // arena or record space ran out: mesh it again later
					++retries;
					queue_stale(key, c, j < b.n_urgent);
					continue;
				}
				uint64_t lo = 0;
				if (res.quad_count) { // one contiguous allocation per chunk
					lo = ~0ull;
					for (uint32_t m = 0; m < res.mesh_count; ++m)
						lo = std::min<uint64_t>(lo, b.h_recs[res.first_mesh + m].first_quad);
					kept.push_back(Range{lo, res.quad_count});
				}
				if (c.has_mesh)
					pool->release(c.mesh.first, c.mesh.count);
				c.has_mesh = true;
				c.mesh = Range{b.arena.first + lo, res.quad_count};
				c.recs.assign(b.h_recs + res.first_mesh, b.h_recs + res.first_mesh + res.mesh_count);
				for (auto &r : c.recs)
					r.first_quad += (uint32_t)b.arena.first;
				c.meshed_epoch = b.epoch;
				++meshed;
				quads += res.quad_count;
				if (cfg.keep_history) {
					HistoryEntry h;
					key_pos(key, h.pos);
					h.epoch = b.epoch;
					for (uint32_t m = 0; m < res.mesh_count; ++m) {
						hcm_mesh_rec r = b.h_recs[res.first_mesh + m];
						const size_t q0 = r.first_quad;
						r.first_quad = (uint32_t)(h.verts.size() / 4);
						h.verts.insert(h.verts.end(), hv.begin() + 4 * q0, hv.begin() + 4 * (q0 + r.quad_count));
						h.idx.insert(h.idx.end(), hi.begin() + 6 * q0, hi.begin() + 6 * (q0 + r.quad_count));
						h.recs.push_back(r);
					}
					std::lock_guard<std::mutex> lk(st_mu);
					history.push_back(std::move(h));
				}
			}
			// the parts of the arena no live mesh holds go back to the pool
			std::sort(kept.begin(), kept.end(), [](const Range &x, const Range &y) { return x.first < y.first; });
			uint64_t at = 0;
			for (const Range &r : kept) {
				pool->release(b.arena.first + at, r.first - at);
				at = r.first + r.count;
			}
			pool->release(b.arena.first + at, b.arena.count - at);
			if (meshed)
				quads_per_chunk = 0.75 * quads_per_chunk + 0.25 * (double(quads) / double(meshed));
			if (retries) {
				quads_per_chunk *= 2.0;
				if (!meshed && b.arena_was_largest)
					return HCM_ERR_POOL_FULL; // not one chunk fits the largest free range
			}
			std::lock_guard<std::mutex> lk(st_mu);
			stats.batches++;
			stats.chunks_meshed += meshed;
			stats.quads_meshed += quads;
			stats.retries += retries;
			stats.gpu_ms += ms;
		}
		applied_done = std::max(applied_done, b.epoch);
		completions_due = true;
		b.busy = false;
		pool_blocked = false; // its arena's unused part and replaced meshes went back to the pool
		return HCM_OK;
	}

	// Runs after a batch finished or requests were applied. Only requests up to applied_done can complete:
	// they lead the epoch-ordered waiting list. A chunk's meshed_epoch only grows (batches retire in launch
	// order), so a request resumes its target check where the last pass stopped.
	void complete_ready() {
		if (!completions_due)
			return;
		completions_due = false;
		// with no data change pending or in flight, the device holds every request's data (a request that
		// changed nothing, such as an edit that sets a voxel to its block, then completes at once)
		if (uploads.empty() && edits.empty() && !any_busy())
			applied_done = epoch;
		const auto now = Clock::now();
		std::vector<hcm_completion> out;
		size_t n_eligible = 0;
		for (auto &r : waiting) {
			if (r->epoch > applied_done)
				break;
			++n_eligible;
			while (r->targets_done < r->targets.size() && world.at(r->targets[r->targets_done]).meshed_epoch >= r->epoch)
				++r->targets_done;
			if (r->targets_done < r->targets.size())
				continue;
			hcm_completion c;
			c.tag = r->tag;
			c.kind = r->kind;
			c.remeshed = (uint32_t)r->targets.size();
			c.epoch = r->epoch;
			c.latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
			out.push_back(c);
			r.reset();
		}
		if (!out.empty())
			waiting.erase(std::remove(waiting.begin(), waiting.begin() + n_eligible, nullptr), waiting.begin() + n_eligible);
		if (!out.empty()) {
			std::lock_guard<std::mutex> lk(out_mu);
			for (auto &c : out)
				done.push_back(c);
			completed += out.size();
			out_cv.notify_all();
		}
	}

	void run() {
		// short naps while batches run (below) should last microseconds, not the default 50 us timer slack
		prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);
		for (;;) {
			bool progress = false;
			std::vector<std::unique_ptr<Request>> got;
			{
				std::unique_lock<std::mutex> lk(in_mu);
				bool pending;
				{
					std::lock_guard<std::mutex> wl(world_mu);
					pending = work_pending();
				}
				if (incoming.empty() && !pending && !stop)
					in_cv.wait_for(lk, std::chrono::microseconds(500));
				if (stop && incoming.empty() && !pending)
					return;
				while (!incoming.empty()) {
					got.push_back(std::move(incoming.front()));
					incoming.pop_front();
				}
			}
			std::unique_lock<std::mutex> wl(world_mu);
			progress = !got.empty();
			for (auto &r : got)
				apply(std::move(r));
			for (;;) { // retire finished batches in launch order
				BatchSlot *oldest = nullptr;
				for (auto &b : slots)
					if (b.busy && (!oldest || b.id < oldest->id))
						oldest = &b;
				if (!oldest)
					break;
				const hipError_t q = hipEventQuery(oldest->ev_done);
				if (q == hipErrorNotReady)
					break;
				const int rc = q == hipSuccess ? finish_batch(*oldest) : HCM_ERR_HIP;
				if (rc != HCM_OK)
					return fail(rc);
				progress = true;
			}
			complete_ready();
			for (;;) {
				uint32_t busy = 0;
				BatchSlot *freeb = nullptr;
				for (auto &b : slots) {
					if (b.busy)
						++busy;
					else if (!freeb)
						freeb = &b;
				}
				if (!freeb || busy >= cfg.max_inflight)
					break;
				const int rc = launch_batch(*freeb);
				if (rc < 0)
					return fail(rc);
				if (rc > 0)
					break;
				progress = true;
			}
			wl.unlock();
			// nothing arrived, finished or launched in this pass: batches are running (an idle server waits
			// on in_cv above). Nap briefly instead of polling their events and the locks in a tight loop; a
			// request submitted meanwhile is taken up after at most one nap.
			if (!progress)
				std::this_thread::sleep_for(std::chrono::microseconds(10));
		}
	}
};

static void free_server(hcm_server *s) {
	for (auto &b : s->slots) {
		for (void *p : {(void *)b.d_nbr, (void *)b.d_edits, (void *)b.d_ctr, (void *)b.d_res, (void *)b.d_recs,
		                (void *)b.d_rslots, (void *)b.d_lctr, (void *)b.d_halo, (void *)b.d_up_blocks,
		                (void *)b.d_up_light, (void *)b.d_up_heights, (void *)b.d_moves})
			if (p)
				(void)hipFree(p);
		for (void *p : {(void *)b.h_nbr, (void *)b.h_edits, (void *)b.h_blocks, (void *)b.h_light, (void *)b.h_res,
		                (void *)b.h_recs, (void *)b.h_ctr, (void *)b.h_heights, (void *)b.h_rslots, (void *)b.h_moves})
			if (p)
				(void)hipHostFree(p);
		for (hipEvent_t e : {b.ev_up, b.ev_k0, b.ev_k1, b.ev_done})
			if (e)
				(void)hipEventDestroy(e);
	}
	for (void *p : {(void *)s->d_blocks, (void *)s->d_light, (void *)s->d_verts, (void *)s->d_idx, (void *)s->d_heights,
	                (void *)s->d_packed, (void *)s->d_levels})
		if (p)
			(void)hipFree(p);
	if (s->s_copy)
		(void)hipStreamDestroy(s->s_copy);
	if (s->s_comp)
		(void)hipStreamDestroy(s->s_comp);
	delete s->pool;
	delete s;
}

template <typename T> static hipError_t dalloc(T **p, size_t count) {
	return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T));
}
template <typename T> static hipError_t halloc(T **p, size_t count) {
	return hipHostMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T));
}

static int server_init(hcm_server *s) {
	const hcm_server_config &c = s->cfg;
	const int r = ctx_info(s->ctx, c.threads, &s->info);
	if (r != HCM_OK)
		return r;
	HCM_CHECK(hipStreamCreateWithFlags(&s->s_copy, hipStreamNonBlocking));
	HCM_CHECK(hipStreamCreateWithFlags(&s->s_comp, hipStreamNonBlocking));
	HCM_CHECK(dalloc(&s->d_blocks, (size_t)c.max_chunks * HCM_CHUNK_VOXELS));
	HCM_CHECK(dalloc(&s->d_light, (size_t)c.max_chunks * HCM_CHUNK_VOXELS));
	if (s->relight()) {
		HCM_CHECK(dalloc(&s->d_heights, (size_t)c.max_chunks * 256));
		HCM_CHECK(dalloc(&s->d_packed, (size_t)c.max_chunks * kLightPlanes * 256));
		HCM_CHECK(dalloc(&s->d_levels, (size_t)c.max_chunks));
	}
	HCM_CHECK(dalloc(&s->d_verts, (size_t)c.pool_quads * 4));
	HCM_CHECK(dalloc(&s->d_idx, (size_t)c.pool_quads * 6));
	s->pool = new RangeAllocator(c.pool_quads);
	s->slots.resize(c.max_inflight);
	for (auto &b : s->slots) {
		HCM_CHECK(dalloc(&b.d_nbr, (size_t)c.max_batch * 27));
		HCM_CHECK(dalloc(&b.d_edits, kMaxEdits));
		HCM_CHECK(dalloc(&b.d_ctr, 1));
		HCM_CHECK(hipMemset(b.d_ctr, 0, sizeof(Counters)));
		HCM_CHECK(dalloc(&b.d_res, c.max_batch));
		HCM_CHECK(dalloc(&b.d_recs, s->rec_cap()));
		HCM_CHECK(halloc(&b.h_nbr, (size_t)c.max_batch * 27));
		HCM_CHECK(halloc(&b.h_edits, kMaxEdits));
		HCM_CHECK(halloc(&b.h_blocks, (size_t)kMaxUploads * HCM_CHUNK_VOXELS));
		HCM_CHECK(dalloc(&b.d_up_blocks, (size_t)kMaxUploads * HCM_CHUNK_VOXELS));
		HCM_CHECK(halloc(&b.h_moves, kMaxUploads));
		HCM_CHECK(dalloc(&b.d_moves, kMaxUploads));
		if (!s->relight()) {
			HCM_CHECK(halloc(&b.h_light, (size_t)kMaxUploads * HCM_CHUNK_VOXELS));
			HCM_CHECK(dalloc(&b.d_up_light, (size_t)kMaxUploads * HCM_CHUNK_VOXELS));
		}
		if (s->relight()) {
			HCM_CHECK(halloc(&b.h_heights, (size_t)kMaxUploads * 256));
			HCM_CHECK(dalloc(&b.d_up_heights, (size_t)kMaxUploads * 256));
			HCM_CHECK(dalloc(&b.d_rslots, (size_t)c.max_batch * 27));
			HCM_CHECK(halloc(&b.h_rslots, (size_t)c.max_batch * 27));
			HCM_CHECK(dalloc(&b.d_lctr, 1));
			HCM_CHECK(hipMemset(b.d_lctr, 0, sizeof(Counters)));
			HCM_CHECK(dalloc(&b.d_halo, (size_t)c.max_batch * kLightJobBytes));
		}
		HCM_CHECK(halloc(&b.h_res, c.max_batch));
		HCM_CHECK(halloc(&b.h_recs, s->rec_cap()));
		HCM_CHECK(halloc(&b.h_ctr, 1));
		HCM_CHECK(hipEventCreateWithFlags(&b.ev_up, hipEventDisableTiming));
		HCM_CHECK(hipEventCreate(&b.ev_k0));
		HCM_CHECK(hipEventCreate(&b.ev_k1));
		HCM_CHECK(hipEventCreateWithFlags(&b.ev_done, hipEventDisableTiming));
	}
	return HCM_OK;
}

extern "C" int hcm_server_create(hcm_ctx *ctx, const hcm_server_config *cfg, hcm_server **out) {
	if (!out)
		return HCM_ERR_INVALID;
	*out = nullptr;
	if (!ctx || !cfg || cfg->max_chunks == 0 || cfg->max_chunks > (1u << 20) || cfg->max_batch == 0 ||
	    cfg->max_batch > (1u << 20) || cfg->max_inflight == 0 || cfg->max_inflight > 16 || cfg->pool_quads == 0 ||
	    cfg->pool_quads >= (1ull << 32) || (cfg->light_mode != HCM_LIGHT_STORE && cfg->light_mode != HCM_LIGHT_REBUILD))
		return HCM_ERR_INVALID;
	hcm_server *s = new hcm_server;
	s->ctx = ctx;
	s->cfg = *cfg;
	const int r = server_init(s);
	if (r != HCM_OK) {
		free_server(s);
		return r;
	}
	s->thread = std::thread([s]() { s->run(); });
	*out = s;
	return HCM_OK;
}

extern "C" void hcm_server_destroy(hcm_server *s) {
	if (!s)
		return;
	{
		std::lock_guard<std::mutex> lk(s->in_mu);
		s->stop = true;
	}
	s->in_cv.notify_all();
	if (s->thread.joinable())
		s->thread.join();
	(void)hipStreamSynchronize(s->s_copy);
	(void)hipStreamSynchronize(s->s_comp);
	free_server(s);
}

static int submit_load(hcm_server *s, int32_t cx, int32_t cy, int32_t cz, const uint16_t *blocks,
                       const uint8_t *light, const int8_t *heights, uint64_t tag) {
	if (cx < -(1 << 20) || cx >= (1 << 20) || cy < -(1 << 20) || cy >= (1 << 20) || cz < -(1 << 20) || cz >= (1 << 20))
		return HCM_ERR_INVALID;
	auto r = std::make_unique<Request>();
	r->tag = tag;
	r->kind = HCM_REQ_LOAD;
	r->pos[0] = cx, r->pos[1] = cy, r->pos[2] = cz;
	r->blocks.assign(blocks, blocks + HCM_CHUNK_VOXELS);
	if (light)
		r->light.assign(light, light + HCM_CHUNK_VOXELS);
	if (heights)
		r->heights.assign(heights, heights + 256);
	r->t_submit = Clock::now();
	{
		std::lock_guard<std::mutex> lk(s->in_mu);
		if (s->error != HCM_OK)
			return s->error;
		const uint64_t key = chunk_key(cx, cy, cz);
		if (!s->accepted.count(key)) {
			if (s->accepted.size() >= s->cfg.max_chunks)
				return HCM_ERR_BUSY;
			const uint32_t slot = (uint32_t)s->accepted.size();
			s->accepted.emplace(key, slot);
		}
		r->slot = s->accepted.at(key);
		++s->submitted;
		s->incoming.push_back(std::move(r));
	}
	s->in_cv.notify_one();
	return HCM_OK;
}

extern "C" int hcm_server_load(hcm_server *s, int32_t cx, int32_t cy, int32_t cz, const uint16_t *blocks,
                               const uint8_t *light, uint64_t tag) {
	if (!s || !blocks || !light || s->relight())
		return HCM_ERR_INVALID;
	return submit_load(s, cx, cy, cz, blocks, light, nullptr, tag);
}

extern "C" int hcm_server_load_sunlit(hcm_server *s, int32_t cx, int32_t cy, int32_t cz, const uint16_t *blocks,
                                      const int8_t *heights, uint64_t tag) {
	if (!s || !blocks || !heights || !s->relight())
		return HCM_ERR_INVALID;
	return submit_load(s, cx, cy, cz, blocks, nullptr, heights, tag);
}

extern "C" int hcm_server_edit(hcm_server *s, int32_t x, int32_t y, int32_t z, uint16_t block, uint8_t light,
                               uint64_t tag) {
	if (!s)
		return HCM_ERR_INVALID;
	auto r = std::make_unique<Request>();
	r->tag = tag;
	r->kind = HCM_REQ_EDIT;
	r->pos[0] = x, r->pos[1] = y, r->pos[2] = z;
	r->block = block;
	r->light1 = light;
	r->t_submit = Clock::now();
	{
		std::lock_guard<std::mutex> lk(s->in_mu);
		if (s->error != HCM_OK)
			return s->error;
		if (!s->accepted.count(chunk_key(floor_div16(x), floor_div16(y), floor_div16(z))))
			return HCM_ERR_INVALID;
		++s->submitted;
		s->incoming.push_back(std::move(r));
	}
	s->in_cv.notify_one();
	return HCM_OK;
}

extern "C" int hcm_server_poll(hcm_server *s, hcm_completion *out, uint32_t cap, uint32_t *n) {
	if (!s || !n || (cap && !out))
		return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(s->out_mu);
	uint32_t k = 0;
	while (k < cap && !s->done.empty()) {
		out[k++] = s->done.front();
		s->done.pop_front();
	}
	*n = k;
	return s->error;
}

extern "C" int hcm_server_drain(hcm_server *s) {
	if (!s)
		return HCM_ERR_INVALID;
	uint64_t target;
	{
		std::lock_guard<std::mutex> lk(s->in_mu);
		target = s->submitted;
	}
	std::unique_lock<std::mutex> lk(s->out_mu);
	while (s->completed < target && s->error == HCM_OK)
		s->out_cv.wait_for(lk, std::chrono::milliseconds(1));
	return s->error;
}

static int copy_meshes(const std::vector<hcm_mesh_rec> &recs, const uint2 *verts, const uint16_t *idx, bool device,
                       hcm_host_result *out) {
	uint64_t nv = 0, ni = 0;
	for (auto &m : recs)
		nv += 4ull * m.quad_count, ni += 6ull * m.quad_count;
	out->vertex_count = (uint32_t)nv, out->index_count = (uint32_t)ni, out->mesh_count = (uint32_t)recs.size();
	if (nv > out->vertex_capacity || ni > out->index_capacity || recs.size() > out->mesh_capacity ||
	    (nv && (!out->vertices || !out->indices)) || (!recs.empty() && !out->meshes))
		return HCM_ERR_CAPACITY;
	uint32_t vb = 0, ib = 0;
	for (size_t m = 0; m < recs.size(); ++m) {
		const hcm_mesh_rec &rc = recs[m];
		hcm_mesh_desc &d = out->meshes[m];
		d.vert_begin = vb, d.vert_count = 4 * rc.quad_count, d.idx_begin = ib, d.idx_count = 6 * rc.quad_count;
		std::memcpy(d.aabb, rc.aabb, sizeof(d.aabb));
		d.transparent = rc.transparent;
		const size_t vbytes = 4ull * rc.quad_count * sizeof(uint2), ibytes = 12ull * rc.quad_count;
		if (device) {
			HCM_CHECK(hipMemcpy(out->vertices + vb, verts + 4ull * rc.first_quad, vbytes, hipMemcpyDeviceToHost));
			HCM_CHECK(hipMemcpy(out->indices + ib, idx + 6ull * rc.first_quad, ibytes, hipMemcpyDeviceToHost));
		} else {
			std::memcpy(out->vertices + vb, verts + 4ull * rc.first_quad, vbytes);
			std::memcpy(out->indices + ib, idx + 6ull * rc.first_quad, ibytes);
		}
		vb += d.vert_count, ib += d.idx_count;
	}
	return HCM_OK;
}

extern "C" int hcm_server_chunk_mesh(hcm_server *s, int32_t cx, int32_t cy, int32_t cz, hcm_host_result *out) {
	if (!s || !out)
		return HCM_ERR_INVALID;
	const int r = hcm_server_drain(s);
	if (r != HCM_OK)
		return r;
	// world_mu keeps the chunk's pool range live while it is copied
	std::lock_guard<std::mutex> wl(s->world_mu);
	auto it = s->world.find(chunk_key(cx, cy, cz));
	if (it == s->world.end() || !it->second.has_mesh)
		return HCM_ERR_INVALID;
	return copy_meshes(it->second.recs, s->d_verts, s->d_idx, true, out);
}

extern "C" int hcm_server_get_stats(hcm_server *s, hcm_server_stats *out) {
	if (!s || !out)
		return HCM_ERR_INVALID;
	hcm_server_stats st;
	{
		std::lock_guard<std::mutex> lk(s->st_mu);
		st = s->stats;
	}
	{
		std::lock_guard<std::mutex> lk(s->in_mu);
		st.requests = s->submitted;
	}
	{
		std::lock_guard<std::mutex> lk(s->out_mu);
		st.completed = s->completed;
	}
	{
		std::lock_guard<std::mutex> lk(s->world_mu);
		st.pool_free_quads = s->pool->free_total();
		st.pool_used_quads = s->pool->capacity() - st.pool_free_quads;
	}
	*out = st;
	return HCM_OK;
}

extern "C" int hcm_server_history_size(hcm_server *s, uint64_t *n) {
	if (!s || !n)
		return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(s->st_mu);
	*n = s->history.size();
	return HCM_OK;
}

extern "C" int hcm_server_history_get(hcm_server *s, uint64_t index, int32_t pos[3], uint64_t *epoch,
                                      hcm_host_result *out) {
	if (!s || !pos || !epoch || !out)
		return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(s->st_mu);
	if (index >= s->history.size())
		return HCM_ERR_INVALID;
	const HistoryEntry &h = s->history[index];
	std::memcpy(pos, h.pos, sizeof(h.pos));
	*epoch = h.epoch;
	return copy_meshes(h.recs, h.verts.data(), h.idx.data(), false, out);
}
