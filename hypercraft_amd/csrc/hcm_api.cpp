// hcm_api.cpp -- host side of the C ABI declared in include/hcm.h.
//
// Owns device memory (registry tables, quad pools, chunk stores, job tables, depth hierarchies), the HIP
// stream each context works on, and the HIP graph that replays the batched pipeline (the light rebuild
// when asked for, then the persistent meshing kernel; both clear their own counters). Kernels live in
// hcm_kernels.hip, hcm_light.hip and hcm_draw.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "hcm_internal.h"

using namespace hcm;

namespace {

thread_local int g_last_hip = 0;

#define HCM_HIP(call)                                                                                              \
	do {                                                                                                           \
		hipError_t e_ = (call);                                                                                    \
		if (e_ != hipSuccess) {                                                                                    \
			g_last_hip = (int)e_;                                                                                  \
			return HCM_ERR_HIP;                                                                                    \
		}                                                                                                          \
	} while (0)

template <typename T> hipError_t dmalloc(T **p, size_t count) {
	return hipMalloc(reinterpret_cast<void **>(p), std::max<size_t>(count, 1) * sizeof(T));
}

// Registry flattened for the kernels (see hcm_internal.h for the word layouts).
struct DeviceTables {
	std::vector<uint8_t> pid_of;
	std::vector<uint2> rows;
	std::vector<uint32_t> pid_vox;
	std::vector<uint32_t> codes;
	std::vector<hcm_custom_mesh> meshes;
	std::vector<uint8_t> tex_flags;
	std::vector<uint8_t> light_info; // [65536]: light level | kLightInfoPass
	uint32_t flow_pid_base = 0, flow_tex = 0, src_id = 0, flow_id = 0;
};

int build_tables(const hcm_registry *r, DeviceTables *t) {
	if (!r || !r->row_of || !r->rows || r->n_rows == 0 || (r->n_meshes && !r->meshes) || (r->n_tex && !r->tex_flags))
		return HCM_ERR_INVALID;
	if (r->n_rows + kFlowPatterns > (uint32_t)kMaxPid || r->n_meshes >= 255)
		return HCM_ERR_INVALID;
	auto tex_flags = [&](uint16_t tex) -> uint32_t {
		uint32_t id = tex & 0x1fffu;
		return id < r->n_tex ? r->tex_flags[id] : 0u;
	};
	// face class of a texture for BlockTexture::Show: empty, opaque, transparent, liquid
	auto tex_class = [&](uint16_t tex) -> uint32_t {
		if ((tex & 0x1fffu) == 0)
			return kClassEmpty;
		uint32_t f = tex_flags(tex);
		if (f & HCM_TEX_LIQUID)
			return kClassLiquid;
		return (f & HCM_TEX_TRANSPARENT) ? kClassTransparent : kClassOpaque;
	};
	// texture codes: code 0 is the empty texture value 0; codes are unique per full texture value
	std::vector<uint16_t> code_tex{0};
	auto code_of = [&](uint16_t tex) -> uint32_t {
		for (size_t i = 0; i < code_tex.size(); ++i)
			if (code_tex[i] == tex)
				return (uint32_t)i;
		code_tex.push_back(tex);
		return (uint32_t)code_tex.size() - 1;
	};
	// BlockTexture::Show (texture/include/texture/BlockTexture.hpp:68-76) on full texture values; texture
	// id 0 counts as transparent (block_texture_transparency.inl, index 0)
	auto show = [&](uint16_t self, uint16_t other) -> bool {
		const bool s_empty = (self & 0x1fffu) == 0, o_empty = (other & 0x1fffu) == 0;
		const bool s_t = s_empty || (tex_flags(self) & HCM_TEX_TRANSPARENT);
		const bool o_t = o_empty || (tex_flags(other) & HCM_TEX_TRANSPARENT);
		const bool s_l = !s_empty && (tex_flags(self) & HCM_TEX_LIQUID);
		const bool o_l = !o_empty && (tex_flags(other) & HCM_TEX_LIQUID);
		if (s_empty || self == other || (!s_t && !o_t) || (s_l && !o_empty))
			return false;
		return !s_t || o_t || o_l;
	};
	auto add_row = [&](const uint16_t tex[6], uint32_t dp, uint32_t ip, uint32_t mesh, uint32_t flags) {
		uint32_t tc[6], cls = 0;
		for (int f = 0; f < 6; ++f) {
			tc[f] = code_of(tex[f]);
			cls |= tex_class(tex[f]) << (2 * f);
		}
		// kRowSelfShow: two neighbouring blocks of this row show a face to each other along some axis
		// (their + and - textures differ and one of them shows against the other), so a halo made of
		// this one block value still has faces
		uint32_t self_show = 0;
		for (int a = 0; a < 3; ++a)
			if (show(tex[2 * a], tex[2 * a + 1]) || show(tex[2 * a + 1], tex[2 * a]))
				self_show = kRowSelfShow;
		const uint32_t pid = (uint32_t)t->rows.size();
		t->rows.push_back(make_uint2(tc[0] | tc[1] << 8 | tc[2] << 16 | tc[3] << 24,
		                             tc[4] | tc[5] << 8 | self_show | mesh << 24));
		t->pid_vox.push_back((dp ? kVoxDirect : 0u) | (ip ? kVoxIndirect : 0u) | cls << kVoxClassShift | flags |
		                     pid << kVoxPidShift);
	};
	t->rows.clear();
	t->pid_vox.clear();
	for (uint32_t i = 0; i < r->n_rows; ++i) {
		const hcm_block_row &row = r->rows[i];
		uint32_t mesh = 0xffu, flags = 0;
		if (row.mesh != HCM_NO_MESH) {
			if (row.mesh >= r->n_meshes)
				return HCM_ERR_INVALID;
			mesh = row.mesh;
			flags |= kVoxMesh;
			if (r->meshes[row.mesh].kind == HCM_MESH_LIQUID_FLOW)
				flags |= kVoxFlow;
			else if (r->meshes[row.mesh].face_count > HCM_MAX_FACES)
				return HCM_ERR_INVALID;
		}
		add_row(row.tex, row.transparency == 2 || row.collision != 1, row.transparency != 0, mesh, flags);
	}
	// dynamic liquid-flow textures: 16 = all faces liquid; else bits right, left, front, back full
	// (the voxel keeps its own light-pass bits; only the face textures come from these rows)
	t->flow_pid_base = r->n_rows;
	const uint16_t ft = r->liquid_flow_tex;
	for (uint32_t p = 0; p < (uint32_t)kFlowPatterns; ++p) {
		uint16_t tex[6];
		if (p == 16)
			for (int f = 0; f < 6; ++f)
				tex[f] = ft;
		else {
			tex[0] = (p & 1) ? ft : 0, tex[1] = (p & 2) ? ft : 0, tex[2] = 0, tex[3] = ft;
			tex[4] = (p & 4) ? ft : 0, tex[5] = (p & 8) ? ft : 0;
		}
		add_row(tex, 0, 0, 0xffu, 0);
	}
	if (code_tex.size() > (size_t)kMaxCodes)
		return HCM_ERR_INVALID;
	t->codes.clear();
	for (uint16_t tex : code_tex) {
		uint32_t f = tex_flags(tex), cf = 0;
		if (f & HCM_TEX_TRANSPARENT)
			cf |= kCodeTransparent;
		if (f & HCM_TEX_TRANS_PASS)
			cf |= kCodeTransPass;
		if (f & HCM_TEX_LIQUID)
			cf |= kCodeLiquid;
		if ((tex & 0x1fffu) == 0)
			cf |= kCodeEmpty;
		t->codes.push_back(uint32_t(tex) | cf << 16);
	}
	t->pid_of.resize(65536 + 256);
	for (uint32_t d = 0; d < 65536; ++d) {
		if (r->row_of[d] >= r->n_rows)
			return HCM_ERR_INVALID;
		t->pid_of[d] = (uint8_t)r->row_of[d];
	}
	// per block id: its row when no meta value changes it (kept in LDS by the kernel), else 0xff
	for (uint32_t id = 0; id < 256; ++id) {
		uint8_t v = t->pid_of[id];
		for (uint32_t meta = 1; meta < 256; ++meta)
			if (t->pid_of[id | meta << 8] != v) {
				v = 0xff;
				break;
			}
		t->pid_of[65536 + id] = v;
	}
	// the light rebuild's view of a block: Block::GetLightLevel, GetIndirectLightPass and
	// GetVerticalLightPass (Block.hpp:69-71)
	t->light_info.resize(65536);
	for (uint32_t d = 0; d < 65536; ++d) {
		const uint32_t lv = r->light_level ? r->light_level[d] : 0u;
		if (lv > 15)
			return HCM_ERR_INVALID;
		const uint32_t tr = r->rows[r->row_of[d]].transparency;
		t->light_info[d] = uint8_t(lv | (tr != 0 ? kLightInfoPass : 0u) | (tr == 2 ? kLightInfoVertical : 0u));
	}
	t->meshes.assign(r->meshes, r->meshes + r->n_meshes);
	t->tex_flags.assign(r->tex_flags, r->tex_flags + r->n_tex);
	t->flow_tex = r->liquid_flow_tex;
	t->src_id = r->liquid_source_id;
	t->flow_id = r->liquid_flow_id;
	return HCM_OK;
}

constexpr uint32_t kPlainPoolQuads = 1u << 17; // > the most quads one 16^3 chunk can produce
static_assert(sizeof(hcm_mesh_info) == 52, "chunk_culling.comp MeshInfo layout (std430)");
static_assert(sizeof(hcm_draw_cmd) == 20, "VkDrawIndexedIndirectCommand layout");
constexpr uint32_t kPlainRecs = 64;
// Workgroup size of launches that mesh one chunk (hcm_mesh_chunk, hcm_mesh_chunk_relight, HCM_RUN_PER_CHUNK):
// a lone workgroup has the GPU to itself, so more lanes per chunk shorten it (cave chunks: 23.9 us at 512
// threads, 27.7 us at 256, 37.3 us at 128). The persistent batch keeps its tuned shape.
constexpr int kChunkLaunchThreads = 512;
// A one-chunk call copies back this many quads with its result, records and vertices in one go; a chunk
// with more takes a second round trip.
constexpr uint32_t kPlainFastQuads = 2048;
constexpr size_t kHaloBytes = HCM_HALO_BLOCKS * 2 + HCM_HALO_LIGHT;
constexpr size_t kPoutRecs = 64, kPoutVerts = 64 + kPlainRecs * sizeof(hcm_mesh_rec);
constexpr size_t kPoutIdx = kPoutVerts + kPlainFastQuads * 4 * sizeof(uint2);
constexpr size_t kPoutBytes = kPoutIdx + kPlainFastQuads * 6 * 2;

// Scratch of hcm_mesh_halos_noalloc: the device memory its kernels use beyond the caller's inputs and
// outputs -- work counters, the job table (job j reads slots 27 j .. 27 j + 26) and the 27-slot stores
// its halos are scattered into -- for at most kNoAllocMaxJobs chunks per call, 2.7 MB. Each GPU holds its
// own copy from the library's load to the end of the process; calls on one GPU share it.
constexpr uint32_t kNoAllocMaxJobs = 8;
constexpr size_t kScratchNbr = (sizeof(Counters) + 255) / 256 * 256;
constexpr size_t kScratchBlocks = kScratchNbr + (kNoAllocMaxJobs * 27 * 4 + 255) / 256 * 256;
constexpr size_t kScratchLight = kScratchBlocks + (size_t)kNoAllocMaxJobs * 27 * HCM_CHUNK_VOXELS * 2;
constexpr size_t kScratchBytes = kScratchLight + (size_t)kNoAllocMaxJobs * 27 * HCM_CHUNK_VOXELS;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];
static_assert(kNoAllocMaxJobs == HCM_NOALLOC_MAX_CHUNKS, "documented limit");

} // namespace

struct hcm_ctx {
	int device = 0;
	int n_cu = 0;
	hipStream_t stream = nullptr;
	std::mutex mu;
	DevReg reg{};
	// device registry buffers
	uint8_t *d_pid_of = nullptr;
	uint2 *d_rows = nullptr;
	uint32_t *d_pid_vox = nullptr;
	uint32_t *d_codes = nullptr;
	hcm_custom_mesh *d_meshes = nullptr;
	uint8_t *d_tex_flags = nullptr;
	uint8_t *d_light_info = nullptr;
	std::vector<uint8_t> light_info; // host copy
	// batched pool
	uint64_t pool_quads = 0;
	uint2 *d_verts = nullptr;
	uint16_t *d_idx = nullptr;
	// plain path
	uint16_t *d_pblocks = nullptr;
	uint8_t *d_plight = nullptr;
	uint32_t *d_pnbr = nullptr;
	Counters *d_pctr = nullptr;
	uint2 *d_pverts = nullptr;
	uint16_t *d_pidx = nullptr;
	hcm_mesh_rec *d_precs = nullptr;
	hcm_job_result *d_pres = nullptr;
	uint16_t *h_pblocks = nullptr; // pinned staging
	uint8_t *h_plight = nullptr;
	// plain path from a packed halo: the halo on the device and in pinned memory, and a pinned copy of
	// the result, the mesh records and the first kPlainFastQuads quads
	uint8_t *d_halo_in = nullptr;
	uint8_t *h_halo_in = nullptr;
	uint8_t *h_pout = nullptr;
	// this device's copy of `scratch` (hcm_mesh_halos_noalloc); host code uses this address, never the
	// array's name, which on the host is only a stand-in
	uint8_t *d_scratch = nullptr;
	// plain path light rebuild
	int8_t *d_pheights = nullptr;
	int8_t *h_pheights = nullptr;
	uint16_t *d_ppacked = nullptr;
	uint32_t *d_plevels = nullptr;
	Counters *d_plctr = nullptr;
	uint8_t *d_phalo = nullptr;
	// depth hierarchy for occlusion culling (hcm_ctx_set_depth)
	float *d_depth = nullptr;
	size_t depth_cap = 0; // floats
	DepthPyramid hiz;
	hcm_tuning tuning{0, 256, {0, 0}};
};

struct hcm_store {
	hcm_ctx *ctx = nullptr;
	uint32_t n_slots = 0;
	uint16_t *d_blocks = nullptr;
	uint8_t *d_light = nullptr;
	int8_t *d_heights = nullptr;  // [n_slots][256] sunlight heights
	uint16_t *d_packed = nullptr; // [n_slots][kLightPlanes][256] light-rebuild inputs, packed per batch run
	uint32_t *d_levels = nullptr; // [n_slots] torchlight levels present
};

struct hcm_batch {
	hcm_ctx *ctx = nullptr;
	hcm_store *store = nullptr;
	uint32_t n_jobs = 0;
	uint32_t rec_cap = 0;
	uint32_t *d_nbr = nullptr;
	Counters *d_ctr = nullptr;
	hcm_mesh_rec *d_recs = nullptr;
	hcm_job_result *d_res = nullptr;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	hipGraph_t graph = nullptr;
	hipGraphExec_t graph_exec = nullptr;
	uint32_t graph_threads = 0, graph_grid = 0;
	uint32_t last_mode = 0;
	bool ran = false;
	bool ctr_clean = false; // the counters are zero: a persistent run needs no reset kernel
	bool launched = false;  // a persistent run has been launched directly: the next one captures the graph
	// light rebuild (HCM_LIGHT_REBUILD)
	uint32_t light_mode = HCM_LIGHT_STORE, graph_light = HCM_LIGHT_STORE;
	uint32_t n_read_slots = 0;
	uint32_t *d_read_slots = nullptr; // the distinct store slots the jobs read
	Counters *d_lctr = nullptr;
	uint8_t *d_halo_light = nullptr;  // [n_jobs][kLightJobBytes]
	bool light_ran = false;
	// draw lists
	int32_t *d_chunk_pos = nullptr;
	hcm_mesh_info *d_infos = nullptr;
	hcm_draw_cmd *d_draws = nullptr; // [2][rec_cap]: opaque, transparent
	uint32_t *d_draw_counts = nullptr;
	uint32_t n_infos = 0;
	bool infos_ready = false;
};

extern "C" int hcm_last_hip_error(void) { return g_last_hip; }

extern "C" int hcm_device_ok(void) {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
		return 0;
	int dev = 0;
	if (hipGetDevice(&dev) != hipSuccess)
		return 0;
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
		return 0;
	return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0 ? 1 : 0;
}

static void dfree(void *p) {
	if (p)
		(void)hipFree(p);
}

static void free_ctx(hcm_ctx *c) {
	if (!c)
		return;
	dfree(c->d_pid_of);
	dfree(c->d_rows);
	dfree(c->d_pid_vox);
	dfree(c->d_codes);
	dfree(c->d_meshes);
	dfree(c->d_tex_flags);
	dfree(c->d_light_info);
	dfree(c->d_verts);
	dfree(c->d_idx);
	dfree(c->d_pblocks);
	dfree(c->d_plight);
	dfree(c->d_pnbr);
	dfree(c->d_pctr);
	dfree(c->d_pverts);
	dfree(c->d_pidx);
	dfree(c->d_precs);
	dfree(c->d_pres);
	if (c->h_pblocks)
		(void)hipHostFree(c->h_pblocks);
	if (c->h_plight)
		(void)hipHostFree(c->h_plight);
	dfree(c->d_halo_in);
	if (c->h_halo_in)
		(void)hipHostFree(c->h_halo_in);
	if (c->h_pout)
		(void)hipHostFree(c->h_pout);
	dfree(c->d_pheights);
	dfree(c->d_ppacked);
	dfree(c->d_plevels);
	dfree(c->d_plctr);
	dfree(c->d_phalo);
	dfree(c->d_depth);
	if (c->h_pheights)
		(void)hipHostFree(c->h_pheights);
	if (c->stream)
		(void)hipStreamDestroy(c->stream);
	delete c;
}

static int create_impl(hcm_ctx *c, const hcm_registry *registry, uint64_t pool_quads) {
	DeviceTables t;
	int r = build_tables(registry, &t);
	if (r != HCM_OK)
		return r;
	HCM_HIP(hipGetDevice(&c->device));
	hipDeviceProp_t prop;
	HCM_HIP(hipGetDeviceProperties(&prop, c->device));
	if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
		return HCM_ERR_NO_DEVICE;
	c->n_cu = prop.multiProcessorCount;
	HCM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));

	HCM_HIP(dmalloc(&c->d_pid_of, t.pid_of.size()));
	HCM_HIP(dmalloc(&c->d_rows, t.rows.size()));
	HCM_HIP(dmalloc(&c->d_pid_vox, t.pid_vox.size()));
	HCM_HIP(dmalloc(&c->d_codes, t.codes.size()));
	HCM_HIP(dmalloc(&c->d_meshes, t.meshes.size()));
	HCM_HIP(dmalloc(&c->d_tex_flags, t.tex_flags.size()));
	HCM_HIP(dmalloc(&c->d_light_info, t.light_info.size()));
	HCM_HIP(hipMemcpy(c->d_light_info, t.light_info.data(), t.light_info.size(), hipMemcpyHostToDevice));
	c->light_info = t.light_info;
	HCM_HIP(hipMemcpy(c->d_pid_of, t.pid_of.data(), t.pid_of.size(), hipMemcpyHostToDevice));
	HCM_HIP(hipMemcpy(c->d_rows, t.rows.data(), t.rows.size() * sizeof(uint2), hipMemcpyHostToDevice));
	HCM_HIP(hipMemcpy(c->d_pid_vox, t.pid_vox.data(), t.pid_vox.size() * 4, hipMemcpyHostToDevice));
	HCM_HIP(hipMemcpy(c->d_codes, t.codes.data(), t.codes.size() * 4, hipMemcpyHostToDevice));
	if (!t.meshes.empty())
		HCM_HIP(hipMemcpy(c->d_meshes, t.meshes.data(), t.meshes.size() * sizeof(hcm_custom_mesh),
		                  hipMemcpyHostToDevice));
	if (!t.tex_flags.empty())
		HCM_HIP(hipMemcpy(c->d_tex_flags, t.tex_flags.data(), t.tex_flags.size(), hipMemcpyHostToDevice));
	c->reg.pid_of = c->d_pid_of;
	c->reg.rows = c->d_rows;
	c->reg.pid_vox = c->d_pid_vox;
	c->reg.codes = c->d_codes;
	c->reg.meshes = c->d_meshes;
	c->reg.tex_flags = c->d_tex_flags;
	c->reg.n_pid = (uint32_t)t.rows.size();
	c->reg.n_codes = (uint32_t)t.codes.size();
	c->reg.n_tex = (uint32_t)t.tex_flags.size();
	c->reg.flow_pid_base = t.flow_pid_base;
	c->reg.flow_tex = t.flow_tex;
	c->reg.src_id = t.src_id;
	c->reg.flow_id = t.flow_id;

	c->pool_quads = std::min<uint64_t>(pool_quads, 0xffffffffull - (1u << 20));
	if (c->pool_quads) {
		HCM_HIP(dmalloc(&c->d_verts, (size_t)c->pool_quads * 4));
		HCM_HIP(dmalloc(&c->d_idx, (size_t)c->pool_quads * 6));
	}
	HCM_HIP(dmalloc(&c->d_pblocks, 27 * HCM_CHUNK_VOXELS));
	HCM_HIP(dmalloc(&c->d_plight, 27 * HCM_CHUNK_VOXELS));
	HCM_HIP(dmalloc(&c->d_pnbr, 27));
	HCM_HIP(dmalloc(&c->d_pctr, 1));
	HCM_HIP(dmalloc(&c->d_pverts, (size_t)kPlainPoolQuads * 4));
	HCM_HIP(dmalloc(&c->d_pidx, (size_t)kPlainPoolQuads * 6));
	HCM_HIP(dmalloc(&c->d_precs, kPlainRecs));
	HCM_HIP(dmalloc(&c->d_pres, 1));
	HCM_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->h_pblocks), 27 * HCM_CHUNK_VOXELS * 2));
	HCM_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->h_plight), 27 * HCM_CHUNK_VOXELS));
	HCM_HIP(dmalloc(&c->d_halo_in, kHaloBytes));
	HCM_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->h_halo_in), kHaloBytes));
	HCM_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->h_pout), kPoutBytes));
	HCM_HIP(dmalloc(&c->d_pheights, 27 * 256));
	HCM_HIP(hipHostMalloc(reinterpret_cast<void **>(&c->h_pheights), 27 * 256));
	HCM_HIP(dmalloc(&c->d_ppacked, 27 * kLightPlanes * 256));
	HCM_HIP(dmalloc(&c->d_plevels, 27));
	HCM_HIP(dmalloc(&c->d_plctr, 1));
	HCM_HIP(dmalloc(&c->d_phalo, kLightJobBytes));
	uint32_t ident[27];
	for (uint32_t i = 0; i < 27; ++i)
		ident[i] = i;
	HCM_HIP(hipMemcpy(c->d_pnbr, ident, sizeof(ident), hipMemcpyHostToDevice));
	// the no-allocation entry point's scratch on this (current) device, and its constant job table; every
	// context on the device writes the same table
	HCM_HIP(hipGetSymbolAddress(reinterpret_cast<void **>(&c->d_scratch), HIP_SYMBOL(scratch)));
	uint32_t jobs[kNoAllocMaxJobs * 27];
	for (uint32_t i = 0; i < kNoAllocMaxJobs * 27; ++i)
		jobs[i] = i;
	HCM_HIP(hipMemcpy(c->d_scratch + kScratchNbr, jobs, sizeof(jobs), hipMemcpyHostToDevice));
	return HCM_OK;
}

extern "C" int hcm_create(const hcm_registry *registry, uint64_t pool_quads, hcm_ctx **out) {
	if (!out)
		return HCM_ERR_INVALID;
	*out = nullptr;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
		return HCM_ERR_NO_DEVICE;
	hcm_ctx *c = new hcm_ctx;
	int r = create_impl(c, registry, pool_quads);
	if (r != HCM_OK) {
		free_ctx(c);
		return r;
	}
	*out = c;
	return HCM_OK;
}

extern "C" void hcm_destroy(hcm_ctx *ctx) {
	if (ctx && ctx->stream)
		(void)hipStreamSynchronize(ctx->stream);
	free_ctx(ctx);
}

extern "C" int hcm_set_tuning(hcm_ctx *ctx, const hcm_tuning *t) {
	if (!ctx || !t)
		return HCM_ERR_INVALID;
	uint32_t th = t->threads ? t->threads : 256;
	if (th != 128 && th != 256 && th != 512)
		return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(ctx->mu);
	ctx->tuning = *t;
	ctx->tuning.threads = th;
	return HCM_OK;
}

extern "C" int hcm_get_tuning(hcm_ctx *ctx, hcm_tuning *t) {
	if (!ctx || !t)
		return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(ctx->mu);
	*t = ctx->tuning;
	return HCM_OK;
}

static MeshParams base_params(hcm_ctx *c) {
	MeshParams p;
	std::memset(&p, 0, sizeof(p));
	p.reg = c->reg;
	return p;
}

// Occupancy-derived persistent grid: workgroups per CU times CUs.
static int persistent_grid(hcm_ctx *c, uint32_t threads) {
	uint32_t per_cu = c->tuning.wg_per_cu;
	if (!per_cu) { // as many as the CU's LDS holds: 128 blocks of 1,280 bytes
		const int lds = mesh_lds_bytes((int)threads);
		per_cu = lds > 0 ? std::max(1, 128 / ((lds + 1279) / 1280)) : 1;
		per_cu = std::min<uint32_t>(per_cu, 2048 / threads);
	}
	return (int)(per_cu * (uint32_t)c->n_cu);
}

int hcm::ctx_info(hcm_ctx *ctx, uint32_t threads, CtxInfo *out) {
	if (!ctx || !out)
		return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(ctx->mu);
	out->reg = ctx->reg;
	out->n_cu = ctx->n_cu;
	out->threads = threads ? threads : ctx->tuning.threads;
	out->grid = (uint32_t)persistent_grid(ctx, out->threads);
	out->light_info = ctx->d_light_info;
	out->light_info_host = ctx->light_info.data();
	return HCM_OK;
}

// Assemble one job's meshes from a device pool into a caller's host buffers.
static int fetch_job(hipStream_t stream, const hcm_job_result &res, const hcm_mesh_rec *d_recs, const uint2 *d_verts,
                     const uint16_t *d_idx, hcm_host_result *out) {
	if (res.status != HCM_OK)
		return res.status;
	std::vector<hcm_mesh_rec> recs(res.mesh_count);
	if (res.mesh_count)
		HCM_HIP(hipMemcpyAsync(recs.data(), d_recs + res.first_mesh, res.mesh_count * sizeof(hcm_mesh_rec),
		                       hipMemcpyDeviceToHost, stream));
	HCM_HIP(hipStreamSynchronize(stream));
	uint64_t nv = 0, ni = 0;
	for (auto &m : recs)
		nv += 4ull * m.quad_count, ni += 6ull * m.quad_count;
	out->vertex_count = (uint32_t)nv;
	out->index_count = (uint32_t)ni;
	out->mesh_count = res.mesh_count;
	if (nv > out->vertex_capacity || ni > out->index_capacity || res.mesh_count > out->mesh_capacity ||
	    (nv && (!out->vertices || !out->indices)) || (res.mesh_count && !out->meshes))
		return HCM_ERR_CAPACITY;
	uint32_t vb = 0, ib = 0;
	for (uint32_t m = 0; m < res.mesh_count; ++m) {
		const hcm_mesh_rec &r = recs[m];
		hcm_mesh_desc &d = out->meshes[m];
		d.vert_begin = vb, d.vert_count = 4 * r.quad_count;
		d.idx_begin = ib, d.idx_count = 6 * r.quad_count;
		std::memcpy(d.aabb, r.aabb, sizeof(d.aabb));
		d.transparent = r.transparent;
		HCM_HIP(hipMemcpyAsync(out->vertices + vb, d_verts + 4ull * r.first_quad, 4ull * r.quad_count * sizeof(uint2),
		                       hipMemcpyDeviceToHost, stream));
		HCM_HIP(hipMemcpyAsync(out->indices + ib, d_idx + 6ull * r.first_quad, 6ull * r.quad_count * 2,
		                       hipMemcpyDeviceToHost, stream));
		vb += d.vert_count, ib += d.idx_count;
	}
	HCM_HIP(hipStreamSynchronize(stream));
	return HCM_OK;
}

// One chunk from the packed halo in ctx->h_halo_in (context lock held): one upload, the halo scattered into
// the plain path's 27 slots on the device, one meshing workgroup, then the result, the mesh records and
// the first kPlainFastQuads quads copied back together; the meshes of such a chunk are assembled from that
// copy after one synchronisation, a larger chunk's are fetched mesh by mesh.
static int mesh_halo_locked(hcm_ctx *ctx, hcm_host_result *out) {
	hipStream_t st = ctx->stream;
	HCM_HIP(hipMemcpyAsync(ctx->d_halo_in, ctx->h_halo_in, kHaloBytes, hipMemcpyHostToDevice, st));
	HCM_HIP(launch_halo_unpack(reinterpret_cast<const uint16_t *>(ctx->d_halo_in), ctx->d_halo_in + HCM_HALO_BLOCKS * 2,
	                           1, ctx->d_pblocks, ctx->d_plight, st));
	HCM_HIP(launch_reset(ctx->d_pctr, st));
	MeshParams p = base_params(ctx);
	p.blocks = ctx->d_pblocks;
	p.light = ctx->d_plight;
	p.nbr = ctx->d_pnbr;
	p.n_jobs = 1;
	p.job_base = 0;
	p.persistent = 0;
	p.pool_cap = kPlainPoolQuads;
	p.ctr = ctx->d_pctr;
	p.pool_verts = ctx->d_pverts;
	p.pool_idx = ctx->d_pidx;
	p.recs = ctx->d_precs;
	p.rec_cap = kPlainRecs;
	p.results = ctx->d_pres;
	HCM_HIP(launch_mesh(p, kChunkLaunchThreads, 1, st));
	uint8_t *po = ctx->h_pout;
	HCM_HIP(hipMemcpyAsync(po, ctx->d_pres, sizeof(hcm_job_result), hipMemcpyDeviceToHost, st));
	HCM_HIP(hipMemcpyAsync(po + kPoutRecs, ctx->d_precs, kPlainRecs * sizeof(hcm_mesh_rec), hipMemcpyDeviceToHost, st));
	HCM_HIP(hipMemcpyAsync(po + kPoutVerts, ctx->d_pverts, kPlainFastQuads * 4 * sizeof(uint2), hipMemcpyDeviceToHost, st));
	HCM_HIP(hipMemcpyAsync(po + kPoutIdx, ctx->d_pidx, kPlainFastQuads * 6 * 2, hipMemcpyDeviceToHost, st));
	HCM_HIP(hipStreamSynchronize(st));
	hcm_job_result res;
	std::memcpy(&res, po, sizeof(res));
	if (res.status != HCM_OK)
		return res.status;
	const hcm_mesh_rec *recs = reinterpret_cast<const hcm_mesh_rec *>(po + kPoutRecs);
	bool fits = res.first_mesh + res.mesh_count <= kPlainRecs;
	for (uint32_t m = 0; fits && m < res.mesh_count; ++m)
		fits = (uint64_t)recs[res.first_mesh + m].first_quad + recs[res.first_mesh + m].quad_count <= kPlainFastQuads;
	if (!fits)
		return fetch_job(st, res, ctx->d_precs, ctx->d_pverts, ctx->d_pidx, out);
	uint64_t nv = 0, ni = 0;
	for (uint32_t m = 0; m < res.mesh_count; ++m)
		nv += 4ull * recs[res.first_mesh + m].quad_count, ni += 6ull * recs[res.first_mesh + m].quad_count;
	out->vertex_count = (uint32_t)nv;
	out->index_count = (uint32_t)ni;
	out->mesh_count = res.mesh_count;
	if (nv > out->vertex_capacity || ni > out->index_capacity || res.mesh_count > out->mesh_capacity ||
	    (nv && (!out->vertices || !out->indices)) || (res.mesh_count && !out->meshes))
		return HCM_ERR_CAPACITY;
	const uint2 *verts = reinterpret_cast<const uint2 *>(po + kPoutVerts);
	const uint16_t *idx = reinterpret_cast<const uint16_t *>(po + kPoutIdx);
	uint32_t vb = 0, ib = 0;
	for (uint32_t m = 0; m < res.mesh_count; ++m) {
		const hcm_mesh_rec &r = recs[res.first_mesh + m];
		hcm_mesh_desc &d = out->meshes[m];
		d.vert_begin = vb, d.vert_count = 4 * r.quad_count;
		d.idx_begin = ib, d.idx_count = 6 * r.quad_count;
		std::memcpy(d.aabb, r.aabb, sizeof(d.aabb));
		d.transparent = r.transparent;
		std::memcpy(out->vertices + vb, verts + 4ull * r.first_quad, 4ull * r.quad_count * sizeof(uint2));
		std::memcpy(out->indices + ib, idx + 6ull * r.first_quad, 6ull * r.quad_count * 2);
		vb += d.vert_count, ib += d.idx_count;
	}
	return HCM_OK;
}

extern "C" int hcm_mesh_halo(hcm_ctx *ctx, const uint16_t *blocks, const uint8_t *light, hcm_host_result *out) {
	if (!ctx || !blocks || !light || !out)
		return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(ctx->mu);
	std::memcpy(ctx->h_halo_in, blocks, HCM_HALO_BLOCKS * 2);
	std::memcpy(ctx->h_halo_in + HCM_HALO_BLOCKS * 2, light, HCM_HALO_LIGHT);
	return mesh_halo_locked(ctx, out);
}

extern "C" int hcm_mesh_halos_noalloc(hcm_ctx *ctx, const uint16_t *halo_blocks, const uint8_t *halo_light, uint32_t n,
                                      uint32_t (*vertices)[2], uint16_t *indices, uint32_t quad_cap, hcm_mesh_rec *recs,
                                      uint32_t rec_cap, hcm_job_result *results, void *stream) {
	if (!ctx || !ctx->d_scratch)
		return HCM_ERR_INVALID;
	if (n == 0)
		return HCM_OK;
	if (n > kNoAllocMaxJobs || !halo_blocks || !halo_light || !vertices || !indices || !recs || !results ||
	    quad_cap == 0 || rec_cap == 0)
		return HCM_ERR_INVALID;
	hipStream_t st = reinterpret_cast<hipStream_t>(stream);
	Counters *ctr = reinterpret_cast<Counters *>(ctx->d_scratch);
	uint16_t *blocks = reinterpret_cast<uint16_t *>(ctx->d_scratch + kScratchBlocks);
	uint8_t *light = ctx->d_scratch + kScratchLight;
	HCM_HIP(launch_halo_unpack(halo_blocks, halo_light, n, blocks, light, st));
	HCM_HIP(launch_reset(ctr, st)); // a kernel, not a memset, so that a captured graph replays it correctly
	MeshParams p = base_params(ctx);
	p.blocks = blocks;
	p.light = light;
	p.nbr = reinterpret_cast<const uint32_t *>(ctx->d_scratch + kScratchNbr);
	p.n_jobs = n;
	p.job_base = 0;
	p.persistent = 0;
	p.pool_cap = quad_cap;
	p.ctr = ctr;
	p.pool_verts = reinterpret_cast<uint2 *>(vertices);
	p.pool_idx = indices;
	p.recs = recs;
	p.rec_cap = rec_cap;
	p.results = results;
	HCM_HIP(launch_mesh(p, kChunkLaunchThreads, (int)n, st));
	return HCM_OK;
}

extern "C" int hcm_mesh_chunk(hcm_ctx *ctx, const uint16_t *const blocks[27], const uint8_t *const light[27],
                              hcm_host_result *out) {
	if (!ctx || !blocks || !light || !out)
		return HCM_ERR_INVALID;
	for (int i = 0; i < 27; ++i)
		if (!blocks[i] || !light[i])
			return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(ctx->mu);
	// the packed halo, row by row: an x-row spans the chunks at cx = -1, 0, 1 (neighbour-slot order of
	// Position.hpp:138-145: L(-1) = 1, L(0) = 2, L(1) = 0)
	uint16_t *hb = reinterpret_cast<uint16_t *>(ctx->h_halo_in);
	uint8_t *hl = ctx->h_halo_in + HCM_HALO_BLOCKS * 2;
	auto lut = [](int v) { return v < 0 ? 1 : (v >= 16 ? 0 : 2); };
	for (int y = -1; y < 17; ++y)
		for (int z = -2; z < 18; ++z) {
			const int yz = 3 * lut(y) + lut(z);
			const size_t row = 16u * (size_t)((z & 15) + 16 * (y & 15));
			uint16_t *dst = hb + 20 * ((z + 2) + 20 * (y + 1));
			const uint16_t *lo = blocks[9 * 1 + yz] + row, *mid = blocks[9 * 2 + yz] + row, *hi = blocks[yz] + row;
			dst[0] = lo[14], dst[1] = lo[15];
			std::memcpy(dst + 2, mid, 32);
			dst[18] = hi[0], dst[19] = hi[1];
			if (z >= -1 && z < 17) {
				uint8_t *ld = hl + 18 * ((z + 1) + 18 * (y + 1));
				const uint8_t *llo = light[9 * 1 + yz] + row, *lmid = light[9 * 2 + yz] + row, *lhi = light[yz] + row;
				ld[0] = llo[15];
				std::memcpy(ld + 1, lmid, 16);
				ld[17] = lhi[0];
			}
		}
	return mesh_halo_locked(ctx, out);
}

extern "C" int hcm_mesh_chunk_relight(hcm_ctx *ctx, const uint16_t *const blocks[27], const int8_t *const heights[27],
                                      hcm_host_result *out, uint8_t *light_out) {
	if (!ctx || !blocks || !heights || !out)
		return HCM_ERR_INVALID;
	for (int i = 0; i < 27; ++i)
		if (!blocks[i] || !heights[i])
			return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(ctx->mu);
	for (int i = 0; i < 27; ++i) {
		std::memcpy(ctx->h_pblocks + (size_t)i * HCM_CHUNK_VOXELS, blocks[i], HCM_CHUNK_VOXELS * 2);
		std::memcpy(ctx->h_pheights + (size_t)i * 256, heights[i], 256);
	}
	hipStream_t st = ctx->stream;
	HCM_HIP(hipMemcpyAsync(ctx->d_pblocks, ctx->h_pblocks, 27 * HCM_CHUNK_VOXELS * 2, hipMemcpyHostToDevice, st));
	HCM_HIP(hipMemcpyAsync(ctx->d_pheights, ctx->h_pheights, 27 * 256, hipMemcpyHostToDevice, st));
	// light rebuild of the one job (slots 0..26 in neighbour-slot order)
	LightParams lp;
	std::memset(&lp, 0, sizeof(lp));
	lp.blocks = ctx->d_pblocks;
	lp.heights = ctx->d_pheights;
	lp.light_info = ctx->d_light_info;
	lp.slots = ctx->d_pnbr;
	lp.n_slots = 27;
	lp.n_jobs = 1;
	lp.packed = ctx->d_ppacked;
	lp.levels = ctx->d_plevels;
	lp.nbr = ctx->d_pnbr;
	lp.ctr = ctx->d_plctr;
	lp.halo_light = ctx->d_phalo;
	HCM_HIP(launch_reset(ctx->d_plctr, st));
	HCM_HIP(launch_light_prep(lp, st));
	HCM_HIP(launch_light(lp, 1, st));
	// meshing over the rebuilt light
	HCM_HIP(launch_reset(ctx->d_pctr, st));
	MeshParams p = base_params(ctx);
	p.blocks = ctx->d_pblocks;
	p.light = ctx->d_plight;
	p.halo_light = ctx->d_phalo;
	p.nbr = ctx->d_pnbr;
	p.n_jobs = 1;
	p.job_base = 0;
	p.persistent = 0;
	p.pool_cap = kPlainPoolQuads;
	p.ctr = ctx->d_pctr;
	p.pool_verts = ctx->d_pverts;
	p.pool_idx = ctx->d_pidx;
	p.recs = ctx->d_precs;
	p.rec_cap = kPlainRecs;
	p.results = ctx->d_pres;
	HCM_HIP(launch_mesh(p, kChunkLaunchThreads, 1, st));
	hcm_job_result res;
	HCM_HIP(hipMemcpyAsync(&res, ctx->d_pres, sizeof(res), hipMemcpyDeviceToHost, st));
	std::vector<uint8_t> rows;
	if (light_out) {
		rows.resize(kLightJobBytes);
		HCM_HIP(hipMemcpyAsync(rows.data(), ctx->d_phalo, kLightJobBytes, hipMemcpyDeviceToHost, st));
	}
	HCM_HIP(hipStreamSynchronize(st));
	if (light_out)
		for (uint32_t r = 0; r < 18 * 18; ++r) // row hz + 18 hy holds x = -1 .. 16
			std::memcpy(light_out + 18 * r, rows.data() + r * kLightRowBytes, 18);
	return fetch_job(st, res, ctx->d_precs, ctx->d_pverts, ctx->d_pidx, out);
}

// ------------------------------------------------------------------------------------------------
// store
// ------------------------------------------------------------------------------------------------
extern "C" int hcm_store_create(hcm_ctx *ctx, uint32_t n_slots, hcm_store **out) {
	if (!ctx || !out || n_slots == 0)
		return HCM_ERR_INVALID;
	*out = nullptr;
	hcm_store *s = new hcm_store;
	s->ctx = ctx;
	s->n_slots = n_slots;
	if (dmalloc(&s->d_blocks, (size_t)n_slots * HCM_CHUNK_VOXELS) != hipSuccess ||
	    dmalloc(&s->d_light, (size_t)n_slots * HCM_CHUNK_VOXELS) != hipSuccess ||
	    dmalloc(&s->d_heights, (size_t)n_slots * 256) != hipSuccess ||
	    hipMemset(s->d_heights, 0, (size_t)n_slots * 256) != hipSuccess) {
		g_last_hip = (int)hipGetLastError();
		hcm_store_destroy(s);
		return HCM_ERR_HIP;
	}
	*out = s;
	return HCM_OK;
}

extern "C" void hcm_store_destroy(hcm_store *s) {
	if (!s)
		return;
	(void)hipStreamSynchronize(s->ctx->stream);
	dfree(s->d_blocks);
	dfree(s->d_light);
	dfree(s->d_heights);
	dfree(s->d_packed);
	dfree(s->d_levels);
	delete s;
}

extern "C" int hcm_store_upload(hcm_store *s, uint32_t first, uint32_t n, const uint16_t *blocks, const uint8_t *light) {
	if (!s || !blocks || !light || (uint64_t)first + n > s->n_slots)
		return HCM_ERR_INVALID;
	hipStream_t st = s->ctx->stream;
	HCM_HIP(hipMemcpyAsync(s->d_blocks + (size_t)first * HCM_CHUNK_VOXELS, blocks, (size_t)n * HCM_CHUNK_VOXELS * 2,
	                       hipMemcpyHostToDevice, st));
	HCM_HIP(hipMemcpyAsync(s->d_light + (size_t)first * HCM_CHUNK_VOXELS, light, (size_t)n * HCM_CHUNK_VOXELS,
	                       hipMemcpyHostToDevice, st));
	HCM_HIP(hipStreamSynchronize(st));
	return HCM_OK;
}

extern "C" int hcm_store_upload_sunlight(hcm_store *s, uint32_t first, uint32_t n, const int8_t *heights) {
	if (!s || !heights || (uint64_t)first + n > s->n_slots)
		return HCM_ERR_INVALID;
	hipStream_t st = s->ctx->stream;
	HCM_HIP(hipMemcpyAsync(s->d_heights + (size_t)first * 256, heights, (size_t)n * 256, hipMemcpyHostToDevice, st));
	HCM_HIP(hipStreamSynchronize(st));
	return HCM_OK;
}

extern "C" int hcm_store_device_ptrs(hcm_store *s, void **blocks, void **light) {
	if (!s)
		return HCM_ERR_INVALID;
	if (blocks)
		*blocks = s->d_blocks;
	if (light)
		*light = s->d_light;
	return HCM_OK;
}

// ------------------------------------------------------------------------------------------------
// batch
// ------------------------------------------------------------------------------------------------
extern "C" int hcm_batch_create(hcm_ctx *ctx, hcm_store *store, const uint32_t *nbr, uint32_t n_jobs,
                                hcm_batch **out) {
	if (!ctx || !store || !nbr || !out || n_jobs == 0 || store->ctx != ctx)
		return HCM_ERR_INVALID;
	*out = nullptr;
	for (uint64_t i = 0; i < 27ull * n_jobs; ++i)
		if (nbr[i] >= store->n_slots)
			return HCM_ERR_INVALID;
	hcm_batch *b = new hcm_batch;
	b->ctx = ctx;
	b->store = store;
	b->n_jobs = n_jobs;
	b->rec_cap = 3 * n_jobs + 1024;
	auto fail = [&]() {
		g_last_hip = (int)hipGetLastError();
		hcm_batch_destroy(b);
		return HCM_ERR_HIP;
	};
	if (dmalloc(&b->d_nbr, 27ull * n_jobs) != hipSuccess || dmalloc(&b->d_ctr, 1) != hipSuccess ||
	    dmalloc(&b->d_recs, b->rec_cap) != hipSuccess || dmalloc(&b->d_res, n_jobs) != hipSuccess ||
	    hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess)
		return fail();
	if (hipMemcpy(b->d_nbr, nbr, 27ull * n_jobs * 4, hipMemcpyHostToDevice) != hipSuccess ||
	    hipMemset(b->d_ctr, 0, sizeof(Counters)) != hipSuccess)
		return fail();
	b->ctr_clean = true;
	// the distinct store slots the jobs read, for the light rebuild's packing pass
	std::vector<uint32_t> slots(nbr, nbr + 27ull * n_jobs);
	std::sort(slots.begin(), slots.end());
	slots.erase(std::unique(slots.begin(), slots.end()), slots.end());
	b->n_read_slots = (uint32_t)slots.size();
	if (dmalloc(&b->d_read_slots, slots.size()) != hipSuccess ||
	    hipMemcpy(b->d_read_slots, slots.data(), slots.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
		return fail();
	*out = b;
	return HCM_OK;
}

extern "C" int hcm_batch_set_light(hcm_batch *b, uint32_t light_mode) {
	if (!b || (light_mode != HCM_LIGHT_STORE && light_mode != HCM_LIGHT_REBUILD))
		return HCM_ERR_INVALID;
	hcm_ctx *c = b->ctx;
	std::lock_guard<std::mutex> lk(c->mu);
	if (light_mode == HCM_LIGHT_REBUILD) {
		hcm_store *st = b->store;
		if (!st->d_packed) {
			HCM_HIP(dmalloc(&st->d_packed, (size_t)st->n_slots * kLightPlanes * 256));
			HCM_HIP(dmalloc(&st->d_levels, st->n_slots));
		}
		if (!b->d_lctr) { // zero once: each light run's last workgroup clears the queues again
			HCM_HIP(dmalloc(&b->d_lctr, 1));
			HCM_HIP(hipMemset(b->d_lctr, 0, sizeof(Counters)));
		}
		if (!b->d_halo_light)
			HCM_HIP(dmalloc(&b->d_halo_light, (size_t)b->n_jobs * kLightJobBytes));
	}
	b->light_mode = light_mode;
	return HCM_OK;
}

extern "C" int hcm_batch_fetch_light(hcm_batch *b, uint32_t job, uint8_t *halo) {
	if (!b || !halo || job >= b->n_jobs || !b->light_ran)
		return HCM_ERR_INVALID;
	std::vector<uint8_t> rows(kLightJobBytes);
	hipStream_t st = b->ctx->stream;
	HCM_HIP(hipMemcpyAsync(rows.data(), b->d_halo_light + (size_t)job * kLightJobBytes, kLightJobBytes,
	                       hipMemcpyDeviceToHost, st));
	HCM_HIP(hipStreamSynchronize(st));
	for (uint32_t r = 0; r < 18 * 18; ++r) // row hz + 18 hy holds x = -1 .. 16
		std::memcpy(halo + 18 * r, rows.data() + r * kLightRowBytes, 18);
	return HCM_OK;
}

extern "C" void hcm_batch_destroy(hcm_batch *b) {
	if (!b)
		return;
	if (b->ctx && b->ctx->stream)
		(void)hipStreamSynchronize(b->ctx->stream);
	if (b->graph_exec)
		(void)hipGraphExecDestroy(b->graph_exec);
	if (b->graph)
		(void)hipGraphDestroy(b->graph);
	if (b->ev0)
		(void)hipEventDestroy(b->ev0);
	if (b->ev1)
		(void)hipEventDestroy(b->ev1);
	dfree(b->d_nbr);
	dfree(b->d_ctr);
	dfree(b->d_recs);
	dfree(b->d_res);
	dfree(b->d_read_slots);
	dfree(b->d_lctr);
	dfree(b->d_halo_light);
	dfree(b->d_chunk_pos);
	dfree(b->d_infos);
	dfree(b->d_draws);
	dfree(b->d_draw_counts);
	delete b;
}

static MeshParams batch_params(hcm_batch *b) {
	hcm_ctx *c = b->ctx;
	MeshParams p = base_params(c);
	p.blocks = b->store->d_blocks;
	p.light = b->store->d_light;
	p.nbr = b->d_nbr;
	p.n_jobs = b->n_jobs;
	p.pool_cap = (uint32_t)c->pool_quads;
	p.ctr = b->d_ctr;
	p.pool_verts = c->d_verts;
	p.pool_idx = c->d_idx;
	p.recs = b->d_recs;
	p.rec_cap = b->rec_cap;
	p.results = b->d_res;
	p.halo_light = b->light_mode == HCM_LIGHT_REBUILD ? b->d_halo_light : nullptr;
	return p;
}

static LightParams light_params(hcm_batch *b) {
	LightParams lp;
	std::memset(&lp, 0, sizeof(lp));
	lp.blocks = b->store->d_blocks;
	lp.heights = b->store->d_heights;
	lp.light_info = b->ctx->d_light_info;
	lp.slots = b->d_read_slots;
	lp.n_slots = b->n_read_slots;
	lp.n_jobs = b->n_jobs;
	lp.packed = b->store->d_packed;
	lp.levels = b->store->d_levels;
	lp.nbr = b->d_nbr;
	lp.ctr = b->d_lctr;
	lp.halo_light = b->d_halo_light;
	return lp;
}

// Light rebuild of every job of the batch: pack the read slots, then the persistent light kernel.
static hipError_t enqueue_light(hcm_batch *b, hipStream_t st) {
	LightParams lp = light_params(b);
	lp.self_reset = 1;
	hipError_t e = launch_light_prep(lp, st);
	if (e == hipSuccess) {
		const int per_cu = std::max(1, std::min(3, (160 * 1024) / std::max(1, light_lds_bytes())));
		e = launch_light(lp, per_cu * b->ctx->n_cu, st);
	}
	return e;
}

// `times` graph replays back to back between the timing events (persistent runs after the first)
static int batch_run(hcm_batch *b, uint32_t mode, uint32_t times) {
	if (!b || (mode != HCM_RUN_PERSISTENT && mode != HCM_RUN_PER_CHUNK))
		return HCM_ERR_INVALID;
	hcm_ctx *c = b->ctx;
	std::lock_guard<std::mutex> lk(c->mu);
	if (!c->pool_quads)
		return HCM_ERR_POOL_FULL;
	hipStream_t st = c->stream;
	const uint32_t threads = c->tuning.threads;
	MeshParams p = batch_params(b);
	if (mode == HCM_RUN_PERSISTENT && !b->launched) {
		// first persistent run: launched directly, since capturing and instantiating a graph costs more
		// than the launches it replaces; the graph is captured when the batch runs again
		if (!b->ctr_clean)
			HCM_HIP(launch_reset(b->d_ctr, st));
		HCM_HIP(hipEventRecord(b->ev0, st));
		if (b->light_mode == HCM_LIGHT_REBUILD)
			HCM_HIP(enqueue_light(b, st));
		p.persistent = 1;
		p.self_reset = 1;
		HCM_HIP(launch_mesh(p, (int)threads, persistent_grid(c, threads), st));
		HCM_HIP(hipEventRecord(b->ev1, st));
		b->ctr_clean = true;
		b->launched = true;
	} else if (mode == HCM_RUN_PERSISTENT) {
		const uint32_t grid = (uint32_t)persistent_grid(c, threads);
		if (!b->graph_exec || b->graph_threads != threads || b->graph_grid != grid || b->graph_light != b->light_mode) {
			if (b->graph_exec)
				(void)hipGraphExecDestroy(b->graph_exec), b->graph_exec = nullptr;
			if (b->graph)
				(void)hipGraphDestroy(b->graph), b->graph = nullptr;
			p.persistent = 1;
			p.self_reset = 1; // the counters start at zero and the run's last workgroup zeroes them again
			HCM_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
			hipError_t e1 = b->light_mode == HCM_LIGHT_REBUILD ? enqueue_light(b, st) : hipSuccess;
			hipError_t e2 = e1 == hipSuccess ? launch_mesh(p, (int)threads, (int)grid, st) : hipSuccess;
			hipError_t e3 = hipStreamEndCapture(st, &b->graph);
			if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess) {
				g_last_hip = (int)(e1 != hipSuccess ? e1 : (e2 != hipSuccess ? e2 : e3));
				return HCM_ERR_HIP;
			}
			HCM_HIP(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
			b->graph_threads = threads;
			b->graph_grid = grid;
			b->graph_light = b->light_mode;
		}
		if (!b->ctr_clean) // after per-chunk launches
			HCM_HIP(launch_reset(b->d_ctr, st));
		HCM_HIP(hipEventRecord(b->ev0, st));
		for (uint32_t k = 0; k < times; ++k)
			HCM_HIP(hipGraphLaunch(b->graph_exec, st));
		HCM_HIP(hipEventRecord(b->ev1, st));
		b->ctr_clean = true;
	} else {
		HCM_HIP(hipEventRecord(b->ev0, st));
		if (b->light_mode == HCM_LIGHT_REBUILD)
			HCM_HIP(enqueue_light(b, st));
		HCM_HIP(launch_reset(b->d_ctr, st));
		p.persistent = 0;
		for (uint32_t j = 0; j < b->n_jobs; ++j) {
			p.job_base = j;
			HCM_HIP(launch_mesh(p, kChunkLaunchThreads, 1, st));
		}
		HCM_HIP(hipEventRecord(b->ev1, st));
		b->ctr_clean = false;
	}
	b->ran = true;
	b->infos_ready = false;
	b->light_ran = b->light_mode == HCM_LIGHT_REBUILD;
	b->last_mode = mode;
	return HCM_OK;
}

extern "C" int hcm_batch_run(hcm_batch *b, uint32_t mode) { return batch_run(b, mode, 1); }

extern "C" int hcm_batch_run_repeat(hcm_batch *b, uint32_t times) {
	if (!b || times == 0)
		return HCM_ERR_INVALID;
	if (!b->launched) { // the first persistent run is a direct launch
		const int r = batch_run(b, HCM_RUN_PERSISTENT, 1);
		if (r != HCM_OK || --times == 0)
			return r;
	}
	return batch_run(b, HCM_RUN_PERSISTENT, times);
}

extern "C" int hcm_batch_sync(hcm_batch *b, float *kernel_ms) {
	if (!b || !b->ran)
		return HCM_ERR_INVALID;
	HCM_HIP(hipEventSynchronize(b->ev1));
	if (kernel_ms)
		HCM_HIP(hipEventElapsedTime(kernel_ms, b->ev0, b->ev1));
	return HCM_OK;
}

extern "C" int hcm_batch_results(hcm_batch *b, hcm_job_result *results) {
	if (!b || !results || !b->ran)
		return HCM_ERR_INVALID;
	hipStream_t st = b->ctx->stream;
	HCM_HIP(hipMemcpyAsync(results, b->d_res, (size_t)b->n_jobs * sizeof(hcm_job_result), hipMemcpyDeviceToHost, st));
	HCM_HIP(hipStreamSynchronize(st));
	return HCM_OK;
}

// The quads and mesh records of the batch's last run: a persistent run's last workgroup copies them to
// total_quads / total_meshes before it clears the cursors; per-chunk launches leave them in the cursors.
static hipError_t run_totals(hcm_batch *b, uint32_t *quads, uint32_t *meshes) {
	Counters ctr;
	hipStream_t st = b->ctx->stream;
	hipError_t e = hipMemcpyAsync(&ctr, b->d_ctr, sizeof(ctr), hipMemcpyDeviceToHost, st);
	if (e == hipSuccess)
		e = hipStreamSynchronize(st);
	if (e != hipSuccess)
		return e;
	const bool persistent = b->last_mode == HCM_RUN_PERSISTENT;
	*quads = persistent ? ctr.total_quads : ctr.pool_cursor;
	*meshes = persistent ? ctr.total_meshes : ctr.mesh_cursor;
	return hipSuccess;
}

extern "C" int hcm_batch_totals(hcm_batch *b, uint64_t *quads, uint64_t *meshes) {
	if (!b || !b->ran)
		return HCM_ERR_INVALID;
	uint32_t nq, nm;
	HCM_HIP(run_totals(b, &nq, &nm));
	if (quads)
		*quads = nq;
	if (meshes)
		*meshes = nm;
	return HCM_OK;
}

extern "C" int hcm_batch_fetch(hcm_batch *b, uint32_t job, hcm_host_result *out) {
	if (!b || !out || job >= b->n_jobs || !b->ran)
		return HCM_ERR_INVALID;
	hipStream_t st = b->ctx->stream;
	hcm_job_result res;
	HCM_HIP(hipMemcpyAsync(&res, b->d_res + job, sizeof(res), hipMemcpyDeviceToHost, st));
	HCM_HIP(hipStreamSynchronize(st));
	return fetch_job(st, res, b->d_recs, b->ctx->d_verts, b->ctx->d_idx, out);
}

extern "C" int hcm_batch_mesh_infos(hcm_batch *b, const int32_t *chunk_pos, hcm_mesh_info *out, uint32_t cap,
                                    uint32_t *n) {
	if (!b || !chunk_pos || !n || !b->ran)
		return HCM_ERR_INVALID;
	hcm_ctx *c = b->ctx;
	std::lock_guard<std::mutex> lk(c->mu);
	hipStream_t st = c->stream;
	if (!b->d_infos) {
		HCM_HIP(dmalloc(&b->d_chunk_pos, 3ull * b->n_jobs));
		HCM_HIP(dmalloc(&b->d_infos, b->rec_cap));
		HCM_HIP(dmalloc(&b->d_draws, 2ull * b->rec_cap));
		HCM_HIP(dmalloc(&b->d_draw_counts, 2));
	}
	uint32_t total_quads, total_meshes;
	HCM_HIP(run_totals(b, &total_quads, &total_meshes));
	HCM_HIP(hipMemcpyAsync(b->d_chunk_pos, chunk_pos, 12ull * b->n_jobs, hipMemcpyHostToDevice, st));
	const uint32_t nr = std::min<uint32_t>(total_meshes, b->rec_cap);
	HCM_HIP(launch_mesh_infos(b->d_recs, nr, b->d_chunk_pos, b->d_infos, st));
	*n = nr;
	b->n_infos = nr;
	b->infos_ready = true;
	if (out) {
		if (nr > cap)
			return HCM_ERR_CAPACITY;
		HCM_HIP(hipMemcpyAsync(out, b->d_infos, (size_t)nr * sizeof(hcm_mesh_info), hipMemcpyDeviceToHost, st));
	}
	HCM_HIP(hipStreamSynchronize(st));
	return HCM_OK;
}

// Culling of the batch's mesh infos with or without the depth-hierarchy test.
static int cull_batch(hcm_batch *b, const float view_position[3], const float view_projection[16], bool occlusion,
                      hcm_draw_cmd *opaque, uint32_t opaque_cap, uint32_t *n_opaque, hcm_draw_cmd *transparent,
                      uint32_t transparent_cap, uint32_t *n_transparent) {
	if (!b || !view_position || !view_projection || !n_opaque || !n_transparent || !b->infos_ready)
		return HCM_ERR_INVALID;
	hcm_ctx *c = b->ctx;
	std::lock_guard<std::mutex> lk(c->mu);
	if (occlusion && !c->hiz.data)
		return HCM_ERR_INVALID;
	hipStream_t st = c->stream;
	CullParams cp;
	std::memcpy(cp.view_pos, view_position, sizeof(cp.view_pos));
	std::memcpy(cp.view_proj, view_projection, sizeof(cp.view_proj));
	if (occlusion)
		cp.hiz = c->hiz;
	HCM_HIP(hipMemsetAsync(b->d_draw_counts, 0, 2 * sizeof(uint32_t), st));
	HCM_HIP(launch_cull(b->d_infos, b->n_infos, cp, b->d_draws, b->d_draws + b->rec_cap, b->d_draw_counts, st));
	uint32_t counts[2];
	HCM_HIP(hipMemcpyAsync(counts, b->d_draw_counts, sizeof(counts), hipMemcpyDeviceToHost, st));
	HCM_HIP(hipStreamSynchronize(st));
	*n_opaque = counts[0];
	*n_transparent = counts[1];
	if ((opaque && counts[0] > opaque_cap) || (transparent && counts[1] > transparent_cap))
		return HCM_ERR_CAPACITY;
	if (opaque && counts[0])
		HCM_HIP(hipMemcpyAsync(opaque, b->d_draws, counts[0] * sizeof(hcm_draw_cmd), hipMemcpyDeviceToHost, st));
	if (transparent && counts[1])
		HCM_HIP(hipMemcpyAsync(transparent, b->d_draws + b->rec_cap, counts[1] * sizeof(hcm_draw_cmd),
		                       hipMemcpyDeviceToHost, st));
	HCM_HIP(hipStreamSynchronize(st));
	return HCM_OK;
}

extern "C" int hcm_batch_cull(hcm_batch *b, const float view_position[3], const float view_projection[16],
                              hcm_draw_cmd *opaque, uint32_t opaque_cap, uint32_t *n_opaque, hcm_draw_cmd *transparent,
                              uint32_t transparent_cap, uint32_t *n_transparent) {
	return cull_batch(b, view_position, view_projection, false, opaque, opaque_cap, n_opaque, transparent,
	                  transparent_cap, n_transparent);
}

extern "C" int hcm_batch_cull_occlusion(hcm_batch *b, const float view_position[3], const float view_projection[16],
                                        hcm_draw_cmd *opaque, uint32_t opaque_cap, uint32_t *n_opaque,
                                        hcm_draw_cmd *transparent, uint32_t transparent_cap, uint32_t *n_transparent) {
	return cull_batch(b, view_position, view_projection, true, opaque, opaque_cap, n_opaque, transparent,
	                  transparent_cap, n_transparent);
}

// Level count of the depth image's chain: myvk::ImageBase::QueryMipLevel (the bit length of width | height).
static uint32_t depth_levels(uint32_t w, uint32_t h) {
	uint32_t n = 0;
	for (uint32_t x = w | h; x; x >>= 1)
		++n;
	return n;
}

extern "C" int hcm_ctx_set_depth(hcm_ctx *ctx, const float *depth, uint32_t width, uint32_t height) {
	if (!ctx || !depth || width == 0 || height == 0 || width > 16384 || height > 16384 || (width | height) < 2)
		return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(ctx->mu);
	DepthPyramid h;
	h.n_levels = depth_levels(width, height);
	size_t total = 0;
	for (uint32_t i = 0, w = width, hh = height; i < h.n_levels; ++i) {
		h.width[i] = w, h.height[i] = hh, h.offset[i] = (uint32_t)total;
		total += (size_t)w * hh;
		w = std::max(w >> 1, 1u), hh = std::max(hh >> 1, 1u);
	}
	hipStream_t st = ctx->stream;
	if (total > ctx->depth_cap) {
		HCM_HIP(hipStreamSynchronize(st));
		dfree(ctx->d_depth);
		ctx->d_depth = nullptr;
		ctx->depth_cap = 0;
		ctx->hiz = DepthPyramid{};
		HCM_HIP(hipMalloc(&ctx->d_depth, total * sizeof(float)));
		ctx->depth_cap = total;
	}
	HCM_HIP(hipMemcpyAsync(ctx->d_depth, depth, (size_t)width * height * sizeof(float), hipMemcpyHostToDevice, st));
	for (uint32_t i = 1; i < h.n_levels; ++i)
		HCM_HIP(launch_depth_level(ctx->d_depth + h.offset[i - 1], h.width[i - 1], h.height[i - 1],
		                           ctx->d_depth + h.offset[i], h.width[i], h.height[i], st));
	HCM_HIP(hipStreamSynchronize(st));
	h.data = ctx->d_depth;
	ctx->hiz = h;
	return HCM_OK;
}

extern "C" int hcm_ctx_depth_level(hcm_ctx *ctx, uint32_t level, float *out, uint64_t cap, uint32_t *width,
                                   uint32_t *height) {
	if (!ctx || !width || !height)
		return HCM_ERR_INVALID;
	std::lock_guard<std::mutex> lk(ctx->mu);
	if (!ctx->hiz.data || level >= ctx->hiz.n_levels)
		return HCM_ERR_INVALID;
	*width = ctx->hiz.width[level];
	*height = ctx->hiz.height[level];
	const uint64_t n = (uint64_t)*width * *height;
	if (!out)
		return HCM_OK;
	if (n > cap)
		return HCM_ERR_CAPACITY;
	HCM_HIP(hipMemcpyAsync(out, ctx->hiz.data + ctx->hiz.offset[level], n * sizeof(float), hipMemcpyDeviceToHost,
	                       ctx->stream));
	HCM_HIP(hipStreamSynchronize(ctx->stream));
	return HCM_OK;
}

extern "C" int hcm_batch_download(hcm_batch *b, hcm_job_result *results, hcm_mesh_rec *recs, uint32_t rec_cap,
                                  uint32_t (*vertices)[2], uint16_t *indices, uint64_t quad_cap) {
	if (!b || !b->ran)
		return HCM_ERR_INVALID;
	hipStream_t st = b->ctx->stream;
	uint32_t total_quads, total_meshes;
	HCM_HIP(run_totals(b, &total_quads, &total_meshes));
	const uint64_t nq = std::min<uint64_t>(total_quads, b->ctx->pool_quads);
	const uint32_t nr = std::min<uint32_t>(total_meshes, b->rec_cap);
	if (results)
		HCM_HIP(hipMemcpyAsync(results, b->d_res, (size_t)b->n_jobs * sizeof(hcm_job_result), hipMemcpyDeviceToHost,
		                       st));
	if (recs) {
		if (nr > rec_cap)
			return HCM_ERR_CAPACITY;
		if (nr)
			HCM_HIP(hipMemcpyAsync(recs, b->d_recs, (size_t)nr * sizeof(hcm_mesh_rec), hipMemcpyDeviceToHost, st));
	}
	if (vertices || indices) {
		if (nq > quad_cap)
			return HCM_ERR_CAPACITY;
		if (nq && vertices)
			HCM_HIP(hipMemcpyAsync(vertices, b->ctx->d_verts, (size_t)nq * 4 * sizeof(uint2), hipMemcpyDeviceToHost, st));
		if (nq && indices)
			HCM_HIP(hipMemcpyAsync(indices, b->ctx->d_idx, (size_t)nq * 12, hipMemcpyDeviceToHost, st));
	}
	HCM_HIP(hipStreamSynchronize(st));
	return HCM_OK;
}
