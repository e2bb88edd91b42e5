// hcm_internal.h -- types shared by the gfx950 kernels (hcm_kernels.hip) and the host API (hcm_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hcm.h"

namespace hcm {

constexpr int kMaxPid = 128;   // registry rows + dynamic liquid-texture rows kept in LDS
constexpr int kMaxCodes = 128; // distinct face textures kept in LDS
constexpr int kFlowPatterns = 17;

// Voxel word staged in LDS per halo voxel: sunlight in bits 0-3 and torchlight in bits 16-19 (the light
// byte's low and high nibbles), and the per-row template below. With the indirect-pass bit at 12, a cell's
// contribution to a light4_init corner sum is the word masked to kVoxLight when it passes light: adding
// such words sums sunlight in bits 0-5, the count of cells in bits 12-13 and torchlight in bits 16-21.
constexpr int kVoxPidShift = 4;            // texture row, 7 bits
constexpr uint32_t kVoxPidMask = 0x7fu;
constexpr uint32_t kVoxDirect = 1u << 11;   // VerticalLightPass || collision != solid (BlockMeshAlgo.hpp:119)
constexpr uint32_t kVoxIndirect = 1u << 12; // transparency != opaque (block/include/block/Block.hpp:70)
constexpr uint32_t kVoxMesh = 1u << 13;     // has a custom mesh
constexpr uint32_t kVoxFlow = 1u << 14;     // liquid flow (dynamic mesh and face textures)
constexpr int kVoxClassShift = 20;          // 2-bit face class per BlockFace f at bit 20 + 2 f
constexpr uint32_t kVoxSunTorch = 0x000f000fu;
constexpr uint32_t kVoxLight = kVoxSunTorch | kVoxIndirect;
// The light byte (sun | torch << 4) moved to the word's light fields.
__host__ __device__ constexpr uint32_t vox_light(uint32_t byte) { return (byte * 0x1001u) & kVoxSunTorch; }
// Face classes for BlockTexture::Show (texture/include/texture/BlockTexture.hpp:68-76)
constexpr uint32_t kClassEmpty = 0, kClassOpaque = 1, kClassTransparent = 2, kClassLiquid = 3;

// Texture-code flags (code word = texture value | flags << 16)
constexpr uint32_t kCodeTransparent = 1u;
constexpr uint32_t kCodeTransPass = 2u;
constexpr uint32_t kCodeLiquid = 4u;
constexpr uint32_t kCodeEmpty = 8u;

// Texture-row flag (in rows[pid].y): a halo made of this row's block alone has visible faces.
constexpr uint32_t kRowSelfShow = 1u << 16;

struct DevReg {
	const uint8_t *pid_of;         // [65536] Block::m_data -> texture row (pid); then [256] by block id:
	                               // the row when the meta byte does not change it, else 0xff
	const uint2 *rows;             // [n_pid]: x = tc0..tc3 (8 bits each),
	                               // y = tc4 | tc5 << 8 | kRowSelfShow | mesh << 24
	const uint32_t *pid_vox;       // [n_pid]: voxel-word template (everything but the light fields)
	const uint32_t *codes;         // [n_codes]
	const hcm_custom_mesh *meshes; // [n_meshes]
	const uint8_t *tex_flags;      // [n_tex] by texture id
	uint32_t n_pid, n_codes, n_tex, flow_pid_base;
	uint32_t flow_tex, src_id, flow_id, pad;
};

// The persistent kernel's job list is cut into kQueues contiguous parts, one per XCD (gfx950 has 8,
// and workgroups are dispatched to them round-robin by workgroup id). A workgroup takes jobs from its
// own XCD's part first, so neighbouring chunks, which share halo data, are read through one L2; then it
// takes jobs from the other parts. Each counter sits on its own 256-byte line: all are hit by
// device-scope atomics, and atomics on one line are serialised where that line lives.
constexpr uint32_t kQueues = 8;

struct Counters {
	struct alignas(256) Line {
		uint32_t v;
	};
	Line queue[kQueues];               // jobs taken from each part of the job list
	alignas(256) uint32_t pool_cursor; // quads allocated in the pool
	alignas(256) uint32_t mesh_cursor; // mesh records allocated
	// self-resetting persistent runs: workgroups done; the last one copies the run's totals below and
	// clears every counter for the next run
	alignas(256) uint32_t done;
	uint32_t total_quads, total_meshes;
};

struct MeshParams {
	DevReg reg;
	const uint16_t *blocks; // store: slot * 4096 + local
	const uint8_t *light;
	const uint32_t *nbr;    // 27 slots per job
	uint32_t n_jobs;
	uint32_t job_base; // per-chunk launches: job = job_base + blockIdx.x
	uint32_t persistent;
	uint32_t pool_cap; // quads
	Counters *ctr;
	uint2 *pool_verts;   // 4 vertices per quad
	uint16_t *pool_idx;  // 6 indices per quad
	hcm_mesh_rec *recs;
	uint32_t rec_cap;
	uint32_t self_reset; // persistent: the last workgroup to finish clears the counters (no reset kernel)
	hcm_job_result *results;
	// [n_jobs][kLightJobBytes]: each job's rebuilt light over its halo, or null: light from the store
	const uint8_t *halo_light;
};

// ---- light rebuild before meshing (client/src/ChunkMeshTask.cpp:48-111) ----
// Per store slot, the inputs of the light rebuild packed as 16-bit x-rows (row z + 16 y, bit x):
// plane 0 the indirect light pass, plane 1 sunlight, planes 2..5 the bits of the torchlight level,
// plane 6 the cells of torchlight level 15.
constexpr int kLightPlanes = 7;
constexpr uint32_t kLightRowBytes = 20;                        // one halo x-row of rebuilt light (18 used)
constexpr uint32_t kLightJobBytes = 18 * 18 * kLightRowBytes; // rows hz + 18 hy
constexpr uint8_t kLightInfoPass = 0x10;     // light_info: level | kLightInfoPass when the block passes light
constexpr uint8_t kLightInfoVertical = 0x20; // ... | kLightInfoVertical when it passes sunlight straight down

struct LightParams {
	const uint16_t *blocks;    // store: slot * 4096 + local
	const int8_t *heights;     // store: slot * 256 + x + 16 z
	const uint8_t *light_info; // [65536] by Block::m_data
	const uint32_t *slots;     // the store slots the batch reads
	uint32_t n_slots;
	uint32_t n_jobs;
	uint16_t *packed;          // [store slots][kLightPlanes][256]
	uint32_t *levels;          // [store slots]: bit l set when a voxel emits torchlight level l
	const uint32_t *nbr;       // 27 slots per job
	Counters *ctr;             // work queues of the persistent light kernel
	uint8_t *halo_light;       // [n_jobs][kLightJobBytes]
	uint32_t self_reset;       // the last workgroup to finish clears the work queues (no reset kernel)
};

// One voxel edit of a device chunk store: slot, chunk-local index, block, light. local >= 4096 sets the
// slot's sunlight height local - 4096 (index x + 16 z) to the low byte of block instead.
struct VoxelEdit {
	uint32_t slot, local, block, light;
};

// Launchers (hcm_kernels.hip)
hipError_t launch_reset(Counters *ctr, hipStream_t stream);
hipError_t launch_mesh(const MeshParams &p, int threads, int grid, hipStream_t stream);
hipError_t launch_halo_unpack(const uint16_t *hblocks, const uint8_t *hlight, uint32_t n, uint16_t *blocks,
                              uint8_t *light, hipStream_t stream);
// staged chunk k (sblocks + 4096 k, and slight + 4096 k or sheights + 256 k; the unused one null) to store
// slot moves[i].y for k = moves[i].x, i < n; all arrays 16-byte aligned
hipError_t launch_upload_scatter(const uint16_t *sblocks, const uint8_t *slight, const int8_t *sheights,
                                 const uint2 *moves, uint32_t n, uint16_t *blocks, uint8_t *light, int8_t *heights,
                                 hipStream_t stream);
hipError_t launch_edits(uint16_t *blocks, uint8_t *light, int8_t *heights, const VoxelEdit *edits, uint32_t n,
                        hipStream_t stream);
int mesh_lds_bytes(int threads);
// Launchers (hcm_draw.hip)
// Depth hierarchy of the renderer's depth image (client/include/client/rg/DepthHierarchyPass.hpp:166-196):
// level 0 is the image, level i >= 1 is max(level i - 1 size >> 1, 1) per axis; the culling shader's
// texture holds levels 1 .. n_levels - 1 as its mips 0 .. n_levels - 2.
constexpr int kMaxDepthLevels = 16; // images up to 16384 x 16384 have at most 15 levels
struct DepthPyramid {
	const float *data = nullptr; // every level, level i at data + offset[i], row y at + y * width[i]
	uint32_t n_levels = 0;
	uint32_t width[kMaxDepthLevels] = {}, height[kMaxDepthLevels] = {}, offset[kMaxDepthLevels] = {};
};
struct CullParams {
	float view_pos[3];
	float view_proj[16]; // column-major
	DepthPyramid hiz;    // data == nullptr: frustum test only
};
hipError_t launch_depth_level(const float *prev, uint32_t pw, uint32_t ph, float *cur, uint32_t cw, uint32_t ch,
                              hipStream_t stream);
hipError_t launch_mesh_infos(const hcm_mesh_rec *recs, uint32_t n, const int32_t *chunk_pos, hcm_mesh_info *out,
                             hipStream_t stream);
hipError_t launch_cull(const hcm_mesh_info *infos, uint32_t n, const CullParams &cp, hcm_draw_cmd *opaque,
                       hcm_draw_cmd *transparent, uint32_t *counts, hipStream_t stream);
// Launchers (hcm_light.hip)
hipError_t launch_light_prep(const LightParams &p, hipStream_t stream);
hipError_t launch_light(const LightParams &p, int grid, hipStream_t stream);
int light_lds_bytes();

// Context internals the serving loop needs (hcm_api.cpp)
struct CtxInfo {
	DevReg reg;
	int n_cu;
	uint32_t threads;
	uint32_t grid;              // persistent grid for `threads`
	const uint8_t *light_info;  // [65536] for the light rebuild
	const uint8_t *light_info_host; // the same table in host memory
};
int ctx_info(hcm_ctx *ctx, uint32_t threads, CtxInfo *out);

} // namespace hcm
