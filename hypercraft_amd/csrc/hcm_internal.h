// hcm_internal.h -- types shared by the gfx950 kernels (hcm_kernels.hip) and the host API (hcm_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hcm.h"

namespace hcm {

constexpr int kMaxPid = 128;   // registry rows + dynamic liquid-texture rows kept in LDS
constexpr int kMaxCodes = 128; // distinct face textures kept in LDS
constexpr int kFlowPatterns = 17;

// Row word y layout: tc4 | tc5 << 8 | flags << 16 | mesh << 24 (mesh 0xFF = none)
constexpr uint32_t kRowDirectPass = 1u;   // VerticalLightPass || collision != solid (BlockMeshAlgo.hpp:119)
constexpr uint32_t kRowIndirectPass = 2u; // transparency != opaque (Block.hpp:70)
constexpr uint32_t kRowHasMesh = 4u;
constexpr uint32_t kRowLiquidFlow = 8u;

// Texture-code flags (code word = texture value | flags << 16)
constexpr uint32_t kCodeTransparent = 1u;
constexpr uint32_t kCodeTransPass = 2u;
constexpr uint32_t kCodeLiquid = 4u;
constexpr uint32_t kCodeEmpty = 8u;

struct DevReg {
	const uint8_t *pid_of;         // [65536] Block::m_data -> pid
	const uint2 *rows;             // [n_pid]: x = tc0..tc3, y = tc4 | tc5 << 8 | flags << 16 | mesh << 24
	const uint32_t *codes;         // [n_codes]
	const hcm_custom_mesh *meshes; // [n_meshes]
	const uint8_t *tex_flags;      // [n_tex] by texture id
	uint32_t n_pid, n_codes, n_tex, flow_pid_base;
	uint32_t flow_tex, src_id, flow_id, pad;
};

struct Counters {
	uint32_t queue;       // persistent work-queue head
	uint32_t pool_cursor; // quads allocated in the pool
	uint32_t mesh_cursor; // mesh records allocated
	uint32_t pad;
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
	uint32_t pad;
	hcm_job_result *results;
};

// Launchers (hcm_kernels.hip)
hipError_t launch_reset(Counters *ctr, hipStream_t stream);
hipError_t launch_mesh(const MeshParams &p, int threads, int grid, hipStream_t stream);
int mesh_lds_bytes(int threads);

} // namespace hcm
