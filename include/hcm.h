/*
 * hcm.h -- C ABI of the MI355X (gfx950) chunk mesher for HyperCraft.
 *
 * Drop-in for the client's chunk-meshing call
 *     BlockMeshAlgo<BlockAlgoConfig<InnerPos1, {0,0,0,S,S,S}, kBlockAlgoSwizzleYZX>>{}.Generate(get_block, get_light)
 * at client/src/ChunkMeshTask.cpp:113-122 (client/include/client/BlockMeshAlgo.hpp:494-521), for S = 16.
 *
 * Conventions shared by every entry point:
 *   - A chunk is 16^3 voxels indexed x + (y*16 + z)*16 (common/include/common/Position.hpp:98-100).
 *   - A block is Block::m_data: id | meta << 8 (block/include/block/Block.hpp:24-33).
 *   - A light byte is sunlight | torchlight << 4 (block/include/block/Light.hpp:18-23).
 *   - The 27 chunks around a meshed chunk are passed in neighbour-slot order
 *     slot = 9*L[cx] + 3*L[cy] + L[cz] with L = {-1 -> 1, 0 -> 2, +1 -> 0}; the chunk itself is slot 26
 *     (Position.hpp:138-145, client/src/ChunkMeshTask.cpp:19-34).
 *   - Output vertices are the reference's packed BlockVertex words
 *     (client/include/client/BlockVertex.hpp:8-18), indices are u16 relative to their mesh, and meshes
 *     come in the reference's order: split meshes first, then opaque, then transparent
 *     (BlockMeshAlgo.hpp:515-520).
 *   - Every call returns an int status (HCM_OK or a negative HCM_ERR_*); nothing throws across the ABI.
 *     The reference has no error channel; statuses replace its silent behaviours.
 */
#ifndef HCM_H
#define HCM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HCM_OK 0
#define HCM_ERR_INVALID (-1)     /* bad argument or registry */
#define HCM_ERR_HIP (-2)         /* a HIP runtime call failed; hcm_last_hip_error() has the code */
#define HCM_ERR_POOL_FULL (-3)   /* the device quad pool cannot hold the output (MeshHandle.hpp:65 drops it) */
#define HCM_ERR_CAPACITY (-4)    /* caller's host buffers too small; the needed counts were written */
#define HCM_ERR_NO_DEVICE (-5)   /* no gfx950 device, or the library was built without one */
#define HCM_ERR_BUSY (-6)        /* serving queue full (back-pressure); retry later */

#define HCM_CHUNK 16
#define HCM_CHUNK_VOXELS 4096
#define HCM_MAX_FACES 16 /* kBlockMeshMaxFaceCount, block/include/block/BlockMesh.hpp:37 */

/* ------------------------------------------------------------------------------------------------- */
/* Block registry: the reference's 65,536-entry BlockProperty table (block/generator/Block.cpp.in:30-32)
 * flattened to distinct rows, with texture flags (texture/internal/generated/block_texture_transparency.inl)
 * and custom meshes (block/include/block/BlockMesh.hpp:15-48).                                       */
/* ------------------------------------------------------------------------------------------------- */
typedef struct hcm_mesh_vertex {
	uint8_t x, y, z; /* 1/16 block units */
	uint8_t ao;      /* 4 = interpolated (BlockMesh.hpp:20) */
} hcm_mesh_vertex;

typedef struct hcm_mesh_face { /* BlockMeshFace, BlockMesh.hpp:21-26 */
	uint8_t axis, light_face, render_face, pad;
	uint16_t texture, pad2; /* BlockTexture::m_data = id | transformation << 13 */
	hcm_mesh_vertex v[4];
} hcm_mesh_face;

#define HCM_MESH_STATIC 0u
#define HCM_MESH_LIQUID_FLOW 1u /* flow_mesh_gen, block/block/public/Liquid.hpp:156-302 */

typedef struct hcm_custom_mesh {
	uint32_t face_count;
	uint32_t kind;        /* HCM_MESH_STATIC or HCM_MESH_LIQUID_FLOW */
	uint32_t flow_height; /* liquid flow: get_flow_height(Variant) (Liquid.hpp:151-154) */
	uint32_t reserved;
	hcm_mesh_face faces[HCM_MAX_FACES];
} hcm_custom_mesh;

typedef struct hcm_block_row { /* the BlockProperty fields the mesher reads (BlockProperty.hpp:17-29) */
	uint16_t tex[6];      /* textures by BlockFace (right, left, top, bottom, front, back) */
	uint8_t transparency; /* 0 opaque, 1 semi-transparent, 2 transparent */
	uint8_t collision;    /* 0 none, 1 solid, 2 water */
	uint16_t mesh;        /* index into meshes; 0xFFFF = none */
} hcm_block_row;

#define HCM_TEX_TRANSPARENT 1u /* BlockTexture::IsTransparent */
#define HCM_TEX_TRANS_PASS 2u  /* BlockTexture::UseTransparentPass */
#define HCM_TEX_LIQUID 4u      /* BlockTexture::IsLiquid */
#define HCM_NO_MESH 0xFFFFu

typedef struct hcm_registry {
	uint32_t n_rows, n_meshes, n_tex;
	const uint16_t *row_of; /* [65536]: Block::m_data -> row */
	const hcm_block_row *rows;
	const hcm_custom_mesh *meshes;
	const uint8_t *tex_flags; /* [n_tex], by texture id */
	uint8_t liquid_source_id, liquid_flow_id; /* WaterTrait::kSourceID / kFlowID (Water.hpp:6) */
	uint16_t liquid_flow_tex;                 /* WaterTrait::kFlowTex */
	/* Block::GetLightLevel by Block::m_data (block/include/block/Block.hpp:71), [65536]; NULL: no block
	 * emits light. Read only by the light rebuild before meshing (HCM_LIGHT_REBUILD). */
	const uint8_t *light_level;
} hcm_registry;

/* The library's built-in copy of the reference registry (the block headers under block/block/,
 * block/register/blocks, texture/register/block_textures). A client that already has hc::block::Block
 * can instead fill an hcm_registry from its own property table. */
const hcm_registry *hcm_registry_default(void);

/* ------------------------------------------------------------------------------------------------- */
/* Context                                                                                           */
/* ------------------------------------------------------------------------------------------------- */
typedef struct hcm_ctx hcm_ctx;

/* Replaces constructing BlockMeshAlgo (one per worker thread, ChunkTaskPool.hpp:193): uploads the
 * registry to the current HIP device and allocates a device quad pool of pool_quads quads
 * (44 bytes each: 4 vertices + 6 indices). The registry is copied. */
int hcm_create(const hcm_registry *registry, uint64_t pool_quads, hcm_ctx **out);
void hcm_destroy(hcm_ctx *ctx);
int hcm_last_hip_error(void);
/* 1 if a gfx950 device is visible and the kernels are loadable, else 0. */
int hcm_device_ok(void);

/* ------------------------------------------------------------------------------------------------- */
/* Plain path: one chunk per call, host buffers (BlockMeshAlgo::Generate)                             */
/* ------------------------------------------------------------------------------------------------- */
typedef struct hcm_mesh_desc { /* one BlockMesh (BlockVertex.hpp:20-25) inside hcm_host_result */
	uint32_t vert_begin, vert_count; /* into vertices (units of vertices) */
	uint32_t idx_begin, idx_count;   /* into indices */
	uint32_t aabb[6];                /* AABB<uint32_t> min xyz, max xyz in 1/16 units (dep/AABB.hpp) */
	uint32_t transparent;
} hcm_mesh_desc;

typedef struct hcm_host_result {
	uint32_t (*vertices)[2]; /* caller-owned; 2 words per vertex */
	uint32_t vertex_capacity;
	uint16_t *indices;
	uint32_t index_capacity;
	hcm_mesh_desc *meshes;
	uint32_t mesh_capacity;
	uint32_t vertex_count, index_count, mesh_count; /* written; on HCM_ERR_CAPACITY: the sizes needed */
} hcm_host_result;

/* Meshes the chunk in slot 26 of blocks/light (27 host arrays of 4096 voxels each), like
 * Generate(get_block, get_light) over the 27-chunk lambdas of ChunkMeshTask.cpp:117-122: gathers the
 * chunk's packed halo from them on the host and meshes it as hcm_mesh_halo does.
 * Thread-safe per context (calls are serialised on the context's stream). */
int hcm_mesh_chunk(hcm_ctx *ctx, const uint16_t *const blocks[27], const uint8_t *const light[27],
                   hcm_host_result *out);

/* Packed halo of one chunk: exactly the voxels the mesher reads around the 16^3 chunk. Blocks over
 * x, z in [-2, 18) and y in [-1, 17) -- the 18^3 halo plus the x/z ring that flowing water reads around
 * halo cells (BlockMeshAlgo.hpp:202-236) -- at index (x+2) + 20 ((z+2) + 20 (y+1)); light over the 18^3
 * halo [-1, 17)^3 at index (x+1) + 18 ((z+1) + 18 (y+1)). 14,400 + 5,832 bytes per chunk. */
#define HCM_HALO_BLOCKS (20 * 20 * 18)
#define HCM_HALO_LIGHT (18 * 18 * 18)

/* Meshes one chunk from its packed halo, like Generate(get_block, get_light)
 * (client/include/client/BlockMeshAlgo.hpp:494-521) at the reference's call site
 * (client/src/ChunkMeshTask.cpp:113-122), whose lambdas are evaluated exactly over these voxels. The
 * halo (20 KB) is the only upload; the results of a chunk of up to 2,048 quads come back with one
 * synchronisation. Same output and errors as hcm_mesh_chunk. */
int hcm_mesh_halo(hcm_ctx *ctx, const uint16_t *blocks, const uint8_t *light, hcm_host_result *out);

/* The whole mesh task of ChunkTaskRunner<kMesh>::Run (client/src/ChunkMeshTask.cpp:41-122): rebuilds light
 * from the 27 chunks' blocks and sunlight heights (int8[256] each, Chunk::m_sunlight_heights, index
 * x + 16 z), then meshes over it. When light_out is not NULL it receives the rebuilt light over the 18^3
 * cells [-1, 17)^3 the mesher reads, index (x+1) + 18 ((z+1) + 18 (y+1)). */
int hcm_mesh_chunk_relight(hcm_ctx *ctx, const uint16_t *const blocks[27], const int8_t *const heights[27],
                           hcm_host_result *out, uint8_t *light_out);

/* ------------------------------------------------------------------------------------------------- */
/* Device-resident chunk store and batched meshing (no counterpart in the reference, which meshes one
 * chunk per task on CPU threads; this is the batched form of the same call).                        */
/* ------------------------------------------------------------------------------------------------- */
typedef struct hcm_store hcm_store;

/* n_slots chunk slots of 4096 u16 blocks + 4096 u8 light bytes in HBM. */
int hcm_store_create(hcm_ctx *ctx, uint32_t n_slots, hcm_store **out);
void hcm_store_destroy(hcm_store *store);
/* Copies n chunks (n*4096 blocks, n*4096 light bytes, host memory) into slots [first, first+n). */
int hcm_store_upload(hcm_store *store, uint32_t first, uint32_t n, const uint16_t *blocks, const uint8_t *light);
/* Device pointers of the slot arrays (blocks: n_slots*4096 u16, light: n_slots*4096 u8). */
int hcm_store_device_ptrs(hcm_store *store, void **blocks, void **light);
/* Copies the sunlight heights of n chunks (int8[n][256], host memory) into slots [first, first+n):
 * Chunk::m_sunlight_heights (client/include/client/Chunk.hpp:65-76, 96), index x + 16 z; a voxel is
 * sunlit when its chunk-local y >= the height. A new store holds 0 everywhere, Chunk's default (every
 * voxel sunlit). Read only by the light rebuild (HCM_LIGHT_REBUILD). */
int hcm_store_upload_sunlight(hcm_store *store, uint32_t first, uint32_t n, const int8_t *heights);

typedef struct hcm_mesh_rec { /* device-side mesh record; 40 bytes (cf. ChunkMeshInfo, ChunkMesh.hpp:12-17) */
	uint32_t aabb[6];
	uint32_t first_quad; /* vertices at pool_vertices[4*first_quad], indices at pool_indices[6*first_quad] */
	uint32_t quad_count;
	uint32_t job;
	uint32_t transparent;
} hcm_mesh_rec;

typedef struct hcm_job_result {
	int32_t status;      /* HCM_OK or HCM_ERR_POOL_FULL */
	uint32_t first_mesh; /* into the batch's mesh records */
	uint32_t mesh_count;
	uint32_t quad_count;
} hcm_job_result;

/* No counterpart in the reference. For callers that cannot allocate -- code under HIP graph capture,
 * tight serving loops -- the form of hcm_mesh_halo that allocates nothing and does not synchronise: n
 * packed halos already in device memory (halo j at halo_blocks + j HCM_HALO_BLOCKS and halo_light +
 * j HCM_HALO_LIGHT) are meshed on `stream` (a hipStream_t; NULL: the default stream) into the caller's
 * device buffers: quads in vertices[4 q ..]/indices[6 q ..] for q < quad_cap, mesh records in
 * recs[0 .. rec_cap) (first_quad into those buffers), per-halo results in results[0 .. n) -- the
 * hcm_batch_download layout. The kernels' other memory (work counters, job table, the 27-slot stores the
 * halos are scattered into) is a fixed device array of the library, one per GPU, held until the process
 * ends: n is at most HCM_NOALLOC_MAX_CHUNKS (HCM_ERR_INVALID beyond). Calls that overlap on one GPU share
 * that array and race; calls on one stream do not. A halo whose output does not fit gets
 * HCM_ERR_POOL_FULL in its result. */
#define HCM_NOALLOC_MAX_CHUNKS 8
int hcm_mesh_halos_noalloc(hcm_ctx *ctx, const uint16_t *halo_blocks, const uint8_t *halo_light, uint32_t n,
                           uint32_t (*vertices)[2], uint16_t *indices, uint32_t quad_cap, hcm_mesh_rec *recs,
                           uint32_t rec_cap, hcm_job_result *results, void *stream);

typedef struct hcm_batch hcm_batch;

#define HCM_RUN_PERSISTENT 0u /* one persistent launch with a work queue, replayed as a HIP graph */
#define HCM_RUN_PER_CHUNK 1u  /* one kernel launch per chunk (the plain per-chunk path) */

/* A batch of n_jobs chunk jobs over a store: nbr holds 27 slot indices per job (host array, copied
 * to the device). Output meshes go to the context's quad pool. */
int hcm_batch_create(hcm_ctx *ctx, hcm_store *store, const uint32_t *nbr, uint32_t n_jobs, hcm_batch **out);
void hcm_batch_destroy(hcm_batch *batch);
/* Enqueues one meshing pass over every job (resets the pool first). Asynchronous. */
int hcm_batch_run(hcm_batch *batch, uint32_t mode);
/* Enqueues `times` persistent passes back to back, as many calls of hcm_batch_run(batch,
 * HCM_RUN_PERSISTENT) would (each pass's output replaces the previous one's); hcm_batch_sync then gives
 * the HIP-event time of all of them but a batch's first pass, which is launched on its own. A throughput
 * form for repeated meshing of resident chunks; no counterpart in the reference. Asynchronous. */
int hcm_batch_run_repeat(hcm_batch *batch, uint32_t times);
/* Waits for the last run; *kernel_ms receives the HIP-event time of its kernels (with HCM_LIGHT_REBUILD
 * the light rebuild and the meshing). */
int hcm_batch_sync(hcm_batch *batch, float *kernel_ms);
/* Per-job results of the last run (host array of n_jobs). */
int hcm_batch_results(hcm_batch *batch, hcm_job_result *results);
/* Copies one job's meshes from the device pool into host buffers (same layout as hcm_mesh_chunk). */
int hcm_batch_fetch(hcm_batch *batch, uint32_t job, hcm_host_result *out);
/* Totals of the last run: quads and mesh records written. */
int hcm_batch_totals(hcm_batch *batch, uint64_t *quads, uint64_t *meshes);
/* Copies everything the last run produced: per-job results (n_jobs), mesh records (up to rec_cap) and
 * the used part of the quad pool (up to quad_cap quads: 4 vertices and 6 indices each). */
int hcm_batch_download(hcm_batch *batch, hcm_job_result *results, hcm_mesh_rec *recs, uint32_t rec_cap,
                       uint32_t (*vertices)[2], uint16_t *indices, uint64_t quad_cap);

/* Where a batch's light comes from. The reference's mesh task rebuilds light before every mesh
 * (client/src/ChunkMeshTask.cpp:48-111): torchlight from the blocks' light levels, sunlight from the
 * chunks' sunlight heights, then propagation through blocks that pass light indirectly, over the
 * chunk's [-15, 31)^3 neighbourhood. */
#define HCM_LIGHT_STORE 0u   /* mesh with the light bytes held in the store (the default) */
#define HCM_LIGHT_REBUILD 1u /* rebuild each job's light from blocks and sunlight heights, then mesh */
int hcm_batch_set_light(hcm_batch *batch, uint32_t light_mode);
/* After a run with HCM_LIGHT_REBUILD: one job's rebuilt light over the 18^3 cells [-1, 17)^3 the mesher
 * reads, index (x+1) + 18 ((z+1) + 18 (y+1)), light bytes sun | torch << 4 (block/include/block/Light.hpp). */
int hcm_batch_fetch_light(hcm_batch *batch, uint32_t job, uint8_t *halo);

/* Draw lists for a renderer. WorldRenderer::PushChunkMesh (client/include/client/WorldRenderer.hpp:69-87)
 * gives every BlockMesh a ChunkMeshInfo (client/include/client/ChunkMesh.hpp:12-17), and the culling pass
 * (client/shader/src/chunk_culling.comp) turns the visible ones into indexed indirect draws. */
typedef struct hcm_mesh_info { /* chunk_culling.comp:8-13 (std430, 52 bytes) */
	uint32_t index_count, first_index, vertex_offset; /* into the context's quad pool */
	float aabb[6];             /* world space: AABB<uint32_t> / 16 + base_position (WorldRenderer.hpp:72-79) */
	int32_t base_position[3];  /* chunk position * 16 */
	uint32_t transparent;
} hcm_mesh_info;
typedef struct hcm_draw_cmd { /* VkDrawIndexedIndirectCommand (chunk_culling.comp:14-16) */
	uint32_t index_count, instance_count, first_index, vertex_offset, first_instance; /* first_instance: info index */
} hcm_draw_cmd;
/* After a run: the mesh info of each of the run's mesh records (in record order; hcm_batch_download's
 * records), for jobs at chunk positions chunk_pos[job][3]. Kept on the device for hcm_batch_cull; copied
 * to out when out is not NULL (cap entries). *n receives the count. */
int hcm_batch_mesh_infos(hcm_batch *batch, const int32_t *chunk_pos, hcm_mesh_info *out, uint32_t cap, uint32_t *n);
/* Frustum culling of the infos from hcm_batch_mesh_infos (chunk_culling.comp:37-88, 92-132, without the
 * depth-hierarchy test, which hcm_batch_cull_occlusion adds): a mesh is culled when all eight AABB corners lie outside one clip plane or behind
 * the eye. view_projection is column-major (GLSL mat4). Writes the draws of the kept opaque and transparent
 * meshes, in no particular order (the shader's atomics); *n_opaque / *n_transparent receive the counts
 * (HCM_ERR_CAPACITY when a list does not fit its cap). */
int hcm_batch_cull(hcm_batch *batch, const float view_position[3], const float view_projection[16],
                   hcm_draw_cmd *opaque, uint32_t opaque_cap, uint32_t *n_opaque, hcm_draw_cmd *transparent,
                   uint32_t transparent_cap, uint32_t *n_transparent);
/* The depth hierarchy the occlusion test samples (client/include/client/rg/DepthHierarchyPass.hpp:166-196,
 * client/shader/src/depth_hierarchy.comp), built on the device from the renderer's depth image: width x
 * height floats, texel (x, y) at depth[y * width + x], 2 <= (width | height), width and height at most
 * 16,384. Level 0 is the image; level i has max(size >> 1, 1) per axis of level i - 1, each texel the
 * largest of the 2 x 2 texels under it, plus the third column / row where the level above has an odd
 * width / height (texels past its edge read 0); myvk's QueryMipLevel gives the level count (the bit
 * length of width | height). The context keeps it until the next call. */
int hcm_ctx_set_depth(hcm_ctx *ctx, const float *depth, uint32_t width, uint32_t height);
/* Level `level` of that hierarchy (for checks): its size, and its texels into out when out is not NULL
 * (HCM_ERR_CAPACITY when they do not fit cap floats). */
int hcm_ctx_depth_level(hcm_ctx *ctx, uint32_t level, float *out, uint64_t cap, uint32_t *width, uint32_t *height);
/* hcm_batch_cull plus the shader's occlusion test (chunk_culling.comp:84-90, 104) against the context's
 * depth hierarchy, whose levels 1 .. n - 1 are the shader's texture mips 0 .. n - 2: a mesh whose corners
 * all have w > 0 is culled when the texel nearest the centre of its clip-space box, at mip
 * ceil(log2(max box extent in uv * max(mip 0 width, height))) clamped to the mips, is less than the box's
 * nearest depth. Clip coordinates use IEEE division and the mip the exact binary logarithm (the shader's
 * own division and log2 are approximate on GPUs). HCM_ERR_INVALID without a depth hierarchy. */
int hcm_batch_cull_occlusion(hcm_batch *batch, const float view_position[3], const float view_projection[16],
                             hcm_draw_cmd *opaque, uint32_t opaque_cap, uint32_t *n_opaque, hcm_draw_cmd *transparent,
                             uint32_t transparent_cap, uint32_t *n_transparent);

/* Launch shape of the meshing kernel. 0 = default. */
typedef struct hcm_tuning {
	uint32_t wg_per_cu; /* persistent workgroups per compute unit */
	uint32_t threads;   /* workgroup size: 128, 256 or 512 */
	uint32_t reserved[2];
} hcm_tuning;
int hcm_set_tuning(hcm_ctx *ctx, const hcm_tuning *t);
int hcm_get_tuning(hcm_ctx *ctx, hcm_tuning *t);

/* ------------------------------------------------------------------------------------------------- */
/* Serving: continuous batching of chunk-load and block-edit requests over device-resident chunks.     */
/* No counterpart in the reference, whose worker threads pop one mesh task per chunk                  */
/* (client/src/ChunkTaskPool.cpp:87-130) after ChunkTaskData<kMesh>::Pop finds all 27 chunks present  */
/* (client/src/ChunkMeshTask.cpp:14-39); an edit there pushes remesh tasks for nearby chunks          */
/* (client/src/ChunkSetBlockTask.cpp:100-119).                                                        */
/* ------------------------------------------------------------------------------------------------- */
typedef struct hcm_server hcm_server;

typedef struct hcm_server_config {
	uint32_t max_chunks;   /* device chunk slots */
	uint32_t max_batch;    /* chunks meshed per launch */
	uint32_t max_inflight; /* batches queued on the device at once */
	uint32_t threads;      /* workgroup size of the meshing kernel (0 = context tuning) */
	uint64_t pool_quads;   /* vertex pool for the meshes kept on the device */
	uint32_t keep_history; /* 1: keep a host copy of every batch's meshes (tests) */
	/* HCM_LIGHT_STORE: loads and edits carry light bytes. HCM_LIGHT_REBUILD: loads carry sunlight heights
	 * (hcm_server_load_sunlit) and every batch rebuilds its chunks' light before meshing, as the
	 * reference's mesh task does (client/src/ChunkMeshTask.cpp:48-111). */
	uint32_t light_mode;
	uint32_t reserved[2];
} hcm_server_config;

#define HCM_REQ_LOAD 1u
#define HCM_REQ_EDIT 2u

typedef struct hcm_completion {
	uint64_t tag;
	uint32_t kind;       /* HCM_REQ_LOAD or HCM_REQ_EDIT */
	uint32_t remeshed;   /* chunks this request waited on */
	uint64_t epoch;      /* requests applied to the world before the last batch it waited on */
	double latency_ms;   /* submit to completion, host clock */
} hcm_completion;

typedef struct hcm_server_stats {
	uint64_t requests, completed, batches, chunks_meshed, quads_meshed, retries;
	uint64_t pool_used_quads, pool_free_quads;
	double gpu_ms; /* sum of the meshing batches' HIP-event times */
} hcm_server_stats;

int hcm_server_create(hcm_ctx *ctx, const hcm_server_config *cfg, hcm_server **out);
void hcm_server_destroy(hcm_server *srv);
/* Loads chunk (cx, cy, cz): copies 4096 blocks + 4096 light bytes to a device slot; the chunk and every
 * loaded neighbour whose 26 neighbours are then all present are (re)meshed. HCM_ERR_BUSY when no slot. */
int hcm_server_load(hcm_server *srv, int32_t cx, int32_t cy, int32_t cz, const uint16_t *blocks,
                    const uint8_t *light, uint64_t tag);
/* HCM_LIGHT_REBUILD servers: loads chunk (cx, cy, cz) with its 16 x 16 sunlight heights (int8, index
 * x + 16 z, Chunk::m_sunlight_heights, client/include/client/Chunk.hpp:96) instead of light bytes.
 * HCM_ERR_INVALID on a HCM_LIGHT_STORE server, as hcm_server_load is on a HCM_LIGHT_REBUILD one. */
int hcm_server_load_sunlit(hcm_server *srv, int32_t cx, int32_t cy, int32_t cz, const uint16_t *blocks,
                           const int8_t *heights, uint64_t tag);
/* Sets the block (and, with HCM_LIGHT_STORE, the light byte) of world voxel (x, y, z) in a loaded chunk
 * and remeshes the chunks whose mesh can change: with HCM_LIGHT_STORE those whose 18^3 halo holds the
 * voxel (the edited chunk and 0..7 neighbours); with HCM_LIGHT_REBUILD those within light distance 14 of
 * it, the reference's rule (ChunkSetBlockTask.cpp:100-107, ChunkMeshTask.inl:9-15). HCM_ERR_INVALID if
 * the chunk is not loaded. */
int hcm_server_edit(hcm_server *srv, int32_t x, int32_t y, int32_t z, uint16_t block, uint8_t light, uint64_t tag);
/* Completed requests since the last poll (up to cap); *n receives the count. */
int hcm_server_poll(hcm_server *srv, hcm_completion *out, uint32_t cap, uint32_t *n);
/* Blocks until every submitted request has completed. Returns the server's sticky error, if any
 * (HCM_ERR_POOL_FULL when the live meshes leave no room in the vertex pool, HCM_ERR_HIP). */
int hcm_server_drain(hcm_server *srv);
/* Drains, then copies the current meshes of a chunk to host buffers (same layout as hcm_mesh_chunk);
 * HCM_ERR_INVALID if the chunk has no mesh. Requests submitted meanwhile from other threads may replace
 * the meshes while they are copied: call it when no other thread submits. */
int hcm_server_chunk_mesh(hcm_server *srv, int32_t cx, int32_t cy, int32_t cz, hcm_host_result *out);
int hcm_server_get_stats(hcm_server *srv, hcm_server_stats *out);

/* With keep_history: number of (batch, chunk) meshing results recorded, and one of them with its epoch
 * (requests applied to the world before its batch) and meshes. */
int hcm_server_history_size(hcm_server *srv, uint64_t *n);
int hcm_server_history_get(hcm_server *srv, uint64_t index, int32_t pos[3], uint64_t *epoch, hcm_host_result *out);

#ifdef __cplusplus
}
#endif

#endif /* HCM_H */
