/*
 * hcm_oracle.h -- CPU oracle for the HyperCraft chunk mesher (TEST INFRASTRUCTURE ONLY).
 *
 * This header is shared by two CPU implementations that check the HIP mesher:
 *   - hcm_oracle.c : a clean-room C restatement of the reference mesher
 *                    (client/include/client/BlockMeshAlgo.hpp:17-522);
 *   - ref_driver.cpp: a thin wrapper that instantiates the real reference
 *                    BlockMeshAlgo<...>::Generate from its own sources
 *                    (built into oracle/_ref/ by oracle/Makefile).
 * Only tests, the smoke check and the benchmark's CPU-baseline leg may load
 * either of them. The product path never calls into this directory.
 *
 * Inputs follow the reference call site (client/src/ChunkMeshTask.cpp:113-122):
 * 27 chunk arrays of 16^3 voxels in neighbour-slot order
 * (common/include/common/Position.hpp:138-145: slot = 9*L[cx] + 3*L[cy] + L[cz],
 * L = {-1 -> 1, 0 -> 2, +1 -> 0}; centre = slot 26), each voxel indexed
 * x + (y*16 + z)*16 (Position.hpp:98-100). Blocks are Block::m_data
 * (id | meta << 8, block/include/block/Block.hpp:24-33); light bytes are
 * sun | torch << 4 (block/include/block/Light.hpp:18-23).
 *
 * Output mirrors std::vector<BlockMesh> (client/include/client/BlockVertex.hpp:20-25)
 * in the reference's emission order (BlockMeshAlgo.hpp:515-520).
 */
#ifndef HCM_ORACLE_H
#define HCM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_CHUNK 16
#define ORC_CHUNK_VOXELS (ORC_CHUNK * ORC_CHUNK * ORC_CHUNK)
#define ORC_MAX_MESHES 64
#define ORC_MAX_FACES 16 /* kBlockMeshMaxFaceCount, block/include/block/BlockMesh.hpp:37 */

/* ---- registry (flattened; exported from the reference by ref_driver.cpp) ---- */
typedef struct orc_vertex {
	uint8_t x, y, z, ao; /* 1/16 block units; ao == 4 means automatic (BlockMesh.hpp:15-25) */
} orc_vertex;

typedef struct orc_face {
	uint8_t axis, light_face, render_face, pad;
	uint16_t texture, pad2; /* BlockTexture::m_data: id | trans << 13 */
	orc_vertex v[4];
} orc_face; /* 24 bytes */

enum { ORC_MESH_STATIC = 0, ORC_MESH_LIQUID_FLOW = 1 };

typedef struct orc_mesh {
	uint32_t face_count;
	uint32_t kind;        /* ORC_MESH_STATIC or ORC_MESH_LIQUID_FLOW */
	uint32_t flow_height; /* liquid flow: get_flow_height(Variant), Liquid.hpp:151-154 */
	uint32_t reserved;
	orc_face faces[ORC_MAX_FACES];
} orc_mesh;

typedef struct orc_row {
	uint16_t tex[6];      /* BlockProperty::textures, indexed by BlockFace */
	uint8_t transparency; /* 0 opaque, 1 semi-transparent, 2 transparent */
	uint8_t collision;    /* 0 none, 1 solid, 2 water */
	uint16_t mesh;        /* index into meshes, 0xFFFF = no custom mesh */
} orc_row;

#define ORC_TEX_TRANSPARENT 1u
#define ORC_TEX_TRANS_PASS 2u
#define ORC_TEX_LIQUID 4u
#define ORC_NO_MESH 0xFFFFu

typedef struct orc_registry {
	uint32_t n_rows, n_meshes, n_tex;
	uint16_t row_of[65536]; /* Block::m_data -> row */
	orc_row *rows;
	orc_mesh *meshes;
	uint8_t tex_flags[8192]; /* by texture id (13 bits) */
	uint8_t liquid_source_id, liquid_flow_id;
	uint16_t liquid_flow_tex;
} orc_registry;

/* ---- mesher output ---- */
typedef struct orc_mesh_out {
	uint32_t n_meshes;
	uint32_t vert_begin[ORC_MAX_MESHES]; /* in vertices */
	uint32_t n_verts[ORC_MAX_MESHES];
	uint32_t idx_begin[ORC_MAX_MESHES]; /* in indices */
	uint32_t n_idx[ORC_MAX_MESHES];
	uint32_t aabb[ORC_MAX_MESHES][6]; /* min xyz, max xyz (dep/AABB.hpp:9-60) */
	uint32_t transparent[ORC_MAX_MESHES];
	uint32_t vert_cap, idx_cap; /* capacities of verts (in vertices) and idx */
	uint32_t *verts;            /* 2 words per vertex (BlockVertex.hpp:8-18) */
	uint16_t *idx;
} orc_mesh_out;

#define ORC_OK 0
#define ORC_ERR_CAPACITY -1
#define ORC_ERR_IO -2
#define ORC_ERR_FORMAT -3

/* Registry fixture file (little-endian):
 *   char magic[8] = "HCREG01"
 *   u32 n_rows, n_meshes, n_tex, reserved
 *   u16 row_of[65536]
 *   orc_row rows[n_rows]          (16 B each)
 *   orc_mesh meshes[n_meshes]     (400 B each)
 *   u8 tex_flags[n_tex]
 *   u8 liquid_source_id, liquid_flow_id; u16 liquid_flow_tex
 */
int orc_registry_load(const char *path, orc_registry *reg);
void orc_registry_free(orc_registry *reg);

/* Clean-room restatement of BlockMeshAlgo<Config>::Generate for a 16^3 bound. */
int orc_mesh_chunk(const orc_registry *reg, const uint16_t *const blocks[27], const uint8_t *const light[27],
                   orc_mesh_out *out);

/* Same, over n independent jobs with a pool of threads (CPU baseline). nbr holds 27 chunk indices per job
 * into the blocks/light arrays (each ORC_CHUNK_VOXELS long). quads_out[j] receives the job's quad count. */
int orc_mesh_many(const orc_registry *reg, const uint16_t *blocks, const uint8_t *light, const uint32_t *nbr,
                  uint32_t n_jobs, uint32_t n_threads, uint64_t *quads_out);

/* ---- light rebuild before meshing (client/src/ChunkMeshTask.cpp:48-111) ---- */
#define ORC_LIGHT_W 46 /* the extended light buffer spans [-15, 31) per axis (ChunkMeshTask.cpp:10-12) */
#define ORC_LIGHT_WVOX (ORC_LIGHT_W * ORC_LIGHT_W * ORC_LIGHT_W)
#define ORC_HALO 18 /* the mesher reads light over [-1, 17) per axis */
#define ORC_HALO_VOX (ORC_HALO * ORC_HALO * ORC_HALO)

/* Rebuilds the mesh task's extended light buffer: torchlight from Block::GetLightLevel, sunlight 15 where
 * y >= the owning chunk's sunlight height, border seeds, then sunlight and torchlight propagation.
 *   heights[27]  per chunk 16 x 16 InnerPos1 sunlight heights, index x + 16 z (Chunk.hpp:65-76, 96)
 *   light_level  u8[65536], Block::GetLightLevel by Block::m_data (Block.hpp:71)
 *   win          out, ORC_LIGHT_WVOX light bytes (sun | torch << 4), index (x+15) + 46 ((z+15) + 46 (y+15))
 *   halo         out (may be NULL), the ORC_HALO_VOX bytes the mesher reads, index (x+1) + 18 ((z+1) + 18 (y+1))
 * Returns ORC_OK or ORC_ERR_CAPACITY (allocation failure). */
int orc_light_rebuild(const orc_registry *reg, const uint8_t *light_level, const uint16_t *const blocks[27],
                      const int8_t *const heights[27], uint8_t *win, uint8_t *halo);

/* Many jobs on a thread pool (the CPU baseline of the whole mesh task, ChunkMeshTask.cpp:41-127):
 * light rebuild per job, and when mesh != 0 also the mesher over the rebuilt light. heights holds 256
 * InnerPos1 per chunk index. halo_out [n_jobs][ORC_HALO_VOX] and quads_out [n_jobs] may be NULL. */
int orc_relight_many(const orc_registry *reg, const uint8_t *light_level, const uint16_t *blocks,
                     const int8_t *heights, const uint32_t *nbr, uint32_t n_jobs, uint32_t n_threads, int mesh,
                     uint8_t *halo_out, uint64_t *quads_out);

#ifdef __cplusplus
}
#endif

#endif
