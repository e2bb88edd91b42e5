/*
 * hcm_workload.h -- seeded synthetic chunk generators for the benchmark and the tests.
 *
 * No counterpart in the reference (its terrain comes from FastNoise2, client/src/DefaultTerrain.cpp,
 * which is outside the meshing path). All generators use integer fixed-point value noise, so the
 * same seed gives the same voxels on every host. Output chunks use the mesher's input layout
 * (x + (y*16 + z)*16 per voxel, Block::m_data blocks, sun | torch << 4 light bytes).
 */
#ifndef HCM_WORKLOAD_H
#define HCM_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HCM_GEN_PLANTS 1u /* cross-shaped grass plants on 5% of grass-block tops */

/* Heightmap terrain: 2D value noise, 4 octaves, world surface height h(x,z) in [8, 40]; stone below h-4,
 * dirt in [h-4, h-1), one grass-block cap at h-1, air above; sunlight 15 at y >= h, 0 below; torch 0. */
void hcm_gen_terrain_chunk(uint64_t seed, int32_t cx, int32_t cy, int32_t cz, uint32_t flags, uint16_t *blocks,
                           uint8_t *light);

/* Cave terrain: 3D fBm thresholded at ~50% solid under a surface heightmap cap in y [200, 240];
 * ~15% of chunks forced all-air and ~10% all-stone (region chunks only: 0 <= cx, cy, cz < 16);
 * sun and torch light from smooth (spatially correlated) random fields. */
void hcm_gen_cave_chunk(uint64_t seed, int32_t cx, int32_t cy, int32_t cz, uint32_t flags, uint16_t *blocks,
                        uint8_t *light);

/* Worst-case density: i.i.d. voxels, 50% air, 35% opaque cubes drawn from the 33 registered opaque
 * cube values, 10% transparent (glass, leaves of 5 variants), 5% still water; light uniform 0-15 per
 * channel. index selects an independent chunk. */
void hcm_gen_random_chunk(uint64_t seed, uint64_t index, uint16_t *blocks, uint8_t *light);

/* The 33 opaque full-cube block values drawn by hcm_gen_random_chunk. */
uint32_t hcm_gen_opaque_values(uint16_t *out, uint32_t cap);

#ifdef __cplusplus
}
#endif

#endif
