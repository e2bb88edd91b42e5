// hcm_workload.cpp -- seeded synthetic chunks (see include/hcm_workload.h). Integer-only arithmetic.
#include <cstdint>
#include <cstring>

#include "../../include/hcm_workload.h"

namespace {

inline uint64_t mix64(uint64_t x) { // splitmix64 finaliser
	x += 0x9e3779b97f4a7c15ull;
	x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
	x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
	return x ^ (x >> 31);
}
inline uint64_t hash4(uint64_t seed, int64_t a, int64_t b, int64_t c) {
	return mix64(seed ^ mix64((uint64_t)a * 0x632be59bd9b4e019ull ^ mix64((uint64_t)b * 0x8cb92ba72f3d8dd7ull ^
	                                                                      mix64((uint64_t)c + 0x1234567ull))));
}
inline int64_t floor_div(int64_t a, int64_t b) { return a >= 0 ? a / b : -((-a + b - 1) / b); }

// lattice value in [0, 65536)
inline int64_t lattice(uint64_t seed, int64_t x, int64_t y, int64_t z) { return (int64_t)(hash4(seed, x, y, z) >> 48); }
// smoothstep of t in [0, 65536] in 16.16 fixed point
inline int64_t smooth(int64_t t) { return (t * t >> 16) * (3 * 65536 - 2 * t) >> 16; }
inline int64_t lerp(int64_t a, int64_t b, int64_t t) { return a + ((b - a) * t >> 16); }

// 2D value noise with the given lattice period, result in [0, 65536)
int64_t noise2(uint64_t seed, int64_t x, int64_t z, int64_t period) {
	int64_t ix = floor_div(x, period), iz = floor_div(z, period);
	int64_t tx = smooth((x - ix * period) * 65536 / period), tz = smooth((z - iz * period) * 65536 / period);
	int64_t a = lerp(lattice(seed, ix, 0, iz), lattice(seed, ix + 1, 0, iz), tx);
	int64_t b = lerp(lattice(seed, ix, 0, iz + 1), lattice(seed, ix + 1, 0, iz + 1), tx);
	return lerp(a, b, tz);
}
int64_t noise3(uint64_t seed, int64_t x, int64_t y, int64_t z, int64_t period) {
	int64_t ix = floor_div(x, period), iy = floor_div(y, period), iz = floor_div(z, period);
	int64_t tx = smooth((x - ix * period) * 65536 / period), ty = smooth((y - iy * period) * 65536 / period),
	        tz = smooth((z - iz * period) * 65536 / period);
	int64_t c[2][2];
	for (int dy = 0; dy < 2; ++dy)
		for (int dz = 0; dz < 2; ++dz)
			c[dy][dz] = lerp(lattice(seed, ix, iy + dy, iz + dz), lattice(seed, ix + 1, iy + dy, iz + dz), tx);
	return lerp(lerp(c[0][0], c[0][1], tz), lerp(c[1][0], c[1][1], tz), ty);
}
// fBm over `octaves` octaves starting at period p0 (halving), amplitudes halving; result in [0, 65536)
int64_t fbm2(uint64_t seed, int64_t x, int64_t z, int64_t p0, int octaves) {
	int64_t sum = 0, norm = 0, amp = 1 << octaves;
	for (int o = 0; o < octaves; ++o, p0 >>= 1, amp >>= 1) {
		sum += amp * noise2(seed + (uint64_t)o * 0x51ed27ull, x, z, p0);
		norm += amp;
	}
	return sum / norm;
}
int64_t fbm3(uint64_t seed, int64_t x, int64_t y, int64_t z, int64_t p0, int octaves) {
	int64_t sum = 0, norm = 0, amp = 1 << octaves;
	for (int o = 0; o < octaves; ++o, p0 >>= 1, amp >>= 1) {
		sum += amp * noise3(seed + (uint64_t)o * 0x51ed27ull, x, y, z, p0);
		norm += amp;
	}
	return sum / norm;
}

inline uint32_t local(int x, int y, int z) { return uint32_t(x) + 16u * (uint32_t(z) + 16u * uint32_t(y)); }

constexpr uint16_t kAir = 0, kStone = 1, kDirt = 3, kGrassBlock = 4, kGrass = 5, kSand = 6, kGravel = 8, kGlass = 9,
                   kWater = 14, kLeaves = 16;

// terrain surface height in [8, 40]: the fBm value, contrast doubled about its middle and clamped
inline int terrain_height(uint64_t seed, int64_t x, int64_t z) {
	int64_t v = (fbm2(seed, x, z, 64, 4) - 32768) * 2 + 32768;
	v = v < 0 ? 0 : (v > 65535 ? 65535 : v);
	return 8 + (int)(v * 33 >> 16);
}

inline bool plant_at(uint64_t seed, int64_t x, int64_t y, int64_t z, uint16_t *out) {
	uint64_t h = hash4(seed ^ 0x706c616e74ull, x, y, z);
	if (h % 100 >= 5)
		return false;
	*out = uint16_t(kGrass | ((h >> 32) % 4) << 8);
	return true;
}

} // namespace

extern "C" void hcm_gen_terrain_chunk(uint64_t seed, int32_t cx, int32_t cy, int32_t cz, uint32_t flags,
                                      uint16_t *blocks, uint8_t *light) {
	for (int z = 0; z < 16; ++z)
		for (int x = 0; x < 16; ++x) {
			const int64_t wx = (int64_t)cx * 16 + x, wz = (int64_t)cz * 16 + z;
			const int h = terrain_height(seed, wx, wz);
			for (int y = 0; y < 16; ++y) {
				const int64_t wy = (int64_t)cy * 16 + y;
				uint16_t b = kAir;
				if (wy < h - 4)
					b = kStone;
				else if (wy < h - 1)
					b = kDirt;
				else if (wy == h - 1)
					b = kGrassBlock;
				else if (wy == h && (flags & HCM_GEN_PLANTS))
					plant_at(seed, wx, wy, wz, &b);
				blocks[local(x, y, z)] = b;
				light[local(x, y, z)] = wy >= h ? 15 : 0;
			}
		}
}

extern "C" void hcm_gen_cave_chunk(uint64_t seed, int32_t cx, int32_t cy, int32_t cz, uint32_t flags,
                                   uint16_t *blocks, uint8_t *light) {
	const bool in_region = cx >= 0 && cx < 16 && cy >= 0 && cy < 16 && cz >= 0 && cz < 16;
	const uint64_t force = hash4(seed ^ 0x666f726365ull, cx, cy, cz) % 100;
	const int mode = in_region ? (force < 15 ? 1 : (force < 25 ? 2 : 0)) : 0; // 1 all air, 2 all stone
	for (int z = 0; z < 16; ++z)
		for (int x = 0; x < 16; ++x) {
			const int64_t wx = (int64_t)cx * 16 + x, wz = (int64_t)cz * 16 + z;
			const int surface = 200 + (int)(fbm2(seed ^ 0x73757266ull, wx, wz, 64, 3) * 41 >> 16);
			for (int y = 0; y < 16; ++y) {
				const int64_t wy = (int64_t)cy * 16 + y;
				uint16_t b = kAir;
				if (mode == 2)
					b = kStone;
				else if (mode == 0 && wy < surface) {
					const bool solid = fbm3(seed ^ 0x63617665ull, wx, wy, wz, 32, 3) >= 32768;
					if (solid) {
						if (wy == surface - 1)
							b = kGrassBlock;
						else if (wy >= surface - 4)
							b = kDirt;
						else {
							const int64_t m = noise3(seed ^ 0x6d6174ull, wx, wy, wz, 16);
							b = m > 56000 ? kGravel : (m < 6000 ? kSand : kStone);
						}
					}
				}
				if (b == kAir && mode == 0 && (flags & HCM_GEN_PLANTS) && wy == surface) {
					// a plant can only stand on the grass cap
					if (fbm3(seed ^ 0x63617665ull, wx, wy - 1, wz, 32, 3) >= 32768)
						plant_at(seed, wx, wy, wz, &b);
				}
				const uint32_t sun = (uint32_t)(noise3(seed ^ 0x73756eull, wx, wy, wz, 8) * 16 >> 16);
				const uint32_t torch = (uint32_t)(noise3(seed ^ 0x746f7263ull, wx, wy, wz, 8) * 16 >> 16);
				blocks[local(x, y, z)] = b;
				light[local(x, y, z)] = uint8_t(sun | torch << 4);
			}
		}
}

static const uint16_t kOpaque33[33] = {
    1, 2, 3,                                                                      // stone, cobblestone, dirt
    4, 4 | 1 << 8, 4 | 2 << 8, 4 | 3 << 8,                                        // grass block variants
    6, 8, 10, 11, 13, 24,                                                         // sand .. glowstone
    17, 17 | 1 << 8, 17 | 2 << 8, 17 | 3 << 8, 17 | 4 << 8,                        // logs, transform 0
    17 | 8 << 8, 17 | 9 << 8, 17 | 10 << 8, 17 | 11 << 8, 17 | 12 << 8,             // logs, transform 1
    17 | 16 << 8, 17 | 17 << 8, 17 | 18 << 8, 17 | 19 << 8, 17 | 20 << 8,           // logs, transform 2
    18, 18 | 1 << 8, 18 | 2 << 8, 18 | 3 << 8, 18 | 4 << 8,                        // planks
};

extern "C" uint32_t hcm_gen_opaque_values(uint16_t *out, uint32_t cap) {
	for (uint32_t i = 0; i < 33 && i < cap; ++i)
		out[i] = kOpaque33[i];
	return 33;
}

extern "C" void hcm_gen_random_chunk(uint64_t seed, uint64_t index, uint16_t *blocks, uint8_t *light) {
	uint64_t state = mix64(seed ^ mix64(index + 0x72616e64ull));
	for (uint32_t i = 0; i < 4096; ++i) {
		state = mix64(state);
		const uint32_t r = (uint32_t)(state % 100);
		const uint64_t pick = state >> 32;
		uint16_t b;
		if (r < 50)
			b = kAir;
		else if (r < 85)
			b = kOpaque33[pick % 33];
		else if (r < 95)
			b = (pick % 6 == 0) ? kGlass : uint16_t(kLeaves | (pick % 5) << 8);
		else
			b = kWater;
		blocks[i] = b;
		light[i] = uint8_t(((state >> 8) & 15u) | ((state >> 16) & 15u) << 4);
	}
}
