// hcm_kernels.hip -- gfx950 chunk-meshing kernels.
//
// One workgroup meshes one 16^3 chunk at a time (a persistent workgroup takes chunk jobs from a work
// queue until it is empty). Per chunk:
//   1. stage the 18^3 block/light halo from the 27 chunk slots into LDS as one word per voxel
//      (light | direct-pass << 8 | indirect-pass << 9 | texture row << 16), plus bitmaps of cells with
//      custom meshes and of liquid-flow cells; a chunk whose halo is one block value exits here;
//   2. liquid-flow cells get their per-cell face textures (the reference's m_dynamic_block_textures);
//   3. faces: one wavefront per (axis, slice, direction) task computes, 64 cells at a time, the face
//      visibility (BlockTexture::Show), the per-vertex AO / smooth light word (light4_init) and, by
//      ballot, the 16x16 row masks the greedy merge needs: visible, equal-to-right, equal-to-below,
//      transparent-pass;
//   4. one lane per task replays the reference's greedy merge on those bit masks (u-major, v-minor,
//      width first, then height) and counts quads per pass; custom-mesh faces are counted per cell;
//   5. block-wide prefix sums give every quad its index in its pass and its position in the
//      reference's global emission order; one atomic per chunk allocates the output range;
//   6. quads are described into an LDS list (in windows) and written densely, one lane per quad,
//      as packed BlockVertex words and u16 indices, with the reference's 65,532-vertex mesh split,
//      per-mesh AABBs and mesh order.
// All arithmetic is integer. References are to client/include/client/BlockMeshAlgo.hpp unless noted.
#include <hip/hip_runtime.h>

#include "hcm_internal.h"

namespace hcm {

constexpr int kHalo = 18;
constexpr int kHaloVox = kHalo * kHalo * kHalo;
constexpr int kTasks = 102; // 3 axes x 17 slices x 2 directions
constexpr int kListCap = 1024;
constexpr uint32_t kSplitQuads = 16383u; // a BlockMesh holds at most 65,532 vertices (:273-279, :397-403)
constexpr int kMaxMeshPerPass = 8;

// Halo layout: x fastest, then z, then y -- the chunk-local order x + 16 * (z + 16 * y) with a 1-voxel rim.
constexpr int kSX = 1, kSZ = kHalo, kSY = kHalo * kHalo;
__host__ __device__ constexpr int axis_stride(int a) { return a == 0 ? kSX : (a == 1 ? kSY : kSZ); }
// Mask axes per slice axis (BlockAlgo.hpp:70-81 with the YZX swizzle): x -> (U=y, V=z), y -> (U=z, V=x),
// z -> (U=y, V=x).
__host__ __device__ constexpr int u_axis(int a) { return a == 1 ? 2 : 1; }
__host__ __device__ constexpr int v_axis(int a) { return a == 0 ? 2 : 0; }

// The light4_init cell pattern (kLookup3v, :75-108). Vertex k of face F reads, in the plane of the
// front cell, a side cell c0 offset along plane axis Q, the corner cell c1 offset along P and Q, and a
// side cell c2 offset along P; the signs of (P, Q) per vertex follow one of two windings. The host-side
// static_asserts below check this against the reference's literal table.
__host__ __device__ constexpr int plane_p(int F) { return (F >> 1) == 0 ? 1 : 0; }
__host__ __device__ constexpr int plane_q(int F) { return (F >> 1) == 2 ? 1 : 2; }
__host__ __device__ constexpr bool winding_a(int F) { return F == 0 || F == 3 || F == 4; }
__host__ __device__ constexpr int corner_sp(int F, int k) {
	return k == 0 ? -1 : (k == 2 ? 1 : (k == 1 ? (winding_a(F) ? 1 : -1) : (winding_a(F) ? -1 : 1)));
}
__host__ __device__ constexpr int corner_sq(int F, int k) {
	return k == 0 ? -1 : (k == 2 ? 1 : (k == 1 ? (winding_a(F) ? -1 : 1) : (winding_a(F) ? 1 : -1)));
}
__host__ __device__ constexpr int normal_sign(int F) { return (F & 1) ? -1 : 1; }
// Offset of (axis components) d[3] in the 3 axes for cell b of vertex k of face F, as (dx, dy, dz).
__host__ __device__ constexpr int corner_comp(int F, int k, int b, int axis) {
	return (axis == (F >> 1) ? normal_sign(F) : 0) +
	       (axis == plane_p(F) && b >= 1 ? corner_sp(F, k) : 0) +
	       (axis == plane_q(F) && b <= 1 ? corner_sq(F, k) : 0);
}
__host__ __device__ constexpr int corner_off(int F, int k, int b) {
	return corner_comp(F, k, b, 0) * kSX + corner_comp(F, k, b, 1) * kSY + corner_comp(F, k, b, 2) * kSZ;
}

namespace check {
// kLookup3v exactly as the reference lists it (:75-108).
constexpr int8_t kLookup3v[6][4][3][3] = {
    {{{1, 0, -1}, {1, -1, -1}, {1, -1, 0}}, {{1, 0, -1}, {1, 1, -1}, {1, 1, 0}}, {{1, 0, 1}, {1, 1, 1}, {1, 1, 0}},
     {{1, 0, 1}, {1, -1, 1}, {1, -1, 0}}},
    {{{-1, 0, -1}, {-1, -1, -1}, {-1, -1, 0}}, {{-1, 0, 1}, {-1, -1, 1}, {-1, -1, 0}},
     {{-1, 0, 1}, {-1, 1, 1}, {-1, 1, 0}}, {{-1, 0, -1}, {-1, 1, -1}, {-1, 1, 0}}},
    {{{0, 1, -1}, {-1, 1, -1}, {-1, 1, 0}}, {{0, 1, 1}, {-1, 1, 1}, {-1, 1, 0}}, {{0, 1, 1}, {1, 1, 1}, {1, 1, 0}},
     {{0, 1, -1}, {1, 1, -1}, {1, 1, 0}}},
    {{{0, -1, -1}, {-1, -1, -1}, {-1, -1, 0}}, {{0, -1, -1}, {1, -1, -1}, {1, -1, 0}},
     {{0, -1, 1}, {1, -1, 1}, {1, -1, 0}}, {{0, -1, 1}, {-1, -1, 1}, {-1, -1, 0}}},
    {{{0, -1, 1}, {-1, -1, 1}, {-1, 0, 1}}, {{0, -1, 1}, {1, -1, 1}, {1, 0, 1}}, {{0, 1, 1}, {1, 1, 1}, {1, 0, 1}},
     {{0, 1, 1}, {-1, 1, 1}, {-1, 0, 1}}},
    {{{0, -1, -1}, {-1, -1, -1}, {-1, 0, -1}}, {{0, 1, -1}, {-1, 1, -1}, {-1, 0, -1}},
     {{0, 1, -1}, {1, 1, -1}, {1, 0, -1}}, {{0, -1, -1}, {1, -1, -1}, {1, 0, -1}}}};
constexpr bool corner_table_matches() {
	for (int F = 0; F < 6; ++F)
		for (int k = 0; k < 4; ++k)
			for (int b = 0; b < 3; ++b)
				for (int a = 0; a < 3; ++a)
					if (corner_comp(F, k, b, a) != kLookup3v[F][k][b][a])
						return false;
	return true;
}
static_assert(corner_table_matches(), "light4 cell pattern must equal kLookup3v");
} // namespace check

__device__ __forceinline__ int slot27(int x, int y, int z) {
	// CmpXYZ2NeighbourIndex (common/include/common/Position.hpp:138-145): L(-1) = 1, L(0) = 2, L(1) = 0
	int cx = x < 0 ? -1 : (x >= 16 ? 1 : 0), cy = y < 0 ? -1 : (y >= 16 ? 1 : 0), cz = z < 0 ? -1 : (z >= 16 ? 1 : 0);
	return ((cx + 2) % 3) * 9 + ((cy + 2) % 3) * 3 + (cz + 2) % 3;
}
__device__ __forceinline__ uint32_t local16(int x, int y, int z) {
	return uint32_t(x & 15) + 16u * (uint32_t(z & 15) + 16u * uint32_t(y & 15));
}
__device__ __forceinline__ int hidx(int x, int y, int z) { return (x + 1) * kSX + (y + 1) * kSY + (z + 1) * kSZ; }

template <int NT> struct Smem {
	uint32_t vox[kHaloVox];
	uint16_t ne[kTasks][16], eh[kTasks][16], ev[kTasks][16], tp[kTasks][16];
	uint16_t cov[kTasks][16];
	uint2 list[kListCap];
	uint2 rows[kMaxPid];
	uint32_t codes[kMaxCodes];
	uint32_t custom_bits[128]; // in-chunk cells with a custom mesh
	uint32_t flow_bits[(kHaloVox + 31) / 32];
	uint32_t aabb[2][kMaxMeshPerPass][6];
	uint32_t events[16][2];
	uint64_t scan[NT / 64 + 1];
	uint32_t slots[27];
	uint32_t job, flags, n_events, status;
	uint32_t n0, n1, cust0, cust1, pool_base, mesh_base, n_mesh0, n_mesh1;
};

constexpr uint32_t kFlagNonUniform = 1u, kFlagCustom = 2u, kFlagFlow = 4u;

// ------------------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------------------
template <int NT> __device__ __forceinline__ uint32_t row_code(const Smem<NT> &s, uint32_t pid, int face) {
	uint2 r = s.rows[pid];
	uint32_t w = face < 4 ? r.x : r.y;
	return (w >> (8 * (face & 3))) & 0xffu;
}

// BlockTexture::Show (texture/include/texture/BlockTexture.hpp:68-76) on texture codes.
template <int NT> __device__ __forceinline__ bool show(const Smem<NT> &s, uint32_t a, uint32_t b) {
	uint32_t fa = s.codes[a] >> 16, fb = s.codes[b] >> 16;
	if ((fa & kCodeEmpty) || a == b)
		return false;
	bool ta = fa & kCodeTransparent, tb = fb & kCodeTransparent;
	if (!ta && !tb)
		return false;
	if ((fa & kCodeLiquid) && !(fb & kCodeEmpty))
		return false;
	return !ta || tb || (fb & kCodeLiquid);
}

// (s << 2) / cnt for cnt in 1..4 and s <= 60, exactly: multiply by ceil(65536 / cnt) and shift.
__device__ __forceinline__ uint32_t div_cnt(uint32_t x, uint32_t cnt) {
	uint32_t r = cnt == 1 ? 65536u : (cnt == 2 ? 32768u : (cnt == 3 ? 21846u : 16384u));
	return (x * r) >> 16;
}

// light4_init (:48-163) for face F of the block at halo index self. Returns the packed Light4:
//   x = sun0 | sun1 << 6 | sun2 << 12 | sun3 << 18 | ao4 << 24,   y = torch0 | ... | torch3 << 18.
template <int F> __device__ __forceinline__ uint2 light4(const uint32_t *vox, int self) {
	constexpr int A = F >> 1;
	const uint32_t vf = vox[self + normal_sign(F) * axis_stride(A)];
	const uint32_t pass3 = (vf >> 8) & 1u;
	const uint32_t fs = vf & 15u, ft = (vf >> 4) & 15u;
	uint32_t lo = 0, hi = 0, ao4 = 0;
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		const uint32_t c0 = vox[self + corner_off(F, k, 0)];
		const uint32_t c1 = vox[self + corner_off(F, k, 1)];
		const uint32_t c2 = vox[self + corner_off(F, k, 2)];
		const uint32_t d0 = (c0 >> 8) & 1u, d1 = (c1 >> 8) & 1u, d2 = (c2 >> 8) & 1u;
		const uint32_t i0 = (c0 >> 9) & 1u, i1 = (c1 >> 9) & 1u, i2 = (c2 >> 9) & 1u;
		uint32_t ao;
		if ((!d0 && !d2) || (!d1 && !pass3))
			ao = 0;
		else {
			int a = 3 - int(!d0) - int(!d1) - int(!d2) - int(!pass3);
			ao = a < 0 ? 0u : uint32_t(a);
		}
		uint32_t cnt = 1, ss = fs, tt = ft;
		if (i0 | i2) {
			cnt += i0 + i1 + i2;
			ss += (i0 ? (c0 & 15u) : 0u) + (i1 ? (c1 & 15u) : 0u) + (i2 ? (c2 & 15u) : 0u);
			tt += (i0 ? ((c0 >> 4) & 15u) : 0u) + (i1 ? ((c1 >> 4) & 15u) : 0u) + (i2 ? ((c2 >> 4) & 15u) : 0u);
		}
		lo |= div_cnt(ss << 2, cnt) << (6 * k);
		hi |= div_cnt(tt << 2, cnt) << (6 * k);
		ao4 |= ao << (2 * k);
	}
	return make_uint2(lo | (ao4 << 24), hi);
}

__device__ __forceinline__ uint2 light4_rt(const uint32_t *vox, int face, int self) {
	switch (face) {
	case 0:
		return light4<0>(vox, self);
	case 1:
		return light4<1>(vox, self);
	case 2:
		return light4<2>(vox, self);
	case 3:
		return light4<3>(vox, self);
	case 4:
		return light4<4>(vox, self);
	default:
		return light4<5>(vox, self);
	}
}

__device__ __forceinline__ uint32_t l4_sun(uint2 l, int k) { return (l.x >> (6 * k)) & 63u; }
__device__ __forceinline__ uint32_t l4_torch(uint2 l, int k) { return (l.y >> (6 * k)) & 63u; }
__device__ __forceinline__ uint32_t l4_ao(uint2 l, int k) { return (l.x >> (24 + 2 * k)) & 3u; }

// Light4::GetFlip (:38-43): sunlight and AO only.
__device__ __forceinline__ bool l4_flip(uint2 l) {
	int d02 = int(l4_ao(l, 0) + 1) * int(l4_sun(l, 0) + 1) - int(l4_ao(l, 2) + 1) * int(l4_sun(l, 2) + 1);
	int d13 = int(l4_ao(l, 1) + 1) * int(l4_sun(l, 1) + 1) - int(l4_ao(l, 3) + 1) * int(l4_sun(l, 3) + 1);
	return abs(d02) < abs(d13);
}

__device__ __forceinline__ uint64_t ballot64(bool b) { return __ballot(b); }

// ------------------------------------------------------------------------------------------------
// block-wide exclusive scan of 64-bit values (two 32-bit counters packed lo | hi << 32)
// ------------------------------------------------------------------------------------------------
template <int NT> __device__ __forceinline__ uint64_t block_excl_scan(Smem<NT> &s, uint64_t v, uint64_t *total) {
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	uint64_t inc = v;
#pragma unroll
	for (int d = 1; d < 64; d <<= 1) {
		uint32_t lo = __shfl_up((uint32_t)inc, d), hi = __shfl_up((uint32_t)(inc >> 32), d);
		if (lane >= d)
			inc += (uint64_t)lo | ((uint64_t)hi << 32);
	}
	if (lane == 63)
		s.scan[wave] = inc;
	__syncthreads();
	if (threadIdx.x == 0) {
		uint64_t run = 0;
		for (int w = 0; w < NT / 64; ++w) {
			uint64_t t = s.scan[w];
			s.scan[w] = run;
			run += t;
		}
		s.scan[NT / 64] = run;
	}
	__syncthreads();
	uint64_t r = s.scan[wave] + inc - v;
	*total = s.scan[NT / 64];
	__syncthreads();
	return r;
}

// ------------------------------------------------------------------------------------------------
// liquid flow (block/block/public/Liquid.hpp:151-302)
// ------------------------------------------------------------------------------------------------
struct Flow {
	uint32_t h;        // corner heights h0 | h1 << 8 | h2 << 16 | h3 << 24 (0 1 / 2 3 over x, z)
	uint32_t liq_side; // bit per neighbour 1, 3, 4, 6 (back, left, right, front) that is a liquid
};

__device__ __forceinline__ uint32_t flow_h(const Flow &f, int i) { return (f.h >> (8 * i)) & 0xffu; }

__device__ __forceinline__ uint16_t global_block(const MeshParams &p, const uint32_t *slots, int x, int y, int z) {
	return p.blocks[(size_t)slots[slot27(x, y, z)] * HCM_CHUNK_VOXELS + local16(x, y, z)];
}

// get_nei_height (:164-169) as a signed integer: the reference returns it as int8_t, so a flow level's
// uint8_t height max(meta * 2, 1) wraps to [-128, 127].
__device__ __forceinline__ int flow_nei_height(uint32_t b, uint32_t src_id, uint32_t flow_id) {
	const uint32_t id = b & 0xffu, meta = b >> 8;
	if (id == src_id)
		return 16;
	if (id == flow_id) {
		const uint32_t h = (meta * 2u > 1u ? meta * 2u : 1u) & 0xffu;
		return h >= 128u ? int(h) - 256 : int(h);
	}
	return id != 0 ? -1 : 0; // WaterTrait::IsObstacle: id != kAir (block/block/public/Water.hpp:13)
}

// get_vh (:181-186): the signed maximum as uint8_t, or 0 when three of the four are empty (height 0)
__device__ __forceinline__ uint32_t flow_vh(int nb, int a, int b, int c) {
	const int m = max(max(nb, a), max(b, c));
	const uint32_t mx = uint32_t(m) & 0xffu;
	const int zeros = int(nb == 0) + int(a == 0) + int(b == 0) + int(c == 0);
	return (mx != 16u && zeros >= 3) ? 0u : mx;
}

__device__ __forceinline__ Flow flow_heights(const MeshParams &p, const uint32_t *slots, int x, int y, int z,
                                             uint32_t nb_height) {
	const uint32_t src = p.reg.src_id, fid = p.reg.flow_id;
	// neighbours (:291-300): 0 (-1,-1) 1 (0,-1) 2 (1,-1) 3 (-1,0) 4 (1,0) 5 (-1,1) 6 (0,1) 7 (1,1) in (x, z)
	const uint32_t b0 = global_block(p, slots, x - 1, y, z - 1), b1 = global_block(p, slots, x, y, z - 1);
	const uint32_t b2 = global_block(p, slots, x + 1, y, z - 1), b3 = global_block(p, slots, x - 1, y, z);
	const uint32_t b4 = global_block(p, slots, x + 1, y, z), b5 = global_block(p, slots, x - 1, y, z + 1);
	const uint32_t b6 = global_block(p, slots, x, y, z + 1), b7 = global_block(p, slots, x + 1, y, z + 1);
	const int n0 = flow_nei_height(b0, src, fid), n1 = flow_nei_height(b1, src, fid);
	const int n2 = flow_nei_height(b2, src, fid), n3 = flow_nei_height(b3, src, fid);
	const int n4 = flow_nei_height(b4, src, fid), n5 = flow_nei_height(b5, src, fid);
	const int n6 = flow_nei_height(b6, src, fid), n7 = flow_nei_height(b7, src, fid);
	const int nb = int(nb_height);
	Flow f;
	f.h = flow_vh(nb, n0, n1, n3) | flow_vh(nb, n1, n2, n4) << 8 | flow_vh(nb, n3, n5, n6) << 16 |
	      flow_vh(nb, n4, n6, n7) << 24;
	auto liq = [&](uint32_t b) { return ((b & 0xffu) == src || (b & 0xffu) == fid) ? 1u : 0u; };
	f.liq_side = liq(b1) | liq(b3) << 1 | liq(b4) << 2 | liq(b6) << 3;
	return f;
}

__device__ __forceinline__ bool flow_all_full(const Flow &f) { return f.h == 0x10101010u; }

// Emitted faces of a flow cell, in order: top, left, right, back, front (:193-274). Returns a 5-bit set.
__device__ __forceinline__ uint32_t flow_face_set(const Flow &f) {
	if (flow_all_full(f))
		return 0;
	const uint32_t h0 = flow_h(f, 0), h1 = flow_h(f, 1), h2 = flow_h(f, 2), h3 = flow_h(f, 3);
	const bool lf = h0 == 16 && h2 == 16, le = h0 == 0 && h2 == 0;
	const bool rf = h1 == 16 && h3 == 16, re = h1 == 0 && h3 == 0;
	const bool ff = h2 == 16 && h3 == 16, fe = h2 == 0 && h3 == 0;
	const bool bf = h0 == 16 && h1 == 16, be = h0 == 0 && h1 == 0;
	uint32_t set = 1u;
	if (!lf && !le && !(f.liq_side & 2u))
		set |= 2u;
	if (!rf && !re && !(f.liq_side & 4u))
		set |= 4u;
	if (!bf && !be && !(f.liq_side & 1u))
		set |= 8u;
	if (!ff && !fe && !(f.liq_side & 8u))
		set |= 16u;
	return set;
}

// Dynamic texture pattern of a flow cell: 16 = all six faces liquid; otherwise bits right, left, front,
// back full (:193-197, :275-280).
__device__ __forceinline__ uint32_t flow_pattern(const Flow &f) {
	if (flow_all_full(f))
		return 16u;
	const uint32_t h0 = flow_h(f, 0), h1 = flow_h(f, 1), h2 = flow_h(f, 2), h3 = flow_h(f, 3);
	return uint32_t(h1 == 16 && h3 == 16) | (uint32_t(h0 == 16 && h2 == 16) << 1) |
	       (uint32_t(h2 == 16 && h3 == 16) << 2) | (uint32_t(h0 == 16 && h1 == 16) << 3);
}

__device__ __forceinline__ hcm_mesh_vertex fv(uint32_t x, uint32_t y, uint32_t z) {
	hcm_mesh_vertex v;
	v.x = uint8_t(x), v.y = uint8_t(y), v.z = uint8_t(z), v.ao = 4;
	return v;
}

// The kind-th face (0 top, 1 left, 2 right, 3 back, 4 front) of a flow cell (:198-274).
__device__ hcm_mesh_face flow_face(const Flow &f, int kind, uint16_t tex) {
	const uint32_t h0 = flow_h(f, 0), h1 = flow_h(f, 1), h2 = flow_h(f, 2), h3 = flow_h(f, 3);
	hcm_mesh_face m;
	m.pad = 0, m.pad2 = 0, m.texture = tex;
	switch (kind) {
	case 0:
		m.axis = 1, m.light_face = 2;
		m.v[0] = fv(0, h0, 0), m.v[1] = fv(0, h2, 16), m.v[2] = fv(16, h3, 16), m.v[3] = fv(16, h1, 0);
		break;
	case 1:
		m.axis = 0, m.light_face = 1;
		m.v[0] = fv(0, h0, 0), m.v[1] = fv(0, 0, 0), m.v[2] = fv(0, 0, 16), m.v[3] = fv(0, h2, 16);
		break;
	case 2:
		m.axis = 0, m.light_face = 0;
		m.v[0] = fv(16, 0, 0), m.v[1] = fv(16, h1, 0), m.v[2] = fv(16, h3, 16), m.v[3] = fv(16, 0, 16);
		break;
	case 3:
		m.axis = 2, m.light_face = 5;
		m.v[0] = fv(0, 0, 0), m.v[1] = fv(0, h0, 0), m.v[2] = fv(16, h1, 0), m.v[3] = fv(16, 0, 0);
		break;
	default:
		m.axis = 2, m.light_face = 4;
		m.v[0] = fv(0, h2, 16), m.v[1] = fv(0, 0, 16), m.v[2] = fv(16, 0, 16), m.v[3] = fv(16, h3, 16);
		break;
	}
	m.render_face = m.light_face;
	return m;
}

// ------------------------------------------------------------------------------------------------
// custom meshes (:198-306)
// ------------------------------------------------------------------------------------------------
struct CustomCell {
	uint32_t count;     // faces emitted
	uint32_t trans;     // bit j: face j goes to the transparent pass
	uint32_t mesh;      // registry mesh index
	uint32_t flow;      // 1 if liquid flow
	Flow fl;
	uint32_t flow_set; // liquid flow: emitted face kinds
};

__device__ __forceinline__ bool tex_trans_pass(const MeshParams &p, uint16_t t) {
	uint32_t id = t & 0x1fffu;
	return id < p.reg.n_tex && (p.reg.tex_flags[id] & HCM_TEX_TRANS_PASS);
}

template <int NT>
__device__ __forceinline__ CustomCell custom_cell(const MeshParams &p, const Smem<NT> &s, int cell) {
	const int x = cell & 15, z = (cell >> 4) & 15, y = cell >> 8;
	const uint16_t b = p.blocks[(size_t)s.slots[26] * HCM_CHUNK_VOXELS + (uint32_t)cell];
	const uint2 row = s.rows[p.reg.pid_of[b]];
	CustomCell c;
	c.mesh = row.y >> 24;
	c.flow = (row.y >> 16) & kRowLiquidFlow ? 1u : 0u;
	c.trans = 0;
	c.flow_set = 0;
	const hcm_custom_mesh &m = p.reg.meshes[c.mesh];
	if (c.flow) {
		c.fl = flow_heights(p, s.slots, x, y, z, m.flow_height);
		c.flow_set = flow_face_set(c.fl);
		c.count = __popc(c.flow_set);
		if (tex_trans_pass(p, uint16_t(p.reg.flow_tex)))
			c.trans = (1u << c.count) - 1u;
	} else {
		c.count = m.face_count > HCM_MAX_FACES ? HCM_MAX_FACES : m.face_count;
		for (uint32_t j = 0; j < c.count; ++j)
			if (tex_trans_pass(p, m.faces[j].texture))
				c.trans |= 1u << j;
	}
	return c;
}

__device__ __forceinline__ hcm_mesh_face custom_face(const MeshParams &p, const CustomCell &c, int j) {
	if (!c.flow)
		return p.reg.meshes[c.mesh].faces[j];
	uint32_t set = c.flow_set;
	for (int i = 0; i < j; ++i)
		set &= set - 1u;
	return flow_face(c.fl, __ffs(set) - 1, uint16_t(p.reg.flow_tex));
}

// ------------------------------------------------------------------------------------------------
// faces of one (axis, slice, direction) task: one wavefront, 4 rounds of 4 mask rows
// ------------------------------------------------------------------------------------------------
template <int F, int NT> __device__ void face_task(Smem<NT> &s, int t, int sl) {
	constexpr int A = F >> 1, INV = F & 1;
	constexpr int SA = axis_stride(A), SU = axis_stride(u_axis(A)), SV = axis_stride(v_axis(A));
	const int lane = threadIdx.x & 63;
	const int r = lane >> 4, v = lane & 15;
	const int layer = INV ? sl : sl - 1; // the block owning the face (:330-339)
	uint32_t prev_lo = 0, prev_hi = 0;
	bool prev_vis = false;
	uint64_t prev_vt = 0;
	for (int q = 0; q < 4; ++q) {
		const int u = 4 * q + r;
		const int self = (layer + 1) * SA + (u + 1) * SU + (v + 1) * SV;
		const uint32_t vs = s.vox[self], vo = s.vox[self + (INV ? -SA : SA)];
		const uint32_t tcs = row_code(s, vs >> 16, F), tco = row_code(s, vo >> 16, F ^ 1);
		const bool vis = show(s, tcs, tco);
		const uint64_t vis_b = ballot64(vis);
		uint32_t lo = 0, hi = 0;
		uint64_t eh_b = 0, ev_b = 0, tp_b = 0;
		if (vis_b) {
			if (vis) {
				uint2 l = light4<F>(s.vox, self);
				lo = l.x;
				hi = l.y | (tcs << 24);
			}
			const uint32_t rlo = __shfl_down(lo, 1, 16), rhi = __shfl_down(hi, 1, 16);
			const uint32_t dlo = __shfl_down(lo, 16), dhi = __shfl_down(hi, 16);
			eh_b = ballot64(vis && v < 15 && lo == rlo && hi == rhi);
			ev_b = ballot64(vis && r < 3 && lo == dlo && hi == dhi);
			tp_b = ballot64(vis && ((s.codes[tcs] >> 16) & kCodeTransPass));
		}
		if (q > 0) {
			// complete the previous round's last row against this round's first row
			const uint32_t ulo = __shfl(lo, v), uhi = __shfl(hi, v);
			prev_vt |= ballot64(prev_vis && r == 3 && prev_lo == ulo && prev_hi == uhi);
			if (lane == 0)
				*reinterpret_cast<uint64_t *>(&s.ev[t][4 * (q - 1)]) = prev_vt;
		}
		if (lane == 0) {
			*reinterpret_cast<uint64_t *>(&s.ne[t][4 * q]) = vis_b;
			*reinterpret_cast<uint64_t *>(&s.eh[t][4 * q]) = eh_b;
			*reinterpret_cast<uint64_t *>(&s.tp[t][4 * q]) = tp_b;
		}
		prev_lo = lo, prev_hi = hi, prev_vis = vis, prev_vt = ev_b;
	}
	if (lane == 0)
		*reinterpret_cast<uint64_t *>(&s.ev[t][12]) = prev_vt;
}

template <int NT> __device__ void face_tasks(Smem<NT> &s) {
	const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
	for (int t = wave; t < kTasks; t += NT / 64) {
		const int A = t / 34, rem = t % 34, sl = rem >> 1, inv = rem & 1;
		if ((sl == 0 && !inv) || (sl == 16 && inv)) { // no block owns these faces (:330, :335)
			if (lane < 16) {
				s.ne[t][lane] = 0, s.eh[t][lane] = 0, s.ev[t][lane] = 0, s.tp[t][lane] = 0;
			}
			continue;
		}
		switch (2 * A + inv) {
		case 0:
			face_task<0, NT>(s, t, sl);
			break;
		case 1:
			face_task<1, NT>(s, t, sl);
			break;
		case 2:
			face_task<2, NT>(s, t, sl);
			break;
		case 3:
			face_task<3, NT>(s, t, sl);
			break;
		case 4:
			face_task<4, NT>(s, t, sl);
			break;
		default:
			face_task<5, NT>(s, t, sl);
			break;
		}
	}
}

// ------------------------------------------------------------------------------------------------
// greedy merge of one task on its bit masks (:357-476): one lane per task
// ------------------------------------------------------------------------------------------------
struct Quad {
	int u, v0, w, h;
	bool trans;
};

template <int NT, typename Fn> __device__ __forceinline__ void greedy(Smem<NT> &s, int t, Fn &&on_quad) {
	for (int u = 0; u < 16; ++u)
		s.cov[t][u] = 0;
	for (int u = 0; u < 16; ++u) {
		const uint32_t ne = s.ne[t][u], eh = s.eh[t][u], tp = s.tp[t][u];
		for (;;) {
			const uint32_t c = s.cov[t][u];
			const uint32_t avail = ne & ~c;
			if (!avail)
				break;
			const int v0 = __ffs(avail) - 1;
			// width: extend along V while the next cell equals and is not yet covered (:364-368)
			const uint32_t run = eh & ~(c >> 1);
			const int w = __ffs(~(run >> v0)) ; // 1 + number of consecutive ones from v0
			const uint32_t M = ((1u << w) - 1u) << v0;
			// height: extend along U while the whole row segment equals and is uncovered (:371-379)
			int h = 1;
			while (u + h < 16) {
				if ((s.ev[t][u + h - 1] & M) != M || (s.cov[t][u + h] & M))
					break;
				++h;
			}
			for (int r2 = u; r2 < u + h; ++r2)
				s.cov[t][r2] |= M;
			on_quad(Quad{u, v0, w, h, ((tp >> v0) & 1u) != 0});
		}
	}
}

// ------------------------------------------------------------------------------------------------
// writing one quad
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void put_quad(const MeshParams &p, uint32_t slot, const uint32_t w0[4], const uint32_t w1[4],
                                         uint32_t base, int pattern) {
	uint4 *vp = reinterpret_cast<uint4 *>(p.pool_verts + 4 * (size_t)slot);
	vp[0] = make_uint4(w0[0], w1[0], w0[1], w1[1]);
	vp[1] = make_uint4(w0[2], w1[2], w0[3], w1[3]);
	// index patterns: 0 = flip (0,1,2, 0,2,3) (:440-446) and custom (:295-301); 1 = no flip (1,2,3, 0,1,3)
	uint32_t i0, i1, i2, i3, i4, i5;
	if (pattern == 0)
		i0 = 0, i1 = 1, i2 = 2, i3 = 0, i4 = 2, i5 = 3;
	else
		i0 = 1, i1 = 2, i2 = 3, i3 = 0, i4 = 1, i5 = 3;
	uint32_t *ip = reinterpret_cast<uint32_t *>(p.pool_idx + 6 * (size_t)slot);
	ip[0] = ((base + i0) & 0xffffu) | (((base + i1) & 0xffffu) << 16);
	ip[1] = ((base + i2) & 0xffffu) | (((base + i3) & 0xffffu) << 16);
	ip[2] = ((base + i4) & 0xffffu) | (((base + i5) & 0xffffu) << 16);
}

// BlockVertex word 1 (client/include/client/BlockVertex.hpp:13-14)
__device__ __forceinline__ uint32_t vword1(uint16_t tex, uint32_t face, uint32_t ao, uint32_t sun, uint32_t torch) {
	return ((uint32_t)(tex & 0x1fffu) - 1u) | ((uint32_t)(tex >> 13) << 10) | (face << 13) | (ao << 16) | (sun << 18) |
	       (torch << 24);
}

template <int NT> __device__ __forceinline__ void aabb_merge(Smem<NT> &s, int pass, int k, uint32_t lo[3], uint32_t hi[3]) {
	uint32_t *a = s.aabb[pass][k];
	atomicMin(&a[0], lo[0]);
	atomicMin(&a[1], lo[1]);
	atomicMin(&a[2], lo[2]);
	atomicMax(&a[3], hi[0]);
	atomicMax(&a[4], hi[1]);
	atomicMax(&a[5], hi[2]);
}

// Mesh split bookkeeping for quad i of a pass: the vertex base its indices use (the reference reads the
// 16-bit vertex count before splitting, so the first quad of a continuation mesh keeps base 65,532).
__device__ __forceinline__ uint32_t index_base(uint32_t i) {
	uint32_t k = i / kSplitQuads, j = i % kSplitQuads;
	return (j == 0 && k > 0) ? 65532u : 4u * j;
}

template <int F, int NT>
__device__ __forceinline__ void emit_regular_f(const MeshParams &p, Smem<NT> &s, uint32_t desc, uint32_t slot, uint32_t pass,
                               uint32_t i) {
	constexpr int A = F >> 1, INV = F & 1, UA = u_axis(A), VA = v_axis(A);
	// (U+1)%3 == V holds for the x and y axes (:384)
	constexpr bool kHeightOnU = (INV != 0) != (A != 2);
	const int t = desc & 127, u = (desc >> 7) & 15, v0 = (desc >> 11) & 15;
	const int w = int((desc >> 15) & 15) + 1, h = int((desc >> 19) & 15) + 1;
	const int sl = (t % 34) >> 1, layer = INV ? sl : sl - 1;
	// block owning the face, and the quad origin on the slice plane (:382)
	const int self = (layer + 1) * axis_stride(A) + (u + 1) * axis_stride(UA) + (v0 + 1) * axis_stride(VA);
	const uint2 l = light4<F>(s.vox, self);
	const uint32_t tc = row_code(s, s.vox[self] >> 16, F);
	const uint16_t tex = uint16_t(s.codes[tc] & 0xffffu);
	auto comp = [&](int axis, int ku, int kv) -> uint32_t {
		// origin + ku * u-vector + kv * v-vector along one axis (:383-390)
		int c = axis == A ? sl : (axis == UA ? u : v0);
		if (kHeightOnU) {
			c += (axis == UA ? ku * h : 0) + (axis == VA ? kv * w : 0);
		} else {
			c += (axis == VA ? ku * w : 0) + (axis == UA ? kv * h : 0);
		}
		return uint32_t(c);
	};
	uint32_t w0[4], w1[4];
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		const int ku = (k == 1 || k == 2), kv = (k == 2 || k == 3);
		w0[k] = (comp(0, ku, kv) << 4) | ((comp(1, ku, kv) << 4) << 10) | ((comp(2, ku, kv) << 4) << 20) |
		        (uint32_t(A) << 30);
		w1[k] = vword1(tex, uint32_t(F), l4_ao(l, k), l4_sun(l, k), l4_torch(l, k));
	}
	put_quad(p, slot, w0, w1, index_base(i), l4_flip(l) ? 0 : 1);
	uint32_t lo[3], hi[3];
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		lo[a] = comp(a, 0, 0) << 4;
		hi[a] = comp(a, 1, 1) << 4;
	}
	aabb_merge(s, pass, i / kSplitQuads, lo, hi);
}

template <int NT>
__device__ __forceinline__ void emit_regular(const MeshParams &p, Smem<NT> &s, uint32_t desc, uint32_t slot, uint32_t pass, uint32_t i) {
	const int t = desc & 127, face = 2 * (t / 34) + (t & 1);
	switch (face) {
	case 0:
		emit_regular_f<0>(p, s, desc, slot, pass, i);
		break;
	case 1:
		emit_regular_f<1>(p, s, desc, slot, pass, i);
		break;
	case 2:
		emit_regular_f<2>(p, s, desc, slot, pass, i);
		break;
	case 3:
		emit_regular_f<3>(p, s, desc, slot, pass, i);
		break;
	case 4:
		emit_regular_f<4>(p, s, desc, slot, pass, i);
		break;
	default:
		emit_regular_f<5>(p, s, desc, slot, pass, i);
		break;
	}
}

__device__ __forceinline__ uint32_t lerp16(uint32_t a, uint32_t b, uint32_t t) { return a * (16u - t) + b * t; }
__device__ __forceinline__ uint32_t interp(uint32_t l0, uint32_t l1, uint32_t l2, uint32_t l3, uint32_t h0, uint32_t h1,
                                           uint32_t h2, uint32_t h3, uint32_t du, uint32_t dv, uint32_t dw) {
	// light4_interpolate (:165-196): trilinear in 1/16 steps, rounded up
	uint32_t lo = lerp16(lerp16(l0, l1, du), lerp16(l3, l2, du), dv);
	uint32_t hi = lerp16(lerp16(h0, h1, du), lerp16(h3, h2, du), dv);
	uint32_t sum = lerp16(lo, hi, dw);
	return sum / 4096u + (sum % 4096u ? 1u : 0u);
}

template <int NT>
__device__ __forceinline__ void emit_custom(const MeshParams &p, Smem<NT> &s, uint32_t desc, uint32_t slot, uint32_t pass, uint32_t i) {
	const int cell = desc & 4095, j = (desc >> 12) & 15;
	const int x = cell & 15, z = (cell >> 4) & 15, y = cell >> 8;
	const CustomCell c = custom_cell(p, s, cell);
	const hcm_mesh_face f = custom_face(p, c, j);
	const int lf = f.light_face, la = lf >> 1, sg = (lf & 1) ? -1 : 1;
	const int pos = hidx(x, y, z);
	// low: the face of the cell behind; high: this cell's own face if the next cell passes light (:259-267)
	const uint2 lo4 = light4_rt(s.vox, lf, pos - sg * axis_stride(la));
	const uint2 hi4 = ((s.vox[pos + sg * axis_stride(la)] >> 9) & 1u) ? light4_rt(s.vox, lf, pos) : lo4;
	int ua = (la + 1) % 3, va = (la + 2) % 3;
	if (lf & 1) {
		int tmp = ua;
		ua = va, va = tmp;
	}
	const uint32_t bx = uint32_t(x) << 4, by = uint32_t(y) << 4, bz = uint32_t(z) << 4;
	uint32_t w0[4], w1[4], lo[3] = {~0u, ~0u, ~0u}, hi[3] = {0, 0, 0};
#pragma unroll
	for (int k = 0; k < 4; ++k) {
		const uint32_t vx = f.v[k].x, vy = f.v[k].y, vz = f.v[k].z;
		auto pick = [&](int a) { return a == 0 ? vx : (a == 1 ? vy : vz); };
		const uint32_t px = bx + vx, py = by + vy, pz = bz + vz;
		lo[0] = min(lo[0], px), lo[1] = min(lo[1], py), lo[2] = min(lo[2], pz);
		hi[0] = max(hi[0], px), hi[1] = max(hi[1], py), hi[2] = max(hi[2], pz);
		const uint32_t du = pick(ua), dv = pick(va);
		uint32_t dw = pick(la) < 16u ? pick(la) : 16u;
		if (lf & 1)
			dw = 16u - dw;
		uint32_t ao = f.v[k].ao;
		if (ao == 4)
			ao = interp(l4_ao(lo4, 0), l4_ao(lo4, 1), l4_ao(lo4, 2), l4_ao(lo4, 3), l4_ao(hi4, 0), l4_ao(hi4, 1),
			            l4_ao(hi4, 2), l4_ao(hi4, 3), du, dv, dw) & 0xffu;
		const uint32_t sun = interp(l4_sun(lo4, 0), l4_sun(lo4, 1), l4_sun(lo4, 2), l4_sun(lo4, 3), l4_sun(hi4, 0),
		                            l4_sun(hi4, 1), l4_sun(hi4, 2), l4_sun(hi4, 3), du, dv, dw) & 0xffu;
		const uint32_t torch = interp(l4_torch(lo4, 0), l4_torch(lo4, 1), l4_torch(lo4, 2), l4_torch(lo4, 3),
		                              l4_torch(hi4, 0), l4_torch(hi4, 1), l4_torch(hi4, 2), l4_torch(hi4, 3), du, dv,
		                              dw) & 0xffu;
		w0[k] = px | (py << 10) | (pz << 20) | (uint32_t(f.axis) << 30);
		w1[k] = vword1(f.texture, f.render_face, ao, sun, torch);
	}
	put_quad(p, slot, w0, w1, index_base(i), 0);
	aabb_merge(s, pass, i / kSplitQuads, lo, hi);
}

// ------------------------------------------------------------------------------------------------
// one chunk
// ------------------------------------------------------------------------------------------------
template <int NT> __device__ __forceinline__ void mesh_chunk(const MeshParams &p, Smem<NT> &s, uint32_t job) {
	const int tid = threadIdx.x;
	constexpr int CPT = HCM_CHUNK_VOXELS / NT; // custom cells per thread
	if (tid < 27)
		s.slots[tid] = p.nbr[27 * (size_t)job + tid];
	for (int i = tid; i < 128; i += NT)
		s.custom_bits[i] = 0;
	for (int i = tid; i < (kHaloVox + 31) / 32; i += NT)
		s.flow_bits[i] = 0;
	if (tid < 2 * kMaxMeshPerPass * 6) {
		uint32_t *a = &s.aabb[0][0][0];
		a[tid] = (tid % 6) < 3 ? 0xffffffffu : 0u; // AABB<uint32_t>{}: min = max(), max = min() (dep/AABB.hpp:11)
	}
	if (tid == 0)
		s.n_events = 0;
	__syncthreads();

	// 1. halo staging
	const uint16_t ref_block = p.blocks[(size_t)s.slots[26] * HCM_CHUNK_VOXELS];
	uint32_t my_flags = 0;
	for (int i = tid; i < kHaloVox; i += NT) {
		const int hx = i % kHalo, rest = i / kHalo, hz = rest % kHalo, hy = rest / kHalo;
		const int x = hx - 1, y = hy - 1, z = hz - 1;
		const size_t g = (size_t)s.slots[slot27(x, y, z)] * HCM_CHUNK_VOXELS + local16(x, y, z);
		const uint16_t b = p.blocks[g];
		const uint32_t lt = p.light[g];
		const uint32_t pid = p.reg.pid_of[b];
		const uint32_t rf = (s.rows[pid].y >> 16) & 0xffu;
		s.vox[i] = lt | ((rf & kRowDirectPass) << 8) | (((rf & kRowIndirectPass) >> 1) << 9) | (pid << 16);
		if (b != ref_block)
			my_flags |= kFlagNonUniform;
		if (rf & kRowHasMesh) {
			const bool inside = (unsigned)x < 16u && (unsigned)y < 16u && (unsigned)z < 16u;
			if (inside) {
				const uint32_t cell = local16(x, y, z);
				atomicOr(&s.custom_bits[cell >> 5], 1u << (cell & 31));
				my_flags |= kFlagCustom;
			}
			if (rf & kRowLiquidFlow) {
				atomicOr(&s.flow_bits[i >> 5], 1u << (i & 31));
				my_flags |= kFlagFlow;
			}
		}
	}
	if (tid == 0)
		s.flags = 0;
	__syncthreads();
	if (my_flags)
		atomicOr(&s.flags, my_flags);
	__syncthreads();
	const uint32_t flags = s.flags;
	if (!(flags & kFlagNonUniform) && !(flags & kFlagCustom)) {
		// one block value everywhere: every face is hidden by an equal neighbour (BlockTexture::Show)
		if (tid == 0)
			p.results[job] = hcm_job_result{HCM_OK, 0u, 0u, 0u};
		return;
	}

	// 2. liquid-flow cells override their face textures with the flow mesh's textures (:221-236)
	if (flags & kFlagFlow) {
		for (int i = tid; i < kHaloVox; i += NT) {
			if (!((s.flow_bits[i >> 5] >> (i & 31)) & 1u))
				continue;
			const int hx = i % kHalo, rest = i / kHalo, hz = rest % kHalo, hy = rest / kHalo;
			const int x = hx - 1, y = hy - 1, z = hz - 1;
			const uint32_t pid = s.vox[i] >> 16;
			const uint32_t mesh = s.rows[pid].y >> 24;
			const Flow f = flow_heights(p, s.slots, x, y, z, p.reg.meshes[mesh].flow_height);
			s.vox[i] = (s.vox[i] & 0xffffu) | ((p.reg.flow_pid_base + flow_pattern(f)) << 16);
		}
		__syncthreads();
	}

	// 3. visible faces, light words and merge masks
	face_tasks(s);

	// 4. counts: custom faces per thread range of cells, greedy quads per task
	uint32_t c_o = 0, c_t = 0;
	if (flags & kFlagCustom) {
		for (int k = 0; k < CPT; ++k) {
			const int cell = tid * CPT + k;
			if (!((s.custom_bits[cell >> 5] >> (cell & 31)) & 1u))
				continue;
			const CustomCell c = custom_cell(p, s, cell);
			const uint32_t nt = __popc(c.trans);
			c_t += nt;
			c_o += c.count - nt;
		}
	}
	__syncthreads(); // masks complete
	uint32_t q_o = 0, q_t = 0;
	if (tid < kTasks)
		greedy(s, tid, [&](const Quad &q) { (q.trans ? q_t : q_o) += 1; });
	uint64_t cust_total, task_total;
	const uint64_t cust_pre = block_excl_scan(s, (uint64_t)c_o | ((uint64_t)c_t << 32), &cust_total);
	const uint64_t task_pre = block_excl_scan(s, (uint64_t)q_o | ((uint64_t)q_t << 32), &task_total);
	const uint32_t cust0 = (uint32_t)cust_total, cust1 = (uint32_t)(cust_total >> 32);
	const uint32_t n0 = cust0 + (uint32_t)task_total, n1 = cust1 + (uint32_t)(task_total >> 32);
	const uint32_t total = n0 + n1;
	if (total == 0) {
		if (tid == 0)
			p.results[job] = hcm_job_result{HCM_OK, 0u, 0u, 0u};
		return;
	}

	// 5. allocation: the chunk's quads [pool_base, +n0) opaque, [+n0, +n0+n1) transparent
	if (tid == 0) {
		const uint32_t nm0 = (n0 + kSplitQuads - 1) / kSplitQuads, nm1 = (n1 + kSplitQuads - 1) / kSplitQuads;
		uint32_t status = HCM_OK;
		const uint32_t base = atomicAdd(&p.ctr->pool_cursor, total);
		if ((uint64_t)base + total > p.pool_cap)
			status = (uint32_t)HCM_ERR_POOL_FULL;
		const uint32_t mbase = atomicAdd(&p.ctr->mesh_cursor, nm0 + nm1);
		if ((uint64_t)mbase + nm0 + nm1 > p.rec_cap)
			status = (uint32_t)HCM_ERR_POOL_FULL;
		s.status = status;
		s.pool_base = base;
		s.mesh_base = mbase;
		s.n_mesh0 = nm0;
		s.n_mesh1 = nm1;
		if (status != HCM_OK)
			p.results[job] = hcm_job_result{(int32_t)status, 0u, 0u, total};
	}
	__syncthreads();
	if (s.status != HCM_OK)
		return;
	const uint32_t pool_base = s.pool_base;

	// 6. describe quads into the LDS list window by window, then write them one lane per quad
	const uint32_t task_o = cust0 + (uint32_t)task_pre, task_t = cust1 + (uint32_t)(task_pre >> 32);
	const uint32_t task_g = cust0 + cust1 + (uint32_t)task_pre + (uint32_t)(task_pre >> 32);
	for (uint32_t w0 = 0; w0 < total; w0 += kListCap) {
		const uint32_t w1 = w0 + kListCap;
		if (flags & kFlagCustom) {
			uint32_t io = (uint32_t)cust_pre, it = (uint32_t)(cust_pre >> 32), g = io + it;
			if (g < w1) {
				for (int k = 0; k < CPT; ++k) {
					const int cell = tid * CPT + k;
					if (!((s.custom_bits[cell >> 5] >> (cell & 31)) & 1u))
						continue;
					const CustomCell c = custom_cell(p, s, cell);
					for (uint32_t j = 0; j < c.count; ++j, ++g) {
						const uint32_t tr = (c.trans >> j) & 1u, i = tr ? it++ : io++;
						if (g >= w0 && g < w1)
							s.list[g - w0] = make_uint2(0x80000000u | uint32_t(cell) | (j << 12), i | (tr << 31));
					}
				}
			}
		}
		if (tid < kTasks) {
			uint32_t io = task_o, it = task_t, g = task_g;
			const uint32_t my_end = task_g + q_o + q_t;
			if (g < w1 && my_end > w0)
				greedy(s, tid, [&](const Quad &q) {
					const uint32_t i = q.trans ? it++ : io++;
					if (g >= w0 && g < w1)
						s.list[g - w0] = make_uint2(uint32_t(tid) | (uint32_t(q.u) << 7) | (uint32_t(q.v0) << 11) |
						                                (uint32_t(q.w - 1) << 15) | (uint32_t(q.h - 1) << 19),
						                            i | (uint32_t(q.trans) << 31));
					++g;
				});
		}
		__syncthreads();
		const uint32_t n = min(total - w0, (uint32_t)kListCap);
		for (uint32_t e = tid; e < n; e += NT) {
			const uint2 ent = s.list[e];
			const uint32_t pass = ent.y >> 31, i = ent.y & 0x7fffffffu;
			const uint32_t slot = pool_base + (pass ? n0 : 0u) + i;
			if (ent.x & 0x80000000u)
				emit_custom(p, s, ent.x, slot, pass, i);
			else
				emit_regular(p, s, ent.x, slot, pass, i);
			if (i > 0 && i % kSplitQuads == 0) { // a split happened right before this quad (:397-403)
				const uint32_t ev = atomicAdd(&s.n_events, 1u);
				if (ev < 16) {
					s.events[ev][0] = w0 + e;
					s.events[ev][1] = (pass << 16) | (i / kSplitQuads - 1u);
				}
			}
		}
		__syncthreads();
	}

	// 7. mesh records in the reference's order: split meshes by time, then opaque, then transparent
	if (tid == 0) {
		const uint32_t ne = min(s.n_events, 16u);
		for (uint32_t a = 1; a < ne; ++a) // insertion sort by emission position
			for (uint32_t b = a; b > 0 && s.events[b - 1][0] > s.events[b][0]; --b) {
				uint32_t t0 = s.events[b][0], t1 = s.events[b][1];
				s.events[b][0] = s.events[b - 1][0], s.events[b][1] = s.events[b - 1][1];
				s.events[b - 1][0] = t0, s.events[b - 1][1] = t1;
			}
		const uint32_t nm0 = s.n_mesh0, nm1 = s.n_mesh1, mbase = s.mesh_base;
		uint32_t m = 0;
		auto put = [&](uint32_t pass, uint32_t k) {
			const uint32_t n_p = pass ? n1 : n0;
			hcm_mesh_rec r;
			for (int a = 0; a < 6; ++a)
				r.aabb[a] = s.aabb[pass][k][a];
			r.first_quad = pool_base + (pass ? n0 : 0u) + k * kSplitQuads;
			r.quad_count = min(kSplitQuads, n_p - k * kSplitQuads);
			r.job = job;
			r.transparent = pass;
			p.recs[mbase + m++] = r;
		};
		for (uint32_t e = 0; e < ne; ++e)
			put(s.events[e][1] >> 16, s.events[e][1] & 0xffffu);
		if (nm0)
			put(0, nm0 - 1);
		if (nm1)
			put(1, nm1 - 1);
		p.results[job] = hcm_job_result{HCM_OK, mbase, m, total};
	}
}

template <int NT> __global__ __launch_bounds__(NT) void mesh_kernel(MeshParams p) {
	__shared__ Smem<NT> s;
	for (int i = threadIdx.x; i < kMaxPid; i += NT)
		s.rows[i] = i < (int)p.reg.n_pid ? p.reg.rows[i] : make_uint2(0, 0);
	for (int i = threadIdx.x; i < kMaxCodes; i += NT)
		s.codes[i] = i < (int)p.reg.n_codes ? p.reg.codes[i] : (kCodeEmpty | kCodeTransparent) << 16;
	if (!p.persistent) {
		__syncthreads();
		const uint32_t job = p.job_base + blockIdx.x;
		if (job < p.n_jobs)
			mesh_chunk(p, s, job);
		return;
	}
	for (;;) {
		if (threadIdx.x == 0)
			s.job = atomicAdd(&p.ctr->queue, 1u);
		__syncthreads();
		const uint32_t job = s.job;
		if (job >= p.n_jobs)
			break;
		mesh_chunk(p, s, job);
		__syncthreads();
	}
}

__global__ void reset_kernel(Counters *ctr) {
	if (threadIdx.x == 0) {
		ctr->queue = 0;
		ctr->pool_cursor = 0;
		ctr->mesh_cursor = 0;
		ctr->pad = 0;
	}
}

hipError_t launch_reset(Counters *ctr, hipStream_t stream) {
	hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, stream, ctr);
	return hipGetLastError();
}

hipError_t launch_mesh(const MeshParams &p, int threads, int grid, hipStream_t stream) {
	switch (threads) {
	case 128:
		hipLaunchKernelGGL(mesh_kernel<128>, dim3(grid), dim3(128), 0, stream, p);
		break;
	case 256:
		hipLaunchKernelGGL(mesh_kernel<256>, dim3(grid), dim3(256), 0, stream, p);
		break;
	case 512:
		hipLaunchKernelGGL(mesh_kernel<512>, dim3(grid), dim3(512), 0, stream, p);
		break;
	default:
		return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

int mesh_lds_bytes(int threads) {
	switch (threads) {
	case 128:
		return (int)sizeof(Smem<128>);
	case 256:
		return (int)sizeof(Smem<256>);
	case 512:
		return (int)sizeof(Smem<512>);
	default:
		return 0;
	}
}

} // namespace hcm
