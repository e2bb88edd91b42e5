// hcm_kernels.hip -- gfx950 chunk-meshing kernels.
//
// One workgroup meshes one 16^3 chunk at a time; a persistent workgroup takes chunk jobs from work
// queues (one per XCD, then the others') until they are empty. Per chunk:
//   1. staging: one thread per 18-voxel x-row of the 18^3 halo loads the row from its chunk slots with
//      vector loads and writes one LDS word per voxel: light byte, direct/indirect light-pass bits,
//      a 2-bit face class per face (empty / opaque / transparent / liquid), custom-mesh and flow bits
//      and the texture row. A chunk without a custom mesh whose halo holds a single block value (whose
//      opposite faces hide each other) or only opaque cubes, or none of whose own voxels has a face, has
//      no quads and stops here; so does one
//      with no visible face after step 3's visibility masks.
//   2. liquid-flow cells take their per-cell face textures (the reference's m_dynamic_block_textures).
//   3. faces, per (axis, slice, direction) task, each a 16x16 mask:
//      - visibility, bit-sliced: one lane per column of an axis keeps its cells' face classes as bit
//        planes, so BlockTexture::Show is a few bit operations (texture codes are compared only for
//        transparent against transparent); ballots transpose the columns into the tasks' mask rows;
//      - face words: visible faces are numbered densely in the reference's scan order and one lane per
//        face evaluates light4_init; the face word is Light4 plus texture code and pass;
//      - link masks: each face is compared with its right and lower neighbours, giving the
//        equal-to-right, equal-to-below and transparent-pass masks.
//   4. greedy merge: one lane per task replays the reference's merge row by row on the bit masks:
//      the quads starting in a row are the row's uncovered visible runs split where the face word
//      changes; each extends downward while every column of it continues (equal and uncovered).
//   5. block-wide prefix sums give every quad its index in its pass and its position in the
//      reference's emission order; one atomic per chunk allocates its output range; quads are then
//      described into an LDS list and written one lane per quad as packed BlockVertex words and u16
//      indices, with the reference's 65,532-vertex mesh split, per-mesh AABBs and mesh order.
// All arithmetic is integer. References are to client/include/client/BlockMeshAlgo.hpp unless noted.
#include <hip/hip_runtime.h>

#include "hcm_internal.h"

namespace hcm {

constexpr int kHalo = 18;
constexpr int kHaloVox = kHalo * kHalo * kHalo;
constexpr int kHaloRows = kHalo * kHalo;
constexpr int kTasks = 102; // 3 axes x 17 slices x 2 directions
// Entries of the face-word stage. 1,472 keeps the 256-thread workgroup's LDS within three per CU
// (static_assert below); the tasks of a chunk with more visible faces are taken in several groups.
constexpr int kListCap = 1472;
constexpr uint32_t kSplitQuads = 16383u; // a BlockMesh holds at most 65,532 vertices (:273-279, :397-403)
constexpr int kMaxMeshPerPass = 8;

// Halo layout: x fastest, then z, then y -- the chunk-local order x + 16 * (z + 16 * y) with a 1-voxel rim.
constexpr int kSX = 1, kSZ = kHalo, kSY = kHalo * kHalo;
__host__ __device__ constexpr int axis_stride(int a) { return a == 0 ? kSX : (a == 1 ? kSY : kSZ); }
// Mask axes per slice axis (BlockAlgo.hpp:70-81 with the YZX swizzle): x -> (U=y, V=z), y -> (U=z, V=x),
// z -> (U=y, V=x).
__host__ __device__ constexpr int u_axis(int a) { return a == 1 ? 2 : 1; }
__host__ __device__ constexpr int v_axis(int a) { return a == 0 ? 2 : 0; }

// The light4_init cell pattern (kLookup3v, :75-108). Vertex k of face F reads, in the plane of the front
// cell, a side cell c0 offset along plane axis Q, the corner cell c1 offset along P and Q, and a side cell
// c2 offset along P; (P, Q) = (y, z) for x faces, (x, z) for y faces, (x, y) for z faces. The signs along
// (P, Q) are (-,-), (+,-), (+,+), (-,+) for faces 0, 3, 4 ("winding a") and (-,-), (-,+), (+,+), (+,-) for
// faces 1, 2, 5. The static_assert below checks this against the reference's literal table.
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
__host__ __device__ constexpr int corner_comp(int F, int k, int b, int axis) {
	return (axis == (F >> 1) ? normal_sign(F) : 0) + (axis == plane_p(F) && b >= 1 ? corner_sp(F, k) : 0) +
	       (axis == plane_q(F) && b <= 1 ? corner_sq(F, k) : 0);
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
// slot27's y and z terms for the halo row (hy, hz) = (y + 1, z + 1): a halo coordinate h contributes 1 at
// h = 0 (-1), 2 at h = 1 .. 16 (0) and 0 at h = 17 (16), two bits per h in kSlotTerm.
constexpr uint64_t slot_term_table() {
	uint64_t c = 0;
	for (int h = 0; h < 18; ++h)
		c |= uint64_t(h == 0 ? 1 : (h == 17 ? 0 : 2)) << (2 * h);
	return c;
}
constexpr uint64_t kSlotTerm = slot_term_table();
static_assert(kSlotTerm == 0x2aaaaaaa9ull, "slot27 halo-coordinate terms");
__device__ __forceinline__ int slot27_yz(int hy, int hz) {
	return 3 * int((kSlotTerm >> (2 * hy)) & 3u) + int((kSlotTerm >> (2 * hz)) & 3u);
}
__device__ __forceinline__ uint32_t local16(int x, int y, int z) {
	return uint32_t(x & 15) + 16u * (uint32_t(z & 15) + 16u * uint32_t(y & 15));
}
__device__ __forceinline__ int hidx(int x, int y, int z) { return (x + 1) * kSX + (y + 1) * kSY + (z + 1) * kSZ; }

template <int NT> struct Smem {
	uint32_t vox[kHaloVox];
	// per task and mask row: visible, equal-to-right, equal-to-below, transparent pass
	alignas(16) uint16_t ne[kTasks][16];
	alignas(16) uint16_t eh[kTasks][16];
	alignas(16) uint16_t ev[kTasks][16];
	alignas(16) uint16_t tp[kTasks][16];
	uint2 key[kListCap];       // face words of the current group of tasks, in dense order
	uint16_t fpre[kTasks + 2]; // prefix of visible faces per task (the dense face order; at most 24,576)
	uint32_t qpre_o[kTasks];   // per task: index of its first opaque quad in the opaque pass
	uint32_t qpre_t[kTasks];   // per task: index of its first transparent quad in the transparent pass
	uint2 rows[kMaxPid];
	uint32_t pid_vox[kMaxPid];
	uint32_t codes[kMaxCodes];
	uint32_t custom_bits[128]; // in-chunk cells with a custom mesh
	uint32_t tmpl_id[256];     // voxel-word template by block id when the meta byte does not change it, else kTmplMeta
	uint32_t aabb[2][kMaxMeshPerPass][6];
	uint32_t events[16][2];
	uint32_t scan_w[NT / 64][4]; // per wavefront: its totals of block_excl_sums
	uint32_t slots[27];
	uint32_t job, flags, n_events, n_link;
	uint32_t wave_ref[NT / 64];
	uint32_t pool_base, mesh_base, group_end;
};

constexpr uint32_t kFlagNonUniform = 1u, kFlagCustom = 2u, kFlagFlow = 4u;
// kFlagMixed: some halo voxel is not an opaque cube (opaque class on all six faces, no custom or flow
// mesh); kFlagInnerFaces: some chunk voxel has a non-empty face class or a custom or flow mesh. A chunk
// without either has no visible face: BlockTexture::Show never shows an empty face, nor an opaque face
// against an opaque one (texture/include/texture/BlockTexture.hpp:68-76).
constexpr uint32_t kFlagMixed = 8u, kFlagInnerFaces = 16u;
constexpr uint32_t kVoxShape = 0xfff00000u | kVoxMesh | kVoxFlow; // the face classes and the mesh and flow bits
constexpr uint32_t kVoxOpaqueCube = 0x555u << kVoxClassShift; // opaque on all six faces

// ------------------------------------------------------------------------------------------------
// small helpers
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t vox_pid(uint32_t vox) { return (vox >> kVoxPidShift) & kVoxPidMask; }

template <int NT> __device__ __forceinline__ uint32_t row_code(const Smem<NT> &s, uint32_t pid, int face) {
	const uint2 r = s.rows[pid];
	const uint32_t w = face < 4 ? r.x : r.y;
	return (w >> (8 * (face & 3))) & 0xffu;
}

// Texture row of a block value: from LDS for most ids, from the full table for meta-dependent ones.
// A template word flag (bit 15 of the voxel word is otherwise unused): the block id's texture row depends
// on the meta byte, so the full table decides.
constexpr uint32_t kTmplMeta = 1u << 15;

template <int NT> __device__ __forceinline__ uint32_t block_pid(const MeshParams &p, const Smem<NT> &s, uint32_t b) {
	const uint32_t q = s.tmpl_id[b & 0xffu];
	return q & kTmplMeta ? p.reg.pid_of[b] : vox_pid(q);
}

__device__ __forceinline__ uint32_t face_class(uint32_t vox, int face) {
	return (vox >> (kVoxClassShift + 2 * face)) & 3u;
}

// AO of a light4_init vertex (:126-130) by side, corner and side cell direct-pass bits and the front cell's:
// two bits per index d0 | d1 << 1 | d2 << 2 | pass3 << 3, 0 when both side cells or the corner and front
// cells block, else 3 minus the number of blocking cells.
constexpr uint32_t ao_table() {
	uint32_t t = 0;
	for (uint32_t i = 0; i < 16; ++i) {
		const int d0 = i & 1, d1 = (i >> 1) & 1, d2 = (i >> 2) & 1, p = (i >> 3) & 1;
		const int a = 3 - !d0 - !d1 - !d2 - !p;
		t |= ((!d0 && !d2) || (!d1 && !p) || a < 0 ? 0u : uint32_t(a)) << (2 * i);
	}
	return t;
}
constexpr uint32_t kAoTable = ao_table();
static_assert(kAoTable == 0xe9849040u, "AO table");

// (sum << 2) / cnt for cnt in 1..4 and sum <= 15 cnt, exactly: (sum * R) >> 7 with R = 512, 256, 171, 128,
// packed 16 bits apart by cnt - 1. Checked below for every possible sum.
constexpr uint64_t kRecip = 0x008000ab01000200ull;
constexpr bool recip_exact() {
	for (uint32_t cnt = 1; cnt <= 4; ++cnt)
		for (uint32_t sum = 0; sum <= 15 * cnt; ++sum)
			if (((sum * uint32_t((kRecip >> (16 * (cnt - 1))) & 0x3ffu)) >> 7) != (sum << 2) / cnt)
				return false;
	return true;
}
static_assert(recip_exact(), "reciprocal table");

// A cell's light in a light4_init sum: its word masked to the light fields and the count bit when it
// passes indirect light (:132-161), else 0.
__device__ __forceinline__ uint32_t lpass(uint32_t c) { return c & kVoxLight & (0u - ((c >> 12) & 1u)); }

// One vertex of light4_init (:114-162), without branches, from its side cells a and b, corner cell c and
// the front cell: the side cells count when they pass indirect light, the corner when it does and a side
// cell does; sunlight and torchlight are averaged as (sum << 2) / count, the front cell always counted.
// Returns sun | torch << 8 | ao << 16.
__device__ __forceinline__ uint32_t light_vertex(uint32_t a, uint32_t c, uint32_t b, uint32_t pf, uint32_t d3) {
	const uint32_t g = 0u - (((a | b) >> 12) & 1u);
	const uint32_t sum = pf + lpass(a) + lpass(b) + (lpass(c) & g); // count - 1 in bits 12-13
	const uint32_t r = uint32_t(kRecip >> ((sum >> 8) & 0x30u)) & 0x3ffu;
	const uint32_t prod = (sum & 0x003f003fu) * r;
	const uint32_t idx = ((a >> 11) & 1u) | ((c >> 10) & 2u) | ((b >> 9) & 4u) | d3;
	return ((prod >> 7) & 63u) | (((prod >> 23) & 63u) << 8) | (((kAoTable >> (2 * idx)) & 3u) << 16);
}

// light4_init (:48-163) for face `face` of the block at halo index self. Returns the packed Light4:
//   x = sun0 | sun1 << 6 | sun2 << 12 | sun3 << 18 | ao4 << 24,   y = torch0 | ... | torch3 << 18.
// The front cell F and its eight neighbours in the face's plane; vertex k's side cells along the plane
// axes Q and P and the corner between them have (P, Q) signs (-,-), then (+,-) or (-,+) by the winding,
// then (+,+), then the other of (+,-) and (-,+).
__device__ __forceinline__ uint2 light4(const uint32_t *vox, int face, int self) {
	const int a = face >> 1;
	const int sn = a == 0 ? kSX : (a == 1 ? kSY : kSZ);
	const int sp = a == 0 ? kSY : kSX; // plane axis P
	const int sq = a == 2 ? kSY : kSZ; // plane axis Q
	const int f = self + ((face & 1) ? -sn : sn);
	const uint32_t F = vox[f];
	const uint32_t qm = vox[f - sq], qp = vox[f + sq], pm = vox[f - sp], pp = vox[f + sp];
	const uint32_t mm = vox[f - sp - sq], mp = vox[f - sp + sq], pmq = vox[f + sp - sq], ppq = vox[f + sp + sq];
	const bool wa = face == 0 || face == 3 || face == 4;
	const uint32_t pf = F & kVoxSunTorch, d3 = ((F >> 11) & 1u) << 3;
	// the corners (+,-) and (-,+): vertex 1 is the first with winding a, else the second; vertex 3 the
	// other (each side cell then enters two vertices through the same expressions)
	const uint32_t vpm = light_vertex(qm, pmq, pp, pf, d3), vmp = light_vertex(qp, mp, pm, pf, d3);
	const uint32_t v0 = light_vertex(qm, mm, pm, pf, d3), v1 = wa ? vpm : vmp;
	const uint32_t v2 = light_vertex(qp, ppq, pp, pf, d3), v3 = wa ? vmp : vpm;
	const uint32_t lo = (v0 & 63u) | (v1 & 63u) << 6 | (v2 & 63u) << 12 | (v3 & 63u) << 18;
	const uint32_t hi = ((v0 >> 8) & 63u) | ((v1 >> 8) & 63u) << 6 | ((v2 >> 8) & 63u) << 12 | ((v3 >> 8) & 63u) << 18;
	const uint32_t ao4 = (v0 >> 16) | (v1 >> 16) << 2 | (v2 >> 16) << 4 | (v3 >> 16) << 6;
	return make_uint2(lo | (ao4 << 24), hi);
}

__device__ __forceinline__ uint32_t l4_sun(uint2 l, int k) { return (l.x >> (6 * k)) & 63u; }
__device__ __forceinline__ uint32_t l4_torch(uint2 l, int k) { return (l.y >> (6 * k)) & 63u; }
__device__ __forceinline__ uint32_t l4_ao(uint2 l, int k) { return (l.x >> (24 + 2 * k)) & 3u; }

// Light4::GetFlip (:38-43): sunlight and AO only.
__device__ __forceinline__ bool l4_flip(uint2 l) {
	const int d02 = int(l4_ao(l, 0) + 1) * int(l4_sun(l, 0) + 1) - int(l4_ao(l, 2) + 1) * int(l4_sun(l, 2) + 1);
	const int d13 = int(l4_ao(l, 1) + 1) * int(l4_sun(l, 1) + 1) - int(l4_ao(l, 3) + 1) * int(l4_sun(l, 3) + 1);
	return abs(d02) < abs(d13);
}

__device__ __forceinline__ uint64_t ballot64(bool b) { return __ballot(b); }

// Minimum / maximum over the wavefront with DPP lane moves (no LDS): swaps within quads, the half-row
// and row mirrors leave every lane of a 16-lane row with the row's result, then the row broadcasts of
// lanes 15 and 31 fold the rows into lane 63. Lanes a move does not write keep the identity.
template <bool kMax> __device__ __forceinline__ uint32_t wave_fold(uint32_t v) {
	constexpr uint32_t id = kMax ? 0u : 0xffffffffu;
	auto op = [](uint32_t a, uint32_t b) { return kMax ? max(a, b) : min(a, b); };
	v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0xb1, 0xf, 0xf, false));  // quad_perm 1,0,3,2
	v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x4e, 0xf, 0xf, false));  // quad_perm 2,3,0,1
	v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x141, 0xf, 0xf, false)); // row_half_mirror
	v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x140, 0xf, 0xf, false)); // row_mirror
	v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x142, 0xa, 0xf, false)); // row_bcast:15
	v = op(v, (uint32_t)__builtin_amdgcn_update_dpp((int)id, (int)v, 0x143, 0xc, 0xf, false)); // row_bcast:31
	return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}

// ------------------------------------------------------------------------------------------------
// block-wide exclusive scan of three 32-bit counters (every thread reads s.scan_w after the barrier; the
// next write to it is a later chunk's, after several more barriers)
// ------------------------------------------------------------------------------------------------
// Inclusive prefix sum over the wavefront by DPP: shifts by 1, 2, 4 and 8 lanes within each row of 16,
// then the last lane of row 0 into row 1 (and of row 2 into row 3), then lane 31 into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_incl_sum(uint32_t v) {
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true); // row_shr:1
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true); // row_shr:2
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true); // row_shr:4
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true); // row_shr:8
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, true); // row_bcast:15
	v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, true); // row_bcast:31
	return v;
}

// Exclusive prefix sums over the workgroup of three 32-bit values per thread, and their totals, with one
// barrier.
template <int NT>
__device__ __forceinline__ void block_excl_sums(Smem<NT> &s, uint32_t v[3], uint32_t tot[3]) {
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
	uint32_t inc[3];
#pragma unroll
	for (int i = 0; i < 3; ++i)
		inc[i] = wave_incl_sum(v[i]);
	if (lane == 63) {
#pragma unroll
		for (int i = 0; i < 3; ++i)
			s.scan_w[wave][i] = inc[i];
	}
	__syncthreads();
#pragma unroll
	for (int i = 0; i < 3; ++i) {
		uint32_t before = 0, all = 0;
#pragma unroll
		for (int w = 0; w < NT / 64; ++w) {
			const uint32_t x = s.scan_w[w][i];
			before += w < wave ? x : 0u;
			all += x;
		}
		v[i] = before + inc[i] - v[i];
		tot[i] = all;
	}
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
	uint32_t count; // faces emitted
	uint32_t trans; // bit j: face j goes to the transparent pass
	uint32_t mesh;  // registry mesh index
	uint32_t flow;  // 1 if liquid flow
	Flow fl;
	uint32_t flow_set; // liquid flow: emitted face kinds
};

__device__ __forceinline__ bool tex_trans_pass(const MeshParams &p, uint16_t t) {
	const uint32_t id = t & 0x1fffu;
	return id < p.reg.n_tex && (p.reg.tex_flags[id] & HCM_TEX_TRANS_PASS);
}

template <int NT> __device__ __forceinline__ CustomCell custom_cell(const MeshParams &p, const Smem<NT> &s, int cell) {
	const int x = cell & 15, z = (cell >> 4) & 15, y = cell >> 8;
	const uint16_t b = p.blocks[(size_t)s.slots[26] * HCM_CHUNK_VOXELS + (uint32_t)cell];
	const uint32_t pid = block_pid(p, s, b);
	CustomCell c;
	c.mesh = s.rows[pid].y >> 24;
	c.flow = (s.pid_vox[pid] & kVoxFlow) ? 1u : 0u;
	c.trans = 0;
	c.flow_set = 0;
	c.fl.h = 0, c.fl.liq_side = 0;
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
// writing one quad
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void put_indices(const MeshParams &p, uint32_t slot, uint32_t base, bool flip) {
	// index patterns: flip (0,1,2, 0,2,3) (:440-446) -- also the custom-mesh pattern (:295-301) --
	// and no flip (1,2,3, 0,1,3) (:453-459)
	const uint32_t b0 = base, b1 = base + 1, b2 = base + 2, b3 = base + 3;
	uint32_t *ip = reinterpret_cast<uint32_t *>(p.pool_idx + 6 * (size_t)slot);
	if (flip) {
		ip[0] = b0 | (b1 << 16);
		ip[1] = b2 | (b0 << 16);
		ip[2] = b2 | (b3 << 16);
	} else {
		ip[0] = b1 | (b2 << 16);
		ip[1] = b3 | (b0 << 16);
		ip[2] = b1 | (b3 << 16);
	}
}

__device__ __forceinline__ void put_quad(const MeshParams &p, uint32_t slot, const uint32_t w0[4], const uint32_t w1[4],
                                         uint32_t base, bool flip) {
	uint4 *vp = reinterpret_cast<uint4 *>(p.pool_verts + 4 * (size_t)slot);
	vp[0] = make_uint4(w0[0], w1[0], w0[1], w1[1]);
	vp[1] = make_uint4(w0[2], w1[2], w0[3], w1[3]);
	put_indices(p, slot, base, flip);
}

// BlockVertex word 1 (client/include/client/BlockVertex.hpp:13-14)
__device__ __forceinline__ uint32_t vword1(uint16_t tex, uint32_t face, uint32_t ao, uint32_t sun, uint32_t torch) {
	return ((uint32_t)(tex & 0x1fffu) - 1u) | ((uint32_t)(tex >> 13) << 10) | (face << 13) | (ao << 16) | (sun << 18) |
	       (torch << 24);
}

template <int NT>
__device__ __forceinline__ void aabb_merge(Smem<NT> &s, int pass, int k, const uint32_t lo[3], const uint32_t hi[3]) {
	uint32_t *a = s.aabb[pass][k];
	atomicMin(&a[0], lo[0]);
	atomicMin(&a[1], lo[1]);
	atomicMin(&a[2], lo[2]);
	atomicMax(&a[3], hi[0]);
	atomicMax(&a[4], hi[1]);
	atomicMax(&a[5], hi[2]);
}

// Regular quad: task t = 34 * axis + 2 * slice + direction, origin (u, v0) on the mask, width w along V,
// height h along U (:357-460), with the face word of its origin face (Light4 and texture code).
template <int NT>
__device__ __forceinline__ void emit_quad(const MeshParams &p, const Smem<NT> &s, int t, int u, int v0, int w, int h,
                                          uint2 word, uint32_t slot, uint32_t vbase, uint32_t lo[3], uint32_t hi[3]) {
	const int A = t / 34, sl = (t % 34) >> 1, inv = t & 1, face = 2 * A + inv;
	const int ua = u_axis(A), va = v_axis(A);
	const uint2 l = make_uint2(word.x, word.y & 0xffffffu);
	const uint16_t tex = uint16_t(s.codes[(word.y >> 24) & 0x7fu] & 0xffffu);
	// quad vectors (:383-390): (U+1)%3 == V holds for the x and y axes
	const bool height_on_u = (inv != 0) != (A != 2);
	int du[3], dv[3], o[3];
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		const int on_u = a == ua, on_v = a == va;
		o[a] = a == A ? sl : (on_u ? u : v0);
		du[a] = height_on_u ? on_u * h : on_v * w;
		dv[a] = height_on_u ? on_v * w : on_u * h;
	}
	// word 0 packs x, y, z (1/16 units, at most 256) in 10-bit fields: the corners are the origin's word
	// plus the packed quad vectors, with no carry between fields
	const auto pack = [](uint32_t x, uint32_t y, uint32_t z) { return (x << 4) | (y << 14) | (z << 24); };
	const uint32_t base = pack(o[0], o[1], o[2]) | (uint32_t(A) << 30);
	const uint32_t pu = pack(du[0], du[1], du[2]), pv = pack(dv[0], dv[1], dv[2]);
	const uint32_t w0[4] = {base, base + pu, base + pu + pv, base + pv};
	uint32_t w1[4];
#pragma unroll
	for (int k = 0; k < 4; ++k)
		w1[k] = vword1(tex, uint32_t(face), l4_ao(l, k), l4_sun(l, k), l4_torch(l, k));
	put_quad(p, slot, w0, w1, vbase, l4_flip(l));
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		lo[a] = uint32_t(o[a]) << 4;
		hi[a] = uint32_t(o[a] + du[a] + dv[a]) << 4;
	}
}

__device__ __forceinline__ uint32_t lerp16(uint32_t a, uint32_t b, uint32_t t) { return a * (16u - t) + b * t; }
__device__ __forceinline__ uint32_t interp(uint32_t l0, uint32_t l1, uint32_t l2, uint32_t l3, uint32_t h0, uint32_t h1,
                                           uint32_t h2, uint32_t h3, uint32_t du, uint32_t dv, uint32_t dw) {
	// light4_interpolate (:165-196): trilinear in 1/16 steps, rounded up
	const uint32_t lo = lerp16(lerp16(l0, l1, du), lerp16(l3, l2, du), dv);
	const uint32_t hi = lerp16(lerp16(h0, h1, du), lerp16(h3, h2, du), dv);
	const uint32_t sum = lerp16(lo, hi, dw);
	return sum / 4096u + (sum % 4096u ? 1u : 0u);
}

template <int NT>
__device__ __forceinline__ void emit_custom(const MeshParams &p, Smem<NT> &s, uint32_t desc, uint32_t slot,
                                         uint32_t vbase, uint32_t lo[3], uint32_t hi[3]) {
	const int cell = desc & 4095, j = (desc >> 12) & 15;
	const int x = cell & 15, z = (cell >> 4) & 15, y = cell >> 8;
	const CustomCell c = custom_cell(p, s, cell);
	const hcm_mesh_face f = custom_face(p, c, j);
	const int lf = f.light_face, la = lf >> 1, sg = (lf & 1) ? -1 : 1;
	const int pos = hidx(x, y, z);
	// low: the face of the cell behind; high: this cell's own face if the next cell passes light (:259-267)
	const uint2 lo4 = light4(s.vox, lf, pos - sg * axis_stride(la));
	const uint2 hi4 = (s.vox[pos + sg * axis_stride(la)] & kVoxIndirect) ? light4(s.vox, lf, pos) : lo4;
	int ua = (la + 1) % 3, va = (la + 2) % 3;
	if (lf & 1) {
		const int tmp = ua;
		ua = va, va = tmp;
	}
	const uint32_t bx = uint32_t(x) << 4, by = uint32_t(y) << 4, bz = uint32_t(z) << 4;
	lo[0] = lo[1] = lo[2] = ~0u;
	hi[0] = hi[1] = hi[2] = 0u;
	// vertices one at a time (custom faces are rare; this keeps the register footprint small), each
	// picked from packed words so that nothing is indexed at run time
	uint32_t vw[4];
#pragma unroll
	for (int k = 0; k < 4; ++k)
		vw[k] = uint32_t(f.v[k].x) | uint32_t(f.v[k].y) << 8 | uint32_t(f.v[k].z) << 16 | uint32_t(f.v[k].ao) << 24;
#pragma unroll 1
	for (int k = 0; k < 4; ++k) {
		const uint32_t w = k == 0 ? vw[0] : (k == 1 ? vw[1] : (k == 2 ? vw[2] : vw[3]));
		const uint32_t vx = w & 0xffu, vy = (w >> 8) & 0xffu, vz = (w >> 16) & 0xffu;
		auto pick = [&](int a) { return a == 0 ? vx : (a == 1 ? vy : vz); };
		const uint32_t px = bx + vx, py = by + vy, pz = bz + vz;
		lo[0] = min(lo[0], px), lo[1] = min(lo[1], py), lo[2] = min(lo[2], pz);
		hi[0] = max(hi[0], px), hi[1] = max(hi[1], py), hi[2] = max(hi[2], pz);
		const uint32_t du = pick(ua), dv = pick(va);
		uint32_t dw = pick(la) < 16u ? pick(la) : 16u;
		if (lf & 1)
			dw = 16u - dw;
		uint32_t ao = w >> 24;
		if (ao == 4)
			ao = interp(l4_ao(lo4, 0), l4_ao(lo4, 1), l4_ao(lo4, 2), l4_ao(lo4, 3), l4_ao(hi4, 0), l4_ao(hi4, 1),
			            l4_ao(hi4, 2), l4_ao(hi4, 3), du, dv, dw) & 0xffu;
		const uint32_t sun = interp(l4_sun(lo4, 0), l4_sun(lo4, 1), l4_sun(lo4, 2), l4_sun(lo4, 3), l4_sun(hi4, 0),
		                            l4_sun(hi4, 1), l4_sun(hi4, 2), l4_sun(hi4, 3), du, dv, dw) & 0xffu;
		const uint32_t torch = interp(l4_torch(lo4, 0), l4_torch(lo4, 1), l4_torch(lo4, 2), l4_torch(lo4, 3),
		                              l4_torch(hi4, 0), l4_torch(hi4, 1), l4_torch(hi4, 2), l4_torch(hi4, 3), du, dv,
		                              dw) & 0xffu;
		p.pool_verts[4 * (size_t)slot + k] = make_uint2(px | (py << 10) | (pz << 20) | (uint32_t(f.axis) << 30),
		                                                vword1(f.texture, f.render_face, ao, sun, torch));
	}
	put_indices(p, slot, vbase, true);
}

// ------------------------------------------------------------------------------------------------
// visibility: one lane per (u, v) column along the slicing axis
// ------------------------------------------------------------------------------------------------
// BlockTexture::Show (texture/include/texture/BlockTexture.hpp:68-76) on the 2-bit face classes
// (empty 0, opaque 1, transparent 2, liquid 3):
//   opaque self:      visible unless the other face is opaque,
//   transparent self: visible if the other is empty or liquid, or transparent with another texture,
//   liquid self:      visible if the other is empty.
// The face between halo positions a and a + 1 is in slice a; with i = class(+ face of a) |
// class(- face of a + 1) << 2, bit i of kShowPlus says whether the + face of a shows and bit i of
// kShowMinus whether the - face of a + 1 does. Transparent against transparent (i = 10) depends on
// the two texture codes and is resolved from the voxels' texture rows.
__host__ __device__ constexpr uint32_t show_class(uint32_t self, uint32_t other) {
	return self == 1 ? (other != 1 ? 1u : 0u) : self == 2 ? ((other == 0 || other == 3) ? 1u : 0u) : self == 3 ? (other == 0 ? 1u : 0u) : 0u;
}
__host__ __device__ constexpr uint32_t show_table(bool plus) {
	uint32_t t = 0;
	for (uint32_t i = 0; i < 16; ++i)
		t |= (plus ? show_class(i & 3u, i >> 2) : show_class(i >> 2, i & 3u)) << i;
	return t;
}
constexpr uint32_t kShowPlus = show_table(true), kShowMinus = show_table(false), kShowTT = 2u | 2u << 2;
static_assert(kShowPlus == 0x620eu && kShowMinus == 0x19d0u, "Show on face classes");

// Only blocks inside the chunk own faces (:330, :335): + faces of halo positions 1..16 (slices 1..16)
// and - faces of positions 1..16 (slices 0..15). Ballots over the wave give the 16x16 masks of the
// (axis, slice, direction) tasks, four mask rows per wave.
template <int A, int NT> __device__ __forceinline__ void visibility_column(Smem<NT> &s, int col) {
	const int lane = threadIdx.x & 63;
	const int u = col >> 4, v = col & 15;
	constexpr int sa = axis_stride(A);
	constexpr int shp = kVoxClassShift + 4 * A; // class of face 2A (+); face 2A + 1 (-) follows
	const uint32_t *cw = s.vox + (u + 1) * axis_stride(u_axis(A)) + (v + 1) * axis_stride(v_axis(A));
	uint32_t c[kHalo];
#pragma unroll
	for (int a = 0; a < kHalo; ++a)
		c[a] = (cw[a * sa] >> shp) & 15u;
	const int q = (col >> 6) & 3; // this wave's group of four mask rows
	uint32_t tt_bits = 0;         // 1 << i over all slices: bit kShowTT marks transparent against transparent
#pragma unroll
	for (int a = 0; a <= 16; ++a) {
		const uint32_t i = (c[a] & 3u) | (c[a + 1] & 12u);
		const uint64_t bp = a >= 1 ? ballot64((kShowPlus >> i) & 1u) : 0ull;
		const uint64_t bm = a <= 15 ? ballot64((kShowMinus >> i) & 1u) : 0ull;
		tt_bits |= 1u << i;
		if (lane == 0) {
			reinterpret_cast<uint64_t *>(s.ne[34 * A + 2 * a])[q] = bp;
			reinterpret_cast<uint64_t *>(s.ne[34 * A + 2 * a + 1])[q] = bm;
		}
	}
	if (ballot64((tt_bits >> kShowTT) & 1u)) { // such faces show when the two texture codes differ (both directions alike)
#pragma unroll 1
		for (int a = 0; a <= 16; ++a) {
			const uint32_t wa = cw[a * sa], wb = cw[(a + 1) * sa];
			const bool d = (((wa >> shp) & 3u) | ((wb >> shp) & 12u)) == kShowTT &&
			               row_code(s, vox_pid(wa), 2 * A) != row_code(s, vox_pid(wb), 2 * A + 1);
			const uint64_t bd = ballot64(d);
			if (bd && lane == 0) {
				if (a >= 1)
					reinterpret_cast<uint64_t *>(s.ne[34 * A + 2 * a])[q] |= bd;
				if (a <= 15)
					reinterpret_cast<uint64_t *>(s.ne[34 * A + 2 * a + 1])[q] |= bd;
			}
		}
	}
}

// Halo index of the block owning face (task t, row u, column v), and its face.
__device__ __forceinline__ int face_owner(int t, int u, int v, int &face) {
	const int A = t / 34, sl = (t % 34) >> 1, inv = t & 1;
	face = 2 * A + inv;
	const int layer = inv ? sl : sl - 1;
	return (layer + 1) * axis_stride(A) + (u + 1) * axis_stride(u_axis(A)) + (v + 1) * axis_stride(v_axis(A));
}

// Task, mask row and column of the visible face with dense index gi (task-major, row-major, as the
// reference scans them), searched among the tasks [t0, t1), without branches: the task by seven halving
// steps over the prefix (the last t with fpre[t] <= gi; fpre[t1] > gi and a group spans under 128 tasks),
// then the task's four 64-bit mask words and the word's four 16-bit rows by comparing the remaining index
// with their running counts, then the column as the rem-th set bit of the row.
template <int NT> __device__ __forceinline__ void face_pos(const Smem<NT> &s, int t0, int t1, uint32_t gi, int &t, int &u, int &v) {
	int lo = t0;
#pragma unroll
	for (int step = 64; step >= 1; step >>= 1) {
		const int m = min(lo + step, t1);
		lo = s.fpre[m] <= gi ? m : lo;
	}
	t = lo;
	uint32_t rem = gi - s.fpre[t];
	const uint64_t *rows = reinterpret_cast<const uint64_t *>(s.ne[t]);
	const uint64_t w0 = rows[0], w1 = rows[1], w2 = rows[2], w3 = rows[3];
	const uint32_t p1 = __popcll(w0), p2 = p1 + __popcll(w1), p3 = p2 + __popcll(w2);
	const bool q1 = rem >= p1, q2 = rem >= p2, q3 = rem >= p3;
	const uint64_t w = q3 ? w3 : (q2 ? w2 : (q1 ? w1 : w0));
	rem -= q3 ? p3 : (q2 ? p2 : (q1 ? p1 : 0u));
	const uint32_t lo32 = uint32_t(w), hi32 = uint32_t(w >> 32);
	const uint32_t r1 = __popc(lo32 & 0xffffu), r2 = __popc(lo32), r3 = r2 + __popc(hi32 & 0xffffu);
	const bool s1 = rem >= r1, s2 = rem >= r2, s3 = rem >= r3;
	uint32_t bits = s3 ? hi32 >> 16 : (s2 ? hi32 & 0xffffu : (s1 ? lo32 >> 16 : lo32 & 0xffffu));
	rem -= s3 ? r3 : (s2 ? r2 : (s1 ? r1 : 0u));
	u = 4 * (int(q1) + int(q2) + int(q3)) + int(s1) + int(s2) + int(s3);
	// the column: the rem-th set bit of the row
	v = 0;
	const uint32_t c8 = __popc(bits & 0xffu);
	if (rem >= c8)
		rem -= c8, bits >>= 8, v += 8;
	const uint32_t c4 = __popc(bits & 0xfu);
	if (rem >= c4)
		rem -= c4, bits >>= 4, v += 4;
	const uint32_t c2 = __popc(bits & 0x3u);
	if (rem >= c2)
		rem -= c2, bits >>= 2, v += 2;
	if (rem >= (bits & 1u))
		v += 1;
}

// Word of the visible face with dense index gi (task-major, row-major, as the reference scans them):
// Light4 (light4_init, :48-163) in x and the low 24 bits of y, the texture code in y bits 24-30 and the
// transparent-pass flag in bit 31. Faces merge only if their words are equal (Light4::operator==, :44,
// and the texture compare, :367-368).
template <int NT>
__device__ __forceinline__ uint2 face_word(const Smem<NT> &s, int t0, int t1, uint32_t gi, uint32_t &desc) {
	int t, u, v;
	face_pos(s, t0, t1, gi, t, u, v);
	desc = uint32_t(t) | uint32_t(u) << 7 | uint32_t(v) << 11;
	int face;
	const int self = face_owner(t, u, v, face);
	const uint2 l = light4(s.vox, face, self);
	const uint32_t tc = row_code(s, vox_pid(s.vox[self]), face);
	const uint32_t tpass = (s.codes[tc] >> 16) & kCodeTransPass ? 1u : 0u;
	return make_uint2(l.x, l.y | (tc << 24) | (tpass << 31));
}

// Equal-to-right, equal-to-below and transparent-pass bits of the visible face with dense index i (local
// to the group; desc = t | u << 7 | v << 11): its right and lower neighbours, when visible, are at
// i + 1 and after the rest of row u and the start of row u + 1 in the dense order.
template <int NT> __device__ __forceinline__ void face_links(Smem<NT> &s, uint32_t i, uint32_t desc) {
	const int t = desc & 127, u = (desc >> 7) & 15, v = (desc >> 11) & 15;
	// the neighbours' words are read whether or not the neighbours are visible (positions clamped into the
	// stage) and compared as whole words, so that only the rare atomics sit behind branches
	const uint32_t ne_u = s.ne[t][u], ne_d = uint32_t(s.ne[t][min(u + 1, 15)]) & (0u - uint32_t(u < 15));
	const uint2 w = s.key[i];
	const uint32_t bit = 1u << (v + 16 * (u & 1));
	const uint32_t ir = min(i + 1, uint32_t(kListCap - 1));
	const uint32_t id = min(i + 1 + __popc(ne_u >> (v + 1)) + __popc(ne_d & ((1u << v) - 1u)), uint32_t(kListCap - 1));
	const uint2 r = s.key[ir], d = s.key[id];
	const bool eq_r = v < 15 && ((ne_u >> (v + 1)) & 1u) && ((r.x ^ w.x) | (r.y ^ w.y)) == 0;
	const bool eq_d = ((ne_d >> v) & 1u) && ((d.x ^ w.x) | (d.y ^ w.y)) == 0;
	if (eq_r)
		atomicOr(reinterpret_cast<uint32_t *>(&s.eh[t][u & ~1]), bit);
	if (eq_d)
		atomicOr(reinterpret_cast<uint32_t *>(&s.ev[t][u & ~1]), bit);
	if (w.y >> 31)
		atomicOr(reinterpret_cast<uint32_t *>(&s.tp[t][u & ~1]), bit);
}

// ------------------------------------------------------------------------------------------------
// greedy merge of one task (:357-476), one lane per task
// ------------------------------------------------------------------------------------------------
// The reference scans rows u ascending and cells v ascending; at an uncovered visible cell it grows the
// quad along V while the next cell has an equal face word and is uncovered, then along U while every
// cell of the next row's segment is equal and uncovered, and marks the covered cells. The quads that start
// in row u are therefore exactly the maximal runs of uncovered visible cells split wherever the face word
// changes, and each run's height depends only on the rows below. Here a row's runs are found with bit
// operations and all of them are extended downward together.
// Moves marked bits from v to v - 1 (spread_down) or v - 1 to v (spread_up) while v is in `link`, i.e.
// while v continues the run of v - 1: it reaches a run's start from any of its cells, and vice versa.
__device__ __forceinline__ uint32_t spread_down(uint32_t d, uint32_t link) {
	uint32_t m = link;
	d |= (d & m) >> 1;
	m &= m >> 1;
	d |= (d & (m << 1)) >> 2;
	m &= m >> 2;
	d |= (d & (m << 3)) >> 4;
	m &= m >> 4;
	d |= (d & (m << 7)) >> 8;
	return d;
}
__device__ __forceinline__ uint32_t spread_up(uint32_t f, uint32_t link) {
	uint32_t m = link;
	f |= (f << 1) & m;
	m &= m << 1;
	f |= (f << 2) & m;
	m &= m << 2;
	f |= (f << 4) & m;
	m &= m << 4;
	f |= (f << 8) & m;
	return f;
}

template <int NT>
__device__ __forceinline__ void greedy_count(Smem<NT> &s, int t, bool active, uint32_t &n_o, uint32_t &n_t) {
	// coverage of the rows below the current one by quads of earlier rows, and the equal-to-below masks,
	// held in registers as windows that shift up by one row per step (window row k = row u + k)
	uint32_t cov[8], ev[8];
	const uint32_t *evs = reinterpret_cast<const uint32_t *>(s.ev[active ? t : 0]);
#pragma unroll
	for (int k = 0; k < 8; ++k) {
		cov[k] = 0;
		ev[k] = active ? evs[k] : 0u;
	}
	n_o = 0, n_t = 0;
	// a row's masks are read one step ahead, so that their LDS latency overlaps the previous row's work
	uint32_t ne_n = active ? s.ne[t][0] : 0u, eh_n = active ? s.eh[t][0] : 0u, tp_n = active ? s.tp[t][0] : 0u;
#pragma unroll 1
	for (int u = 0; u < 16; ++u) {
		const uint32_t ne = ne_n, eh = eh_n, tp = tp_n;
		if (u < 15 && active) // row u + 1 is not written before this step writes row u
			ne_n = s.ne[t][u + 1], eh_n = s.eh[t][u + 1], tp_n = s.tp[t][u + 1];
		const uint32_t S = ne & ~(cov[0] & 0xffffu); // uncovered visible cells of row u
		const uint32_t link = S & ((S & eh) << 1);   // v continues the run of v - 1
		const uint32_t starts = S & ~link;
		if (active) { // row u's results replace its (already read) inputs; ev's rows are in registers
			s.ev[t][u] = uint16_t(S);
			s.eh[t][u] = uint16_t(starts & ~tp); // opaque quads start here
			s.tp[t][u] = uint16_t(starts & tp);  // transparent quads start here
		}
		if (__ballot(S != 0)) {
			uint32_t alive = starts, cols = S;
#pragma unroll
			for (int k = 1; k <= 16; ++k) {
				// row u + k, if it exists: every column of a surviving run must continue into it
				if (!__ballot(alive != 0))
					break;
				uint32_t ok = 0; // rows past 15 read as empty from the shifted windows
				if (k < 16) {
					const uint32_t evr = (ev[(k - 1) >> 1] >> (16 * ((k - 1) & 1))) & 0xffffu;
					const uint32_t cvr = (cov[k >> 1] >> (16 * (k & 1))) & 0xffffu;
					ok = evr & ~cvr;
				}
				const uint32_t bad = cols & ~ok;
				if (bad) {
					const uint32_t dead = spread_down(bad, link) & alive;
					alive &= ~dead;
					cols &= ~spread_up(dead, link);
				}
				if (k < 16)
					cov[k >> 1] |= cols << (16 * (k & 1));
			}
			const uint32_t nt = __popc(starts & tp);
			n_t += nt;
			n_o += __popc(starts) - nt;
		}
		// shift the windows up by one row
#pragma unroll
		for (int k = 0; k < 8; ++k) {
			const uint32_t nc = k < 7 ? cov[k + 1] : 0u, nv = k < 7 ? ev[k + 1] : 0u;
			cov[k] = (cov[k] >> 16) | (nc << 16);
			ev[k] = (ev[k] >> 16) | (nv << 16);
		}
	}
}

// ------------------------------------------------------------------------------------------------
// emission of one quad
// ------------------------------------------------------------------------------------------------
// The end of the longest run of tasks from t0 whose visible faces fit the face-word stage (one task, at
// most 256 faces, always fits). Thread 0 searches; a barrier publishes the result.
template <int NT> __device__ __forceinline__ int group_end(Smem<NT> &s, int t0) {
	if (s.fpre[kTasks] - s.fpre[t0] <= (uint32_t)kListCap) // the rest fits (the usual case): every thread knows
		return kTasks;
	if (threadIdx.x == 0) {
		int lo = t0 + 1, hi = kTasks; // largest t1 with fpre[t1] - fpre[t0] <= kListCap
		while (lo < hi) {
			const int mid = (lo + hi + 1) >> 1;
			if (s.fpre[mid] - s.fpre[t0] <= (uint32_t)kListCap)
				lo = mid;
			else
				hi = mid - 1;
		}
		s.group_end = lo;
	}
	__syncthreads();
	return s.group_end;
}

// Where quad i of a pass goes: its pool slot, the base vertex of its mesh, the mesh of the pass (mk) and
// its index in that mesh (mj). The reference reads the 16-bit vertex count before splitting, so the first
// quad of a continuation mesh keeps base 65,532 (:273-279, :397-403).
__device__ __forceinline__ void quad_slot(uint32_t pass, uint32_t i, uint32_t n0, uint32_t pool_base, uint32_t &slot,
                                          uint32_t &vbase, uint32_t &mk, uint32_t &mj) {
	slot = pool_base + (pass ? n0 : 0u) + i;
	mk = i / kSplitQuads;
	mj = i - mk * kSplitQuads;
	vbase = (mj == 0 && mk > 0) ? 65532u : 4u * mj;
}

// A written quad's box into its mesh's AABB (per lane for the first mesh of a pass, LDS atomics for the
// others) and, when it starts a continuation mesh, a split event at its emission position epos.
// A quad's box into the lane's box of the first mesh of its pass (static register indices only): the
// other pass's box takes the empty box (lower corner all ones, upper zero), which changes nothing, and the
// transparent pass's box is left alone in chunks without transparent quads (any_tr, the same in every lane).
__device__ __forceinline__ void box_add(uint32_t box[2][6], uint32_t pass, const uint32_t lo[3], const uint32_t hi[3],
                                        bool any_tr = true) {
	const uint32_t m = 0u - pass; // all ones for the transparent pass
#pragma unroll
	for (int a = 0; a < 3; ++a) {
		box[0][a] = min(box[0][a], lo[a] | m);
		box[0][3 + a] = max(box[0][3 + a], hi[a] & ~m);
	}
	if (any_tr) {
#pragma unroll
		for (int a = 0; a < 3; ++a) {
			box[1][a] = min(box[1][a], lo[a] | ~m);
			box[1][3 + a] = max(box[1][3 + a], hi[a] & m);
		}
	}
}

template <int NT>
__device__ __forceinline__ void quad_done(Smem<NT> &s, uint32_t pass, uint32_t mk, uint32_t mj, uint32_t epos,
                                          const uint32_t lo[3], const uint32_t hi[3], uint32_t box[2][6]) {
	if (mk == 0)
		box_add(box, pass, lo, hi);
	else
		aabb_merge(s, pass, mk, lo, hi);
	if (mj == 0 && mk > 0) { // a split happened right before this quad (:397-403)
		const uint32_t ev = atomicAdd(&s.n_events, 1u);
		if (ev < 16) {
			s.events[ev][0] = epos;
			s.events[ev][1] = (pass << 16) | (mk - 1u);
		}
	}
}

// The quad that starts at the visible face at desc (task t | row u << 7 | column v << 11) with face word
// `word`, if one does. Per
// task, ne still holds the visible cells and the counting pass left the opaque quad starts in eh, the
// transparent ones in tp and the uncovered visible cells S in ev; the covered cells are C = ne & ~S. A
// quad from (u, v0) spans S's run from v0 to the next start along V, and covers rows u .. u + h - 1 of its
// columns, so below it column v0 is in C exactly for rows u + 1 .. u + h - 1: the cell under its last row
// is either not covered or starts a quad. Its index in its pass follows the pass's starts in the task's
// earlier rows and columns (the reference's order: rows, then columns). kSplit: some pass of the chunk
// has more than kSplitQuads quads; without, every quad is in the first mesh of its pass.
template <bool kSplit, int NT>
__device__ __forceinline__ void emit_face(const MeshParams &p, Smem<NT> &s, uint32_t desc, uint2 word, uint32_t n0,
                                          uint32_t pool_base, uint32_t box[2][6], bool any_tr) {
	const int t = desc & 127, u = (desc >> 7) & 15, v = (desc >> 11) & 15;
	const uint32_t so = s.eh[t][u], st = s.tp[t][u], starts = so | st;
	if (!((starts >> v) & 1u))
		return;
	const uint32_t tr = (st >> v) & 1u, below = (1u << v) - 1u;
	const uint32_t link = uint32_t(s.ev[t][u]) & ~starts;
	const int w = __ffs(~(link >> (v + 1)));
	// most quads are one row high: the covered cells of the row below are tested before the loop (row 15
	// has none; its index is clamped so that the load needs no branch)
	const int ub = min(u + 1, 15);
	const uint32_t cov_below = uint32_t(s.ne[t][ub] & ~s.ev[t][ub]) & (0u - uint32_t(u < 15));
	int h = 1;
	if ((cov_below >> v) & 1u) {
		h = 2;
#pragma unroll 1
		for (int r = u + 2; r < 16 && ((uint32_t(s.ne[t][r] & ~s.ev[t][r]) >> v) & 1u); ++r)
			++h;
	}
	// the pass's starts in rows 0 .. u - 1: words of two rows each, word k whole when k < u / 2 (k = 7
	// never: u <= 15), and the low row of word u / 2 when u is odd
	const uint32_t *m32 = reinterpret_cast<const uint32_t *>(tr ? s.tp[t] : s.eh[t]);
	const uint4 wa = reinterpret_cast<const uint4 *>(m32)[0], wb = reinterpret_cast<const uint4 *>(m32)[1];
	const uint32_t wk[7] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z};
	uint32_t ord = __popc((tr ? st : so) & below);
	const int half = u >> 1;
#pragma unroll
	for (int k = 0; k < 7; ++k)
		ord += __popc(k < half ? wk[k] : 0u);
	ord += __popc(m32[half] & (0u - uint32_t(u & 1)) & 0xffffu);
	const uint32_t i = (tr ? s.qpre_t[t] : s.qpre_o[t]) + ord;
	uint32_t slot, vbase, mk, mj, lo[3], hi[3];
	if (!kSplit) {
		emit_quad(p, s, t, u, v, w, h, word, pool_base + (tr ? n0 : 0u) + i, 4u * i, lo, hi);
		box_add(box, tr, lo, hi, any_tr);
		return;
	}
	quad_slot(tr, i, n0, pool_base, slot, vbase, mk, mj);
	emit_quad(p, s, t, u, v, w, h, word, slot, vbase, lo, hi);
	uint32_t epos = 0;
	if (mj == 0 && mk > 0) { // emission position, only needed for a split: both passes' starts before it
		const uint64_t *o64 = reinterpret_cast<const uint64_t *>(s.eh[t]);
		const uint64_t *t64 = reinterpret_cast<const uint64_t *>(s.tp[t]);
		epos = s.qpre_o[t] + s.qpre_t[t] + __popc(starts & below); // custom quads, then the earlier tasks'
		for (int q = 0; q < 4; ++q) {
			const int nb = 16 * (u - 4 * q);
			const uint64_t m = nb >= 64 ? ~0ull : (nb > 0 ? (1ull << nb) - 1ull : 0ull);
			epos += __popcll(o64[q] & m) + __popcll(t64[q] & m);
		}
	}
	quad_done(s, tr, mk, mj, epos, lo, hi, box);
}

// The regular quads of a chunk, one lane per visible face (a quad per face that starts one): first the
// last group of tasks, whose face words are still in key (keep: the lane's face positions, two 15-bit
// positions to a register, faces tid + 2 k NT and tid + (2 k + 1) NT in register k), then the earlier
// groups with their words computed again.
template <bool kSplit, int NT>
__device__ __forceinline__ void emit_regular(const MeshParams &p, Smem<NT> &s, uint32_t *keep, int t0_last, uint32_t n0,
                                             uint32_t pool_base, uint32_t box[2][6], bool any_tr) {
	const int tid = threadIdx.x;
	constexpr int kMine = (kListCap + NT - 1) / NT, kPairs = (kMine + 1) / 2;
	{
		const uint32_t g0 = s.fpre[t0_last], n = s.fpre[kTasks] - g0;
#pragma unroll 1
		for (int k = 0; k < kPairs; ++k) {
			const uint32_t i0 = tid + 2 * k * NT, i1 = i0 + NT;
			if (i0 < n)
				emit_face<kSplit>(p, s, keep[0] & 0xffffu, s.key[i0], n0, pool_base, box, any_tr);
			if (i1 < n)
				emit_face<kSplit>(p, s, keep[0] >> 16, s.key[i1], n0, pool_base, box, any_tr);
#pragma unroll
			for (int j = 0; j + 1 < kPairs; ++j)
				keep[j] = keep[j + 1];
		}
	}
#pragma unroll 1
	for (int t0 = 0; t0 < t0_last;) {
		__syncthreads(); // key is read by the previous group's lanes
		const int t1 = group_end(s, t0);
		const uint32_t g0 = s.fpre[t0], n = s.fpre[t1] - g0;
#pragma unroll 1
		for (int k = 0; k < kPairs; ++k) {
			const uint32_t i0 = tid + 2 * k * NT, i1 = i0 + NT;
			uint32_t d0 = 0, d1 = 0;
			if (i0 < n)
				s.key[i0] = face_word(s, t0, t1, g0 + i0, d0);
			if (i1 < n)
				s.key[i1] = face_word(s, t0, t1, g0 + i1, d1);
#pragma unroll
			for (int j = 0; j + 1 < kPairs; ++j)
				keep[j] = keep[j + 1];
			keep[kPairs - 1] = d0 | d1 << 16;
		}
		__syncthreads();
#pragma unroll 1
		for (int k = 0; k < kPairs; ++k) {
			const uint32_t i0 = tid + 2 * k * NT, i1 = i0 + NT;
			if (i0 < n)
				emit_face<kSplit>(p, s, keep[0] & 0xffffu, s.key[i0], n0, pool_base, box, any_tr);
			if (i1 < n)
				emit_face<kSplit>(p, s, keep[0] >> 16, s.key[i1], n0, pool_base, box, any_tr);
#pragma unroll
			for (int j = 0; j + 1 < kPairs; ++j)
				keep[j] = keep[j + 1];
		}
		t0 = t1;
	}
}

// ------------------------------------------------------------------------------------------------
// one chunk
// ------------------------------------------------------------------------------------------------
// Per-chunk LDS state a job starts from (the slot list is loaded with the job).
template <int NT> __device__ __forceinline__ void chunk_init(Smem<NT> &s) {
	const int tid = threadIdx.x;
	for (int i = tid; i < 128; i += NT)
		s.custom_bits[i] = 0;
	if (tid < 2 * kMaxMeshPerPass * 6) {
		uint32_t *a = &s.aabb[0][0][0];
		a[tid] = (tid % 6) < 3 ? 0xffffffffu : 0u; // AABB<uint32_t>{}: min = max(), max = min() (dep/AABB.hpp:11)
	}
	if (tid == 0)
		s.n_events = 0, s.flags = 0, s.n_link = 0;
}

// One job; s.slots and the state of chunk_init are in place (a barrier after them).
template <int NT> __device__ __forceinline__ void mesh_chunk(const MeshParams &p, Smem<NT> &s, uint32_t job) {
	const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
	constexpr int CPT = HCM_CHUNK_VOXELS / NT; // custom cells per thread

	// 1. halo staging: one thread per x-row (hy, hz) of 18 voxels. A voxel's word is its light fields and
	// the template its block id's byte looks up in LDS (the full table only for the few ids whose meta byte
	// changes the texture row); the row's flags come from an AND and an OR over its templates and from
	// comparing packed block pairs, the custom-mesh cells only when some template has a mesh.
	uint32_t my_flags = 0, wave_ref = 0xffffffffu;
	for (int row = tid; row < kHaloRows; row += NT) {
		const int hz = row % kHalo, hy = row / kHalo, y = hy - 1, z = hz - 1;
		const size_t lrow = 16u * (uint32_t(z & 15) + 16u * uint32_t(y & 15));
		const int syz = slot27_yz(hy, hz); // slot27(x, y, z) = 9 * (x term) + syz: 18 at x = 0, 9 at -1, 0 at 16
		const size_t cm = (size_t)s.slots[18 + syz] * HCM_CHUNK_VOXELS + lrow;
		const size_t cl = (size_t)s.slots[9 + syz] * HCM_CHUNK_VOXELS + lrow + 15;
		const size_t cr = (size_t)s.slots[syz] * HCM_CHUNK_VOXELS + lrow;
		const uint4 b0 = *reinterpret_cast<const uint4 *>(p.blocks + cm);
		const uint4 b1 = *reinterpret_cast<const uint4 *>(p.blocks + cm + 8);
		const uint32_t bw[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
		const uint32_t bl = p.blocks[cl], br = p.blocks[cr];
		uint32_t lw[5];
		if (p.halo_light) { // the job's rebuilt light, one 20-byte record per halo x-row
			const uint32_t *hl = reinterpret_cast<const uint32_t *>(p.halo_light + (size_t)job * kLightJobBytes +
			                                                        row * kLightRowBytes);
#pragma unroll
			for (int k = 0; k < 5; ++k)
				lw[k] = hl[k];
		} else { // the same layout from the store: x = -1 in byte 0, x = 0 .. 15 in bytes 1 .. 16, x = 16 in 17
			const uint4 lm = *reinterpret_cast<const uint4 *>(p.light + cm);
			const uint32_t ll = p.light[cl], lr = p.light[cr];
			lw[0] = ll | (lm.x << 8);
			lw[1] = (lm.x >> 24) | (lm.y << 8);
			lw[2] = (lm.y >> 24) | (lm.z << 8);
			lw[3] = (lm.z >> 24) | (lm.w << 8);
			lw[4] = (lm.w >> 24) | (lr << 8);
		}
		// one reference value per wavefront (its first row's first voxel): a halo of one block value has
		// every voxel equal to it and all wavefronts' references equal
		const uint32_t v0 = __builtin_amdgcn_readfirstlane(bl);
		if (row < NT)
			wave_ref = v0;
		const uint32_t ref2 = wave_ref | (wave_ref << 16);
		uint32_t diff = (bl ^ wave_ref) | (br ^ wave_ref);
#pragma unroll
		for (int k = 0; k < 8; ++k)
			diff |= bw[k] ^ ref2;
		const bool nonuniform = diff != 0;
		uint32_t tm[18];
		tm[0] = s.tmpl_id[bl & 0xffu];
		tm[17] = s.tmpl_id[br & 0xffu];
#pragma unroll
		for (int x = 0; x < 16; ++x)
			tm[x + 1] = s.tmpl_id[(bw[x >> 1] >> (16 * (x & 1))) & 0xffu];
		uint32_t any_meta = 0;
#pragma unroll
		for (int x = 0; x < 18; ++x)
			any_meta |= tm[x];
		if (any_meta & kTmplMeta) { // ids whose texture row depends on the meta byte
			// every voxel's row from the full table (any 16-bit value indexes it), issued as one batch of
			// loads and then one of LDS lookups, so the row waits once on memory and not once per voxel
			uint32_t full[18];
#pragma unroll
			for (int x = 0; x < 18; ++x) {
				const uint32_t b = x == 0 ? bl : (x == 17 ? br : (bw[(x - 1) >> 1] >> (16 * ((x - 1) & 1))) & 0xffffu);
				full[x] = p.reg.pid_of[b];
			}
#pragma unroll
			for (int x = 0; x < 18; ++x) {
				const uint32_t v = s.pid_vox[full[x]];
				tm[x] = tm[x] & kTmplMeta ? v : tm[x];
			}
		}
		uint32_t and_all = ~0u, or_inner = 0;
		uint32_t *dst = s.vox + row * kHalo;
#pragma unroll
		for (int x = 0; x < 18; ++x) {
			dst[x] = vox_light((lw[x >> 2] >> (8 * (x & 3))) & 0xffu) | tm[x];
			and_all &= tm[x];
			if (x >= 1 && x <= 16)
				or_inner |= tm[x];
		}
		const uint32_t or_all = or_inner | tm[0] | tm[17];
		const bool inside_yz = (unsigned)y < 16u && (unsigned)z < 16u;
		// every voxel an opaque cube: the shape bits that the cube has set are set in all, the others in none
		const bool mixed = (and_all & kVoxShape & kVoxOpaqueCube) != (kVoxShape & kVoxOpaqueCube) ||
		                   (or_all & kVoxShape & ~kVoxOpaqueCube) != 0;
		if (nonuniform)
			my_flags |= kFlagNonUniform;
		if (mixed)
			my_flags |= kFlagMixed;
		if (inside_yz && (or_inner & kVoxShape))
			my_flags |= kFlagInnerFaces;
		if (or_all & kVoxFlow)
			my_flags |= kFlagFlow;
		if (inside_yz && (or_inner & kVoxMesh)) {
			uint32_t mesh_bits = 0;
#pragma unroll
			for (int x = 0; x < 16; ++x)
				mesh_bits |= ((tm[x + 1] / kVoxMesh) & 1u) << x;
			const uint32_t cell0 = 16u * (uint32_t(z) + 16u * uint32_t(y));
			atomicOr(&s.custom_bits[cell0 >> 5], mesh_bits << (cell0 & 31));
			my_flags |= kFlagCustom;
		}
	}
	// the wavefront's flags, bit by bit (kFlagNonUniform .. kFlagInnerFaces), then one lane's LDS atomic:
	// an atomic from every lane with a different value is expanded into a scalar loop over the lanes
	uint32_t wave_flags = 0;
#pragma unroll
	for (int b = 0; b < 5; ++b)
		wave_flags |= ballot64((my_flags >> b) & 1u) ? 1u << b : 0u;
	if (lane == 0) {
		if (wave_flags)
			atomicOr(&s.flags, wave_flags);
		s.wave_ref[wave] = wave_ref; // 0xffffffff: the wavefront staged no row
	}
	__syncthreads();
	uint32_t flags = s.flags;
#pragma unroll
	for (int w = 1; w < NT / 64; ++w)
		if (s.wave_ref[w] != 0xffffffffu && s.wave_ref[w] != s.wave_ref[0])
			flags |= kFlagNonUniform;

	// one block value everywhere hides every face when the block's opposite textures do not show to each
	// other (true of every registered block; a registry may define one whose do)
	const bool uniform_hidden =
	    !(flags & kFlagNonUniform) && !(s.rows[block_pid(p, s, s.wave_ref[0])].y & kRowSelfShow);
	if (!(flags & kFlagCustom) && (uniform_hidden || !(flags & kFlagMixed) || !(flags & kFlagInnerFaces))) {
		// one block value everywhere whose faces hide each other, opaque cubes only, or no chunk voxel
		// with a face: no quads
		if (tid == 0)
			p.results[job] = hcm_job_result{HCM_OK, 0u, 0u, 0u};
		return;
	}

	// 2. liquid-flow cells take the flow mesh's face textures (:221-236)
	if (flags & kFlagFlow) {
		constexpr uint32_t kKeep = kVoxSunTorch | kVoxDirect | kVoxIndirect;
		for (int i = tid; i < kHaloVox; i += NT) {
			if (!(s.vox[i] & kVoxFlow)) // each thread rewrites only its own cells
				continue;
			const int hx = i % kHalo, rest = i / kHalo, hz = rest % kHalo, hy = rest / kHalo;
			const uint32_t pid = vox_pid(s.vox[i]);
			const uint32_t mesh = s.rows[pid].y >> 24;
			const Flow f = flow_heights(p, s.slots, hx - 1, hy - 1, hz - 1, p.reg.meshes[mesh].flow_height);
			const uint32_t tmpl = s.pid_vox[p.reg.flow_pid_base + flow_pattern(f)];
			s.vox[i] = (s.vox[i] & kKeep) | (tmpl & ~kKeep);
		}
		__syncthreads();
	}

	// 3. faces: visibility masks by bit slices; then, for each group of tasks whose visible faces fit the
	// stage, the face words (one lane per face, in dense order) and the link masks (one lane per row)
	for (int col = tid; col < 256; col += NT)
		visibility_column<0>(s, col);
	for (int col = tid; col < 256; col += NT)
		visibility_column<1>(s, col);
	for (int col = tid; col < 256; col += NT)
		visibility_column<2>(s, col);
	__syncthreads();
	if (wave == 0) { // exclusive prefix of visible faces over the tasks: lane l counts tasks l and l + 64,
	                 // packed in one word (a chunk has at most 24,576 visible faces), and one DPP scan
		uint32_t c = 0;
		if (lane < kTasks) {
			const uint64_t *r = reinterpret_cast<const uint64_t *>(s.ne[lane]);
			c = __popcll(r[0]) + __popcll(r[1]) + __popcll(r[2]) + __popcll(r[3]);
		}
		if (lane + 64 < kTasks) {
			const uint64_t *r = reinterpret_cast<const uint64_t *>(s.ne[lane + 64]);
			c |= uint32_t(__popcll(r[0]) + __popcll(r[1]) + __popcll(r[2]) + __popcll(r[3])) << 16;
		}
		const uint32_t inc = wave_incl_sum(c), ex = inc - c;
		const uint32_t tot = (uint32_t)__builtin_amdgcn_readlane((int)inc, 63), tot_a = tot & 0xffffu;
		if (lane < kTasks)
			s.fpre[lane] = ex & 0xffffu;
		if (lane + 64 < kTasks)
			s.fpre[lane + 64] = tot_a + (ex >> 16);
		if (lane == 63)
			s.fpre[kTasks] = tot_a + (tot >> 16);
	}
	__syncthreads();
	if (s.fpre[kTasks] == 0 && !(flags & kFlagCustom)) { // no visible face and no custom mesh: no quads
		if (tid == 0)
			p.results[job] = hcm_job_result{HCM_OK, 0u, 0u, 0u};
		return;
	}
	// face words per group of tasks, then the group's link masks; the last group's words stay in key. A
	// lane keeps the positions of its faces (at most kListCap / NT of them, 15 bits each, two to a
	// register) for the links and the emission, in a register queue shifted once per two faces (static
	// register indices)
	constexpr int kMine = (kListCap + NT - 1) / NT, kPairs = (kMine + 1) / 2;
	int t0_last = 0;
	uint32_t keep[kPairs]; // the positions of this lane's faces of the last group, for the emission
#pragma unroll 1
	for (int t0 = 0; t0 < kTasks;) {
		const int t1 = group_end(s, t0);
		const uint32_t g0 = s.fpre[t0], n = s.fpre[t1] - g0;
		for (int k = tid; k < (t1 - t0) * 8; k += NT) { // clear the group's link masks
			const int t = t0 + (k >> 3), q = k & 7;
			reinterpret_cast<uint32_t *>(s.eh[t])[q] = 0;
			reinterpret_cast<uint32_t *>(s.ev[t])[q] = 0;
			reinterpret_cast<uint32_t *>(s.tp[t])[q] = 0;
		}
		uint32_t mine[kPairs] = {};
#pragma unroll 1
		for (int k = 0; k < kPairs; ++k) {
			const uint32_t i0 = tid + 2 * k * NT, i1 = i0 + NT;
			uint32_t d0 = 0, d1 = 0;
			if (i0 < n)
				s.key[i0] = face_word(s, t0, t1, g0 + i0, d0);
			if (i1 < n)
				s.key[i1] = face_word(s, t0, t1, g0 + i1, d1);
#pragma unroll
			for (int j = 0; j + 1 < kPairs; ++j)
				mine[j] = mine[j + 1];
			mine[kPairs - 1] = d0 | d1 << 16;
		}
#pragma unroll
		for (int j = 0; j < kPairs; ++j)
			keep[j] = mine[j];
		__syncthreads();
#pragma unroll 1
		for (int k = 0; k < kPairs; ++k) {
			const uint32_t i0 = tid + 2 * k * NT, i1 = i0 + NT;
			if (i0 < n)
				face_links(s, i0, mine[0] & 0xffffu);
			if (i1 < n)
				face_links(s, i1, mine[0] >> 16);
#pragma unroll
			for (int j = 0; j + 1 < kPairs; ++j)
				mine[j] = mine[j + 1];
		}
		__syncthreads();
		t0_last = t0;
		t0 = t1;
	}

	// 4. counting: custom faces per thread range of cells, greedy quads per task (one lane per task)
	uint32_t c_o = 0, c_t = 0, q_o = 0, q_t = 0;
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
	// A task whose equal neighbouring faces are all side by side (no equal-to-below bit) has one quad per
	// run of equal faces along V: a visible face starts one unless it continues its left neighbour, and no
	// cell is covered from above. One whose equal faces are all above each other (no equal-to-right bit)
	// has one quad per run along U: a face starts one unless it continues the face above, and the cells
	// below a start are covered by its quad. With neither, every visible face is a 1 x 1 quad. These follow
	// from the masks directly (rows are 16-bit fields of the 64-bit words, row u + 4 q at bit 16 u, and a
	// row's equal-to-right bit 15 is never set). Tasks with both are listed and merged by one lane each, in
	// as few wavefronts as they fill. The list passes through qpre_t and the counts through qpre_o
	// (o | t << 16) before either holds its own values.
	if (tid < kTasks) { // the link masks are complete: the group loop ends with a barrier
		const int t = tid;
		const uint64_t *eh64 = reinterpret_cast<const uint64_t *>(s.eh[t]);
		const uint64_t *ev64 = reinterpret_cast<const uint64_t *>(s.ev[t]);
		uint64_t eh[4], ev[4], anyh = 0, anyv = 0;
#pragma unroll
		for (int q = 0; q < 4; ++q) {
			eh[q] = eh64[q], ev[q] = ev64[q];
			anyh |= eh[q], anyv |= ev[q];
		}
		if (anyh && anyv) {
			s.qpre_t[atomicAdd(&s.n_link, 1u)] = uint32_t(t);
		} else {
			uint32_t no = 0, nt = 0;
#pragma unroll
			for (int q = 0; q < 4; ++q) {
				uint64_t *tp64 = reinterpret_cast<uint64_t *>(s.tp[t]) + q;
				const uint64_t ne = reinterpret_cast<const uint64_t *>(s.ne[t])[q], tp = *tp64 & ne;
				// equal-to-below bits of the row above each row: the word shifted up one row, with the last
				// row of the previous word
				const uint64_t above = (ev[q] << 16) | (q > 0 ? ev[q - 1] >> 48 : 0ull);
				const uint64_t st = ne & ~(eh[q] << 1) & ~above;     // quad starts
				reinterpret_cast<uint64_t *>(s.ev[t])[q] = anyv ? st : ne; // uncovered visible cells
				reinterpret_cast<uint64_t *>(s.eh[t])[q] = st & ~tp;     // opaque quads start here
				*tp64 = st & tp;                                         // transparent quads start here
				no += __popcll(st & ~tp);
				nt += __popcll(st & tp);
			}
			s.qpre_o[t] = no | nt << 16;
		}
	}
	__syncthreads();
	{
		const uint32_t n_link = s.n_link;
		if (wave * 64 < (int)n_link) {
			const bool active = tid < (int)n_link;
			const int t = active ? (int)s.qpre_t[tid] : 0;
			uint32_t no, nt;
			greedy_count(s, t, active, no, nt);
			if (active)
				s.qpre_o[t] = no | nt << 16;
		}
	}
	__syncthreads();
	if (tid < kTasks) {
		const uint32_t c = s.qpre_o[tid];
		q_o = c & 0xffffu, q_t = c >> 16;
	}
	// 5. block-wide prefix sums and allocation: the chunk's output is [pool_base, +n0) opaque and
	// [+n0, +n0+n1) transparent; custom quads come first in each pass, then the tasks' quads in task order
	// (a chunk's regular quads per pass are at most its 24,576 faces, so both passes' counts share a word)
	uint32_t scan[3] = {q_o | q_t << 16, c_o, c_t}, tot[3];
	block_excl_sums(s, scan, tot);
	const uint64_t cust_pre = (uint64_t)scan[1] | ((uint64_t)scan[2] << 32);
	const uint32_t cust0 = tot[1], cust1 = tot[2];
	const uint32_t n0 = cust0 + (tot[0] & 0xffffu), n1 = cust1 + (tot[0] >> 16);
	const uint32_t total = n0 + n1;
	if (total == 0) {
		if (tid == 0)
			p.results[job] = hcm_job_result{HCM_OK, 0u, 0u, 0u};
		return;
	}
	if (tid < kTasks) {
		s.qpre_o[tid] = cust0 + (scan[0] & 0xffffu);
		s.qpre_t[tid] = cust1 + (scan[0] >> 16);
	}
	// allocation: the quad range and the mesh records, one atomic each, issued by two lanes of one
	// instruction so that both are in flight together
	const uint32_t nm0 = (n0 + kSplitQuads - 1) / kSplitQuads, nm1 = (n1 + kSplitQuads - 1) / kSplitQuads;
	if (tid < 2) {
		const uint32_t base = atomicAdd(tid == 0 ? &p.ctr->pool_cursor : &p.ctr->mesh_cursor, tid == 0 ? total : nm0 + nm1);
		if (tid == 0)
			s.pool_base = base;
		else
			s.mesh_base = base;
	}
	__syncthreads();
	const uint32_t pool_base = s.pool_base;
	if ((uint64_t)pool_base + total > p.pool_cap || (uint64_t)s.mesh_base + nm0 + nm1 > p.rec_cap) {
		if (tid == 0)
			p.results[job] = hcm_job_result{(int32_t)HCM_ERR_POOL_FULL, 0u, 0u, total};
		return;
	}

	// 6. emission. AABB<uint32_t>::Merge (dep/AABB.hpp:52-59) of the first mesh of each pass per lane,
	// reduced at the end; later meshes of a pass (past 16,383 quads) by LDS atomics
	uint32_t box[2][6];
#pragma unroll
	for (int a = 0; a < 6; ++a)
		box[0][a] = box[1][a] = a < 3 ? ~0u : 0u;
	// regular quads, one lane per visible face (a quad per face that starts one)
	const bool any_tr = __builtin_amdgcn_readfirstlane(int(n1)) != 0; // the same in every lane
	if (n0 > kSplitQuads || n1 > kSplitQuads)
		emit_regular<true>(p, s, keep, t0_last, n0, pool_base, box, any_tr);
	else
		emit_regular<false>(p, s, keep, t0_last, n0, pool_base, box, any_tr);
	// custom quads, after the regular ones (their positions in the output are fixed by the prefix sums): in
	// windows of the face-word stage, the threads list their cells' quads (cells in YZX order, faces in
	// mesh order, each thread its range of cells), then one lane per listed quad writes it
	if (flags & kFlagCustom) {
		const uint32_t n_cust = cust0 + cust1;
#pragma unroll 1
		for (uint32_t w0 = 0; w0 < n_cust; w0 += kListCap) {
			__syncthreads(); // key is free: the regular quads and the previous window are written
			const uint32_t w1 = w0 + kListCap;
			uint32_t io = (uint32_t)cust_pre, it = (uint32_t)(cust_pre >> 32);
			if (io + it < w1 && io + it + c_o + c_t > w0) {
				for (int k = 0; k < CPT && io + it < w1; ++k) {
					const int cell = tid * CPT + k;
					if (!((s.custom_bits[cell >> 5] >> (cell & 31)) & 1u))
						continue;
					const CustomCell c = custom_cell(p, s, cell);
					for (uint32_t j = 0; j < c.count; ++j) {
						const uint32_t tr = (c.trans >> j) & 1u, g = io + it, i = tr ? it++ : io++;
						if (g >= w0 && g < w1)
							s.key[g - w0] = make_uint2(uint32_t(cell) | (j << 12) | (tr << 31), i);
					}
				}
			}
			__syncthreads();
			const uint32_t n = min(n_cust - w0, (uint32_t)kListCap);
			for (uint32_t e = tid; e < n; e += NT) {
				const uint2 ent = s.key[e];
				const uint32_t tr = ent.x >> 31;
				uint32_t slot, vbase, mk, mj, lo[3], hi[3];
				quad_slot(tr, ent.y, n0, pool_base, slot, vbase, mk, mj);
				emit_custom(p, s, ent.x & 0xffffu, slot, vbase, lo, hi);
				quad_done(s, tr, mk, mj, w0 + e, lo, hi, box);
			}
		}
	}

	// the passes' boxes over the wavefront (a pass without quads is skipped: its boxes are empty)
	if (n0) {
#pragma unroll
		for (int a = 0; a < 6; ++a)
			box[0][a] = a < 3 ? wave_fold<false>(box[0][a]) : wave_fold<true>(box[0][a]);
	}
	if (n1) {
#pragma unroll
		for (int a = 0; a < 6; ++a)
			box[1][a] = a < 3 ? wave_fold<false>(box[1][a]) : wave_fold<true>(box[1][a]);
	}
	if (lane == 0 && (n0 | n1)) {
		if (n0)
			aabb_merge(s, 0, 0, box[0], box[0] + 3);
		if (n1)
			aabb_merge(s, 1, 0, box[1], box[1] + 3);
	}
	__syncthreads();

	// 7. mesh records in the reference's order: split meshes by time, then opaque, then transparent
	if (tid == 0) {
		const uint32_t ne = min(s.n_events, 16u);
		for (uint32_t a = 1; a < ne; ++a) // insertion sort by emission position
			for (uint32_t b = a; b > 0 && s.events[b - 1][0] > s.events[b][0]; --b) {
				const uint32_t e0 = s.events[b][0], e1 = s.events[b][1];
				s.events[b][0] = s.events[b - 1][0], s.events[b][1] = s.events[b - 1][1];
				s.events[b - 1][0] = e0, s.events[b - 1][1] = e1;
			}
		const uint32_t mbase = s.mesh_base;
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

template <int NT> __global__ __launch_bounds__(NT, NT >= 512 ? 2 : 3) void mesh_kernel(MeshParams p) {
	__shared__ Smem<NT> s;
	for (int i = threadIdx.x; i < kMaxPid; i += NT) {
		const bool ok = i < (int)p.reg.n_pid;
		s.rows[i] = ok ? p.reg.rows[i] : make_uint2(0, 0);
		s.pid_vox[i] = ok ? p.reg.pid_vox[i] : (uint32_t(i) << kVoxPidShift);
	}
	for (int i = threadIdx.x; i < kMaxCodes; i += NT)
		s.codes[i] = i < (int)p.reg.n_codes ? p.reg.codes[i] : (kCodeEmpty | kCodeTransparent) << 16;
	for (int i = threadIdx.x; i < 256; i += NT)
	{
		const uint32_t q = p.reg.pid_of[65536 + i];
		s.tmpl_id[i] = q == 0xffu ? kTmplMeta : p.reg.pid_vox[q];
	}
	if (!p.persistent) {
		const uint32_t job = p.job_base + blockIdx.x;
		if (job < p.n_jobs && threadIdx.x < 27)
			s.slots[threadIdx.x] = p.nbr[27 * (size_t)job + threadIdx.x];
		chunk_init(s);
		__syncthreads();
		if (job < p.n_jobs)
			mesh_chunk(p, s, job);
		return;
	}
	// part q of the job list is [q * n / kQueues, (q + 1) * n / kQueues); a part is empty once its
	// counter reaches its length, and the workgroup moves on to the next part
	const uint32_t home = blockIdx.x % kQueues;
	uint32_t tried = 0; // lane 0 of wavefront 0 only
	const int lane = threadIdx.x & 63;
	for (;;) {
		// wavefront 0 claims the next job and loads its slot list while the others set up its LDS state
		if (threadIdx.x < 64) {
			uint32_t job = ~0u;
			if (lane == 0) {
				while (job == ~0u && tried < kQueues) {
					const uint32_t q = (home + tried) % kQueues;
					const uint32_t lo = uint32_t(uint64_t(q) * p.n_jobs / kQueues);
					const uint32_t len = uint32_t(uint64_t(q + 1) * p.n_jobs / kQueues) - lo;
					const uint32_t k = atomicAdd(&p.ctr->queue[q].v, 1u);
					if (k < len)
						job = lo + k;
					else
						++tried;
				}
				s.job = job;
			}
			job = (uint32_t)__builtin_amdgcn_readfirstlane((int)job);
			if (job < p.n_jobs && lane < 27)
				s.slots[lane] = p.nbr[27 * (size_t)job + lane];
		}
		chunk_init(s);
		__syncthreads();
		const uint32_t job = s.job;
		if (job >= p.n_jobs)
			break;
		mesh_chunk(p, s, job);
		__syncthreads();
	}
	// every claim and allocation of this workgroup precedes its increment of `done`, so the last
	// workgroup to increment it sees final totals and may clear the counters for the next run
	if (p.self_reset && threadIdx.x == 0) {
		__threadfence();
		if (atomicAdd(&p.ctr->done, 1u) == gridDim.x - 1) {
			p.ctr->total_quads = atomicAdd(&p.ctr->pool_cursor, 0u);
			p.ctr->total_meshes = atomicAdd(&p.ctr->mesh_cursor, 0u);
			for (uint32_t q = 0; q < kQueues; ++q)
				p.ctr->queue[q].v = 0;
			p.ctr->pool_cursor = 0;
			p.ctr->mesh_cursor = 0;
			p.ctr->done = 0;
		}
	}
}

__global__ void reset_kernel(Counters *ctr) {
	if (threadIdx.x == 0) {
		for (uint32_t q = 0; q < kQueues; ++q)
			ctr->queue[q].v = 0;
		ctr->pool_cursor = 0;
		ctr->mesh_cursor = 0;
		ctr->done = 0;
	}
}

// Applies voxel edits in order of their index; edits to one voxel within a launch are resolved by the
// host, which keeps only the last one. An edit with local >= 4096 sets sunlight height local - 4096 of
// its slot to the low byte of `block` (heights may be null when no such edit is sent).
__global__ void edit_kernel(uint16_t *blocks, uint8_t *light, int8_t *heights, const VoxelEdit *edits, uint32_t n) {
	const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n) {
		const VoxelEdit e = edits[i];
		if (e.local >= HCM_CHUNK_VOXELS) {
			heights[(size_t)e.slot * 256 + (e.local - HCM_CHUNK_VOXELS)] = int8_t(uint8_t(e.block));
			return;
		}
		const size_t g = (size_t)e.slot * HCM_CHUNK_VOXELS + e.local;
		blocks[g] = uint16_t(e.block);
		light[g] = uint8_t(e.light);
	}
}

hipError_t launch_edits(uint16_t *blocks, uint8_t *light, int8_t *heights, const VoxelEdit *edits, uint32_t n,
                        hipStream_t stream) {
	if (n == 0)
		return hipSuccess;
	hipLaunchKernelGGL(edit_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, blocks, light, heights, edits, n);
	return hipGetLastError();
}

// Copies staged chunk uploads into their store slots, one workgroup per chunk: move k copies staged
// chunk moves[k].x to slot moves[k].y -- its 4096 blocks, and its 4096 light bytes or its 256 sunlight
// heights, whichever staging array is given.
__global__ void upload_scatter_kernel(const uint4 *sblocks, const uint4 *slight, const uint4 *sheights,
                                      const uint2 *moves, uint4 *blocks, uint4 *light, uint4 *heights) {
	const uint2 m = moves[blockIdx.x];
	const uint32_t t = threadIdx.x;
	constexpr uint32_t kB = HCM_CHUNK_VOXELS * 2 / 16, kL = HCM_CHUNK_VOXELS / 16, kH = 256 / 16;
	blocks[(size_t)m.y * kB + t] = sblocks[(size_t)m.x * kB + t];
	blocks[(size_t)m.y * kB + 256 + t] = sblocks[(size_t)m.x * kB + 256 + t];
	if (slight)
		light[(size_t)m.y * kL + t] = slight[(size_t)m.x * kL + t];
	if (sheights && t < kH)
		heights[(size_t)m.y * kH + t] = sheights[(size_t)m.x * kH + t];
}
static_assert(HCM_CHUNK_VOXELS * 2 / 16 == 512 && HCM_CHUNK_VOXELS / 16 == 256, "one 256-thread workgroup per chunk");

hipError_t launch_upload_scatter(const uint16_t *sblocks, const uint8_t *slight, const int8_t *sheights,
                                 const uint2 *moves, uint32_t n, uint16_t *blocks, uint8_t *light, int8_t *heights,
                                 hipStream_t stream) {
	if (n == 0)
		return hipSuccess;
	hipLaunchKernelGGL(upload_scatter_kernel, dim3(n), dim3(256), 0, stream, reinterpret_cast<const uint4 *>(sblocks),
	                   reinterpret_cast<const uint4 *>(slight), reinterpret_cast<const uint4 *>(sheights), moves,
	                   reinterpret_cast<uint4 *>(blocks), reinterpret_cast<uint4 *>(light),
	                   reinterpret_cast<uint4 *>(heights));
	return hipGetLastError();
}

// Scatters n packed chunk halos (include/hcm.h, HCM_HALO_BLOCKS / HCM_HALO_LIGHT; halo j at j times those
// sizes) into 27 store slots each, job j's chunks at slots 27 j .. 27 j + 26 in neighbour-slot order, where
// the mesher reads them; cells of the slots outside the halo are not written (the mesher never reads them).
constexpr int kUnpackBlocksPerHalo = (HCM_HALO_BLOCKS + 255) / 256;
__global__ void halo_unpack_kernel(const uint16_t *hblocks, const uint8_t *hlight, uint16_t *blocks, uint8_t *light) {
	const int job = blockIdx.x / kUnpackBlocksPerHalo;
	const int i = (blockIdx.x % kUnpackBlocksPerHalo) * 256 + threadIdx.x;
	const size_t base = (size_t)job * 27;
	if (i < HCM_HALO_BLOCKS) {
		const int x = i % 20 - 2, z = (i / 20) % 20 - 2, y = i / 400 - 1;
		blocks[(base + slot27(x, y, z)) * HCM_CHUNK_VOXELS + local16(x, y, z)] = hblocks[(size_t)job * HCM_HALO_BLOCKS + i];
	}
	if (i < HCM_HALO_LIGHT) {
		const int x = i % 18 - 1, z = (i / 18) % 18 - 1, y = i / 324 - 1;
		light[(base + slot27(x, y, z)) * HCM_CHUNK_VOXELS + local16(x, y, z)] = hlight[(size_t)job * HCM_HALO_LIGHT + i];
	}
}

hipError_t launch_halo_unpack(const uint16_t *hblocks, const uint8_t *hlight, uint32_t n, uint16_t *blocks,
                              uint8_t *light, hipStream_t stream) {
	if (n == 0)
		return hipSuccess;
	hipLaunchKernelGGL(halo_unpack_kernel, dim3(n * kUnpackBlocksPerHalo), dim3(256), 0, stream, hblocks, hlight, blocks,
	                   light);
	return hipGetLastError();
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

// gfx950 allocates a CU's 160 KB of LDS in 1,280-byte blocks: three 256-thread workgroups per CU need
// at most 42 blocks each (a 54,208-byte build ran at two per CU).
static_assert(sizeof(Smem<256>) <= 42 * 1280, "three 256-thread workgroups per CU");

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
