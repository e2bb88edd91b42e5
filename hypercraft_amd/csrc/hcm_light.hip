// hcm_light.hip -- gfx950 kernels of the light rebuild before meshing (client/src/ChunkMeshTask.cpp:48-111).
//
// The reference's mesh task rebuilds light over its chunk's [-15, 31)^3 neighbourhood: torchlight from
// the blocks' light levels, sunlight 15 at and above each column's sunlight height, then a first-in
// first-out propagation (client/include/client/BlockLightAlgo.hpp:44-73) that lowers the level by one
// per step into blocks that pass light indirectly, pruned where a level cannot reach the bound
// [-1, 17)^3 (BlockLightAlgo.hpp:32-42 with ChunkMeshTask.inl:53-56). The mesher reads only the 18^3
// cells of that bound. Along a path that ends in the bound with level >= 1, the level exceeds the
// distance to the bound at every cell, so the pruning never changes those cells: there the result is
// the max-plus fixpoint
//     L(v) = max(L0(v), max over the six neighbours u of L(u) - 1, when v passes light),
// whatever order the propagation runs in. This file computes it level by level on bit masks. With G_k
// the cells of level >= k,
//     G_15 = sources of level 15,   G_k = G_{k+1} | sources of level k | (dilate(G_{k+1}) & passable),
// one 64-bit word per x-row of the 46-wide window; at step k only rows within distance k - 1 of the
// bound can still reach it. A cell's level is the largest k with the cell in G_k.
//
// Two kernels:
//   light_prep_kernel  one workgroup per store slot the batch reads: packs its blocks' light pass, its
//                      sunlit cells (from the sunlight heights), its torchlight levels and its level-15
//                      torchlight sources into 16-bit x-rows (3.5 KB per slot);
//   light_kernel       one workgroup per job (persistent, per-XCD work queues, three per CU): assembles
//                      the 46^3 window from the 27 slots' packed rows in LDS, runs the sunlight and
//                      torchlight levels, and writes the 18^3 rebuilt light bytes (sun | torch << 4) the
//                      mesher reads.
#include <hip/hip_runtime.h>

#include "hcm_internal.h"

namespace hcm {

namespace {

constexpr int kW = 46;                 // window [-15, 31) per axis
constexpr int kWRows = kW * kW;        // x-rows, index wz + 46 wy
constexpr int kH = 18;                 // the mesher's halo [-1, 17) per axis
constexpr int kHRows = kH * kH;        // halo x-rows, index hz + 18 hy
constexpr int kHOff = 14;              // window index of halo index 0
constexpr uint64_t kWMask = (1ull << kW) - 1ull;
constexpr int kLightThreads = 512;
constexpr int kIn = kW - 2;          // the level steps update rows y, z in [-14, 30): window rows 1 .. 44
constexpr int kInRows = kIn * kIn;

// 53.4 KB: three workgroups per CU
struct LightSmem {
	uint64_t pass[kInRows]; // cells that pass light indirectly (Block::GetIndirectLightPass, Block.hpp:70)
	uint64_t g[2][kWRows];  // G_{k+1} and G_k
	uint64_t lv012[kHRows]; // the channel's level of each halo cell: bit planes 0, 1, 2 at bits 0, 18, 36
	uint32_t lv3[kHRows];   // and bit plane 3
	uint32_t slots[27];
	uint32_t changed[3]; // per level step (rotating): some row of the step changed
	uint32_t levels;     // torchlight levels present in the 27 chunks
	uint32_t job;
};

// CmpXYZ2NeighbourIndex (common/include/common/Position.hpp:138-145): L(-1) = 1, L(0) = 2, L(1) = 0
__device__ __forceinline__ int nbr_slot(int x, int y, int z) {
	const int cx = x < 0 ? -1 : (x >= 16 ? 1 : 0), cy = y < 0 ? -1 : (y >= 16 ? 1 : 0), cz = z < 0 ? -1 : (z >= 16 ? 1 : 0);
	return ((cx + 2) % 3) * 9 + ((cy + 2) % 3) * 3 + (cz + 2) % 3;
}

// The packed rows of window row (wy, wz): x in [-15, 0) from the -x chunk, [0, 16) from the centre
// column, [16, 31) from the +x chunk.
struct RowSrc {
	const uint16_t *l, *c, *r;
};
__device__ __forceinline__ RowSrc row_src(const LightParams &p, const LightSmem &s, int wy, int wz) {
	const int y = wy - 15, z = wz - 15;
	const uint32_t lr = uint32_t(z & 15) + 16u * uint32_t(y & 15);
	const size_t stride = (size_t)kLightPlanes * 256;
	return {p.packed + s.slots[nbr_slot(-1, y, z)] * stride + lr, p.packed + s.slots[nbr_slot(0, y, z)] * stride + lr,
	        p.packed + s.slots[nbr_slot(16, y, z)] * stride + lr};
}
__device__ __forceinline__ uint64_t plane_row(const RowSrc &r, int q) {
	const uint64_t a = r.l[q * 256], b = r.c[q * 256], c = r.r[q * 256];
	return (a >> 1) | (b << 15) | ((c & 0x7fffu) << 31);
}
// Cells of the row whose torchlight level is exactly k.
__device__ __forceinline__ uint64_t level_row(const RowSrc &r, uint32_t k) {
	uint64_t m = kWMask;
#pragma unroll
	for (int b = 0; b < 4; ++b) {
		const uint64_t t = plane_row(r, 2 + b);
		m &= (k >> b) & 1u ? t : ~t;
	}
	return m;
}

// One channel: G_15 is in s.g[0]; records each halo cell's level in s.lv012 / s.lv3. A step that adds
// no cell to its rows leaves G unchanged for every later step (each later step reads only rows of the
// earlier step), so the channel ends there unless sources enter at a lower level.
template <int NT>
__device__ __forceinline__ void propagate(const LightParams &p, LightSmem &s, bool inject) {
	const int tid = threadIdx.x;
	if (tid < 3)
		s.changed[tid] = 0;
	for (int hr = tid; hr < kHRows; hr += NT) {
		const int hy = hr / kH, hz = hr % kH;
		const uint64_t bits = (s.g[0][(hz + kHOff) + kW * (hy + kHOff)] >> kHOff) & 0x3ffffu;
		s.lv012[hr] = bits | (bits << 18) | (bits << 36); // level 15
		s.lv3[hr] = uint32_t(bits);
	}
	__syncthreads();
	int cur = 0;
#pragma unroll 1
	for (uint32_t k = 14; k >= 1; --k) {
		const int lo = 15 - int(k), n = 16 + 2 * int(k);
		const uint32_t inv = ((1u << 20) + uint32_t(n) - 1u) / uint32_t(n); // i / n = (i * inv) >> 20, i < n * n
		const bool inj = inject && ((s.levels >> k) & 1u);
		const uint64_t *gi = s.g[cur];
		uint64_t *go = s.g[cur ^ 1];
		if (tid == 0)
			s.changed[(k + 2) % 3] = 0; // the flag of step k - 1, last read before the previous barrier
		bool changed = false;
		for (int i = tid; i < n * n; i += NT) {
			const int qy = int((uint32_t(i) * inv) >> 20), wy = lo + qy, wz = lo + (i - qy * n);
			const int r = wz + kW * wy;
			const uint64_t g = gi[r];
			const uint64_t d = (g << 1) | (g >> 1) | gi[r - 1] | gi[r + 1] | gi[r - kW] | gi[r + kW];
			uint64_t nw = g | (d & s.pass[(wz - 1) + kIn * (wy - 1)]);
			if (inj)
				nw |= level_row(row_src(p, s, wy, wz), k);
			go[r] = nw;
			changed |= nw != g;
			const int hy = wy - kHOff, hz = wz - kHOff;
			if ((unsigned)hy < (unsigned)kH && (unsigned)hz < (unsigned)kH) {
				const uint64_t add = ((nw & ~g) >> kHOff) & 0x3ffffu;
				if (add) {
					const int hr = hz + kH * hy;
					const uint64_t m =
					    ((k & 1u) ? add : 0ull) | ((k & 2u) ? add << 18 : 0ull) | ((k & 4u) ? add << 36 : 0ull);
					if (m)
						s.lv012[hr] |= m;
					if (k & 8u)
						s.lv3[hr] |= uint32_t(add);
				}
			}
		}
		if (__ballot(changed) && (tid & 63) == 0)
			s.changed[k % 3] = 1;
		__syncthreads();
		cur ^= 1;
		if (!s.changed[k % 3] && !(inject && (s.levels & ((1u << k) - 1u))))
			break;
	}
}

// Wavefront 0: the slot list of `job` and the torchlight levels present in its 27 chunks, into LDS.
__device__ __forceinline__ void light_slots(const LightParams &p, LightSmem &s, uint32_t job) {
	const int lane = threadIdx.x & 63;
	uint32_t slot = 0, lv = 0;
	if (lane < 27) {
		slot = p.nbr[27 * (size_t)job + lane];
		s.slots[lane] = slot;
		lv = p.levels[slot];
	}
#pragma unroll
	for (int d = 1; d < 64; d <<= 1)
		lv |= __shfl_xor(lv, d);
	if (lane == 0)
		s.levels = lv & ~1u;
}

// One job; s.slots and s.levels are in place (a barrier after them). Every window row is read from the
// packed planes once: its light pass and sunlit cells into LDS, its level-15 torchlight sources into
// the reading thread's registers until the torchlight channel starts.
template <int NT> __device__ __forceinline__ void light_job(const LightParams &p, LightSmem &s, uint32_t job) {
	static_assert(NT >= kHRows, "one halo x-row per thread");
	constexpr int kRowsPerThread = (kWRows + NT - 1) / NT;
	const int tid = threadIdx.x;
	const bool torch = s.levels != 0, torch15 = (s.levels >> 15) & 1u;
	uint64_t src15[kRowsPerThread];
#pragma unroll
	for (int k = 0; k < kRowsPerThread; ++k) {
		const int r = tid + k * NT;
		src15[k] = 0;
		if (r < kWRows) {
			const int wy = r / kW, wz = r % kW;
			const RowSrc src = row_src(p, s, wy, wz);
			if (unsigned(wy - 1) < unsigned(kIn) && unsigned(wz - 1) < unsigned(kIn))
				s.pass[(wz - 1) + kIn * (wy - 1)] = plane_row(src, 0);
			s.g[0][r] = plane_row(src, 1); // sunlight: 15 at and above the sunlight height (Chunk.hpp:73-79)
			if (torch15)
				src15[k] = plane_row(src, 6);
		}
	}
	__syncthreads();
	// light bytes sun | torch << 4 (block/include/block/Light.hpp), one 20-byte record per halo x-row,
	// written once: a thread's halo row's sunlight nibbles wait in its registers for the torchlight ones
	const int hr = tid;
	uint32_t w[5] = {0, 0, 0, 0, 0};
	auto nibbles = [&](int shift) {
		if (hr < kHRows) {
			const uint64_t l012 = s.lv012[hr];
			const uint32_t pl[4] = {uint32_t(l012) & 0x3ffffu, uint32_t(l012 >> 18) & 0x3ffffu,
			                        uint32_t(l012 >> 36) & 0x3ffffu, s.lv3[hr]};
#pragma unroll
			for (int x = 0; x < kH; ++x) {
				uint32_t v = 0;
#pragma unroll
				for (int b = 0; b < 4; ++b)
					v |= ((pl[b] >> x) & 1u) << b;
				w[x >> 2] |= v << (8 * (x & 3) + shift);
			}
		}
	};
	propagate<NT>(p, s, false);
	nibbles(0);
	if (torch) {
		__syncthreads(); // every halo row's sunlight level is read
#pragma unroll
		for (int k = 0; k < kRowsPerThread; ++k)
			if (tid + k * NT < kWRows)
				s.g[0][tid + k * NT] = src15[k];
		__syncthreads();
		propagate<NT>(p, s, true);
		nibbles(4);
	}
	if (hr < kHRows) {
		uint32_t *o = reinterpret_cast<uint32_t *>(p.halo_light + (size_t)job * kLightJobBytes) + hr * (kLightRowBytes / 4);
#pragma unroll
		for (int q = 0; q < 5; ++q)
			o[q] = w[q];
	}
}

} // namespace

__global__ __launch_bounds__(256) void light_prep_kernel(LightParams p) {
	__shared__ uint32_t s_levels;
	const uint32_t slot = p.slots[blockIdx.x];
	const int row = threadIdx.x, y = row >> 4, z = row & 15;
	if (row == 0)
		s_levels = 0;
	const uint16_t *src = p.blocks + (size_t)slot * HCM_CHUNK_VOXELS + row * 16;
	const uint4 b0 = *reinterpret_cast<const uint4 *>(src), b1 = *reinterpret_cast<const uint4 *>(src + 8);
	const uint4 hv = *reinterpret_cast<const uint4 *>(p.heights + (size_t)slot * 256 + 16 * z);
	const uint32_t bw[8] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
	const uint32_t hw[4] = {hv.x, hv.y, hv.z, hv.w};
	uint32_t info[16];
#pragma unroll
	for (int x = 0; x < 16; ++x)
		info[x] = p.light_info[(bw[x >> 1] >> (16 * (x & 1))) & 0xffffu];
	uint32_t pass = 0, sun = 0, t0 = 0, t1 = 0, t2 = 0, t3 = 0, t15 = 0, lv = 0;
#pragma unroll
	for (int x = 0; x < 16; ++x) {
		const int h = int(int8_t(hw[x >> 2] >> (8 * (x & 3))));
		sun |= uint32_t(y >= h) << x; // Chunk::GetSunlight (Chunk.hpp:73-76)
		pass |= ((info[x] & kLightInfoPass) ? 1u : 0u) << x;
		const uint32_t l = info[x] & 15u;
		t0 |= (l & 1u) << x, t1 |= ((l >> 1) & 1u) << x, t2 |= ((l >> 2) & 1u) << x, t3 |= ((l >> 3) & 1u) << x;
		t15 |= (l == 15u ? 1u : 0u) << x;
		lv |= 1u << l;
	}
	uint16_t *dst = p.packed + (size_t)slot * kLightPlanes * 256 + row;
	dst[0] = uint16_t(pass);
	dst[256] = uint16_t(sun);
	dst[512] = uint16_t(t0);
	dst[768] = uint16_t(t1);
	dst[1024] = uint16_t(t2);
	dst[1280] = uint16_t(t3);
	dst[1536] = uint16_t(t15);
	__syncthreads();
	if (lv & ~1u)
		atomicOr(&s_levels, lv & ~1u);
	__syncthreads();
	if (row == 0)
		p.levels[slot] = s_levels;
}

template <int NT> __global__ __launch_bounds__(NT, 3) void light_kernel(LightParams p) {
	__shared__ LightSmem s;
	// part q of the job list is [q * n / kQueues, (q + 1) * n / kQueues): a workgroup takes jobs from
	// its own XCD's part first, then from the others
	const uint32_t home = blockIdx.x % kQueues;
	uint32_t tried = 0; // lane 0 of wavefront 0 only
	for (;;) {
		if (threadIdx.x < 64) { // wavefront 0 claims the next job and loads its slot list and levels
			uint32_t job = ~0u;
			if (threadIdx.x == 0) {
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
			job = (uint32_t)__shfl((int)job, 0);
			if (job < p.n_jobs)
				light_slots(p, s, job);
		}
		__syncthreads();
		const uint32_t job = s.job;
		if (job >= p.n_jobs)
			break;
		light_job<NT>(p, s, job);
		__syncthreads();
	}
	// as in the mesh kernel: the last workgroup to finish clears the queues for the next run
	if (p.self_reset && threadIdx.x == 0) {
		__threadfence();
		if (atomicAdd(&p.ctr->done, 1u) == gridDim.x - 1) {
			for (uint32_t q = 0; q < kQueues; ++q)
				p.ctr->queue[q].v = 0;
			p.ctr->done = 0;
		}
	}
}

hipError_t launch_light_prep(const LightParams &p, hipStream_t stream) {
	if (p.n_slots == 0)
		return hipSuccess;
	hipLaunchKernelGGL(light_prep_kernel, dim3(p.n_slots), dim3(256), 0, stream, p);
	return hipGetLastError();
}

hipError_t launch_light(const LightParams &p, int grid, hipStream_t stream) {
	hipLaunchKernelGGL(light_kernel<kLightThreads>, dim3(grid), dim3(kLightThreads), 0, stream, p);
	return hipGetLastError();
}

int light_lds_bytes() {
	return (int)sizeof(LightSmem);
}

} // namespace hcm
