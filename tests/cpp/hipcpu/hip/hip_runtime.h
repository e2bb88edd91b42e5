// A CPU build of this project's HIP sources, for debugging kernels without a GPU (test infrastructure;
// the product library is built for gfx950 only). Every lane of a workgroup runs as a coroutine on its own
// stack, one workgroup at a time. A lane runs until it reaches a barrier or a wavefront operation
// (ballot, shuffle, readlane, DPP); an operation completes when every lane of the wavefront that has not
// finished is waiting, over the lanes waiting at that same call site -- the lanes a GPU would have active
// there -- and __syncthreads completes when every unfinished lane of the workgroup waits at a barrier.
// LDS is a static object shared by the workgroup's lanes; device memory is host memory; streams run
// synchronously and a captured graph replays the launches recorded into it.
// Limit: lanes do not reconverge. A wavefront operation inside a branch that only some lanes take lets the
// other lanes run on to the next operation after the branch, which can then complete over part of the
// wavefront (a fold that reads the missing lanes as stale values), where a GPU would run it only after
// the branch. The kernels keep wavefront operations out of such branches, or make them reached by every
// lane of the wavefront.
#pragma once

#include <ucontext.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)
#define HIP_SYMBOL(x) (&(x))

struct uint2 {
	uint32_t x, y;
};
struct uint4 {
	uint32_t x, y, z, w;
};
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }
struct dim3 {
	uint32_t x, y, z;
	constexpr dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};

namespace hipcpu {

enum LaneState { kRunnable, kAtBarrier, kAtWaveOp, kDone };
enum WaveOp { kBallot, kShfl, kShflUp, kShflXor, kReadLane, kReadFirst, kDpp };

struct Lane {
	ucontext_t ctx;
	uint32_t tid = 0;
	LaneState state = kRunnable;
	const void *site = nullptr;
	WaveOp op = kBallot;
	uint64_t val = 0, arg = 0, arg2 = 0, result = 0;
	std::unique_ptr<char[]> stack;
};

struct Block {
	dim3 bid, bdim, gdim;
	std::vector<Lane> lanes;
	ucontext_t sched;
	Lane *cur = nullptr;
	const std::function<void()> *body = nullptr;
};

Block *current();
void launch(dim3 grid, dim3 block, const std::function<void()> &body);
// parks the current lane at a wavefront operation or a barrier until it completes
uint64_t wait_wave(const void *site, WaveOp op, uint64_t val, uint64_t arg = 0, uint64_t arg2 = 0);
void wait_barrier();

inline dim3 thread_idx() { return dim3(current()->cur->tid, 0, 0); }
inline dim3 block_idx() { return current()->bid; }
inline dim3 block_dim() { return current()->bdim; }
inline dim3 grid_dim() { return current()->gdim; }

} // namespace hipcpu

#define threadIdx (::hipcpu::thread_idx())
#define blockIdx (::hipcpu::block_idx())
#define blockDim (::hipcpu::block_dim())
#define gridDim (::hipcpu::grid_dim())

#define __syncthreads() ::hipcpu::wait_barrier()
#define __threadfence() ((void)0)
namespace hipcpu {
// wavefront operations: a call site is told apart by the return address into its caller (each
// inlined copy of a caller has its own); the last argument is unused
__attribute__((noinline)) uint64_t site_ballot(bool p, const void *site);
__attribute__((noinline)) uint32_t site_shfl(uint32_t v, int src, const void *site);
__attribute__((noinline)) uint32_t site_shfl_up(uint32_t v, uint32_t d, const void *site);
__attribute__((noinline)) uint32_t site_shfl_xor(uint32_t v, uint32_t m, const void *site);
__attribute__((noinline)) uint32_t site_readlane(uint32_t v, uint32_t l, const void *site);
__attribute__((noinline)) uint32_t site_readfirst(uint32_t v, const void *site);
__attribute__((noinline)) uint32_t site_dpp(uint32_t old, uint32_t src, uint32_t ctrl, uint32_t row_mask, uint32_t bank_mask, const void *site);
} // namespace hipcpu

#define __ballot(p) ::hipcpu::site_ballot((p), nullptr)
#define __shfl(v, s) ::hipcpu::site_shfl((uint32_t)(v), (int)(s), nullptr)
#define __shfl_up(v, d) ::hipcpu::site_shfl_up((uint32_t)(v), (uint32_t)(d), nullptr)
#define __shfl_xor(v, m) ::hipcpu::site_shfl_xor((uint32_t)(v), (uint32_t)(m), nullptr)
#define __builtin_amdgcn_readlane(v, l) ((int)::hipcpu::site_readlane((uint32_t)(v), (uint32_t)(l), nullptr))
#define __builtin_amdgcn_readfirstlane(v) ((int)::hipcpu::site_readfirst((uint32_t)(v), nullptr))
#define __builtin_amdgcn_update_dpp(old, src, ctrl, rm, bm, bc) \
	((int)::hipcpu::site_dpp((uint32_t)(old), (uint32_t)(src), (ctrl), (rm), (bm), nullptr))

// integer intrinsics
inline int __popc(uint32_t v) { return __builtin_popcount(v); }
inline int __popcll(uint64_t v) { return __builtin_popcountll(v); }
inline int __ffs(uint32_t v) { return __builtin_ffs((int)v); }
inline int __ffs(int v) { return __builtin_ffs(v); }
inline int __clz(int v) { return v ? __builtin_clz((unsigned)v) : 32; }
inline int __clzll(long long v) { return v ? __builtin_clzll((unsigned long long)v) : 64; }
inline int __ffsll(unsigned long long v) { return __builtin_ffsll((long long)v); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline int __ffsll(unsigned long v) { return __builtin_ffsll((long long)v); }
inline int __ffsll(long v) { return __builtin_ffsll((long long)v); }
using std::isinf;
using std::isnan;
inline uint32_t min(uint32_t a, uint32_t b) { return a < b ? a : b; }
inline uint32_t max(uint32_t a, uint32_t b) { return a > b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline uint64_t min(uint64_t a, uint64_t b) { return a < b ? a : b; }
inline uint64_t max(uint64_t a, uint64_t b) { return a > b ? a : b; }

// atomics: lanes switch only at synchronisation points, so a plain update is atomic
template <typename T> inline T atomicAdd(T *p, T v) { T o = *p; *p = o + v; return o; }
template <typename T> inline T atomicOr(T *p, T v) { T o = *p; *p = o | v; return o; }
template <typename T> inline T atomicMin(T *p, T v) { T o = *p; *p = o < v ? o : v; return o; }
template <typename T> inline T atomicMax(T *p, T v) { T o = *p; *p = o > v ? o : v; return o; }

// the device library's correctly rounded float operations (build with -ffp-contract=off)
inline float __ocml_mul_rte_f32(float a, float b) { return a * b; }
inline float __ocml_add_rte_f32(float a, float b) { return a + b; }
inline float __ocml_sub_rte_f32(float a, float b) { return a - b; }

// ---- runtime API ----
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNoDevice = 100,
       hipErrorNotReady = 600 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1,
                            hipStreamCaptureModeRelaxed = 2 };
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2 };
struct ihipStream_t;
typedef ihipStream_t *hipStream_t;
struct ihipEvent_t;
typedef ihipEvent_t *hipEvent_t;
struct ihipGraph_t;
typedef ihipGraph_t *hipGraph_t;
typedef ihipGraph_t *hipGraphExec_t;
struct hipGraphNode_t_;
typedef hipGraphNode_t_ *hipGraphNode_t;
struct hipDeviceProp_t {
	char name[256];
	char gcnArchName[256];
	int multiProcessorCount;
	size_t totalGlobalMem;
	int major, minor;
};

extern "C" {
hipError_t hipMalloc(void **p, size_t n);
}
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc(reinterpret_cast<void **>(p), n); }
extern "C" {
hipError_t hipFree(void *p);
hipError_t hipHostMalloc(void **p, size_t n, unsigned flags = 0);
hipError_t hipHostFree(void *p);
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t st = nullptr);
hipError_t hipMemset(void *d, int v, size_t n);
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st = nullptr);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned f);
hipError_t hipStreamCreate(hipStream_t *s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned f);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *n, char *log, size_t sz);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned f);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDevice(int *d);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int d);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
hipError_t hipGetSymbolAddress(void **p, const void *symbol);
const char *hipGetErrorString(hipError_t e);
}

namespace hipcpu {
// a launch: runs now, or is recorded when the stream is being captured
void submit(hipStream_t s, const std::function<void()> &work);
} // namespace hipcpu

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...)                                                \
	::hipcpu::submit((stream), [=]() { ::hipcpu::launch(dim3(grid), dim3(block), [=]() { kernel(__VA_ARGS__); }); })
