// A CPU stand-in for the few HIP runtime calls the serving loop (hcm_serve.cpp) makes, so that the
// server's host code -- request queues, scheduler thread, world model, pool allocator -- builds with g++
// under AddressSanitizer, UndefinedBehaviorSanitizer and ThreadSanitizer (tests/cpp/serve_check.cpp).
// "Device" memory is host memory, copies are synchronous, and an event reports "not ready" to its
// first two queries after each record, so batches stay in flight across scheduler passes. The kernels
// are replaced by tests/cpp/serve_stub_device.cpp.
#pragma once

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#define __host__
#define __device__

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2, hipErrorNotReady = 600 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum { hipStreamNonBlocking = 1, hipEventDisableTiming = 2 };

struct ihipStream_t {
	int unused;
};
typedef ihipStream_t *hipStream_t;
struct ihipEvent_t {
	std::atomic<int> pending{0};
};
typedef ihipEvent_t *hipEvent_t;

struct uint2 {
	uint32_t x, y;
};
inline uint2 make_uint2(uint32_t x, uint32_t y) { return uint2{x, y}; }

inline hipError_t hipMalloc(void **p, size_t n) { // 256-byte aligned and zeroed, like device allocations here
	const size_t m = (n + 255) / 256 * 256;
	*p = std::aligned_alloc(256, m ? m : 256);
	if (*p)
		std::memset(*p, 0, m ? m : 256);
	return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void *p) {
	std::free(p);
	return hipSuccess;
}
inline hipError_t hipHostMalloc(void **p, size_t n) { return hipMalloc(p, n); }
inline hipError_t hipHostFree(void *p) { return hipFree(p); }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
	if (n)
		std::memmove(d, s, n);
	return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t) {
	return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemset(void *d, int v, size_t n) {
	std::memset(d, v, n);
	return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
	*s = new ihipStream_t{0};
	return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t s) {
	delete s;
	return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) {
	*e = new ihipEvent_t;
	return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t *e) { return hipEventCreateWithFlags(e, 0); }
inline hipError_t hipEventDestroy(hipEvent_t e) {
	delete e;
	return hipSuccess;
}
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
	e->pending = 2;
	return hipSuccess;
}
inline hipError_t hipEventQuery(hipEvent_t e) {
	if (e->pending.load() > 0) {
		--e->pending;
		return hipErrorNotReady;
	}
	return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t e) {
	e->pending = 0;
	return hipSuccess;
}
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
	*ms = 0.01f;
	return hipSuccess;
}
