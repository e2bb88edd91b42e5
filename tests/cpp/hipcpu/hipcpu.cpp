// The CPU build's lane scheduler and runtime calls (see hip/hip_runtime.h).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>

namespace hipcpu {

static Block *g_block = nullptr;
constexpr size_t kStack = 256 * 1024;

Block *current() { return g_block; }

static void lane_entry(uint32_t lo, uint32_t hi) {
	Block *b = reinterpret_cast<Block *>((uintptr_t(hi) << 32) | uintptr_t(lo));
	(*b->body)();
	b->cur->state = kDone;
	swapcontext(&b->cur->ctx, &b->sched);
}

static void yield() { swapcontext(&g_block->cur->ctx, &g_block->sched); }

uint64_t wait_wave(const void *site, WaveOp op, uint64_t val, uint64_t arg, uint64_t arg2) {
	Lane *l = g_block->cur;
	l->state = kAtWaveOp;
	l->site = site, l->op = op, l->val = val, l->arg = arg, l->arg2 = arg2;
	yield();
	return l->result;
}

void wait_barrier() {
	g_block->cur->state = kAtBarrier;
	yield();
}

// DPP source lane for the controls the kernels use; -1: no source (the lane keeps its old value)
static int dpp_src(uint32_t ctrl, int lane) {
	const int row = lane >> 4, in = lane & 15;
	if (ctrl <= 0xffu) // quad_perm
		return (lane & ~3) | int((ctrl >> (2 * (lane & 3))) & 3u);
	switch (ctrl) {
	case 0x140: return (row << 4) | (15 - in);                  // row_mirror
	case 0x141: return (row << 4) | ((in & 8) | (7 - (in & 7))); // row_half_mirror
	case 0x111: case 0x112: case 0x113: case 0x114: case 0x115: case 0x116: case 0x117: case 0x118:
	case 0x119: case 0x11a: case 0x11b: case 0x11c: case 0x11d: case 0x11e: case 0x11f: { // row_shr:n
		const int n = int(ctrl & 15u);
		return in >= n ? lane - n : -1;
	}
	case 0x142: return row >= 1 ? ((row - 1) << 4) | 15 : -1;    // row_bcast:15
	case 0x143: return row >= 2 ? 31 : -1;                       // row_bcast:31
	default:
		std::fprintf(stderr, "hipcpu: DPP control 0x%x not emulated\n", ctrl);
		std::abort();
	}
}

// Completes the operation at the site of the first waiting lane of the wavefront over all lanes waiting
// there. Returns false if some unfinished lane of the wavefront is still runnable.
static bool resolve_wave(Block *b, int w) {
	const int n = (int)b->lanes.size();
	const int lo = 64 * w, hi = std::min(n, lo + 64);
	Lane *first = nullptr;
	for (int i = lo; i < hi; ++i) {
		Lane &l = b->lanes[i];
		if (l.state == kRunnable)
			return false;
		if (l.state == kAtWaveOp && !first)
			first = &l;
	}
	if (!first)
		return false;
	const void *site = first->site;
	const WaveOp op = first->op;
	bool active[64] = {};
	uint64_t v[64] = {};
	uint64_t ballot = 0;
	int first_lane = -1;
	for (int i = lo; i < hi; ++i) {
		Lane &l = b->lanes[i];
		if (l.state == kAtWaveOp && l.site == site) {
			if (l.op != op) {
				std::fprintf(stderr, "hipcpu: two operations at one call site\n");
				std::abort();
			}
			active[i - lo] = true;
			v[i - lo] = l.val;
			if (op == kBallot && l.val)
				ballot |= 1ull << (i - lo);
			if (first_lane < 0)
				first_lane = i - lo;
		}
	}
	for (int i = lo; i < hi; ++i) {
		Lane &l = b->lanes[i];
		const int k = i - lo;
		if (!active[k])
			continue;
		switch (op) {
		case kBallot: l.result = ballot; break;
		case kShfl: { const int s = int(l.arg) & 63; l.result = active[s] ? v[s] : 0; break; }
		case kShflUp: { const int s = k - int(l.arg); l.result = s >= 0 && active[s] ? v[s] : v[k]; break; }
		case kShflXor: { const int s = (k ^ int(l.arg)) & 63; l.result = active[s] ? v[s] : v[k]; break; }
		case kReadLane: { const int s = int(l.arg) & 63; l.result = active[s] ? v[s] : 0; break; }
		case kReadFirst: l.result = v[first_lane]; break;
		case kDpp: {
			const uint32_t ctrl = uint32_t(l.arg), row_mask = uint32_t(l.arg2) & 15u, bank_mask = uint32_t(l.arg2 >> 8) & 15u;
			const uint32_t old = uint32_t(l.val >> 32);
			const int s = dpp_src(ctrl, k);
			const bool write = ((row_mask >> (k >> 4)) & 1u) && ((bank_mask >> ((k >> 2) & 3)) & 1u);
			l.result = write && s >= 0 && active[s] ? uint32_t(v[s]) : old;
			break;
		}
		}
		l.state = kRunnable;
	}
	return true;
}

void launch(dim3 grid, dim3 block, const std::function<void()> &body) {
	const int n = int(block.x * block.y * block.z);
	Block b;
	b.bdim = block, b.gdim = grid;
	b.body = &body;
	b.lanes.resize(n);
	for (auto &l : b.lanes)
		l.stack.reset(new char[kStack]);
	Block *saved = g_block;
	for (uint32_t bx = 0; bx < grid.x; ++bx) {
		b.bid = dim3(bx, 0, 0);
		g_block = &b;
		for (int i = 0; i < n; ++i) {
			Lane &l = b.lanes[i];
			l.tid = uint32_t(i);
			l.state = kRunnable;
			getcontext(&l.ctx);
			l.ctx.uc_stack.ss_sp = l.stack.get();
			l.ctx.uc_stack.ss_size = kStack;
			l.ctx.uc_link = nullptr;
			const uintptr_t bp = reinterpret_cast<uintptr_t>(&b);
			makecontext(&l.ctx, (void (*)())lane_entry, 2, uint32_t(bp), uint32_t(bp >> 32));
		}
		for (;;) {
			bool ran = false;
			for (int i = 0; i < n; ++i)
				if (b.lanes[i].state == kRunnable) {
					b.cur = &b.lanes[i];
					swapcontext(&b.sched, &b.lanes[i].ctx);
					ran = true;
				}
			bool all_done = true, all_barrier = true;
			for (auto &l : b.lanes) {
				all_done &= l.state == kDone;
				all_barrier &= l.state == kDone || l.state == kAtBarrier;
			}
			if (all_done)
				break;
			bool progress = false;
			for (int w = 0; w * 64 < n; ++w)
				progress |= resolve_wave(&b, w);
			if (progress || ran)
				continue;
			if (all_barrier) {
				for (auto &l : b.lanes)
					if (l.state == kAtBarrier)
						l.state = kRunnable;
				continue;
			}
			std::fprintf(stderr, "hipcpu: deadlock in block %u\n", bx);
			std::abort();
		}
	}
	g_block = saved;
}

uint64_t site_ballot(bool p, const void *) { return wait_wave(__builtin_return_address(0), kBallot, p ? 1 : 0); }
uint32_t site_shfl(uint32_t v, int src, const void *) {
	return (uint32_t)wait_wave(__builtin_return_address(0), kShfl, v, (uint64_t)src);
}
uint32_t site_shfl_up(uint32_t v, uint32_t d, const void *) {
	return (uint32_t)wait_wave(__builtin_return_address(0), kShflUp, v, d);
}
uint32_t site_shfl_xor(uint32_t v, uint32_t m, const void *) {
	return (uint32_t)wait_wave(__builtin_return_address(0), kShflXor, v, m);
}
uint32_t site_readlane(uint32_t v, uint32_t l, const void *) {
	return (uint32_t)wait_wave(__builtin_return_address(0), kReadLane, v, l);
}
uint32_t site_readfirst(uint32_t v, const void *) { return (uint32_t)wait_wave(__builtin_return_address(0), kReadFirst, v); }
uint32_t site_dpp(uint32_t old, uint32_t src, uint32_t ctrl, uint32_t row_mask, uint32_t bank_mask, const void *) {
	return (uint32_t)wait_wave(__builtin_return_address(0), kDpp, (uint64_t(old) << 32) | src, ctrl,
	                           row_mask | (bank_mask << 8));
}

// ---- streams, events, graphs ----
struct Graph {
	std::vector<std::function<void()>> work;
};
static std::mutex g_mu;
static std::map<hipStream_t, Graph *> g_capture;
static int g_stream_id = 0;

void submit(hipStream_t s, const std::function<void()> &work) {
	{
		std::lock_guard<std::mutex> lk(g_mu);
		auto it = g_capture.find(s);
		if (it != g_capture.end()) {
			it->second->work.push_back(work);
			return;
		}
	}
	work();
}

} // namespace hipcpu

using namespace hipcpu;

extern "C" {

hipError_t hipMalloc(void **p, size_t n) {
	const size_t m = (n + 255) / 256 * 256;
	*p = std::aligned_alloc(256, m ? m : 256);
	if (*p)
		std::memset(*p, 0, m ? m : 256);
	return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) {
	std::free(p);
	return hipSuccess;
}
hipError_t hipHostMalloc(void **p, size_t n, unsigned) { return hipMalloc(p, n); }
hipError_t hipHostFree(void *p) { return hipFree(p); }
hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
	if (n)
		std::memmove(d, s, n);
	return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t st) {
	submit(st, [=] { if (n) std::memmove(d, s, n); });
	return hipSuccess;
}
hipError_t hipMemset(void *d, int v, size_t n) {
	std::memset(d, v, n);
	return hipSuccess;
}
hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t st) {
	submit(st, [=] { std::memset(d, v, n); });
	return hipSuccess;
}
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
	std::lock_guard<std::mutex> lk(g_mu);
	*s = reinterpret_cast<hipStream_t>(uintptr_t(++g_stream_id));
	return hipSuccess;
}
hipError_t hipStreamCreate(hipStream_t *s) { return hipStreamCreateWithFlags(s, 0); }
hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
	std::lock_guard<std::mutex> lk(g_mu);
	g_capture[s] = new Graph;
	return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
	std::lock_guard<std::mutex> lk(g_mu);
	*g = reinterpret_cast<hipGraph_t>(g_capture.at(s));
	g_capture.erase(s);
	return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
	*e = reinterpret_cast<hipGraphExec_t>(new Graph(*reinterpret_cast<Graph *>(g)));
	return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s) {
	for (auto &w : reinterpret_cast<Graph *>(e)->work)
		submit(s, w);
	return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
	delete reinterpret_cast<Graph *>(g);
	return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
	delete reinterpret_cast<Graph *>(e);
	return hipSuccess;
}
struct ihipEvent_t {
	int unused;
};
hipError_t hipEventCreate(hipEvent_t *e) {
	*e = new ihipEvent_t{0};
	return hipSuccess;
}
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
hipError_t hipEventDestroy(hipEvent_t e) {
	delete e;
	return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventQuery(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) {
	*ms = 0.001f;
	return hipSuccess;
}
hipError_t hipGetDeviceCount(int *n) {
	*n = 1;
	return hipSuccess;
}
hipError_t hipGetDevice(int *d) {
	*d = 0;
	return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
	std::memset(p, 0, sizeof(*p));
	std::strcpy(p->name, "CPU build of the gfx950 kernels");
	std::strcpy(p->gcnArchName, "gfx950:sramecc+:xnack-");
	p->multiProcessorCount = 2;
	p->totalGlobalMem = size_t(64) << 30;
	p->major = 9, p->minor = 5;
	return hipSuccess;
}
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetLastError() { return hipSuccess; }
hipError_t hipGetSymbolAddress(void **p, const void *symbol) {
	*p = const_cast<void *>(symbol);
	return hipSuccess;
}
const char *hipGetErrorString(hipError_t) { return "hipcpu"; }

} // extern "C"
