// hcm_range_alloc.h -- first-fit allocator of quad ranges, used by the serving loop (hcm_serve.cpp) to
// give each batch an arena of the vertex pool and each chunk's meshes their range. Host code only (no HIP),
// so the CPU checks build it with the sanitizers.
#pragma once

#include <algorithm>
#include <cstdint>
#include <iterator>
#include <map>

namespace hcm {

struct Range {
	uint64_t first = 0, count = 0;
};

// First-fit allocator of quad ranges in [0, capacity): free ranges by start, neighbours merged on release.
class RangeAllocator {
public:
	explicit RangeAllocator(uint64_t cap) : m_cap(cap), m_free_total(cap) {
		if (cap)
			m_free[0] = cap;
	}
	bool alloc(uint64_t n, uint64_t *first) {
		for (auto it = m_free.begin(); it != m_free.end(); ++it)
			if (it->second >= n) {
				*first = it->first;
				const uint64_t rest = it->second - n, at = it->first + n;
				m_free.erase(it);
				if (rest)
					m_free[at] = rest;
				m_free_total -= n;
				return true;
			}
		return false;
	}
	void release(uint64_t first, uint64_t n) {
		if (!n)
			return;
		m_free_total += n;
		auto it = m_free.emplace(first, n).first;
		auto nx = std::next(it);
		if (nx != m_free.end() && it->first + it->second == nx->first) {
			it->second += nx->second;
			m_free.erase(nx);
		}
		if (it != m_free.begin()) {
			auto pv = std::prev(it);
			if (pv->first + pv->second == it->first) {
				pv->second += it->second;
				m_free.erase(it);
			}
		}
	}
	uint64_t largest() const {
		uint64_t m = 0;
		for (auto &f : m_free)
			m = std::max(m, f.second);
		return m;
	}
	uint64_t free_total() const { return m_free_total; }
	uint64_t capacity() const { return m_cap; }

private:
	uint64_t m_cap, m_free_total;
	std::map<uint64_t, uint64_t> m_free; // first -> count
};

} // namespace hcm
