// Size-keyed cache of allocations, independent of the allocator behind it (device memory,
// pinned host memory, or plain malloc in the CPU unit test of this class).
//
// A released buffer is kept and handed out again to a request it fits (smallest cached buffer
// that is large enough and at most twice the request); requests are rounded up to 2 MiB
// granules. Buffers beyond `cap` cached bytes are returned to the allocator, largest first.
// Only cached buffers are ever returned to the allocator: a buffer in use stays valid until its
// owner releases it.
//
// Releasing a pointer the cache does not hold as in use (released twice, or never handed out)
// changes nothing and is counted in bad_releases(): the caller has a stale pointer, which the
// counter makes visible instead of silently handing a live buffer to a second owner.
#pragma once

#include <cstddef>
#include <functional>
#include <iterator>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace dmk {

struct PoolCounters {
	size_t used_bytes   = 0; // handed out and not yet released
	size_t cached_bytes = 0; // released and kept for reuse
	size_t allocs       = 0; // allocator calls (cache misses)
	size_t frees        = 0; // buffers returned to the allocator
	size_t reuses       = 0; // requests served from the cache
	size_t bad_releases = 0; // releases of pointers not held as in use
};

class PoolCache {
public:
	using AllocFn = std::function<void*(size_t)>; // nullptr on failure
	using FreeFn  = std::function<void(void*)>;

	PoolCache(AllocFn alloc, FreeFn release, size_t cap, size_t granule = size_t(2) << 20)
	    : alloc_(std::move(alloc))
	    , free_fn_(std::move(release))
	    , cap_(cap)
	    , granule_(granule)
	{ }

	// Never returns nullptr: throws std::runtime_error when the allocator fails even after the
	// cache was emptied.
	void* alloc(size_t bytes)
	{
		bytes = round(bytes);
		{
			std::lock_guard<std::mutex> lock(m_);
			auto                        it = free_.lower_bound(bytes);
			if (it != free_.end() && it->first <= 2 * bytes) {
				void* p = it->second;
				c_.cached_bytes -= it->first;
				c_.used_bytes += it->first;
				c_.reuses++;
				used_[p] = it->first;
				free_.erase(it);
				return p;
			}
		}
		void* p = alloc_(bytes);
		if (!p) {
			trim(0); // give the cache back and try once more
			p = alloc_(bytes);
		}
		if (!p)
			throw std::runtime_error("pool: allocation of " + std::to_string(bytes)
			                         + " bytes failed");
		std::lock_guard<std::mutex> lock(m_);
		used_[p] = bytes;
		c_.used_bytes += bytes;
		c_.allocs++;
		return p;
	}

	void release(void* p)
	{
		if (!p)
			return;
		{
			std::lock_guard<std::mutex> lock(m_);
			auto                        it = used_.find(p);
			if (it == used_.end()) {
				c_.bad_releases++;
				return;
			}
			free_.emplace(it->second, p);
			c_.cached_bytes += it->second;
			c_.used_bytes -= it->second;
			used_.erase(it);
		}
		trim(cap_);
	}

	bool owns(void* p)
	{
		std::lock_guard<std::mutex> lock(m_);
		return used_.count(p) != 0;
	}

	// Size of the buffer p was handed out with (0 when p is not in use).
	size_t size_of(void* p)
	{
		std::lock_guard<std::mutex> lock(m_);
		auto                        it = used_.find(p);
		return it == used_.end() ? 0 : it->second;
	}

	PoolCounters counters()
	{
		std::lock_guard<std::mutex> lock(m_);
		return c_;
	}

	// Returns cached buffers to the allocator, largest first, until at most `limit` bytes stay
	// cached. Buffers in use are never touched.
	void trim(size_t limit)
	{
		std::vector<void*> drop;
		{
			std::lock_guard<std::mutex> lock(m_);
			while (c_.cached_bytes > limit && !free_.empty()) {
				auto it = std::prev(free_.end());
				c_.cached_bytes -= it->first;
				c_.frees++;
				drop.push_back(it->second);
				free_.erase(it);
			}
		}
		for (void* q : drop)
			free_fn_(q);
	}

private:
	size_t round(size_t b) const
	{
		return b == 0 ? granule_ : ((b + granule_ - 1) / granule_) * granule_;
	}

	AllocFn                            alloc_;
	FreeFn                             free_fn_;
	const size_t                       cap_;
	const size_t                       granule_;
	std::mutex                         m_;
	std::multimap<size_t, void*>       free_;
	std::unordered_map<void*, size_t>  used_;
	PoolCounters                       c_;
};

} // namespace dmk
