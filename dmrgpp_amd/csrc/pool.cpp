#include "pool.h"

#include "pool_cache.h"

#include <algorithm>
#include <atomic>

namespace dmk {

namespace {

// Cached buffers are left to the process teardown: freeing them from a static destructor could
// run after the HIP runtime has shut down (the caches are never destroyed).
PoolCache& device_pool()
{
	static PoolCache* p = new PoolCache(
	    [](size_t bytes) -> void* {
		    void* q = nullptr;
		    return hipMalloc(&q, bytes) == hipSuccess ? q : ((void)hipGetLastError(), nullptr);
	    },
	    [](void* q) { (void)hipFree(q); }, size_t(48) << 30);
	return *p;
}

PoolCache& host_pool()
{
	static PoolCache* p = new PoolCache(
	    [](size_t bytes) -> void* {
		    void* q = nullptr;
		    return hipHostMalloc(&q, bytes, hipHostMallocDefault) == hipSuccess
		               ? q
		               : ((void)hipGetLastError(), nullptr);
	    },
	    [](void* q) { (void)hipHostFree(q); }, size_t(8) << 30);
	return *p;
}

} // namespace

void* pool_device_alloc(size_t bytes) { return device_pool().alloc(bytes); }
void  pool_device_free(void* p) { device_pool().release(p); }
bool  pool_device_owns(void* p) { return device_pool().owns(p); }
void* pool_host_alloc(size_t bytes) { return host_pool().alloc(bytes); }
void  pool_host_free(void* p) { host_pool().release(p); }
PoolCounters pool_device_counters() { return device_pool().counters(); }
PoolCounters pool_host_counters() { return host_pool().counters(); }

void parallel_for(size_t n, const std::function<void(size_t)>& fn, unsigned threads)
{
	if (threads == 0)
		threads = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
	threads = unsigned(std::min<size_t>(threads, n));
	if (threads <= 1) {
		for (size_t i = 0; i < n; ++i)
			fn(i);
		return;
	}
	// Dynamic scheduling: the items differ in cost (blocks of very different sizes).
	std::atomic<size_t>      next { 0 };
	std::vector<std::thread> pool;
	pool.reserve(threads);
	for (unsigned t = 0; t < threads; ++t)
		pool.emplace_back([&] {
			for (size_t i = next.fetch_add(1); i < n; i = next.fetch_add(1))
				fn(i);
		});
	for (auto& th : pool)
		th.join();
}

} // namespace dmk
