// CPU unit test of the allocation cache behind the device and pinned-host pools
// (dmrgpp_amd/csrc/pool_cache.h), on malloc: reuse rules, the cap, that buffers in use are never
// returned to the allocator, that a stale release is counted and changes nothing, and that
// concurrent alloc / release keeps the books.
#include "pool_cache.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <set>
#include <thread>
#include <vector>

namespace {
int failures = 0;
#define CHECK(c)                                                             \
	do {                                                                     \
		if (!(c)) {                                                          \
			std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #c);      \
			++failures;                                                      \
		}                                                                    \
	} while (0)

std::set<void*>   live; // buffers the allocator has handed out and not taken back
std::atomic<long> alloc_calls { 0 };
std::mutex        live_m;

dmk::PoolCache make(size_t cap, size_t granule = 1024, long fail_after = -1)
{
	return dmk::PoolCache(
	    [fail_after](size_t b) -> void* {
		    if (fail_after >= 0 && alloc_calls.load() >= fail_after)
			    return nullptr;
		    ++alloc_calls;
		    void*                       p = std::malloc(b);
		    std::lock_guard<std::mutex> l(live_m);
		    live.insert(p);
		    return p;
	    },
	    [](void* p) {
		    {
			    std::lock_guard<std::mutex> l(live_m);
			    if (!live.erase(p))
				    std::printf("FAILED: allocator got back a buffer it does not hold\n"), ++failures;
		    }
		    std::free(p);
	    },
	    cap, granule);
}
} // namespace

int main()
{
	{ // reuse: a released buffer serves a request that fits (within 2x), not a larger one
		auto  pc = make(1 << 20);
		void* a  = pc.alloc(1500); // 2 granules
		pc.release(a);
		void* b = pc.alloc(1100); // 2 granules again: reused
		CHECK(b == a);
		void* c = pc.alloc(5000); // larger: fresh
		CHECK(c != a);
		pc.release(b);
		void* d = pc.alloc(100); // 1 granule: the cached 2-granule buffer fits within 2x
		CHECK(d == a);
		pc.release(c);
		pc.release(d);
		const auto k = pc.counters();
		CHECK(k.used_bytes == 0);
		CHECK(k.reuses == 2);
		CHECK(k.bad_releases == 0);
		pc.trim(0);
	}
	{ // a second release of a buffer while it sits in the cache, or of a pointer never handed
	  // out, is counted and changes nothing. (A second release after the buffer went to a new
	  // owner cannot be told from the new owner's own release: owners clear their pointer on
	  // release, pool_device_release / pool_host_release, so that it cannot happen.)
		auto  pc    = make(1 << 20);
		void* first = pc.alloc(1024);
		pc.release(first);
		pc.release(first); // stale: the buffer is cached, not in use
		CHECK(pc.counters().bad_releases == 1);
		CHECK(pc.counters().cached_bytes == 1024);
		void* second = pc.alloc(1024); // handed out once, not twice
		void* third  = pc.alloc(1024);
		CHECK(second == first);
		CHECK(third != second);
		pc.release(second);
		pc.release(third);
		pc.release(reinterpret_cast<void*>(0x1234)); // never handed out
		CHECK(pc.counters().bad_releases == 2);
		pc.trim(0);
	}
	{ // the cap returns cached buffers, largest first, and never a buffer in use
		auto               pc = make(4096, 1024);
		std::vector<void*> keep;
		for (int i = 0; i < 4; ++i)
			keep.push_back(pc.alloc(size_t(1024) * size_t(i + 1)));
		pc.release(keep[3]); // 4 KiB cached: at the cap
		CHECK(pc.counters().cached_bytes == 4096);
		pc.release(keep[2]); // 7 KiB > cap: the largest (4 KiB) goes back
		CHECK(pc.counters().cached_bytes == 3072);
		CHECK(pc.counters().frees == 1);
		for (int i = 0; i < 2; ++i)
			CHECK(live.count(keep[size_t(i)]) == 1); // in use: still allocated
		pc.trim(0);
		CHECK(pc.counters().cached_bytes == 0);
		CHECK(pc.owns(keep[0]) && pc.owns(keep[1]));
		pc.release(keep[0]);
		pc.release(keep[1]);
		pc.trim(0);
	}
	{ // allocator failure: the cache is emptied and the request retried, then an exception
		alloc_calls = 0;
		auto  pc    = make(1 << 20, 1024, 1); // the second allocator call fails
		void* a     = pc.alloc(1024);
		pc.release(a);
		bool thrown = false;
		try {
			(void)pc.alloc(8192); // does not fit the cached buffer; allocator refuses
		} catch (const std::runtime_error&) {
			thrown = true;
		}
		CHECK(thrown);
		CHECK(pc.counters().cached_bytes == 0); // trimmed before the retry
	}
	{ // concurrent alloc / release from 8 threads: the books balance and no buffer is shared
		alloc_calls = 0;
		auto                     pc = make(1 << 16, 256);
		std::atomic<int>         overlap { 0 };
		std::vector<std::thread> th;
		for (int t = 0; t < 8; ++t)
			th.emplace_back([&pc, &overlap, t] {
				for (int i = 0; i < 2000; ++i) {
					const size_t b = size_t(256) * size_t(1 + (i * 7 + t) % 5);
					auto*        p = static_cast<unsigned char*>(pc.alloc(b));
					for (size_t k = 0; k < b; ++k)
						p[k] = static_cast<unsigned char>(t);
					for (size_t k = 0; k < b; ++k)
						if (p[k] != static_cast<unsigned char>(t))
							++overlap;
					pc.release(p);
				}
			});
		for (auto& x : th)
			x.join();
		CHECK(overlap.load() == 0);
		const auto k = pc.counters();
		CHECK(k.used_bytes == 0);
		CHECK(k.bad_releases == 0);
		pc.trim(0);
	}
	CHECK(live.empty());
	if (failures == 0)
		std::printf("pool cache: all checks passed\n");
	return failures == 0 ? 0 : 1;
}
