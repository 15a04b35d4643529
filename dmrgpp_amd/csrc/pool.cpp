#include "pool.h"

#include <atomic>

#include "dmk_internal.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace dmk {

namespace {

class Pool {
public:
	Pool(bool host, size_t cap)
	    : host_(host)
	    , cap_(cap)
	{ }

	void* alloc(size_t bytes)
	{
		bytes = round(bytes);
		{
			std::lock_guard<std::mutex> lock(m_);
			auto                        it = free_.lower_bound(bytes);
			if (it != free_.end() && it->first <= 2 * bytes) {
				void* p = it->second;
				cached_ -= it->first;
				used_[p] = it->first;
				free_.erase(it);
				return p;
			}
		}
		void*      p = nullptr;
		hipError_t e = host_ ? hipHostMalloc(&p, bytes, hipHostMallocDefault) : hipMalloc(&p, bytes);
		if (e != hipSuccess) {
			// Give the cache back to the runtime and try once more.
			trim(0);
			e = host_ ? hipHostMalloc(&p, bytes, hipHostMallocDefault) : hipMalloc(&p, bytes);
		}
		if (e != hipSuccess)
			throw std::runtime_error(std::string(host_ ? "hipHostMalloc" : "hipMalloc") + ": "
			                         + hipGetErrorString(e));
		std::lock_guard<std::mutex> lock(m_);
		used_[p] = bytes;
		return p;
	}

	void release(void* p)
	{
		if (!p)
			return;
		{
			std::lock_guard<std::mutex> lock(m_);
			auto                        it = used_.find(p);
			if (it == used_.end())
				return;
			free_.emplace(it->second, p);
			cached_ += it->second;
			used_.erase(it);
		}
		trim(cap_);
	}

	// Cached buffers are left to the process teardown: freeing them from a static destructor
	// could run after the HIP runtime has shut down.
	~Pool() = default;

	bool owns(void* p)
	{
		std::lock_guard<std::mutex> lock(m_);
		return used_.count(p) != 0;
	}

private:
	static size_t round(size_t b)
	{
		const size_t g = size_t(2) << 20; // 2 MiB granules
		return b == 0 ? g : ((b + g - 1) / g) * g;
	}

	// Frees cached buffers, largest first, until at most `limit` bytes stay cached.
	void trim(size_t limit)
	{
		std::vector<void*> drop;
		{
			std::lock_guard<std::mutex> lock(m_);
			while (cached_ > limit && !free_.empty()) {
				auto it = std::prev(free_.end());
				cached_ -= it->first;
				drop.push_back(it->second);
				free_.erase(it);
			}
		}
		for (void* p : drop) {
			if (host_)
				(void)hipHostFree(p);
			else
				(void)hipFree(p);
		}
	}

	const bool                        host_;
	const size_t                      cap_;
	std::mutex                        m_;
	std::multimap<size_t, void*>      free_;
	std::unordered_map<void*, size_t> used_;
	size_t                            cached_ = 0;
};

Pool& device_pool()
{
	static Pool p(false, size_t(48) << 30);
	return p;
}

Pool& host_pool()
{
	static Pool p(true, size_t(8) << 30);
	return p;
}

} // namespace

void* pool_device_alloc(size_t bytes) { return device_pool().alloc(bytes); }
void  pool_device_free(void* p) { device_pool().release(p); }
bool  pool_device_owns(void* p) { return device_pool().owns(p); }
void* pool_host_alloc(size_t bytes) { return host_pool().alloc(bytes); }
void  pool_host_free(void* p) { host_pool().release(p); }

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
