// Process-wide caches of device and pinned host allocations.
//
// A DMRG sweep creates one plan per step with buffers of nearly the same sizes, so freed
// buffers are kept and handed out again (smallest cached buffer that is large enough and at
// most twice the request) instead of paying hipMalloc / hipHostMalloc and their frees every
// step. Buffers are returned only after the owner has synchronized its stream.
#pragma once

#include <hip/hip_runtime.h>

#include "pool_cache.h"

#include <cstddef>
#include <functional>
#include <thread>
#include <vector>

namespace dmk {

void* pool_device_alloc(size_t bytes);
void  pool_device_free(void* p);
bool  pool_device_owns(void* p);
void* pool_host_alloc(size_t bytes); // pinned
void  pool_host_free(void* p);
PoolCounters pool_device_counters();
PoolCounters pool_host_counters();

// Releases p and clears the owner's pointer, so that a later release of the same member (an
// error path, a destructor) cannot hand a buffer that has gone to another owner back to the
// cache.
template <class T> void pool_device_release(T*& p)
{
	pool_device_free(static_cast<void*>(p));
	p = nullptr;
}
template <class T> void pool_host_release(T*& p)
{
	pool_host_free(static_cast<void*>(p));
	p = nullptr;
}

// Runs fn(i) for i in [0, n) on up to `threads` host threads (items handed out one at a time).
void parallel_for(size_t n, const std::function<void(size_t)>& fn, unsigned threads = 0);

} // namespace dmk
