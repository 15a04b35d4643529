// Process-wide caches of device and pinned host allocations.
//
// A DMRG sweep creates one plan per step with buffers of nearly the same sizes, so freed
// buffers are kept and handed out again (smallest cached buffer that is large enough and at
// most twice the request) instead of paying hipMalloc / hipHostMalloc and their frees every
// step. Buffers are returned only after the owner has synchronized its stream.
#pragma once

#include <hip/hip_runtime.h>

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

// Runs fn(i) for i in [0, n) on up to `threads` host threads (items handed out one at a time).
void parallel_for(size_t n, const std::function<void(size_t)>& fn, unsigned threads = 0);

} // namespace dmk
