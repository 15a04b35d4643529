// The static device scratch of the allocation-free matvec (noalloc_scratch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace dmk {

// Size of each GPU's static scratch array: the largest scratch one allocation-free matvec may
// use (its stage-1 products, split-k partials and input panels together).
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;

// Device address of the current device's copy of the array (looked up with
// hipGetSymbolAddress on the first call on each device, with that device current, then kept).
void* noalloc_scratch_device_ptr();

// Clears `bytes` bytes at p (16-byte aligned and sized) with a fill kernel on stream s (no
// hipMemsetAsync, so a caller may capture the call into a HIP graph).
void launch_fill_zero(void* p, size_t bytes, hipStream_t s);

} // namespace dmk
