// Statically allocated device scratch of the allocation-free matvec entry point
// (dmk_apply_noalloc_d): the stage-1 products, split-k partials and input panels of one
// matvec live in this file-scope array instead of plan-owned pool buffers, so the call makes
// no hipMalloc / hipFree at all (for callers under HIP graph capture or in loops that must not
// allocate).
//
// Every GPU the library runs on holds its own copy of the array (kNoallocScratchBytes) from
// the first use of the library's code on that device until the process ends. Host code never
// uses the array's name as an address: the device address of each GPU's copy comes from
// hipGetSymbolAddress, looked up once per device with that device current, and is what the
// kernels receive.
#include "dmk_internal.h"
#include "noalloc_scratch.h"

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

namespace dmk {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex    g_scratch_mutex;
void*         g_scratch_ptr[kMaxDevices] = {};

// Clears n 16-byte words at p.
__global__ __launch_bounds__(256) void fill_zero_kernel(uint4* __restrict__ p, int64_t n)
{
	const uint4 z = make_uint4(0u, 0u, 0u, 0u);
	for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
		p[i] = z;
}

} // namespace

void* noalloc_scratch_device_ptr()
{
	int dev = 0;
	DMK_HIP_CHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		throw std::runtime_error("dmk_apply_noalloc_d: device ordinal " + std::to_string(dev)
		                         + " out of range");
	std::lock_guard<std::mutex> lock(g_scratch_mutex);
	if (!g_scratch_ptr[dev]) {
		void*            ptr = nullptr;
		const hipError_t e   = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
		if (e != hipSuccess || !ptr)
			throw std::runtime_error(std::string("dmk_apply_noalloc_d: hipGetSymbolAddress of the "
			                                     "static scratch failed: ")
			                         + hipGetErrorString(e));
		g_scratch_ptr[dev] = ptr;
	}
	return g_scratch_ptr[dev];
}

void launch_fill_zero(void* p, size_t bytes, hipStream_t s)
{
	if (bytes == 0)
		return;
	if ((bytes % 16) != 0 || (reinterpret_cast<uintptr_t>(p) % 16) != 0)
		throw std::runtime_error("fill_zero: region must be 16-byte aligned and sized");
	const int64_t words = int64_t(bytes / 16);
	int64_t       b     = (words + 255) / 256;
	if (b > 4096)
		b = 4096;
	hipLaunchKernelGGL(fill_zero_kernel, dim3(unsigned(b < 1 ? 1 : b)), dim3(256), 0, s,
	                   static_cast<uint4*>(p), words);
	DMK_HIP_CHECK(hipGetLastError());
}

} // namespace dmk
