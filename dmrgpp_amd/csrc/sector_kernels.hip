// Gather / scatter between a sector-ordered vector and the plugin's patch order, on the device
// (InitKronHamiltonian::copyIn, InitKronHamiltonian.h:158-221, and InitKronBase::copyOut,
// InitKronBase.h:315-360, which do it on the host around every matvec).
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

namespace dmk {

namespace {
unsigned grid_of(int64_t n)
{
	int64_t b = (n + 255) / 256;
	if (b > 4096)
		b = 4096;
	return unsigned(b < 1 ? 1 : b);
}
} // namespace

// y[ip] = v[map[ip]].
__global__ __launch_bounds__(256) void sector_gather_kernel(const double* __restrict__ v,
                                                            const int64_t* __restrict__ map,
                                                            int64_t n, double* __restrict__ y)
{
	for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
		y[i] = v[map[i]];
}

// v[map[ip]] += x[ip] (map is a bijection, so no two threads meet).
__global__ __launch_bounds__(256) void sector_scatter_add_kernel(const double* __restrict__ x,
                                                                 const int64_t* __restrict__ map,
                                                                 int64_t n, double* __restrict__ v)
{
	for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
		v[map[i]] += x[i];
}

void launch_sector_gather(const double* v, const int64_t* map, int64_t n, double* y, hipStream_t s)
{
	hipLaunchKernelGGL(sector_gather_kernel, dim3(grid_of(n)), dim3(256), 0, s, v, map, n, y);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_sector_scatter_add(const double* x, const int64_t* map, int64_t n, double* v,
                               hipStream_t s)
{
	hipLaunchKernelGGL(sector_scatter_add_kernel, dim3(grid_of(n)), dim3(256), 0, s, x, map, n, v);
	DMK_HIP_CHECK(hipGetLastError());
}

} // namespace dmk
