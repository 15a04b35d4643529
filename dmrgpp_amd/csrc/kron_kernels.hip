// The split-k reduction and the FP64 -> FP32 cast of the matvec (panel_kernels.hip holds the
// matvec's GEMM kernels, which also serve the variable-size batched GEMM).
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <type_traits>

namespace dmk {

// Sums split-k partials in a fixed order (s = 0, 1, ...) into the output, accumulating in TO.
// Each thread owns kReduceLanes elements of the chunk (stride 256) and keeps their loads of one
// piece in flight together.
constexpr int kReduceLanes = 8;
static_assert(kReduceChunk == 256 * kReduceLanes, "reduce chunk / block mismatch");

template <typename T, typename TO>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const Reduce* __restrict__ red,
                                                            const ReduceChunk* __restrict__ chunks,
                                                            const T* __restrict__ ws,
                                                            int64_t wsstride, TO* __restrict__ out,
                                                            int64_t ystride, int beta)
{
	const ReduceChunk ch = chunks[blockIdx.x];
	const Reduce      rd = red[ch.r];
	ws += int64_t(blockIdx.y) * wsstride + rd.wsoff + ch.start;
	out += int64_t(blockIdx.y) * ystride + rd.yoff + ch.start;
	TO v[kReduceLanes];
#pragma unroll
	for (int j = 0; j < kReduceLanes; ++j)
		v[j] = 0;
	for (int s = 0; s < rd.nsplit; ++s) {
		const T* p = ws + int64_t(s) * rd.count;
#pragma unroll
		for (int j = 0; j < kReduceLanes; ++j) {
			const int e = int(threadIdx.x) + 256 * j;
			if (e < ch.len)
				v[j] += TO(p[e]);
		}
	}
#pragma unroll
	for (int j = 0; j < kReduceLanes; ++j) {
		const int e = int(threadIdx.x) + 256 * j;
		if (e < ch.len)
			out[e] = beta ? v[j] + out[e] : v[j];
	}
}

__global__ void cast_f64_to_f32_kernel(const double* __restrict__ x, float* __restrict__ y, int64_t n)
{
	int64_t i      = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (; i < n; i += stride)
		y[i] = float(x[i]);
}

// Host-side launchers ------------------------------------------------------------------------

template <typename T, typename TO>
void launch_splitk_reduce(const Reduce* red, const ReduceChunk* chunks, int nchunks, const T* ws,
                          int64_t wsstride, TO* out, int64_t ystride, int nvec, int beta,
                          hipStream_t stream)
{
	if (nchunks <= 0 || nvec <= 0)
		return;
	const dim3 grid { unsigned(nchunks), unsigned(nvec), 1u };
	hipLaunchKernelGGL((splitk_reduce_kernel<T, TO>), grid, dim3(256), 0, stream, red, chunks, ws,
	                   wsstride, out, ystride, beta);
}

template void launch_splitk_reduce<double, double>(const Reduce*, const ReduceChunk*, int,
                                                   const double*, int64_t, double*, int64_t, int,
                                                   int, hipStream_t);
template void launch_splitk_reduce<float, double>(const Reduce*, const ReduceChunk*, int,
                                                  const float*, int64_t, double*, int64_t, int,
                                                  int, hipStream_t);

void launch_cast_f64_to_f32(const double* x, float* y, int64_t n, hipStream_t stream)
{
	if (n <= 0)
		return;
	int64_t blocks = (n + 255) / 256;
	if (blocks > 4096)
		blocks = 4096;
	hipLaunchKernelGGL(cast_f64_to_f32_kernel, dim3(unsigned(blocks)), dim3(256), 0, stream, x, y, n);
}

} // namespace dmk
