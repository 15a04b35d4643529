// Vector kernels of the device-resident Lanczos recurrence.
//
// Reductions are deterministic: a fixed grid of kRedBlocks workgroups each folds a fixed slice
// pattern (grid-stride) into one partial, wave64 shuffles and an LDS tree reduce inside the
// workgroup, and a single workgroup folds the partials in index order. The same n always gives
// bitwise the same result.
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

namespace dmk {

constexpr int kRedBlocks  = 1024;
constexpr int kRedThreads = 256;

__device__ inline double wave_sum(double v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_down(v, off, 64);
	return v;
}

// Sum over the workgroup of v; the result is valid in thread 0.
__device__ inline double block_sum(double v)
{
	__shared__ double part[kRedThreads / 64];
	v = wave_sum(v);
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	if (lane == 0)
		part[wave] = v;
	__syncthreads();
	double s = 0;
	if (threadIdx.x == 0)
		for (int w = 0; w < kRedThreads / 64; ++w)
			s += part[w];
	__syncthreads();
	return s;
}

__global__ __launch_bounds__(kRedThreads) void dot_partial_kernel(const double* __restrict__ x,
                                                                  const double* __restrict__ y,
                                                                  int64_t n, double* __restrict__ partial)
{
	double  s      = 0;
	int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
		s += x[i] * y[i];
	s = block_sum(s);
	if (threadIdx.x == 0)
		partial[blockIdx.x] = s;
}

// out[0] = sum(partial[0..np)); mode 1 writes sqrt instead; mode 2 writes sqrt to out[0] and its
// inverse (0 when the norm is 0) to out[1].
__global__ __launch_bounds__(kRedThreads) void reduce_final_kernel(const double* __restrict__ partial,
                                                                   int np, double* __restrict__ out,
                                                                   int mode)
{
	double s = 0;
	for (int i = threadIdx.x; i < np; i += blockDim.x)
		s += partial[i];
	s = block_sum(s);
	if (threadIdx.x == 0) {
		if (mode == 0) {
			out[0] = s;
		} else {
			const double r = sqrt(s);
			out[0]         = r;
			if (mode == 2)
				out[1] = (r > 0) ? 1.0 / r : 0.0;
		}
	}
}

// v2 = v2 - alpha * v1 - beta * v0 (alpha, beta read from device memory), and the partial sums
// of |v2|^2 (LanczosVectors::oneStepDecomposition, LanczosVectors.h:212-228).
__global__ __launch_bounds__(kRedThreads) void lanczos_update_kernel(
    double* __restrict__ v2, const double* __restrict__ v1, const double* __restrict__ v0,
    const double* __restrict__ alpha, const double* __restrict__ beta, int64_t n,
    double* __restrict__ partial)
{
	const double a      = *alpha;
	const double b      = *beta;
	double       s      = 0;
	int64_t      stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const double v = v2[i] - a * v1[i] - b * v0[i];
		v2[i]          = v;
		s += v * v;
	}
	s = block_sum(s);
	if (threadIdx.x == 0)
		partial[blockIdx.x] = s;
}

// x *= *scal.
__global__ void scale_kernel(double* __restrict__ x, const double* __restrict__ scal, int64_t n)
{
	const double s      = *scal;
	int64_t      stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
		x[i] *= s;
}

// y = x * (*scal).
__global__ void scale_copy_kernel(double* __restrict__ y, const double* __restrict__ x,
                                  const double* __restrict__ scal, int64_t n)
{
	const double s      = *scal;
	int64_t      stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
		y[i] = x[i] * s;
}

// y -= (*coef) * x.
__global__ void axpy_neg_kernel(double* __restrict__ y, const double* __restrict__ x,
                                const double* __restrict__ coef, int64_t n)
{
	const double c      = *coef;
	int64_t      stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
		y[i] -= c * x[i];
}

// z[i] = sum_j c[j] * V[i + j*ld], j = 0..m-1 (LanczosVectors::excitedVector, LanczosVectors.h:
// 192-201, as one pass over the basis instead of m axpys).
__global__ void combine_kernel(double* __restrict__ z, const double* __restrict__ V, int64_t ld,
                               const double* __restrict__ c, int m, int64_t n)
{
	int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		double s = 0;
		// Unrolled so that eight basis loads are in flight at once (the sum order is unchanged).
#pragma unroll 8
		for (int j = 0; j < m; ++j)
			s += c[j] * V[i + int64_t(j) * ld];
		z[i] = s;
	}
}

// Partial sums of |hx - e x|^2 (the residual norm of an eigenpair estimate).
__global__ __launch_bounds__(kRedThreads) void residual_kernel(const double* __restrict__ hx,
                                                               const double* __restrict__ x,
                                                               double e, int64_t n,
                                                               double* __restrict__ partial)
{
	double  s      = 0;
	int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const double v = hx[i] - e * x[i];
		s += v * v;
	}
	s = block_sum(s);
	if (threadIdx.x == 0)
		partial[blockIdx.x] = s;
}

__global__ void copy_kernel(double* __restrict__ y, const double* __restrict__ x, int64_t n)
{
	int64_t stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
		y[i] = x[i];
}

// Step-indexed kernels of the captured Lanczos step: the current step j is read from device
// memory (jdev), so one captured graph serves every step of its ring phase.

// base[*jdev + offset] = sum(partial[0..np)).
__global__ __launch_bounds__(kRedThreads) void reduce_final_at_kernel(const double* __restrict__ partial,
                                                                      int np, double* __restrict__ base,
                                                                      const int* __restrict__ jdev,
                                                                      int offset)
{
	double s = 0;
	for (int i = threadIdx.x; i < np; i += blockDim.x)
		s += partial[i];
	s = block_sum(s);
	if (threadIdx.x == 0)
		base[*jdev + offset] = s;
}

// v2 = v2 - A[j] v1 - B[j] v0, partial sums of |v2|^2.
__global__ __launch_bounds__(kRedThreads) void lanczos_update_at_kernel(
    double* __restrict__ v2, const double* __restrict__ v1, const double* __restrict__ v0,
    const double* __restrict__ A, const double* __restrict__ B, const int* __restrict__ jdev,
    int64_t n, double* __restrict__ partial)
{
	const int    j      = *jdev;
	const double a      = A[j];
	const double b      = B[j];
	double       s      = 0;
	int64_t      stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const double v = v2[i] - a * v1[i] - b * v0[i];
		v2[i]          = v;
		s += v * v;
	}
	s = block_sum(s);
	if (threadIdx.x == 0)
		partial[blockIdx.x] = s;
}

// v2 *= nrm[1] and basis column j+1 = v2.
__global__ void finish_step_kernel(double* __restrict__ v2, double* __restrict__ basis, int64_t n,
                                   const double* __restrict__ nrm, const int* __restrict__ jdev)
{
	const double  inv    = nrm[1];
	double*       col    = basis + int64_t(*jdev + 1) * n;
	int64_t       stride = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
		const double v = v2[i] * inv;
		v2[i]          = v;
		col[i]         = v;
	}
}

// report = (A[j], |v2|); B[j+1] = |v2| when j+1 < maxn; j += 1.
__global__ void step_scalars_kernel(const double* __restrict__ A, double* __restrict__ B,
                                    const double* __restrict__ nrm, int* __restrict__ jdev,
                                    int maxn, double* __restrict__ report)
{
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		const int j = *jdev;
		report[0]   = A[j];
		report[1]   = nrm[0];
		if (j + 1 < maxn)
			B[j + 1] = nrm[0];
		*jdev = j + 1;
	}
}

// Host launchers ------------------------------------------------------------------------------

namespace {
unsigned grid_for(int64_t n)
{
	int64_t b = (n + 255) / 256;
	if (b > 8192)
		b = 8192;
	if (b < 1)
		b = 1;
	return unsigned(b);
}
} // namespace

int reduction_blocks() { return kRedBlocks; }

void launch_dot(const double* x, const double* y, int64_t n, double* partial, double* out,
                int mode, hipStream_t s)
{
	hipLaunchKernelGGL(dot_partial_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, x, y, n,
	                   partial);
	hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kRedThreads), 0, s, partial, kRedBlocks,
	                   out, mode);
}

void launch_lanczos_update(double* v2, const double* v1, const double* v0, const double* alpha,
                           const double* beta, int64_t n, double* partial, double* norm_out,
                           hipStream_t s)
{
	hipLaunchKernelGGL(lanczos_update_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, v2, v1,
	                   v0, alpha, beta, n, partial);
	hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kRedThreads), 0, s, partial, kRedBlocks,
	                   norm_out, 2);
}

void launch_scale(double* x, const double* scal, int64_t n, hipStream_t s)
{
	hipLaunchKernelGGL(scale_kernel, dim3(grid_for(n)), dim3(256), 0, s, x, scal, n);
}

void launch_scale_copy(double* y, const double* x, const double* scal, int64_t n, hipStream_t s)
{
	hipLaunchKernelGGL(scale_copy_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, x, scal, n);
}

void launch_axpy_neg(double* y, const double* x, const double* coef, int64_t n, hipStream_t s)
{
	hipLaunchKernelGGL(axpy_neg_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, x, coef, n);
}

void launch_combine(double* z, const double* V, int64_t ld, const double* c, int m, int64_t n,
                    hipStream_t s)
{
	hipLaunchKernelGGL(combine_kernel, dim3(grid_for(n)), dim3(256), 0, s, z, V, ld, c, m, n);
}

void launch_residual(const double* hx, const double* x, double e, int64_t n, double* partial,
                     double* out, hipStream_t s)
{
	hipLaunchKernelGGL(residual_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, hx, x, e, n,
	                   partial);
	hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kRedThreads), 0, s, partial, kRedBlocks,
	                   out, 1);
}

void launch_dot_at(const double* x, const double* y, int64_t n, double* partial, double* base,
                   const int* jdev, int offset, hipStream_t s)
{
	hipLaunchKernelGGL(dot_partial_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, x, y, n,
	                   partial);
	hipLaunchKernelGGL(reduce_final_at_kernel, dim3(1), dim3(kRedThreads), 0, s, partial,
	                   kRedBlocks, base, jdev, offset);
}

void launch_lanczos_update_at(double* v2, const double* v1, const double* v0, const double* A,
                              const double* B, const int* jdev, int64_t n, double* partial,
                              double* norm_out, hipStream_t s)
{
	hipLaunchKernelGGL(lanczos_update_at_kernel, dim3(kRedBlocks), dim3(kRedThreads), 0, s, v2, v1,
	                   v0, A, B, jdev, n, partial);
	hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kRedThreads), 0, s, partial, kRedBlocks,
	                   norm_out, 2);
}

void launch_finish_step(double* v2, double* basis, int64_t n, const double* nrm, const int* jdev,
                        hipStream_t s)
{
	hipLaunchKernelGGL(finish_step_kernel, dim3(grid_for(n)), dim3(256), 0, s, v2, basis, n, nrm,
	                   jdev);
}

void launch_step_scalars(const double* A, double* B, const double* nrm, int* jdev, int maxn,
                         double* report, hipStream_t s)
{
	hipLaunchKernelGGL(step_scalars_kernel, dim3(1), dim3(64), 0, s, A, B, nrm, jdev, maxn, report);
}

void launch_copy(double* y, const double* x, int64_t n, hipStream_t s)
{
	hipLaunchKernelGGL(copy_kernel, dim3(grid_for(n)), dim3(256), 0, s, y, x, n);
}

} // namespace dmk
