// Vector kernels of the batched Krylov propagator: several target vectors advance through their
// own Lanczos recurrences in lockstep, so each kernel works on all targets at once
// (blockIdx.y = target). A target vector is complex, stored split: the real parts of its n
// elements, then the imaginary parts (the real embedding, length 2n).
//
// Reductions are deterministic per target: a fixed grid of kKryBlocks workgroups per target,
// each folding a fixed grid-stride slice, then one workgroup per target folding the partials in
// index order. A target's results do not depend on which other targets share the launch.
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

namespace dmk {

constexpr int kKryBlocks  = 256;
constexpr int kKryThreads = 256;
constexpr int kKryMaxTimes = 8;

int krylov_reduction_blocks() { return kKryBlocks; }
int krylov_max_times() { return kKryMaxTimes; }

namespace {

__device__ inline double kry_wave_sum(double v)
{
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_down(v, off, 64);
	return v;
}

// Sum over the workgroup (kKryThreads threads) of v; the result is valid in thread 0.
__device__ inline double kry_block_sum(double v)
{
	__shared__ double part[kKryThreads / 64];
	v = kry_wave_sum(v);
	if ((threadIdx.x & 63) == 0)
		part[threadIdx.x >> 6] = v;
	__syncthreads();
	double s = 0;
	if (threadIdx.x == 0)
		for (int w = 0; w < kKryThreads / 64; ++w)
			s += part[w];
	__syncthreads();
	return s;
}

} // namespace

// Split copy of interleaved complex inputs: out_t[i] = Re z_t[i], out_t[n + i] = Im z_t[i].
__global__ __launch_bounds__(kKryThreads) void kry_deinterleave_kernel(
    const double* __restrict__ z, int64_t ldz, double* __restrict__ out, int64_t ldout, int64_t n)
{
	const double* zt = z + 2 * ldz * int64_t(blockIdx.y);
	double*       ot = out + ldout * int64_t(blockIdx.y);
	int64_t       st = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += st) {
		const double2 v = reinterpret_cast<const double2*>(zt)[i];
		ot[i]           = v.x;
		ot[n + i]       = v.y;
	}
}

// Interleaved copy of split vectors (inverse of kry_deinterleave_kernel).
__global__ __launch_bounds__(kKryThreads) void kry_interleave_kernel(
    const double* __restrict__ in, int64_t ldin, double* __restrict__ z, int64_t ldz, int64_t n)
{
	const double* it = in + ldin * int64_t(blockIdx.y);
	double*       zt = z + 2 * ldz * int64_t(blockIdx.y);
	int64_t       st = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += st)
		reinterpret_cast<double2*>(zt)[i] = make_double2(it[i], it[n + i]);
}

// partial[t * kKryBlocks + b] = block b's share of x_t . y_t over len elements.
__global__ __launch_bounds__(kKryThreads) void kry_dot_kernel(const double* __restrict__ x,
                                                              const double* __restrict__ y,
                                                              int64_t stride, int64_t len,
                                                              double* __restrict__ partial)
{
	const double* xt = x + stride * int64_t(blockIdx.y);
	const double* yt = y + stride * int64_t(blockIdx.y);
	double        s  = 0;
	int64_t       st = int64_t(kKryBlocks) * kKryThreads;
	for (int64_t i = int64_t(blockIdx.x) * kKryThreads + threadIdx.x; i < len; i += st)
		s += xt[i] * yt[i];
	s = kry_block_sum(s);
	if (threadIdx.x == 0)
		partial[int64_t(blockIdx.y) * kKryBlocks + blockIdx.x] = s;
}

// out[t * ostride] = sum of target t's partials; with sqrt_inv set, out holds sqrt of the sum and
// out[t * ostride + 1] its inverse (0 for a zero norm).
__global__ __launch_bounds__(kKryThreads) void kry_reduce_kernel(const double* __restrict__ partial,
                                                                 double* __restrict__ out,
                                                                 int ostride, int sqrt_inv)
{
	const double* p = partial + int64_t(blockIdx.x) * kKryBlocks;
	double        s = 0;
	for (int i = threadIdx.x; i < kKryBlocks; i += kKryThreads)
		s += p[i];
	s = kry_block_sum(s);
	if (threadIdx.x == 0) {
		double* o = out + int64_t(blockIdx.x) * ostride;
		if (sqrt_inv) {
			const double r = sqrt(s);
			o[0]           = r;
			o[1]           = (r > 0) ? 1.0 / r : 0.0;
		} else {
			o[0] = s;
		}
	}
}

// w_t -= alpha_t v1_t + beta_t v0_t with alpha_t = A[t * cap + j], beta_t = B[t * cap + j], and
// the partial sums of |w_t|^2 (LanczosVectors::oneStepDecomposition, LanczosVectors.h:212-228).
__global__ __launch_bounds__(kKryThreads) void kry_update_kernel(
    double* __restrict__ w, const double* __restrict__ v1, const double* __restrict__ v0,
    const double* __restrict__ A, const double* __restrict__ B, int j, int cap, int64_t stride,
    int64_t len, double* __restrict__ partial)
{
	const int64_t t  = blockIdx.y;
	const double  a  = A[t * cap + j];
	const double  b  = B[t * cap + j];
	double*       wt = w + stride * t;
	const double* v1t = v1 + stride * t;
	const double* v0t = v0 + stride * t;
	double        s  = 0;
	int64_t       st = int64_t(kKryBlocks) * kKryThreads;
	for (int64_t i = int64_t(blockIdx.x) * kKryThreads + threadIdx.x; i < len; i += st) {
		const double v = wt[i] - a * v1t[i] - b * v0t[i];
		wt[i]          = v;
		s += v * v;
	}
	s = kry_block_sum(s);
	if (threadIdx.x == 0)
		partial[t * kKryBlocks + blockIdx.x] = s;
}

// beta_{j+1} of target t from its norm partials. A target whose new norm is negligible against
// its running estimate of |H| (max over steps of |alpha| + beta_j + beta_{j+1}) has reached an
// invariant subspace: its beta is stored as 0 and the inverse as 0, so its later vectors are
// zero and its tridiagonal matrix decouples there.
__global__ __launch_bounds__(kKryThreads) void kry_beta_kernel(const double* __restrict__ partial,
                                                               const double* __restrict__ A,
                                                               double* __restrict__ B,
                                                               double* __restrict__ HN,
                                                               double* __restrict__ INV, int j,
                                                               int cap)
{
	const int64_t t = blockIdx.x;
	const double* p = partial + t * kKryBlocks;
	double        s = 0;
	for (int i = threadIdx.x; i < kKryBlocks; i += kKryThreads)
		s += p[i];
	s = kry_block_sum(s);
	if (threadIdx.x == 0) {
		const double beta = sqrt(s);
		const double hn   = fmax(HN[t], fabs(A[t * cap + j]) + B[t * cap + j] + beta);
		HN[t]             = hn;
		const bool   live = beta > 1e-13 * hn;
		if (j + 1 < cap)
			B[t * cap + j + 1] = live ? beta : 0.0;
		INV[t] = live ? 1.0 / beta : 0.0;
	}
}

// x_t *= s[t * sstride].
__global__ __launch_bounds__(kKryThreads) void kry_scale_kernel(double* __restrict__ x,
                                                                const double* __restrict__ s,
                                                                int sstride, int64_t stride,
                                                                int64_t len)
{
	const double f  = s[int64_t(blockIdx.y) * sstride];
	double*      xt = x + stride * int64_t(blockIdx.y);
	int64_t      st = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < len; i += st)
		xt[i] *= f;
}

// Complex combinations of the Krylov vectors, all times of a target in one pass over its basis:
//   out_{t,s} = sum_j c_{t,s,j} V_{j,t}
// with complex c and split complex V (multiplying by i maps (re, im) to (-im, re)). Written
// interleaved: out + 2 * ldout * (t * ntimes + s). coef holds (re, im) pairs at
// coef[2 * ((t * ntimes + s) * m + j)]; vptr[j] is the block of step j (targets at vstride).
__global__ __launch_bounds__(kKryThreads) void kry_combine_kernel(
    double* __restrict__ out, int64_t ldout, const double* const* __restrict__ vptr,
    int64_t vstride, const double* __restrict__ coef, int m, int ntimes, int64_t n)
{
	const int64_t t  = blockIdx.y;
	const double* ct = coef + 2 * t * int64_t(ntimes) * m;
	int64_t       st = int64_t(gridDim.x) * blockDim.x;
	for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += st) {
		double re[kKryMaxTimes], im[kKryMaxTimes];
#pragma unroll
		for (int s = 0; s < kKryMaxTimes; ++s)
			re[s] = im[s] = 0.0;
		for (int j = 0; j < m; ++j) {
			const double* v  = vptr[j] + vstride * t;
			const double  vr = v[i];
			const double  vi = v[n + i];
#pragma unroll
			for (int s = 0; s < kKryMaxTimes; ++s) {
				if (s < ntimes) {
					const double cr = ct[2 * (int64_t(s) * m + j)];
					const double ci = ct[2 * (int64_t(s) * m + j) + 1];
					re[s] += cr * vr - ci * vi;
					im[s] += cr * vi + ci * vr;
				}
			}
		}
		for (int s = 0; s < ntimes; ++s) {
			double* o = out + 2 * ldout * (t * ntimes + s);
			reinterpret_cast<double2*>(o)[i] = make_double2(re[s], im[s]);
		}
	}
}

namespace {
unsigned grid_for(int64_t n)
{
	int64_t b = (n + kKryThreads - 1) / kKryThreads;
	if (b > 4096)
		b = 4096;
	return unsigned(b < 1 ? 1 : b);
}
} // namespace

void launch_kry_deinterleave(const double* z, int64_t ldz, double* out, int64_t ldout, int64_t n,
                             int nvec, hipStream_t s)
{
	dim3 grid{ grid_for(n), unsigned(nvec) };
	kry_deinterleave_kernel<<<grid, kKryThreads, 0, s>>>(z, ldz, out, ldout, n);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_kry_interleave(const double* in, int64_t ldin, double* z, int64_t ldz, int64_t n,
                           int nvec, hipStream_t s)
{
	dim3 grid{ grid_for(n), unsigned(nvec) };
	kry_interleave_kernel<<<grid, kKryThreads, 0, s>>>(in, ldin, z, ldz, n);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_kry_dot(const double* x, const double* y, int64_t stride, int64_t len, int nvec,
                    double* partial, double* out, int ostride, int sqrt_inv, hipStream_t s)
{
	dim3 grid{ unsigned(kKryBlocks), unsigned(nvec) };
	kry_dot_kernel<<<grid, kKryThreads, 0, s>>>(x, y, stride, len, partial);
	DMK_HIP_CHECK(hipGetLastError());
	kry_reduce_kernel<<<unsigned(nvec), kKryThreads, 0, s>>>(partial, out, ostride, sqrt_inv);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_kry_update_beta(double* w, const double* v1, const double* v0, const double* A,
                            double* B, double* HN, double* INV, int j, int cap, int64_t stride,
                            int64_t len, int nvec, double* partial, hipStream_t s)
{
	dim3 grid{ unsigned(kKryBlocks), unsigned(nvec) };
	kry_update_kernel<<<grid, kKryThreads, 0, s>>>(w, v1, v0, A, B, j, cap, stride, len, partial);
	DMK_HIP_CHECK(hipGetLastError());
	kry_beta_kernel<<<unsigned(nvec), kKryThreads, 0, s>>>(partial, A, B, HN, INV, j, cap);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_kry_scale(double* x, const double* sc, int sstride, int64_t stride, int64_t len,
                      int nvec, hipStream_t s)
{
	dim3 grid{ grid_for(len), unsigned(nvec) };
	kry_scale_kernel<<<grid, kKryThreads, 0, s>>>(x, sc, sstride, stride, len);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_kry_combine(double* out, int64_t ldout, const double* const* vptr, int64_t vstride,
                        const double* coef, int m, int ntimes, int64_t n, int nvec, hipStream_t s)
{
	dim3 grid{ grid_for(n), unsigned(nvec) };
	kry_combine_kernel<<<grid, kKryThreads, 0, s>>>(out, ldout, vptr, vstride, coef, m, ntimes,
	                                                n);
	DMK_HIP_CHECK(hipGetLastError());
}

} // namespace dmk
