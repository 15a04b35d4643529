// Vector kernels of the Davidson eigensolver and its preconditioners: diag(H) from the plan's
// term diagonals, block dot products against the basis, Ritz vector and residual, the diagonal
// and the MLP corrections, and the harvest of (features, target) training pairs.
//
// Reductions are deterministic: fixed grids, fixed grid-stride slices, partials folded in index
// order by one workgroup per result.
#include "dmk_internal.h"
#include "mlp.h"

#include <hip/hip_runtime.h>

namespace dmk {

namespace {
constexpr int kDT     = 256;
constexpr int kDBlk   = 512; // partial sums per reduction
unsigned      grid_n(int64_t n)
{
	int64_t b = (n + kDT - 1) / kDT;
	if (b > 4096)
		b = 4096;
	return unsigned(b < 1 ? 1 : b);
}

__device__ inline double dv_block_sum(double v)
{
	__shared__ double part[kDT / 64];
#pragma unroll
	for (int off = 32; off > 0; off >>= 1)
		v += __shfl_down(v, off, 64);
	if ((threadIdx.x & 63) == 0)
		part[threadIdx.x >> 6] = v;
	__syncthreads();
	double s = 0;
	if (threadIdx.x == 0)
		for (int w = 0; w < kDT / 64; ++w)
			s += part[w];
	__syncthreads();
	return s;
}
} // namespace

int davidson_reduction_blocks() { return kDBlk; }

// diag(H)[e]: element e of patch I is (r, l) = ((e - poff[I]) % R_I, (e - poff[I]) / R_I) and
// gets sum over the patch's (I, I, k) terms of diagA_k[l] * diagB_k[r].
__global__ __launch_bounds__(kDT) void plan_diag_kernel(const int64_t* __restrict__ poff,
                                                        const int64_t* __restrict__ toff,
                                                        const int32_t* __restrict__ pL,
                                                        const int32_t* __restrict__ pR,
                                                        const int32_t* __restrict__ tbegin,
                                                        const double* __restrict__ vals, int np,
                                                        double* __restrict__ out, int64_t n)
{
	for (int64_t e = int64_t(blockIdx.x) * kDT + threadIdx.x; e < n; e += int64_t(gridDim.x) * kDT) {
		int lo = 0, hi = np; // poff[lo] <= e < poff[hi]
		while (hi - lo > 1) {
			const int mid = (lo + hi) >> 1;
			if (poff[mid] <= e)
				lo = mid;
			else
				hi = mid;
		}
		const int     I  = lo;
		const int64_t q  = e - poff[I];
		const int     R  = pR[I];
		const int     L  = pL[I];
		const int     r  = int(q % R);
		const int     l  = int(q / R);
		double        s  = 0;
		for (int t = tbegin[I]; t < tbegin[I + 1]; ++t)
			s += vals[toff[t] + l] * vals[toff[t] + L + r];
		out[e] = s;
	}
}

void launch_plan_diagonal(const void* tables, int np, int64_t nterms, int64_t nvals,
                          const int64_t* /*unused*/, double* out, int64_t n, hipStream_t s)
{
	const char*    p      = static_cast<const char*>(tables);
	const int64_t* poff   = reinterpret_cast<const int64_t*>(p);
	p += size_t(np + 1) * 8;
	const int64_t* toff   = reinterpret_cast<const int64_t*>(p);
	p += size_t(nterms) * 8;
	const int32_t* pL     = reinterpret_cast<const int32_t*>(p);
	p += size_t(np) * 4;
	const int32_t* pR     = reinterpret_cast<const int32_t*>(p);
	p += size_t(np) * 4;
	const int32_t* tbegin = reinterpret_cast<const int32_t*>(p);
	p += size_t(np + 1) * 4;
	const size_t used = size_t(p - static_cast<const char*>(tables));
	if (used % 8)
		p += 8 - used % 8;
	const double* vals = reinterpret_cast<const double*>(p);
	(void)nvals;
	hipLaunchKernelGGL(plan_diag_kernel, dim3(grid_n(n)), dim3(kDT), 0, s, poff, toff, pL, pR,
	                   tbegin, vals, np, out, n);
	DMK_HIP_CHECK(hipGetLastError());
}

// partial[j * kDBlk + b] = block b's share of V_j . t for j < k (grid (kDBlk, k)).
__global__ __launch_bounds__(kDT) void multidot_kernel(const double* __restrict__ V, int64_t ldv,
                                                       const double* __restrict__ t, int64_t n,
                                                       double* __restrict__ partial)
{
	const double* v = V + ldv * int64_t(blockIdx.y);
	double        s = 0;
	// Unrolled so that four pairs of loads are in flight at once (the sum order is unchanged).
#pragma unroll 4
	for (int64_t i = int64_t(blockIdx.x) * kDT + threadIdx.x; i < n; i += int64_t(kDBlk) * kDT)
		s += v[i] * t[i];
	s = dv_block_sum(s);
	if (threadIdx.x == 0)
		partial[int64_t(blockIdx.y) * kDBlk + blockIdx.x] = s;
}

// out[j] = sum of partials of j (one workgroup per j); mode 1 also writes sqrt to out[j].
__global__ __launch_bounds__(kDT) void dv_reduce_kernel(const double* __restrict__ partial,
                                                        double* __restrict__ out, int mode)
{
	const double* p = partial + int64_t(blockIdx.x) * kDBlk;
	double        s = 0;
	for (int i = threadIdx.x; i < kDBlk; i += kDT)
		s += p[i];
	s = dv_block_sum(s);
	if (threadIdx.x == 0)
		out[blockIdx.x] = mode ? sqrt(s) : s;
}

// t -= sum_j c[j] V_j.
__global__ __launch_bounds__(kDT) void sub_combine_kernel(double* __restrict__ t,
                                                          const double* __restrict__ V, int64_t ldv,
                                                          const double* __restrict__ c, int k,
                                                          int64_t n)
{
	for (int64_t i = int64_t(blockIdx.x) * kDT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kDT) {
		double s = t[i];
		// Unrolled so that eight basis loads are in flight at once (the sum order is unchanged).
#pragma unroll 8
		for (int j = 0; j < k; ++j)
			s -= c[j] * V[int64_t(j) * ldv + i];
		t[i] = s;
	}
}

// u = V s, hu = W s, r = hu - theta u; partial sums of |r|^2 (slot 0) and |u|^2 (slot 1).
__global__ __launch_bounds__(kDT) void ritz_kernel(const double* __restrict__ V,
                                                   const double* __restrict__ W, int64_t ld,
                                                   const double* __restrict__ sv, int k,
                                                   double theta, int64_t n, double* __restrict__ u,
                                                   double* __restrict__ hu, double* __restrict__ r,
                                                   double* __restrict__ partial)
{
	double rr = 0, uu = 0;
	for (int64_t i = int64_t(blockIdx.x) * kDT + threadIdx.x; i < n; i += int64_t(kDBlk) * kDT) {
		double a = 0, b = 0;
		// Unrolled so that eight pairs of basis loads are in flight at once (the sums' order is
		// unchanged).
#pragma unroll 8
		for (int j = 0; j < k; ++j) {
			a += sv[j] * V[int64_t(j) * ld + i];
			b += sv[j] * W[int64_t(j) * ld + i];
		}
		const double ri = b - theta * a;
		u[i]            = a;
		hu[i]           = b;
		r[i]            = ri;
		rr += ri * ri;
		uu += a * a;
	}
	rr = dv_block_sum(rr);
	uu = dv_block_sum(uu);
	if (threadIdx.x == 0) {
		partial[blockIdx.x]         = rr;
		partial[kDBlk + blockIdx.x] = uu;
	}
}

// t = r / (theta - D) with the denominator kept at least floor in magnitude (Davidson's
// diagonal correction).
__global__ __launch_bounds__(kDT) void precond_diag_kernel(const double* __restrict__ r,
                                                           const double* __restrict__ D,
                                                           double theta, double floor_,
                                                           int64_t n, double* __restrict__ t)
{
	for (int64_t i = int64_t(blockIdx.x) * kDT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kDT) {
		double d = theta - D[i];
		if (fabs(d) < floor_)
			d = (d < 0) ? -floor_ : floor_;
		t[i] = r[i] / d;
	}
}

// Olsen's form of the diagonal correction, t = M^-1 (eps u - r) with M = diag(H) - theta and
// eps = (u . M^-1 r) / (u . M^-1 u), which makes t orthogonal to u (Davidson's plain
// M^-1 r can stall when M is close to a multiple of the identity). Pass 1: partial sums of
// u . M^-1 r (slot 0) and u . M^-1 u (slot 1); |M| is kept at least floor.
__global__ __launch_bounds__(kDT) void olsen_dots_kernel(const double* __restrict__ r,
                                                         const double* __restrict__ u,
                                                         const double* __restrict__ D,
                                                         double theta, double floor_, int64_t n,
                                                         double* __restrict__ partial)
{
	double a = 0, b = 0;
	for (int64_t i = int64_t(blockIdx.x) * kDT + threadIdx.x; i < n; i += int64_t(kDBlk) * kDT) {
		double d = D[i] - theta;
		if (fabs(d) < floor_)
			d = (d < 0) ? -floor_ : floor_;
		a += u[i] * r[i] / d;
		b += u[i] * u[i] / d;
	}
	a = dv_block_sum(a);
	b = dv_block_sum(b);
	if (threadIdx.x == 0) {
		partial[blockIdx.x]         = a;
		partial[kDBlk + blockIdx.x] = b;
	}
}

// Pass 2: t = (eps u - r) / M with eps = sums[0] / sums[1] (0 when the denominator is 0).
__global__ __launch_bounds__(kDT) void olsen_apply_kernel(const double* __restrict__ r,
                                                          const double* __restrict__ u,
                                                          const double* __restrict__ D,
                                                          double theta, double floor_,
                                                          const double* __restrict__ sums,
                                                          int64_t n, double* __restrict__ t)
{
	const double eps = (sums[1] != 0.0) ? sums[0] / sums[1] : 0.0;
	for (int64_t i = int64_t(blockIdx.x) * kDT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kDT) {
		double d = D[i] - theta;
		if (fabs(d) < floor_)
			d = (d < 0) ? -floor_ : floor_;
		t[i] = (eps * u[i] - r[i]) / d;
	}
}

void launch_precond_olsen(const double* r, const double* u, const double* D, double theta,
                          double floor_, int64_t n, const double* sums, double* t, hipStream_t s)
{
	hipLaunchKernelGGL(olsen_apply_kernel, dim3(grid_n(n)), dim3(kDT), 0, s, r, u, D, theta, floor_,
	                   sums, n, t);
	DMK_HIP_CHECK(hipGetLastError());
}

namespace {
// Olsen's correction of element i, t_ols = (eps u - r) / M with M = D - theta kept at least
// floor in magnitude.
__device__ inline double olsen_elem(double Di, double ri, double ui, double theta, double floor_,
                                    double eps)
{
	double d = Di - theta;
	if (fabs(d) < floor_)
		d = (d < 0) ? -floor_ : floor_;
	return (eps * ui - ri) / d;
}

// The four features of element i: (D_i - theta) / sD, r_i / rms_r, u_i / rms_u and Olsen's
// correction in units of rms_r / sD, clipped to [-10, 10].
__device__ inline void features(double Di, double ri, double ui, double theta, double sD,
                                double rms_r, double rms_u, double tols_unit, float f[kMlpIn])
{
	f[0] = float((Di - theta) / sD);
	f[1] = float(ri / rms_r);
	f[2] = float(ui / rms_u);
	f[3] = float(fmin(fmax(tols_unit, -10.0), 10.0));
}

__device__ inline float mlp_eval(const float* __restrict__ P, const float x[kMlpIn])
{
	float h1[kMlpHidden];
#pragma unroll
	for (int o = 0; o < kMlpHidden; ++o) {
		float a = P[kMlpOffB1 + o];
#pragma unroll
		for (int i = 0; i < kMlpIn; ++i)
			a += P[kMlpOffW1 + o * kMlpIn + i] * x[i];
		h1[o] = tanhf(a);
	}
	float out = P[kMlpOffB3];
	for (int o = 0; o < kMlpHidden; ++o) {
		float a = P[kMlpOffB2 + o];
#pragma unroll 16
		for (int i = 0; i < kMlpHidden; ++i)
			a += P[kMlpOffW2 + o * kMlpHidden + i] * h1[i];
		out += P[kMlpOffW3 + o] * tanhf(a);
	}
	return out;
}
} // namespace

// t_i = t_ols,i + MLP(features_i) * rms_r / sD: the network learns a correction on top of
// Olsen's diagonal one, in units of rms_r / sD (its output layer starts at zero, so an
// untrained network is exactly the diagonal method). eps = sums[0] / sums[1] from the Olsen
// dot products of this iterate.
__global__ __launch_bounds__(kDT) void precond_mlp_kernel(const float* __restrict__ params,
                                                          const double* __restrict__ r,
                                                          const double* __restrict__ u,
                                                          const double* __restrict__ D,
                                                          double theta, double floor_,
                                                          const double* __restrict__ sums,
                                                          double sD, double rms_r, double rms_u,
                                                          int64_t n, double* __restrict__ t)
{
	__shared__ float P[kMlpParams];
	for (int i = threadIdx.x; i < kMlpParams; i += kDT)
		P[i] = params[i];
	__syncthreads();
	const double eps  = (sums[1] != 0.0) ? sums[0] / sums[1] : 0.0;
	const double unit = rms_r / sD;
	for (int64_t i = int64_t(blockIdx.x) * kDT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kDT) {
		const double tols = olsen_elem(D[i], r[i], u[i], theta, floor_, eps);
		float        f[kMlpIn];
		features(D[i], r[i], u[i], theta, sD, rms_r, rms_u, tols / unit, f);
		t[i] = tols + double(mlp_eval(P, f)) * unit;
	}
}

// Features of `count` sampled elements i_q = (phase + q * stride) mod n of one iterate
// (x_out row major, kMlpIn per sample) and Olsen's correction there in units (tols_out).
__global__ __launch_bounds__(kDT) void harvest_features_kernel(
    const double* __restrict__ r, const double* __restrict__ u, const double* __restrict__ D,
    double theta, double floor_, const double* __restrict__ sums, double sD, double rms_r,
    double rms_u, int64_t n, int64_t phase, int64_t stride, int64_t count,
    float* __restrict__ x_out, float* __restrict__ tols_out)
{
	const double eps  = (sums[1] != 0.0) ? sums[0] / sums[1] : 0.0;
	const double unit = rms_r / sD;
	for (int64_t q = int64_t(blockIdx.x) * kDT + threadIdx.x; q < count; q += int64_t(gridDim.x) * kDT) {
		const int64_t i    = (phase + q * stride) % n;
		const double  tols = olsen_elem(D[i], r[i], u[i], theta, floor_, eps) / unit;
		float         f[kMlpIn];
		features(D[i], r[i], u[i], theta, sD, rms_r, rms_u, tols, f);
#pragma unroll
		for (int c = 0; c < kMlpIn; ++c)
			x_out[q * kMlpIn + c] = f[c];
		tols_out[q] = float(tols);
	}
}

// Targets of the sampled elements once the eigenvector psi is known: the exact correction
// t* = psi / (psi . u) - u of the iterate u minus Olsen's correction, in units of rms_r / sD,
// clipped to [-20, 20].
__global__ __launch_bounds__(kDT) void harvest_targets_kernel(
    const double* __restrict__ psi, const double* __restrict__ u, const double* __restrict__ pu,
    const float* __restrict__ tols, double unit_inv, int64_t n, int64_t phase, int64_t stride,
    int64_t count, float* __restrict__ t_out)
{
	const double inv = 1.0 / pu[0];
	for (int64_t q = int64_t(blockIdx.x) * kDT + threadIdx.x; q < count; q += int64_t(gridDim.x) * kDT) {
		const int64_t i = (phase + q * stride) % n;
		double        y = (psi[i] * inv - u[i]) * unit_inv - double(tols[q]);
		y               = fmin(fmax(y, -20.0), 20.0);
		t_out[q]        = float(y);
	}
}

__global__ __launch_bounds__(kDT) void fill_kernel(double* __restrict__ x, double v, int64_t n)
{
	for (int64_t i = int64_t(blockIdx.x) * kDT + threadIdx.x; i < n; i += int64_t(gridDim.x) * kDT)
		x[i] = v;
}

void launch_fill(double* x, double v, int64_t n, hipStream_t s)
{
	hipLaunchKernelGGL(fill_kernel, dim3(grid_n(n)), dim3(kDT), 0, s, x, v, n);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_multidot(const double* V, int64_t ldv, const double* t, int64_t n, int k,
                     double* partial, double* out, hipStream_t s)
{
	if (k <= 0)
		return;
	hipLaunchKernelGGL(multidot_kernel, dim3(kDBlk, unsigned(k)), dim3(kDT), 0, s, V, ldv, t, n,
	                   partial);
	hipLaunchKernelGGL(dv_reduce_kernel, dim3(unsigned(k)), dim3(kDT), 0, s, partial, out, 0);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_sub_combine(double* t, const double* V, int64_t ldv, const double* c, int k, int64_t n,
                        hipStream_t s)
{
	if (k <= 0)
		return;
	hipLaunchKernelGGL(sub_combine_kernel, dim3(grid_n(n)), dim3(kDT), 0, s, t, V, ldv, c, k, n);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_ritz(const double* V, const double* W, int64_t ld, const double* sv, int k,
                 double theta, int64_t n, double* u, double* hu, double* r, double* partial,
                 double* out2, hipStream_t s)
{
	hipLaunchKernelGGL(ritz_kernel, dim3(kDBlk), dim3(kDT), 0, s, V, W, ld, sv, k, theta, n, u, hu,
	                   r, partial);
	hipLaunchKernelGGL(dv_reduce_kernel, dim3(2), dim3(kDT), 0, s, partial, out2, 1);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_precond_diag(const double* r, const double* D, double theta, double floor_, int64_t n,
                         double* t, hipStream_t s)
{
	hipLaunchKernelGGL(precond_diag_kernel, dim3(grid_n(n)), dim3(kDT), 0, s, r, D, theta, floor_,
	                   n, t);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_precond_mlp(const float* params, const double* r, const double* u, const double* D,
                        double theta, double floor_, const double* sums, double sD, double rms_r,
                        double rms_u, int64_t n, double* t, hipStream_t s)
{
	unsigned g = grid_n(n);
	if (g > 2048)
		g = 2048;
	hipLaunchKernelGGL(precond_mlp_kernel, dim3(g), dim3(kDT), 0, s, params, r, u, D, theta, floor_,
	                   sums, sD, rms_r, rms_u, n, t);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_olsen_sums(const double* r, const double* u, const double* D, double theta,
                       double floor_, int64_t n, double* partial, double* sums, hipStream_t s)
{
	hipLaunchKernelGGL(olsen_dots_kernel, dim3(kDBlk), dim3(kDT), 0, s, r, u, D, theta, floor_, n,
	                   partial);
	hipLaunchKernelGGL(dv_reduce_kernel, dim3(2), dim3(kDT), 0, s, partial, sums, 0);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_harvest_features(const double* r, const double* u, const double* D, double theta,
                             double floor_, const double* sums, double sD, double rms_r,
                             double rms_u, int64_t n, int64_t phase, int64_t stride, int64_t count,
                             float* x_out, float* tols_out, hipStream_t s)
{
	if (count <= 0)
		return;
	hipLaunchKernelGGL(harvest_features_kernel, dim3(grid_n(count)), dim3(kDT), 0, s, r, u, D,
	                   theta, floor_, sums, sD, rms_r, rms_u, n, phase, stride, count, x_out,
	                   tols_out);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_harvest_targets(const double* psi, const double* u, const double* pu, const float* tols,
                            double unit_inv, int64_t n, int64_t phase, int64_t stride,
                            int64_t count, float* t_out, hipStream_t s)
{
	if (count <= 0)
		return;
	hipLaunchKernelGGL(harvest_targets_kernel, dim3(grid_n(count)), dim3(kDT), 0, s, psi, u, pu,
	                   tols, unit_inv, n, phase, stride, count, t_out);
	DMK_HIP_CHECK(hipGetLastError());
}

} // namespace dmk
