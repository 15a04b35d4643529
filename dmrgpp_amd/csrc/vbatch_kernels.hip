// Staging and scatter kernels of the variable-size batched GEMM (dmk_gemm_vbatch_{d,z}).
//
// A batch entry C = alpha op(A) op(B) + beta C runs in three launches: the operands are staged
// into one workspace as column-major matrices with zero-padded leading dimensions (transposed,
// conjugated, split into real and imaginary parts as the entry asks), the panel MFMA GEMM
// multiplies the staged operands into a workspace result, and the scatter applies alpha and
// beta into the caller's C. Each
// launch takes a descriptor list; each descriptor is one chunk of one matrix.
#include "dmk_internal.h"
#include "vbatch.h"

#include <hip/hip_runtime.h>

namespace dmk {

// A pointer the caller passed in a descriptor, as global memory: global (not flat) loads, which
// the compiler counts on the vector-memory counter alone and so can keep several in flight.
__device__ inline const __attribute__((address_space(1))) double* global_ptr(const double* p)
{
	return (const __attribute__((address_space(1))) double*)p;
}

__global__ __launch_bounds__(256) void vb_copy_kernel(const VbCopy* __restrict__ jobs,
                                                      double* __restrict__ ws)
{
	const VbCopy jb = jobs[blockIdx.x];
	const bool   tr = (jb.flags & VB_TRANS) != 0;
	const bool   cx = (jb.flags & VB_CPLX) != 0;
	const int    im = (jb.flags & VB_IMAG) ? 1 : 0;
	const double sg = (jb.flags & VB_NEG) ? -1.0 : 1.0;
	// (i, j) of element e, advanced by the block size without a 64-bit division per element;
	// kVbUnroll elements per thread are loaded before any is stored, so that as many loads are
	// in flight.
	constexpr int kVbUnroll = 8;
	const int64_t end       = jb.start + jb.len;
	const int64_t e0        = jb.start + threadIdx.x;
	int64_t       j         = e0 / jb.ldd;
	int           i         = int(e0 - j * jb.ldd);
	for (int64_t e = e0; e < end; e += kVbUnroll * int64_t(blockDim.x)) {
		// Loads without branches (an element outside the block, in the padding rows or past the
		// chunk, reads element 0 and is replaced by zero), so that the compiler counts them.
		double v[kVbUnroll];
		bool   ok[kVbUnroll];
#pragma unroll
		for (int u = 0; u < kVbUnroll; ++u) {
			ok[u]              = e + int64_t(u) * blockDim.x < end && i < jb.rows;
			const int64_t idx  = tr ? (j + int64_t(i) * jb.ld) : (i + j * int64_t(jb.ld));
			const int64_t safe = ok[u] ? idx : 0;
			v[u]               = global_ptr(jb.src)[cx ? 2 * safe + im : safe];
			i += int(blockDim.x);
			if (i >= jb.ldd) {
				j += i / jb.ldd;
				i %= jb.ldd;
			}
		}
		if (e + int64_t(kVbUnroll - 1) * blockDim.x < end) {
#pragma unroll
			for (int u = 0; u < kVbUnroll; ++u)
				ws[jb.dst + e + int64_t(u) * blockDim.x] = ok[u] ? sg * v[u] : 0.0;
		} else {
#pragma unroll
			for (int u = 0; u < kVbUnroll; ++u)
				if (e + int64_t(u) * blockDim.x < end)
					ws[jb.dst + e + int64_t(u) * blockDim.x] = ok[u] ? sg * v[u] : 0.0;
		}
	}
}

__global__ __launch_bounds__(256) void vb_scatter_kernel(const VbOut* __restrict__ jobs,
                                                         const double* __restrict__ cw, double ar,
                                                         double ai, double br, double bi,
                                                         int beta_zero)
{
	const VbOut jb = jobs[blockIdx.x];
	// As vb_copy_kernel: kVbUnroll elements per thread, all loads before the stores.
	constexpr int kVbUnroll = 8;
	const int64_t end       = jb.start + jb.len;
	const int64_t e0        = jb.start + threadIdx.x;
	int64_t       j         = e0 / jb.rows;
	int           i         = int(e0 - j * jb.rows);
	for (int64_t e = e0; e < end; e += kVbUnroll * int64_t(blockDim.x)) {
		int64_t addr[kVbUnroll];
		double  cr[kVbUnroll], ci[kVbUnroll], c[kVbUnroll], d[kVbUnroll];
#pragma unroll
		for (int u = 0; u < kVbUnroll; ++u) {
			const int64_t eu = e + int64_t(u) * blockDim.x;
			addr[u]          = i + j * int64_t(jb.ldc);
			cr[u] = ci[u] = c[u] = d[u] = 0.0;
			if (eu < end) {
				cr[u] = cw[jb.src_re + eu];
				if (jb.cplx)
					ci[u] = cw[jb.src_im + eu];
				if (!beta_zero) {
					c[u] = jb.cplx ? jb.dst[2 * addr[u]] : jb.dst[addr[u]];
					if (jb.cplx)
						d[u] = jb.dst[2 * addr[u] + 1];
				}
			}
			i += int(blockDim.x);
			if (i >= jb.rows) {
				j += i / jb.rows;
				i %= jb.rows;
			}
		}
#pragma unroll
		for (int u = 0; u < kVbUnroll; ++u) {
			if (e + int64_t(u) * blockDim.x >= end)
				continue;
			if (!jb.cplx) {
				double v = ar * cr[u];
				if (!beta_zero)
					v += br * c[u];
				jb.dst[addr[u]] = v;
			} else {
				double xr = ar * cr[u] - ai * ci[u];
				double xi = ar * ci[u] + ai * cr[u];
				if (!beta_zero) {
					xr += br * c[u] - bi * d[u];
					xi += br * d[u] + bi * c[u];
				}
				jb.dst[2 * addr[u]]     = xr;
				jb.dst[2 * addr[u] + 1] = xi;
			}
		}
	}
}

void launch_vb_copy(const VbCopy* jobs, int njobs, double* ws, hipStream_t s)
{
	if (njobs <= 0)
		return;
	hipLaunchKernelGGL(vb_copy_kernel, dim3(unsigned(njobs)), dim3(256), 0, s, jobs, ws);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_vb_scatter(const VbOut* jobs, int njobs, const double* cw, double ar, double ai,
                       double br, double bi, int beta_zero, hipStream_t s)
{
	if (njobs <= 0)
		return;
	hipLaunchKernelGGL(vb_scatter_kernel, dim3(unsigned(njobs)), dim3(256), 0, s, jobs, cw, ar,
	                   ai, br, bi, beta_zero);
	DMK_HIP_CHECK(hipGetLastError());
}

} // namespace dmk
