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

__global__ __launch_bounds__(256) void vb_copy_kernel(const VbCopy* __restrict__ jobs,
                                                      double* __restrict__ ws)
{
	const VbCopy jb = jobs[blockIdx.x];
	const bool   tr = (jb.flags & VB_TRANS) != 0;
	const bool   cx = (jb.flags & VB_CPLX) != 0;
	const int    im = (jb.flags & VB_IMAG) ? 1 : 0;
	const double sg = (jb.flags & VB_NEG) ? -1.0 : 1.0;
	// (i, j) of element e, advanced by the block size without a 64-bit division per element.
	const int64_t e0 = jb.start + threadIdx.x;
	int64_t       j  = e0 / jb.ldd;
	int           i  = int(e0 - j * jb.ldd);
	for (int64_t e = e0; e < jb.start + jb.len; e += blockDim.x) {
		double v = 0.0; // the padding rows [rows, ldd)
		if (i < jb.rows) {
			const int64_t idx = tr ? (j + int64_t(i) * jb.ld) : (i + j * int64_t(jb.ld));
			v                 = sg * (cx ? jb.src[2 * idx + im] : jb.src[idx]);
		}
		ws[jb.dst + e] = v;
		i += int(blockDim.x);
		if (i >= jb.ldd) {
			j += i / jb.ldd;
			i %= jb.ldd;
		}
	}
}

__global__ __launch_bounds__(256) void vb_scatter_kernel(const VbOut* __restrict__ jobs,
                                                         const double* __restrict__ cw, double ar,
                                                         double ai, double br, double bi,
                                                         int beta_zero)
{
	const VbOut jb = jobs[blockIdx.x];
	const int64_t e0 = jb.start + threadIdx.x;
	int64_t       j  = e0 / jb.rows;
	int           i  = int(e0 - j * jb.rows);
	for (int64_t e = e0; e < jb.start + jb.len; e += blockDim.x) {
		const int64_t addr = i + j * int64_t(jb.ldc);
		if (!jb.cplx) {
			double v = ar * cw[jb.src_re + e];
			if (!beta_zero)
				v += br * jb.dst[addr];
			jb.dst[addr] = v;
		} else {
			const double cr = cw[jb.src_re + e];
			const double ci = cw[jb.src_im + e];
			double       xr = ar * cr - ai * ci;
			double       xi = ar * ci + ai * cr;
			if (!beta_zero) {
				const double c = jb.dst[2 * addr];
				const double d = jb.dst[2 * addr + 1];
				xr += br * c - bi * d;
				xi += br * d + bi * c;
			}
			jb.dst[2 * addr]     = xr;
			jb.dst[2 * addr + 1] = xi;
		}
		i += int(blockDim.x);
		if (i >= jb.rows) {
			j += i / jb.rows;
			i %= jb.rows;
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
