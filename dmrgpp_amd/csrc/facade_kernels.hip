// Kernels of the matvec facade (facade.cpp): the stored-matrix product x += H y over CSR
// (PsimagLite CrsMatrix::matrixVectorProduct, which MatrixVectorKron calls on its matrixStored_,
// MatrixVectorKron.h:146-147) and the accumulation x += t of the Kron path's product.
#include "dmk_internal.h"

#include <hip/hip_runtime.h>

namespace dmk {

namespace {

// One wavefront per row: lanes stride the row's nonzeros, then a fixed-order butterfly sum, so
// the result does not depend on timing. CPLX: values and vectors are interleaved complex.
template <bool CPLX>
__global__ __launch_bounds__(256) void csr_apply_kernel(const int64_t* __restrict__ rowptr,
                                                        const int32_t* __restrict__ col,
                                                        const double* __restrict__ val,
                                                        int64_t rows, const double* __restrict__ y,
                                                        double* __restrict__ x)
{
	const int     lane = threadIdx.x & 63;
	const int64_t row  = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
	if (row >= rows)
		return;
	double re = 0.0, im = 0.0;
	for (int64_t e = rowptr[row] + lane; e < rowptr[row + 1]; e += 64) {
		const int c = col[e];
		if (CPLX) {
			const double vr = val[2 * e], vi = val[2 * e + 1];
			const double yr = y[2 * int64_t(c)], yi = y[2 * int64_t(c) + 1];
			re += vr * yr - vi * yi;
			im += vr * yi + vi * yr;
		} else {
			re += val[e] * y[c];
		}
	}
#pragma unroll
	for (int off = 32; off > 0; off >>= 1) {
		re += __shfl_xor(re, off, 64);
		if (CPLX)
			im += __shfl_xor(im, off, 64);
	}
	if (lane == 0) {
		if (CPLX) {
			x[2 * row] += re;
			x[2 * row + 1] += im;
		} else {
			x[row] += re;
		}
	}
}

__global__ __launch_bounds__(256) void axpy_one_kernel(double* __restrict__ x,
                                                       const double* __restrict__ y, int64_t n)
{
	for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256)
		x[i] += y[i];
}

} // namespace

void launch_csr_apply(const int64_t* rowptr, const int32_t* col, const double* val, int64_t rows,
                      int cplx, const double* y, double* x, hipStream_t s)
{
	if (rows <= 0)
		return;
	const unsigned blocks = unsigned((rows + 3) / 4);
	if (cplx)
		hipLaunchKernelGGL(csr_apply_kernel<true>, dim3(blocks), dim3(256), 0, s, rowptr, col,
		                   val, rows, y, x);
	else
		hipLaunchKernelGGL(csr_apply_kernel<false>, dim3(blocks), dim3(256), 0, s, rowptr, col,
		                   val, rows, y, x);
	DMK_HIP_CHECK(hipGetLastError());
}

void launch_axpy_one(double* x, const double* y, int64_t n, hipStream_t s)
{
	if (n <= 0)
		return;
	int64_t b = (n + 255) / 256;
	if (b > 4096)
		b = 4096;
	hipLaunchKernelGGL(axpy_one_kernel, dim3(unsigned(b)), dim3(256), 0, s, x, y, n);
	DMK_HIP_CHECK(hipGetLastError());
}

} // namespace dmk
