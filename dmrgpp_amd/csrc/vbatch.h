// Variable-size batched GEMM over caller pointers (the dmrg_Xgemm_vbatch boundary of the
// reference plugin, GPUPlugin/dmrg_vbatch.cpp:129-652), built on the K-major panel kernel.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace dmk {

enum VbFlags : int32_t
{
	VB_TRANS = 1, // destination (i, j) = source (j, i)
	VB_CPLX  = 2, // source is interleaved complex
	VB_IMAG  = 4, // take the imaginary part (else the real part)
	VB_NEG   = 8, // negate
};

// One chunk [start, start + len) of a column-major destination matrix (rows x ?, leading
// dimension ldd >= rows, zeros in rows [rows, ldd)) at element offset dst of the staging
// workspace, read from src with leading dimension ld.
struct VbCopy {
	const double* src;
	int64_t       dst;
	int64_t       start;
	int64_t       len;
	int32_t       rows;
	int32_t       ld;
	int32_t       flags;
	int32_t       ldd;
};

// One chunk of a result scattered into the caller's C (rows x ?, leading dimension ldc):
// C = alpha * W + beta * C with W the compact workspace result (real part at src_re, imaginary
// part at src_im for complex).
struct VbOut {
	double* dst;
	int64_t src_re;
	int64_t src_im;
	int64_t start;
	int64_t len;
	int32_t rows;
	int32_t ldc;
	int32_t cplx;
	int32_t pad;
};

void launch_vb_copy(const VbCopy* jobs, int njobs, double* ws, hipStream_t s);
void launch_vb_scatter(const VbOut* jobs, int njobs, const double* cw, double ar, double ai,
                       double br, double bi, int beta_zero, hipStream_t s);

// C_e = alpha op(A_e) op(B_e) + beta C_e for every entry e of every group (m, n, k, lda, ldb,
// ldc per group; pointers per entry, device memory, expanded in group order). cplx selects
// interleaved complex scalars (alpha, beta then point to (re, im) pairs). Synchronous.
void gemm_vbatch(bool cplx, char transa, char transb, const int* m, const int* n, const int* k,
                 const double* alpha, const double* const* a, const int* lda,
                 const double* const* b, const int* ldb, const double* beta, double* const* c,
                 const int* ldc, size_t group_count, const int* group_size);

} // namespace dmk
