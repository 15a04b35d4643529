// Eigen-decomposition of the reduced density matrix on the GPU: the diagonalisation step of the
// DMRG truncation (DiagBlockDiagMatrix::diagonalise, DiagBlockDiagMatrix.h:73-88, which calls
// PsimagLite::diag(block, eigs, 'V') on every symmetry block, BlockDiagonalMatrix.h:317-322).
//
// Every block is copied to the device, diagonalised by rocSOLVER's divide-and-conquer dsyevd
// (eigenvalues ascending, orthonormal eigenvectors in the columns, as LAPACK dsyev returns
// them) and copied back. The blocks share one device buffer sized for the largest; a rocBLAS
// handle and a stream are kept per device for the life of the process.
#include "dmk_internal.h"
#include "pool.h"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

namespace dmk {

namespace {

struct SolverContext {
	rocblas_handle handle = nullptr;
	hipStream_t    stream = nullptr;
};

std::mutex                 g_mutex;
std::vector<SolverContext> g_contexts; // per device, created on first use

void check_rocblas(rocblas_status st, const char* what)
{
	if (st != rocblas_status_success)
		throw std::runtime_error(std::string("hip: ") + what + " failed (rocblas status "
		                         + std::to_string(int(st)) + ")");
}

SolverContext& context_for_current_device()
{
	int dev = 0;
	DMK_HIP_CHECK(hipGetDevice(&dev));
	if (int(g_contexts.size()) <= dev)
		g_contexts.resize(dev + 1);
	SolverContext& c = g_contexts[dev];
	if (!c.handle) {
		DMK_HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
		check_rocblas(rocblas_create_handle(&c.handle), "rocblas_create_handle");
		check_rocblas(rocblas_set_stream(c.handle, c.stream), "rocblas_set_stream");
	}
	return c;
}

} // namespace

void density_diag(int nblocks, const int64_t* n, double* const* blocks, double* eigs)
{
	if (nblocks < 0)
		throw std::invalid_argument("dmk_density_diag_d: negative block count");
	if (nblocks == 0)
		return;
	if (!n || !blocks || !eigs)
		throw std::invalid_argument("dmk_density_diag_d: null argument");
	int64_t nmax = 0;
	for (int b = 0; b < nblocks; ++b) {
		if (n[b] < 0 || n[b] > (int64_t(1) << 15))
			throw std::invalid_argument("dmk_density_diag_d: block size out of range (0..32768)");
		if (n[b] > 0 && !blocks[b])
			throw std::invalid_argument("dmk_density_diag_d: null block");
		nmax = std::max(nmax, n[b]);
	}
	if (nmax == 0)
		return;

	std::lock_guard<std::mutex> lock(g_mutex);
	SolverContext&              ctx = context_for_current_device();
	const size_t bytes = sizeof(double) * size_t(nmax) * size_t(nmax + 2) + 64;
	char*        buf   = static_cast<char*>(pool_device_alloc(bytes));
	double*      dA    = reinterpret_cast<double*>(buf);
	double*      dW    = dA + nmax * nmax;
	double*      dE    = dW + nmax;
	rocblas_int* dInfo = reinterpret_cast<rocblas_int*>(dE + nmax);
	try {
		int64_t off = 0;
		for (int b = 0; b < nblocks; ++b) {
			const int64_t m = n[b];
			if (m == 0)
				continue;
			const size_t mat = sizeof(double) * size_t(m) * size_t(m);
			DMK_HIP_CHECK(hipMemcpyAsync(dA, blocks[b], mat, hipMemcpyHostToDevice, ctx.stream));
			check_rocblas(rocsolver_dsyevd(ctx.handle, rocblas_evect_original, rocblas_fill_upper,
			                               rocblas_int(m), dA, rocblas_int(m), dW, dE, dInfo),
			              "rocsolver_dsyevd");
			rocblas_int info = 0;
			DMK_HIP_CHECK(hipMemcpyAsync(blocks[b], dA, mat, hipMemcpyDeviceToHost, ctx.stream));
			DMK_HIP_CHECK(hipMemcpyAsync(eigs + off, dW, sizeof(double) * size_t(m),
			                             hipMemcpyDeviceToHost, ctx.stream));
			DMK_HIP_CHECK(hipMemcpyAsync(&info, dInfo, sizeof(rocblas_int), hipMemcpyDeviceToHost,
			                             ctx.stream));
			DMK_HIP_CHECK(hipStreamSynchronize(ctx.stream));
			if (info != 0)
				throw std::runtime_error("dmk_density_diag_d: dsyevd did not converge on block "
				                         + std::to_string(b) + " (info "
				                         + std::to_string(int(info)) + ")");
			off += m;
		}
	} catch (...) {
		// Drain the stream before the buffer goes back to the pool; the first error is the
		// one reported.
		(void)hipStreamSynchronize(ctx.stream);
		pool_device_free(buf);
		throw;
	}
	pool_device_free(buf);
}

} // namespace dmk
