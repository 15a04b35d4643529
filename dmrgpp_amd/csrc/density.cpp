// Eigen-decomposition of the reduced density matrix on the GPU: the diagonalisation step of the
// DMRG truncation (DiagBlockDiagMatrix::diagonalise, DiagBlockDiagMatrix.h:73-88, which calls
// PsimagLite::diag(block, eigs, 'V') on every symmetry block, BlockDiagonalMatrix.h:317-322).
//
// Every block is copied to the device, diagonalised by rocSOLVER's divide-and-conquer dsyevd
// (eigenvalues ascending, orthonormal eigenvectors in the columns, as LAPACK dsyev returns
// them) and copied back. A dsyevd of n ~ 1000 is a long chain of small, latency-bound kernels
// (the tridiagonal reduction) that fills a fraction of the chip, so the blocks run
// concurrently: up to kSolverStreams host threads, each with its own rocBLAS handle and stream,
// take the blocks largest first. Handles and streams are kept per device for the life of the
// process; every block has its own region of one device buffer.
//
// The SVD variant (DensityMatrixSvd::ParallelSvd::doTask, DensityMatrixSvd.h:382-410, which
// calls PsimagLite::Svd with jobz 'A' on every group's wave-function matrix) runs rocSOLVER
// dgesvd the same way: the left singular vectors are the density matrix's eigenvectors and the
// squared singular values its eigenvalues.
#include "dmk_internal.h"
#include "pool.h"

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <exception>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace dmk {

namespace {

constexpr int kSolverStreams = 4;

struct SolverContext {
	rocblas_handle handle = nullptr;
	hipStream_t    stream = nullptr;
};

std::mutex                                                g_mutex;
std::vector<std::array<SolverContext, kSolverStreams>> g_contexts; // per device, on first use

void check_rocblas(rocblas_status st, const char* what)
{
	if (st != rocblas_status_success)
		throw std::runtime_error(std::string("hip: ") + what + " failed (rocblas status "
		                         + std::to_string(int(st)) + ")");
}

// The first `count` contexts of the current device, created on first use.
SolverContext* contexts_for_current_device(int count)
{
	int dev = 0;
	DMK_HIP_CHECK(hipGetDevice(&dev));
	if (int(g_contexts.size()) <= dev)
		g_contexts.resize(dev + 1);
	for (int w = 0; w < count; ++w) {
		SolverContext& c = g_contexts[dev][w];
		if (!c.stream)
			DMK_HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
		if (!c.handle) {
			check_rocblas(rocblas_create_handle(&c.handle), "rocblas_create_handle");
			check_rocblas(rocblas_set_stream(c.handle, c.stream), "rocblas_set_stream");
		}
	}
	return g_contexts[dev].data();
}

// One block: upload, dsyevd, download of the eigenvectors, eigenvalues and status.
void diag_one(const SolverContext& ctx, int b, int64_t m, double* host, double* eigs, char* dev)
{
	const size_t mat   = sizeof(double) * size_t(m) * size_t(m);
	double*      dA    = reinterpret_cast<double*>(dev);
	double*      dW    = dA + m * m;
	double*      dE    = dW + m;
	rocblas_int* dInfo = reinterpret_cast<rocblas_int*>(dE + m);
	DMK_HIP_CHECK(hipMemcpyAsync(dA, host, mat, hipMemcpyHostToDevice, ctx.stream));
	check_rocblas(rocsolver_dsyevd(ctx.handle, rocblas_evect_original, rocblas_fill_upper,
	                               rocblas_int(m), dA, rocblas_int(m), dW, dE, dInfo),
	              "rocsolver_dsyevd");
	rocblas_int info = 0;
	DMK_HIP_CHECK(hipMemcpyAsync(host, dA, mat, hipMemcpyDeviceToHost, ctx.stream));
	DMK_HIP_CHECK(hipMemcpyAsync(eigs, dW, sizeof(double) * size_t(m), hipMemcpyDeviceToHost,
	                             ctx.stream));
	DMK_HIP_CHECK(hipMemcpyAsync(&info, dInfo, sizeof(rocblas_int), hipMemcpyDeviceToHost,
	                             ctx.stream));
	DMK_HIP_CHECK(hipStreamSynchronize(ctx.stream));
	if (info != 0)
		throw std::runtime_error("dmk_density_diag_d: dsyevd did not converge on block "
		                         + std::to_string(b) + " (info " + std::to_string(int(info))
		                         + ")");
}

} // namespace

namespace {

// Runs one(ctx, b, region) for the blocks in `order` (largest first) on up to kSolverStreams
// host threads, each with its own context; region is the block's part of one device buffer
// (offsets doff, `total` bytes). Every stream is drained before the buffer goes back to the
// pool; the error of the lowest-numbered failing block is the one reported (else a worker's).
template <typename Fn>
void for_blocks_concurrently(const char* what, int nblocks, const std::vector<int>& order,
                             const std::vector<size_t>& doff, size_t total, Fn one)
{
	std::lock_guard<std::mutex> lock(g_mutex);
	const int      nw   = std::min(kSolverStreams, int(order.size()));
	SolverContext* ctxs = contexts_for_current_device(nw);
	int            dev  = 0;
	DMK_HIP_CHECK(hipGetDevice(&dev));
	char* buf = static_cast<char*>(pool_device_alloc(total));

	std::atomic<size_t>             next { 0 };
	std::atomic<bool>               failed { false };
	// One slot per block (each block is run by one worker) plus one per worker for the
	// worker's own set-up failure: no slot has two writers.
	std::vector<std::exception_ptr> errs(static_cast<size_t>(nblocks), nullptr);
	std::vector<std::exception_ptr> werrs(static_cast<size_t>(nw), nullptr);
	auto work = [&](int w) {
		if (w > 0 && hipSetDevice(dev) != hipSuccess) {
			failed   = true;
			werrs[size_t(w)] = std::make_exception_ptr(
			    std::runtime_error(std::string(what) + ": hipSetDevice failed in a worker"));
			return;
		}
		for (size_t i = next.fetch_add(1); i < order.size() && !failed; i = next.fetch_add(1)) {
			const int b = order[i];
			try {
				one(ctxs[w], b, buf + doff[size_t(b)]);
			} catch (...) {
				errs[size_t(b)] = std::current_exception();
				failed          = true;
			}
		}
		(void)hipStreamSynchronize(ctxs[w].stream);
	};
	std::vector<std::thread> pool;
	try {
		for (int w = 1; w < nw; ++w)
			pool.emplace_back(work, w);
	} catch (...) {
		// No more threads: the ones started and this one take all the blocks.
	}
	work(0);
	for (auto& t : pool)
		t.join();
	pool_device_free(buf);
	for (const auto& e : errs)
		if (e)
			std::rethrow_exception(e);
	for (const auto& e : werrs)
		if (e)
			std::rethrow_exception(e);
}

size_t region_bytes(size_t doubles) { return (sizeof(double) * doubles + 64 + 255) & ~size_t(255); }

// One block of the SVD: upload, dgesvd (all left singular vectors, no right ones), download of
// U, the singular values and the status.
void svd_one(const SolverContext& ctx, int b, int64_t r, int64_t c, const double* host,
             double* u, double* s, char* dev)
{
	const int64_t k  = std::min(r, c);
	double*       dA = reinterpret_cast<double*>(dev);
	double*       dU = dA + r * c;
	double*       dS = dU + r * r;
	double*       dE = dS + k;
	rocblas_int*  dInfo = reinterpret_cast<rocblas_int*>(dE + k);
	DMK_HIP_CHECK(hipMemcpyAsync(dA, host, sizeof(double) * size_t(r) * size_t(c),
	                             hipMemcpyHostToDevice, ctx.stream));
	check_rocblas(rocsolver_dgesvd(ctx.handle, rocblas_svect_all, rocblas_svect_none,
	                               rocblas_int(r), rocblas_int(c), dA, rocblas_int(r), dS, dU,
	                               rocblas_int(r), nullptr, 1, dE, rocblas_outofplace, dInfo),
	              "rocsolver_dgesvd");
	rocblas_int info = 0;
	DMK_HIP_CHECK(hipMemcpyAsync(u, dU, sizeof(double) * size_t(r) * size_t(r),
	                             hipMemcpyDeviceToHost, ctx.stream));
	DMK_HIP_CHECK(hipMemcpyAsync(s, dS, sizeof(double) * size_t(k), hipMemcpyDeviceToHost,
	                             ctx.stream));
	DMK_HIP_CHECK(hipMemcpyAsync(&info, dInfo, sizeof(rocblas_int), hipMemcpyDeviceToHost,
	                             ctx.stream));
	DMK_HIP_CHECK(hipStreamSynchronize(ctx.stream));
	if (info != 0)
		throw std::runtime_error("dmk_density_svd_d: dgesvd did not converge on block "
		                         + std::to_string(b) + " (info " + std::to_string(int(info))
		                         + ")");
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
	std::vector<int64_t> eoff(size_t(nblocks) + 1, 0); // offsets of the blocks' eigenvalues
	std::vector<size_t>  doff(size_t(nblocks), 0);     // offsets of the blocks' device regions
	std::vector<int>     order;
	size_t               total = 0;
	for (int b = 0; b < nblocks; ++b) {
		if (n[b] < 0 || n[b] > (int64_t(1) << 15))
			throw std::invalid_argument("dmk_density_diag_d: block size out of range (0..32768)");
		if (n[b] > 0 && !blocks[b])
			throw std::invalid_argument("dmk_density_diag_d: null block");
		eoff[b + 1] = eoff[b] + n[b];
		if (n[b] > 0) {
			order.push_back(b);
			doff[b] = total;
			total += region_bytes(size_t(n[b]) * size_t(n[b] + 2)); // A, W, E, info
		}
	}
	if (order.empty())
		return;
	// Largest first, so that the longest solves start at once.
	std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return n[a] > n[b]; });
	for_blocks_concurrently("dmk_density_diag_d", nblocks, order, doff, total,
	                        [&](const SolverContext& ctx, int b, char* region) {
		                        diag_one(ctx, b, n[b], blocks[b], eigs + eoff[b], region);
	                        });
}

void density_svd(int nblocks, const int64_t* rows, const int64_t* cols,
                 const double* const* blocks, double* const* u, double* s)
{
	if (nblocks < 0)
		throw std::invalid_argument("dmk_density_svd_d: negative block count");
	if (nblocks == 0)
		return;
	if (!rows || !cols || !blocks || !u || !s)
		throw std::invalid_argument("dmk_density_svd_d: null argument");
	std::vector<int64_t> soff(size_t(nblocks) + 1, 0); // offsets of the singular values
	std::vector<size_t>  doff(size_t(nblocks), 0);
	std::vector<int>     order;
	size_t               total = 0;
	for (int b = 0; b < nblocks; ++b) {
		const int64_t r = rows[b], c = cols[b];
		if (r < 0 || c < 0 || r > (int64_t(1) << 15) || c > (int64_t(1) << 15))
			throw std::invalid_argument("dmk_density_svd_d: block size out of range (0..32768)");
		if (r > 0 && !u[b])
			throw std::invalid_argument("dmk_density_svd_d: null output block");
		if (r > 0 && c > 0 && !blocks[b])
			throw std::invalid_argument("dmk_density_svd_d: null block");
		const int64_t k = std::min(r, c);
		soff[b + 1]     = soff[b] + k;
		if (k > 0) {
			order.push_back(b);
			doff[b] = total;
			total += region_bytes(size_t(r) * size_t(c) + size_t(r) * size_t(r) + 2 * size_t(k));
		}
	}
	// A block with no columns has no singular values: its left singular vectors are any
	// orthonormal basis, the identity here (rocSOLVER is not called for it).
	for (int b = 0; b < nblocks; ++b)
		if (rows[b] > 0 && cols[b] == 0)
			for (int64_t j = 0; j < rows[b]; ++j)
				for (int64_t i = 0; i < rows[b]; ++i)
					u[b][i + j * rows[b]] = i == j ? 1.0 : 0.0;
	if (order.empty())
		return;
	std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
		return rows[a] * cols[a] * std::min(rows[a], cols[a])
		     > rows[b] * cols[b] * std::min(rows[b], cols[b]);
	});
	for_blocks_concurrently("dmk_density_svd_d", nblocks, order, doff, total,
	                        [&](const SolverContext& ctx, int b, char* region) {
		                        svd_one(ctx, b, rows[b], cols[b], blocks[b], u[b],
		                                s + soff[b], region);
	                        });
}

} // namespace dmk
