// Eigen-decomposition of the reduced density matrix on the GPU: the diagonalisation step of the
// DMRG truncation (DiagBlockDiagMatrix::diagonalise, DiagBlockDiagMatrix.h:73-88, which calls
// PsimagLite::diag(block, eigs, 'V') on every symmetry block, BlockDiagonalMatrix.h:317-322).
//
// Every block is copied to the device, diagonalised by rocSOLVER's divide-and-conquer dsyevd
// (eigenvalues ascending, orthonormal eigenvectors in the columns, as LAPACK dsyev returns
// them) and copied back. A dsyevd of n ~ 1000 is a long chain of small, latency-bound kernels
// (the tridiagonal reduction) that fills a fraction of the chip, so blocks are solved together:
// blocks of similar size (down to half the largest of their class) form one strided batch,
// each padded to the class's size N with a diagonal block d I below its spectrum
// (d = -(sqrt(2) |A|_F + 1)), so that one chain of batched kernels serves the whole class. The
// padded matrix is block diagonal, so its eigenpairs are those of A and N - n pairs (d, unit
// vectors of the padding): ascending, the padding's come first and A's eigenvectors are the
// last n columns, zero in the padding rows (checked; a block whose padding rows are not clean
// is solved again on its own). Classes, and single blocks, run concurrently: up to kSolverStreams
// host threads, each with its own rocBLAS handle and stream, take them largest first. Handles
// and streams are kept per device for the life of the process; every class has its own region
// of one device buffer.
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
#include <cmath>
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

// One block: copy in (from host or device memory), dsyevd, download of the eigenvectors (to
// out, host), eigenvalues and status.
void diag_one(const SolverContext& ctx, int b, int64_t m, const double* src, bool on_device,
              double* out, double* eigs, char* dev)
{
	const size_t mat   = sizeof(double) * size_t(m) * size_t(m);
	double*      dA    = reinterpret_cast<double*>(dev);
	double*      dW    = dA + m * m;
	double*      dE    = dW + m;
	rocblas_int* dInfo = reinterpret_cast<rocblas_int*>(dE + m);
	DMK_HIP_CHECK(hipMemcpyAsync(dA, src, mat,
	                             on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
	                             ctx.stream));
	check_rocblas(rocsolver_dsyevd(ctx.handle, rocblas_evect_original, rocblas_fill_upper,
	                               rocblas_int(m), dA, rocblas_int(m), dW, dE, dInfo),
	              "rocsolver_dsyevd");
	rocblas_int info = 0;
	DMK_HIP_CHECK(hipMemcpyAsync(out, dA, mat, hipMemcpyDeviceToHost, ctx.stream));
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

// A class of blocks (indices `members`, sizes up to N) as one padded strided batch: each block
// goes to the top left of its N x N slot (the slots zeroed first), the padding diagonal gets
// d_b = -(sqrt(2) |A_b|_F + 1), one rocsolver_dsyevd_strided_batched solves the class, and each
// block gets back the last n_b eigenpairs (eigenvectors into vecs[b], column major, leading
// dimension n_b; eigenvalues at eigs + eoff[b]). The blocks come from host memory (src[b]) or, when
// on_device, from device memory. Every host transfer is one contiguous copy through a compact
// staging area per slot (strided copies between pageable host memory and the device go row by
// row); the strided placement into and out of the slots is device to device. Returns the
// members whose result is not clean (status, padding eigenvalues not d_b, or eigenvectors not
// zero in the padding rows), to be solved on their own.
std::vector<int> diag_class(const SolverContext& ctx, const std::vector<int>& members, int64_t N,
                            const int64_t* n, const double* const* src, bool on_device,
                            double* const* vecs, const std::vector<int64_t>& eoff, double* eigs,
                            char* dev)
{
	const size_t nb    = members.size();
	const size_t slot  = size_t(N) * size_t(N);
	const size_t pitch = size_t(N) * sizeof(double);
	double*      dA    = reinterpret_cast<double*>(dev);
	double*      dW    = dA + nb * slot;
	double*      dE    = dW + nb * size_t(N);
	double*      dC    = dE + nb * size_t(N); // compact staging, one slot's size per block
	rocblas_int* dInfo = reinterpret_cast<rocblas_int*>(dC + nb * slot);
	DMK_HIP_CHECK(hipMemsetAsync(dA, 0, nb * slot * sizeof(double), ctx.stream));
	std::vector<double> pads(nb, -1.0);
	std::vector<double> padrow(size_t(N), 0.0);
	for (size_t s = 0; s < nb; ++s) {
		const int     b    = members[s];
		const int64_t m    = n[b];
		double*       cs   = dC + s * slot;
		const double* from = src[b];
		if (!on_device) {
			DMK_HIP_CHECK(hipMemcpyAsync(cs, from, size_t(m * m) * sizeof(double),
			                             hipMemcpyHostToDevice, ctx.stream));
			from = cs;
		}
		// |A|_F of the whole stored block on the device: sqrt 2 times it bounds the spectrum
		// of the symmetric matrix that the upper triangle defines.
		double nrm = 0.0;
		check_rocblas(rocblas_dnrm2(ctx.handle, rocblas_int(m * m), from, 1, &nrm),
		              "rocblas_dnrm2");
		pads[s] = -(std::sqrt(2.0) * nrm + 1.0);
		double* slotA = dA + s * slot;
		DMK_HIP_CHECK(hipMemcpy2DAsync(slotA, pitch, from, size_t(m) * sizeof(double),
		                               size_t(m) * sizeof(double), size_t(m),
		                               hipMemcpyDeviceToDevice, ctx.stream));
		if (N > m) {
			// The padding diagonal, elements (j, j) for j = m .. N-1 (stride N + 1), from a
			// short host row staged behind the block (the staging holds N^2 >= m^2 + N - m).
			std::fill(padrow.begin(), padrow.begin() + (N - m), pads[s]);
			double* pstage = cs + m * m;
			DMK_HIP_CHECK(hipMemcpy(pstage, padrow.data(), size_t(N - m) * sizeof(double),
			                        hipMemcpyHostToDevice));
			DMK_HIP_CHECK(hipMemcpy2DAsync(slotA + m + m * N, size_t(N + 1) * sizeof(double),
			                               pstage, sizeof(double), sizeof(double),
			                               size_t(N - m), hipMemcpyDeviceToDevice,
			                               ctx.stream));
		}
	}
	check_rocblas(rocsolver_dsyevd_strided_batched(
	                  ctx.handle, rocblas_evect_original, rocblas_fill_upper, rocblas_int(N), dA,
	                  rocblas_int(N), rocblas_stride(slot), dW, rocblas_stride(N), dE,
	                  rocblas_stride(N), dInfo, rocblas_int(nb)),
	              "rocsolver_dsyevd_strided_batched");
	// First the statuses, the padding eigenvalues and the padding rows of the kept eigenvectors
	// (which must be zero); the caller's outputs are written only for clean results.
	std::vector<rocblas_int>         info(nb, 0);
	std::vector<std::vector<double>> wpad(nb), vpad(nb);
	for (size_t s = 0; s < nb; ++s) {
		const int64_t m = n[members[s]], p = N - m;
		wpad[s].resize(size_t(p));
		vpad[s].resize(size_t(p * m));
		if (p > 0) {
			double* cs = dC + s * slot;
			DMK_HIP_CHECK(hipMemcpy2DAsync(cs, size_t(p) * sizeof(double),
			                               dA + s * slot + m + p * N, pitch,
			                               size_t(p) * sizeof(double), size_t(m),
			                               hipMemcpyDeviceToDevice, ctx.stream));
			DMK_HIP_CHECK(hipMemcpyAsync(vpad[s].data(), cs, size_t(p * m) * sizeof(double),
			                             hipMemcpyDeviceToHost, ctx.stream));
			DMK_HIP_CHECK(hipMemcpyAsync(wpad[s].data(), dW + s * size_t(N),
			                             size_t(p) * sizeof(double), hipMemcpyDeviceToHost,
			                             ctx.stream));
		}
	}
	DMK_HIP_CHECK(hipMemcpyAsync(info.data(), dInfo, nb * sizeof(rocblas_int),
	                             hipMemcpyDeviceToHost, ctx.stream));
	DMK_HIP_CHECK(hipStreamSynchronize(ctx.stream));
	std::vector<int> redo;
	for (size_t s = 0; s < nb; ++s) {
		const int     b = members[s];
		const int64_t m = n[b], p = N - m;
		bool          clean = info[s] == 0;
		for (double w : wpad[s])
			clean = clean && std::fabs(w - pads[s]) <= 1e-12 * std::fabs(pads[s]);
		for (double v : vpad[s])
			clean = clean && std::fabs(v) <= 1e-12;
		if (!clean) {
			redo.push_back(b);
			continue;
		}
		// A's eigenpairs: the last m eigenvalues, rows 0 .. m-1 of columns p .. N-1.
		double* cs = dC + s * slot;
		DMK_HIP_CHECK(hipMemcpy2DAsync(cs, size_t(m) * sizeof(double), dA + s * slot + p * N,
		                               pitch, size_t(m) * sizeof(double), size_t(m),
		                               hipMemcpyDeviceToDevice, ctx.stream));
		DMK_HIP_CHECK(hipMemcpyAsync(vecs[b], cs, size_t(m * m) * sizeof(double),
		                             hipMemcpyDeviceToHost, ctx.stream));
		DMK_HIP_CHECK(hipMemcpyAsync(eigs + eoff[b], dW + s * size_t(N) + p,
		                             size_t(m) * sizeof(double), hipMemcpyDeviceToHost,
		                             ctx.stream));
	}
	DMK_HIP_CHECK(hipStreamSynchronize(ctx.stream));
	return redo;
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

namespace {

// The eigenpairs of nblocks symmetric blocks (n[b] x n[b], compact column major at src[b], in
// host memory or, on_device, in device memory): eigenvectors to the host arrays vecs[b],
// eigenvalues ascending to eigs (block after block).
void diag_blocks(const char* what, int nblocks, const int64_t* n, const double* const* src,
                 bool on_device, double* const* vecs, double* eigs)
{
	if (nblocks < 0)
		throw std::invalid_argument(std::string(what) + ": negative block count");
	if (nblocks == 0)
		return;
	if (!n || !src || !vecs || !eigs)
		throw std::invalid_argument(std::string(what) + ": null argument");
	std::vector<int64_t> eoff(size_t(nblocks) + 1, 0); // offsets of the blocks' eigenvalues
	std::vector<int>     by_size;
	for (int b = 0; b < nblocks; ++b) {
		if (n[b] < 0 || n[b] > (int64_t(1) << 15))
			throw std::invalid_argument(std::string(what)
			                            + ": block size out of range (0..32768)");
		if (n[b] > 0 && (!src[b] || !vecs[b]))
			throw std::invalid_argument(std::string(what) + ": null block");
		eoff[b + 1] = eoff[b] + n[b];
		if (n[b] > 0)
			by_size.push_back(b);
	}
	if (by_size.empty())
		return;
	// Classes, largest first: a class takes the blocks down to half the size of its first one,
	// at most kClassMax of them (a class of one block is solved without padding).
	constexpr size_t kClassMax = 16;
	std::stable_sort(by_size.begin(), by_size.end(), [&](int a, int b) { return n[a] > n[b]; });
	std::vector<std::vector<int>> classes;
	for (int b : by_size) {
		if (classes.empty() || classes.back().size() == kClassMax
		    || 2 * n[b] < n[classes.back().front()])
			classes.emplace_back();
		classes.back().push_back(b);
	}
	std::vector<size_t> doff(classes.size(), 0); // offsets of the classes' device regions
	std::vector<int>    order(classes.size());
	size_t              total = 0;
	for (size_t c = 0; c < classes.size(); ++c) {
		const size_t N  = size_t(n[classes[c].front()]);
		const size_t nb = classes[c].size();
		order[c]        = int(c);
		doff[c]         = total;
		total += region_bytes(nb * N * (2 * N + 2)); // A, W, E, staging of every slot, statuses
	}
	for_blocks_concurrently(what, int(classes.size()), order, doff, total,
	                        [&](const SolverContext& ctx, int c, char* region) {
		                        const std::vector<int>& cl     = classes[size_t(c)];
		                        std::vector<int>        single = cl;
		                        if (cl.size() > 1)
			                        single = diag_class(ctx, cl, n[cl.front()], n, src,
			                                            on_device, vecs, eoff, eigs, region);
		                        // The class's region is free again once diag_class returns.
		                        for (int b : single)
			                        diag_one(ctx, b, n[b], src[b], on_device, vecs[b],
			                                 eigs + eoff[b], region);
	                        });
}

} // namespace

void density_diag(int nblocks, const int64_t* n, double* const* blocks, double* eigs)
{
	diag_blocks("dmk_density_diag_d", nblocks, n, blocks, false, blocks, eigs);
}

void density_diag_device(int nblocks, const int64_t* n, const double* const* dblocks,
                         double* const* vecs, double* eigs)
{
	diag_blocks("dmk_density_diag_dev_d", nblocks, n, dblocks, true, vecs, eigs);
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
