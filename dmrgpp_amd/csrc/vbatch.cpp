// Variable-size batched GEMM: dmrg_Xgemm_vbatch (GPUPlugin/dmrg_vbatch.cpp:129-652) on the
// K-major panel kernel of the matvec (panel_kernels.hip).
//
// The reference expands its groups into per-entry arrays and hands them to MAGMA
// *_vbatched_max_nocheck (or a BLAS loop). Here every entry becomes one task of the panel
// kernel (two for complex scalars), after its operands are staged into one workspace in the
// panel kernel's K-major layout: the task computes C^T = op(B)^T op(A)^T, whose left operand is
// op(B) column major (each column contiguous along k) and whose right operand is op(A)^T column
// major (each row of op(A) contiguous along k), both with leading dimensions padded to a
// multiple of kPanelPad and zero filled; C^T written by rows is C by columns. Staging makes transposes, conjugation and
// complex scalars a property of the copy, so the GEMM launch only ever sees real K-major
// operands, addressed as offsets into one allocation. Complex products run as real ones:
//   Re C = Ar Br + (-Ai) Bi,   Im C = Ar Bi + Ai Br   (two segments per task).
// The scatter then forms alpha W + beta C in the caller's C (C is not read when beta == 0, as
// in BLAS). The call is synchronous, like the reference's.
#include "vbatch.h"

#include "dmk_internal.h"
#include "pool.h"

#include <algorithm>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

namespace dmk {

template <typename T, typename TO>
void launch_panel_gemm(int bm, int bn, int variant, int depth, int kchunk, const PTask* tasks,
                       const PSeg* segs, const Tile* tiles, int ntiles, const T* const* bases,
                       const PStrides& vs, TO* out_y, int64_t ystride, T* out_ws,
                       int64_t wsstride, int nvec, int beta, hipStream_t stream,
                       const PanelLaunchExtra& ex = PanelLaunchExtra {});

namespace {

struct Entry {
	int           m, n, k, lda, ldb, ldc;
	const double* a;
	const double* b;
	double*       c;
};

// op code of a BLAS transpose character: 0 = N, 1 = T, 2 = C (conjugate transpose).
int trans_code(char t, const char* which)
{
	switch (t) {
	case 'N':
	case 'n':
		return 0;
	case 'T':
	case 't':
		return 1;
	case 'C':
	case 'c':
		return 2;
	default:
		throw std::invalid_argument(std::string("gemm_vbatch: bad ") + which + " '" + t + "'");
	}
}

constexpr int64_t kChunk = 8192; // elements per staging / scatter workgroup (32 per thread)

// Appends copy jobs staging op(X) (rows x cols) from src into the workspace at dst, column
// major with the leading dimension padded (panel_ld(rows)) and the padding zero filled.
// Returns the staged size in elements.
int64_t stage(std::vector<VbCopy>& jobs, const double* src, int ld, int rows, int cols,
              int64_t dst, int flags)
{
	const int     ldd   = int(panel_ld(rows));
	const int64_t total = int64_t(ldd) * cols;
	for (int64_t s = 0; s < total; s += kChunk) {
		VbCopy j{};
		j.src   = src;
		j.dst   = dst;
		j.start = s;
		j.len   = std::min(kChunk, total - s);
		j.rows  = rows;
		j.ld    = ld;
		j.flags = flags;
		j.ldd   = ldd;
		jobs.push_back(j);
	}
	return total;
}

template <typename T> T* upload(const std::vector<T>& v, std::vector<void*>& owned)
{
	if (v.empty())
		return nullptr;
	void* p = pool_device_alloc(v.size() * sizeof(T));
	owned.push_back(p);
	DMK_HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
	return static_cast<T*>(p);
}

} // namespace

void gemm_vbatch(bool cplx, char transa, char transb, const int* m, const int* n, const int* k,
                 const double* alpha, const double* const* a, const int* lda,
                 const double* const* b, const int* ldb, const double* beta, double* const* c,
                 const int* ldc, size_t group_count, const int* group_size)
{
	const int ta = trans_code(transa, "transa");
	const int tb = trans_code(transb, "transb");
	if (group_count == 0)
		return;
	if (!m || !n || !k || !alpha || !a || !lda || !b || !ldb || !beta || !c || !ldc || !group_size)
		throw std::invalid_argument("gemm_vbatch: null argument");
	std::vector<Entry> ents;
	size_t             idx = 0;
	for (size_t g = 0; g < group_count; ++g) {
		if (group_size[g] < 0 || m[g] < 0 || n[g] < 0 || k[g] < 0)
			throw std::invalid_argument("gemm_vbatch: negative size in group " + std::to_string(g));
		const int rowsA = ta ? k[g] : m[g];
		const int rowsB = tb ? n[g] : k[g];
		if (lda[g] < std::max(1, rowsA) || ldb[g] < std::max(1, rowsB) || ldc[g] < std::max(1, m[g]))
			throw std::invalid_argument("gemm_vbatch: leading dimension too small in group "
			                            + std::to_string(g));
		for (int i = 0; i < group_size[g]; ++i, ++idx) {
			Entry e{ m[g], n[g], k[g], lda[g], ldb[g], ldc[g], a[idx], b[idx], c[idx] };
			if (e.m == 0 || e.n == 0)
				continue;
			if (!e.c || (e.k > 0 && (!e.a || !e.b)))
				throw std::invalid_argument("gemm_vbatch: null matrix pointer at entry "
				                            + std::to_string(idx));
			ents.push_back(e);
		}
	}
	if (ents.empty())
		return;
	const bool alpha_zero = alpha[0] == 0.0 && (!cplx || alpha[1] == 0.0);
	const bool beta_zero  = beta[0] == 0.0 && (!cplx || beta[1] == 0.0);
	const int  cx         = cplx ? VB_CPLX : 0;

	std::vector<VbCopy> copies;
	std::vector<VbOut>  outs;
	std::vector<PTask>  tasks;
	std::vector<PSeg>   segs;
	std::vector<double> cost;
	int64_t             wsz = 0, csz = 0;
	// Segment of C^T(n x m) += L(n x k) R(k x m): L = staged op(B) part at lo, R = staged
	// op(A)^T part at ro, both K-major with leading dimension panel_ld(k).
	auto seg = [&](int64_t lo, int64_t ro, int kk) {
		PSeg s{};
		s.loff  = lo;
		s.roff  = ro;
		s.lbase = PBASE_APACK;
		s.rbase = PBASE_APACK;
		s.ldl   = int32_t(panel_ld(kk));
		s.ldr   = int32_t(panel_ld(kk));
		s.k     = kk;
		segs.push_back(s);
	};
	// Task C^T (n x m) written by rows of length m: element (j, i) at coff + i + j m, i.e. C
	// (m x n) column major at coff.
	auto task = [&](int64_t coff, int mm, int nn, int kk, int sb, int se) {
		PTask t{};
		t.coff      = coff;
		t.m         = nn;
		t.n         = mm;
		t.rsplit    = nn;
		t.sm        = mm;
		t.seg_begin = sb;
		t.seg_end   = se;
		t.out_is_y  = 1;
		tasks.push_back(t);
		cost.push_back(double(mm) * nn * std::max(1, kk) * std::max(1, se - sb));
	};
	for (const Entry& e : ents) {
		const int64_t mn   = int64_t(e.m) * e.n;
		const bool    mult = e.k > 0 && !alpha_zero;
		// op(A)^T is staged (k x m): transposed unless op(A) already transposes; op(B) (k x n)
		// as it is. Conjugation negates the imaginary part.
		const int fa = (ta ? 0 : VB_TRANS) | cx;
		const int fb = (tb ? VB_TRANS : 0) | cx;
		if (!cplx) {
			int64_t ao = 0, bo = 0;
			if (mult) {
				ao = wsz;
				wsz += stage(copies, e.a, e.lda, e.k, e.m, wsz, fa);
				bo = wsz;
				wsz += stage(copies, e.b, e.ldb, e.k, e.n, wsz, fb);
			}
			const int sb = int(segs.size());
			if (mult)
				seg(bo, ao, e.k);
			task(csz, e.m, e.n, e.k, sb, int(segs.size()));
			VbOut o{};
			o.dst    = e.c;
			o.src_re = csz;
			o.src_im = -1;
			o.rows   = e.m;
			o.ldc    = e.ldc;
			o.cplx   = 0;
			for (int64_t s = 0; s < mn; s += kChunk) {
				o.start = s;
				o.len   = std::min(kChunk, mn - s);
				outs.push_back(o);
			}
			csz += mn;
		} else {
			// Ar, Ai, -Ai (as op(A)^T, k x m), then Br, Bi (k x n); conj of A or B flips the
			// imaginary sign.
			const int na = (ta == 2) ? VB_NEG : 0;
			const int nb = (tb == 2) ? VB_NEG : 0;
			int64_t   ar = 0, ai = 0, nai = 0, br = 0, bi = 0;
			if (mult) {
				ar = wsz;
				wsz += stage(copies, e.a, e.lda, e.k, e.m, wsz, fa);
				ai = wsz;
				wsz += stage(copies, e.a, e.lda, e.k, e.m, wsz, fa | VB_IMAG | na);
				nai = wsz;
				wsz += stage(copies, e.a, e.lda, e.k, e.m, wsz, fa | VB_IMAG | (na ^ VB_NEG));
				br = wsz;
				wsz += stage(copies, e.b, e.ldb, e.k, e.n, wsz, fb);
				bi = wsz;
				wsz += stage(copies, e.b, e.ldb, e.k, e.n, wsz, fb | VB_IMAG | nb);
			}
			const int64_t cr = csz, ci = csz + mn;
			int           sb = int(segs.size());
			if (mult) {
				seg(br, ar, e.k);
				seg(bi, nai, e.k);
			}
			task(cr, e.m, e.n, e.k, sb, int(segs.size()));
			sb = int(segs.size());
			if (mult) {
				seg(bi, ar, e.k);
				seg(br, ai, e.k);
			}
			task(ci, e.m, e.n, e.k, sb, int(segs.size()));
			VbOut o{};
			o.dst    = e.c;
			o.src_re = cr;
			o.src_im = ci;
			o.rows   = e.m;
			o.ldc    = e.ldc;
			o.cplx   = 1;
			for (int64_t s = 0; s < mn; s += kChunk) {
				o.start = s;
				o.len   = std::min(kChunk, mn - s);
				outs.push_back(o);
			}
			csz += 2 * mn;
		}
	}
	// 64 x 64 output tiles, largest tasks first.
	std::vector<int> order(tasks.size());
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[size_t(x)] > cost[size_t(y)]; });
	std::vector<Tile> tiles;
	for (int t : order)
		for (int m0 = 0; m0 < std::max(tasks[size_t(t)].m, 1); m0 += 64)
			for (int n0 = 0; n0 < std::max(tasks[size_t(t)].n, 1); n0 += 64)
				tiles.push_back(Tile{ t, m0, n0, 0 });
	if (tiles.size() > size_t(0x7fffffff) || copies.size() > size_t(0x7fffffff)
	    || outs.size() > size_t(0x7fffffff))
		throw std::runtime_error("gemm_vbatch: batch too large for one launch");
	for (const PSeg& s : segs)
		if (!panel_ld_fits(s.ldl) || !panel_ld_fits(s.ldr))
			throw std::runtime_error("gemm_vbatch: a contraction is too long for the panel kernel");

	std::vector<void*> owned;
	try {
		double* ws = static_cast<double*>(pool_device_alloc(size_t(std::max<int64_t>(wsz, 1)) * sizeof(double)));
		owned.push_back(ws);
		double* cw = static_cast<double*>(pool_device_alloc(size_t(std::max<int64_t>(csz, 1)) * sizeof(double)));
		owned.push_back(cw);
		VbCopy* d_copies = upload(copies, owned);
		VbOut*  d_outs   = upload(outs, owned);
		PTask*  d_tasks  = upload(tasks, owned);
		PSeg*   d_segs   = upload(segs, owned);
		Tile*   d_tiles  = upload(tiles, owned);
		hipStream_t s = nullptr;
		launch_vb_copy(d_copies, int(copies.size()), ws, s);
		const double*  bases[PBASE_COUNT] = { ws, ws, ws, ws, ws };
		const PStrides vs { 0, 0, 0 };
		launch_panel_gemm<double, double>(64, 64, 0, 2, 1, d_tasks, d_segs, d_tiles,
		                                  int(tiles.size()), bases, vs, cw, 0, nullptr, 0, 1, 0, s);
		DMK_HIP_CHECK(hipGetLastError());
		launch_vb_scatter(d_outs, int(outs.size()), cw, alpha[0], cplx ? alpha[1] : 0.0, beta[0],
		                  cplx ? beta[1] : 0.0, beta_zero ? 1 : 0, s);
		DMK_HIP_CHECK(hipStreamSynchronize(s));
	} catch (...) {
		(void)hipDeviceSynchronize();
		for (void* p : owned)
			pool_device_free(p);
		throw;
	}
	for (void* p : owned)
		pool_device_free(p);
}

} // namespace dmk
