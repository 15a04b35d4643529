// Host planner of the superblock Kron matvec.
//
// Input: the BatchedGemm plugin arguments (GPUPlugin/BatchedGemm.h:21-28): noperator, npatches,
// left/right patch sizes (indexed by patch), and for every (I, J, k) a column-major block
// A_k(I,J) (L_I x L_J) and B_k(I,J) (R_I x R_J) at index I + J*np + k*np*np, or nullptr.
//
// What it does, against the reference plugin:
//  * nonzero scan per block, NULL meaning zero, a term counting when both factors are nonzero
//    (setup_nC.cpp:47-124);
//  * packs the factors once into HBM-resident buffers (setup_sparse_batch.cpp:149-343 packs into
//    managed memory): A panels per output patch in (J outer, k inner) order as the reference,
//    B blocks row-stacked per (I, J) so stage 1 is one GEMM per (I, J) instead of one per term;
//  * detects identity factors (the H_L (x) I and I (x) H_R connections of
//    InitKronHamiltonian.h:242-256) and contracts them directly in stage 2;
//  * builds the two grouped-GEMM launches once; every later matvec only launches them (the
//    reference rebuilds its descriptor arrays and BX scratch on every call,
//    apply_Htarget_sparse_not_ready.cpp:99-276).
#include "kron_plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>

namespace dmk {

template <typename T, typename TO>
void launch_grouped_gemm(const Task* tasks, const Seg* segs, const Tile* tiles, int ntiles,
                         const T* const* bases, int64_t xstride, int64_t bxstride, TO* out,
                         int64_t ostride, int nvec, int beta, hipStream_t stream);
void launch_cast_f64_to_f32(const double* x, float* y, int64_t n, hipStream_t stream);

namespace {

constexpr int kTileM = 64;
constexpr int kTileN = 64;

int64_t count_nnz(const double* p, int64_t rows, int64_t cols, int64_t ld)
{
	if (!p)
		return 0;
	int64_t nnz = 0;
	for (int64_t j = 0; j < cols; ++j) {
		const double* c = p + j * ld;
		for (int64_t i = 0; i < rows; ++i)
			nnz += (c[i] != 0.0);
	}
	return nnz;
}

bool is_identity(const double* p, int64_t rows, int64_t cols, int64_t ld)
{
	if (!p || rows != cols)
		return false;
	for (int64_t j = 0; j < cols; ++j) {
		const double* c = p + j * ld;
		for (int64_t i = 0; i < rows; ++i)
			if (c[i] != (i == j ? 1.0 : 0.0))
				return false;
	}
	return true;
}

template <typename X> X* dev_alloc_copy(const std::vector<X>& h)
{
	if (h.empty())
		return nullptr;
	X* d = nullptr;
	DMK_HIP_CHECK(hipMalloc(&d, h.size() * sizeof(X)));
	DMK_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(X), hipMemcpyHostToDevice));
	return d;
}

// Orders tasks by decreasing cost and lists their 64x64 tiles in that order (a flop-sorted
// work list: the dispatcher hands out workgroups in blockIdx order, so the largest tiles
// start first and the small ones fill the tail).
void build_tiles(StageTables& st, const std::vector<double>& cost)
{
	std::vector<int> order(st.tasks.size());
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
	st.tiles.clear();
	for (int t : order) {
		const Task& tk = st.tasks[t];
		for (int m0 = 0; m0 < std::max(tk.m, 1); m0 += kTileM)
			for (int n0 = 0; n0 < std::max(tk.n, 1); n0 += kTileN)
				st.tiles.push_back(Tile { t, m0, n0, 0 });
	}
}

} // namespace

void StageTables::upload()
{
	d_tasks = dev_alloc_copy(tasks);
	d_segs  = dev_alloc_copy(segs);
	d_tiles = dev_alloc_copy(tiles);
}

void StageTables::release()
{
	(void)hipFree(d_tasks);
	(void)hipFree(d_segs);
	(void)hipFree(d_tiles);
	d_tasks = nullptr;
	d_segs  = nullptr;
	d_tiles = nullptr;
}

Plan::Plan(int64_t noperator, int64_t npatches, const size_t* Lsz, const size_t* Rsz,
           const double* const* Amatrix, const size_t* ldA, const double* const* Bmatrix,
           const size_t* ldB)
{
	if (npatches <= 0 || noperator <= 0)
		throw std::runtime_error("dmk_create: npatches and noperator must be positive");
	if (!Lsz || !Rsz || !Amatrix || !Bmatrix || !ldA || !ldB)
		throw std::runtime_error("dmk_create: null argument array");
	const int64_t np = npatches;
	const int64_t nC = noperator;
	std::vector<int64_t> L(np), R(np);
	for (int64_t I = 0; I < np; ++I) {
		L[I] = int64_t(Lsz[I]);
		R[I] = int64_t(Rsz[I]);
		if (L[I] < 1 || R[I] < 1)
			throw std::runtime_error("dmk_create: patch sizes must be >= 1 (patch "
			                         + std::to_string(I) + ")");
		if (L[I] > (int64_t(1) << 30) || R[I] > (int64_t(1) << 30))
			throw std::runtime_error("dmk_create: patch size too large");
	}
	off_.assign(np + 1, 0);
	for (int64_t I = 0; I < np; ++I)
		off_[I + 1] = off_[I] + L[I] * R[I];
	stats_.n         = off_[np];
	stats_.npatches  = np;
	stats_.noperator = nC;
	for (int64_t I = 0; I < np; ++I) {
		stats_.max_patch_left  = std::max(stats_.max_patch_left, L[I]);
		stats_.max_patch_right = std::max(stats_.max_patch_right, R[I]);
	}

	auto idx = [&](int64_t I, int64_t J, int64_t k) { return size_t(I + J * np + k * np * np); };

	// Term classification (setup_nC semantics + identity detection).
	enum Kind : char
	{
		NONE = 0,
		GENERAL,
		IDENT_B, // B is the identity: Y_I += X_J * A^T
		IDENT_A, // A is the identity: Y_I += B * X_J
	};
	std::vector<char> kind(size_t(np * np * nC), NONE);
	for (int64_t k = 0; k < nC; ++k)
		for (int64_t J = 0; J < np; ++J)
			for (int64_t I = 0; I < np; ++I) {
				const size_t  id = idx(I, J, k);
				const double* a  = Amatrix[id];
				const double* b  = Bmatrix[id];
				if (!a || !b)
					continue;
				if (int64_t(ldA[id]) < L[I] || int64_t(ldB[id]) < R[I])
					throw std::runtime_error("dmk_create: leading dimension smaller than rows");
				const int64_t nnzA = count_nnz(a, L[I], L[J], int64_t(ldA[id]));
				const int64_t nnzB = count_nnz(b, R[I], R[J], int64_t(ldB[id]));
				if (nnzA == 0 || nnzB == 0)
					continue;
				stats_.nterms++;
				const double RI = double(R[I]), RJ = double(R[J]), LI = double(L[I]),
				             LJ = double(L[J]);
				stats_.flops_plugin += 2.0 * RI * RJ * LJ + 2.0 * RI * LJ * LI;
				if (is_identity(b, R[I], R[J], int64_t(ldB[id]))) {
					kind[id] = IDENT_B;
					stats_.nterms_identity++;
					stats_.flops_algorithmic += 2.0 * RI * LI * LJ;
					stats_.bytes_compulsory += 8.0 * LI * LJ;
				} else if (is_identity(a, L[I], L[J], int64_t(ldA[id]))) {
					kind[id] = IDENT_A;
					stats_.nterms_identity++;
					stats_.flops_algorithmic += 2.0 * RI * RJ * LI;
					stats_.bytes_compulsory += 8.0 * RI * RJ;
				} else {
					kind[id] = GENERAL;
					const double f1 = 2.0 * RI * RJ * LJ + 2.0 * RI * LJ * LI;
					const double f2 = 2.0 * RJ * LJ * LI + 2.0 * RI * RJ * LI;
					stats_.flops_algorithmic += std::min(f1, f2);
					stats_.bytes_compulsory += 8.0 * (LI * LJ + RI * RJ);
				}
			}
	stats_.bytes_compulsory += 8.0 * 2.0 * double(stats_.n);

	// Packing. A panel of patch I: L_I x K_I, (J outer, k inner) general terms; then the
	// identity-B terms' A blocks. B stack of (I, J): q*R_I x R_J for its q general terms;
	// then identity-A terms' B blocks.
	std::vector<double> apack, bpack;
	std::vector<int64_t> apanel_off(np, 0), K(np, 0), bx_off(np + 1, 0);
	for (int64_t I = 0; I < np; ++I) {
		for (int64_t J = 0; J < np; ++J)
			for (int64_t k = 0; k < nC; ++k)
				if (kind[idx(I, J, k)] == GENERAL)
					K[I] += L[J];
		bx_off[I + 1] = bx_off[I] + R[I] * K[I];
	}
	stats_.bx_elems = bx_off[np];

	st1_.tasks.clear();
	st1_.segs.clear();
	st2_.tasks.clear();
	st2_.segs.clear();
	std::vector<double> cost1, cost2;

	for (int64_t I = 0; I < np; ++I) {
		// A panel for the stage-2 GEMM of patch I.
		apanel_off[I] = int64_t(apack.size());
		apack.resize(apack.size() + size_t(L[I] * K[I]));
		int64_t col = 0;
		for (int64_t J = 0; J < np; ++J) {
			// Stage-1 task for (I, J): the row-stacked B blocks of its general terms.
			std::vector<int64_t> ks;
			for (int64_t k = 0; k < nC; ++k)
				if (kind[idx(I, J, k)] == GENERAL)
					ks.push_back(k);
			if (ks.empty())
				continue;
			const int64_t q       = int64_t(ks.size());
			const int64_t boff    = int64_t(bpack.size());
			const int64_t ldstack = q * R[I];
			bpack.resize(bpack.size() + size_t(ldstack * R[J]));
			for (int64_t t = 0; t < q; ++t) {
				const size_t  id = idx(I, J, ks[t]);
				const double* a  = Amatrix[id];
				const double* b  = Bmatrix[id];
				double*       ad = apack.data() + apanel_off[I] + (col + t * L[J]) * L[I];
				for (int64_t c = 0; c < L[J]; ++c)
					std::memcpy(ad + c * L[I], a + c * int64_t(ldA[id]), size_t(L[I]) * sizeof(double));
				double* bd = bpack.data() + boff + t * R[I];
				for (int64_t c = 0; c < R[J]; ++c)
					std::memcpy(bd + c * ldstack, b + c * int64_t(ldB[id]),
					            size_t(R[I]) * sizeof(double));
			}
			Task tk {};
			tk.coff      = bx_off[I] + col * R[I];
			tk.m         = int32_t(q * R[I]);
			tk.n         = int32_t(L[J]);
			tk.ldc       = int32_t(R[I]);
			tk.rsplit    = int32_t(R[I]);
			tk.seg_begin = int32_t(st1_.segs.size());
			Seg s {};
			s.lbase = BASE_BPACK;
			s.loff  = boff;
			s.ldl   = int32_t(ldstack);
			s.rbase = BASE_X;
			s.roff  = off_[J];
			s.ldr   = int32_t(R[J]);
			s.k     = int32_t(R[J]);
			s.rtrans = 0;
			st1_.segs.push_back(s);
			tk.seg_end  = int32_t(st1_.segs.size());
			tk.out_is_y = 0;
			st1_.tasks.push_back(tk);
			const double f = 2.0 * double(q * R[I]) * double(R[J]) * double(L[J]);
			cost1.push_back(f);
			stats_.flops_executed += f;
			stats_.flops_stage1 += f;
			col += q * L[J];
		}

		// Stage-2 task for patch I.
		Task tk {};
		tk.coff      = off_[I];
		tk.m         = int32_t(R[I]);
		tk.n         = int32_t(L[I]);
		tk.ldc       = int32_t(R[I]);
		tk.rsplit    = int32_t(R[I]);
		tk.seg_begin = int32_t(st2_.segs.size());
		tk.out_is_y  = 1;
		double cost  = 0;
		if (K[I] > 0) {
			Seg s {};
			s.lbase  = BASE_BX;
			s.loff   = bx_off[I];
			s.ldl    = int32_t(R[I]);
			s.rbase  = BASE_APACK;
			s.roff   = apanel_off[I];
			s.ldr    = int32_t(L[I]);
			s.k      = int32_t(K[I]);
			s.rtrans = 1;
			st2_.segs.push_back(s);
			cost += 2.0 * double(R[I]) * double(L[I]) * double(K[I]);
			stats_.bytes_bx += 2.0 * 8.0 * double(R[I]) * double(K[I]);
		}
		for (int64_t J = 0; J < np; ++J)
			for (int64_t k = 0; k < nC; ++k) {
				const size_t id = idx(I, J, k);
				if (kind[id] == IDENT_B) {
					// Y_I += X_J * A^T: L = X_J (R_J x L_J), R = A (L_I x L_J) transposed.
					const int64_t aoff = int64_t(apack.size());
					apack.resize(apack.size() + size_t(L[I] * L[J]));
					const double* a = Amatrix[id];
					for (int64_t c = 0; c < L[J]; ++c)
						std::memcpy(apack.data() + aoff + c * L[I], a + c * int64_t(ldA[id]),
						            size_t(L[I]) * sizeof(double));
					Seg s {};
					s.lbase  = BASE_X;
					s.loff   = off_[J];
					s.ldl    = int32_t(R[J]);
					s.rbase  = BASE_APACK;
					s.roff   = aoff;
					s.ldr    = int32_t(L[I]);
					s.k      = int32_t(L[J]);
					s.rtrans = 1;
					st2_.segs.push_back(s);
					cost += 2.0 * double(R[I]) * double(L[I]) * double(L[J]);
				} else if (kind[id] == IDENT_A) {
					// Y_I += B * X_J: L = B (R_I x R_J), R = X_J (R_J x L_J).
					const int64_t boff = int64_t(bpack.size());
					bpack.resize(bpack.size() + size_t(R[I] * R[J]));
					const double* b = Bmatrix[id];
					for (int64_t c = 0; c < R[J]; ++c)
						std::memcpy(bpack.data() + boff + c * R[I], b + c * int64_t(ldB[id]),
						            size_t(R[I]) * sizeof(double));
					Seg s {};
					s.lbase  = BASE_BPACK;
					s.loff   = boff;
					s.ldl    = int32_t(R[I]);
					s.rbase  = BASE_X;
					s.roff   = off_[J];
					s.ldr    = int32_t(R[J]);
					s.k      = int32_t(R[J]);
					s.rtrans = 0;
					st2_.segs.push_back(s);
					cost += 2.0 * double(R[I]) * double(R[J]) * double(L[I]);
				}
			}
		tk.seg_end = int32_t(st2_.segs.size());
		st2_.tasks.push_back(tk);
		cost2.push_back(cost + 1.0); // empty tasks still write zeros
		stats_.flops_executed += cost;
		stats_.flops_stage2 += cost;
	}

	build_tiles(st1_, cost1);
	build_tiles(st2_, cost2);
	stats_.stage1_tasks = int64_t(st1_.tasks.size());
	stats_.stage2_tasks = int64_t(st2_.tasks.size());
	stats_.stage1_tiles = int64_t(st1_.tiles.size());
	stats_.stage2_tiles = int64_t(st2_.tiles.size());
	stats_.apack_elems  = int64_t(apack.size());
	stats_.bpack_elems  = int64_t(bpack.size());
	if (stats_.n > (int64_t(1) << 31) * 64)
		throw std::runtime_error("dmk_create: superblock too large");

	DMK_HIP_CHECK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
	for (auto& e : ev_)
		DMK_HIP_CHECK(hipEventCreate(&e));
	d_apack_ = dev_alloc_copy(apack);
	d_bpack_ = dev_alloc_copy(bpack);
	st1_.upload();
	st2_.upload();
	ensure_scratch(1);
}

Plan::~Plan()
{
	if (stream_)
		(void)hipStreamSynchronize(stream_);
	(void)hipFree(d_apack_);
	(void)hipFree(d_bpack_);
	(void)hipFree(d_apack32_);
	(void)hipFree(d_bpack32_);
	(void)hipFree(d_bx_);
	(void)hipFree(d_x32_);
	st1_.release();
	st2_.release();
	for (auto& e : ev_)
		if (e)
			(void)hipEventDestroy(e);
	if (stream_)
		(void)hipStreamDestroy(stream_);
}

void Plan::ensure_scratch(int nvec)
{
	const size_t elem  = (precision_ == PREC_FP64) ? sizeof(double) : sizeof(float);
	const size_t bytes = size_t(std::max<int64_t>(stats_.bx_elems, 1)) * elem * size_t(nvec);
	if (bytes > bx_bytes_) {
		DMK_HIP_CHECK(hipStreamSynchronize(stream_));
		(void)hipFree(d_bx_);
		d_bx_     = nullptr;
		bx_bytes_ = 0;
		DMK_HIP_CHECK(hipMalloc(&d_bx_, bytes));
		bx_bytes_ = bytes;
	}
	if (precision_ != PREC_FP64 && nvec > x32_cap_) {
		DMK_HIP_CHECK(hipStreamSynchronize(stream_));
		(void)hipFree(d_x32_);
		d_x32_   = nullptr;
		x32_cap_ = 0;
		DMK_HIP_CHECK(hipMalloc(&d_x32_, size_t(std::max<int64_t>(stats_.n, 1)) * sizeof(float)
		                                      * size_t(nvec)));
		x32_cap_ = nvec;
	}
}

void Plan::ensure_fp32()
{
	if (d_apack32_ || d_bpack32_)
		return;
	auto convert = [&](const double* src, int64_t n, float** dst) {
		if (n <= 0)
			return;
		DMK_HIP_CHECK(hipMalloc(dst, size_t(n) * sizeof(float)));
		launch_cast_f64_to_f32(src, *dst, n, stream_);
	};
	convert(d_apack_, stats_.apack_elems, &d_apack32_);
	convert(d_bpack_, stats_.bpack_elems, &d_bpack32_);
	DMK_HIP_CHECK(hipStreamSynchronize(stream_));
}

void Plan::set_precision(int p)
{
	if (p != PREC_FP64 && p != PREC_FP32)
		throw std::runtime_error("dmk_set_precision: unsupported precision mode "
		                         + std::to_string(p));
	precision_ = p;
	if (p == PREC_FP32)
		ensure_fp32();
	ensure_scratch(1);
}

void Plan::apply(const double* vin, int64_t ldin, double* vout, int64_t ldout, int nvec, int beta)
{
	if (nvec < 1)
		throw std::runtime_error("dmk_apply: nvec must be >= 1");
	if (!vin || !vout)
		throw std::runtime_error("dmk_apply: null vector");
	if (nvec > 1 && (ldin < stats_.n || ldout < stats_.n))
		throw std::runtime_error("dmk_apply: vector stride smaller than n");
	ensure_scratch(nvec);
	const int64_t bxs = stats_.bx_elems;
	if (precision_ == PREC_FP64) {
		const double* bases[BASE_COUNT] = { d_apack_, d_bpack_, vin, static_cast<double*>(d_bx_) };
		launch_grouped_gemm<double, double>(st1_.d_tasks, st1_.d_segs, st1_.d_tiles,
		                                    int(st1_.tiles.size()), bases, ldin, bxs,
		                                    static_cast<double*>(d_bx_), bxs, nvec, 0, stream_);
		launch_grouped_gemm<double, double>(st2_.d_tasks, st2_.d_segs, st2_.d_tiles,
		                                    int(st2_.tiles.size()), bases, ldin, bxs, vout, ldout,
		                                    nvec, beta, stream_);
	} else {
		for (int v = 0; v < nvec; ++v)
			launch_cast_f64_to_f32(vin + v * ldin, d_x32_ + v * stats_.n, stats_.n, stream_);
		const float* bases[BASE_COUNT] = { d_apack32_, d_bpack32_, d_x32_, static_cast<float*>(d_bx_) };
		launch_grouped_gemm<float, float>(st1_.d_tasks, st1_.d_segs, st1_.d_tiles,
		                                  int(st1_.tiles.size()), bases, stats_.n, bxs,
		                                  static_cast<float*>(d_bx_), bxs, nvec, 0, stream_);
		launch_grouped_gemm<float, double>(st2_.d_tasks, st2_.d_segs, st2_.d_tiles,
		                                   int(st2_.tiles.size()), bases, stats_.n, bxs, vout, ldout,
		                                   nvec, beta, stream_);
	}
	DMK_HIP_CHECK(hipGetLastError());
}

void Plan::profile(const double* vin, double* vout, int iters, double* ms1, double* ms2,
                   double* mstot)
{
	if (iters < 1)
		throw std::runtime_error("dmk_profile: iters must be >= 1");
	double s1 = 0, s2 = 0, tot = 0;
	ensure_scratch(1);
	const int64_t bxs = stats_.bx_elems;
	for (int it = 0; it < iters; ++it) {
		DMK_HIP_CHECK(hipEventRecord(ev_[0], stream_));
		if (precision_ == PREC_FP64) {
			const double* bases[BASE_COUNT] = { d_apack_, d_bpack_, vin, static_cast<double*>(d_bx_) };
			launch_grouped_gemm<double, double>(st1_.d_tasks, st1_.d_segs, st1_.d_tiles,
			                                    int(st1_.tiles.size()), bases, 0, bxs,
			                                    static_cast<double*>(d_bx_), 0, 1, 0, stream_);
			DMK_HIP_CHECK(hipEventRecord(ev_[1], stream_));
			launch_grouped_gemm<double, double>(st2_.d_tasks, st2_.d_segs, st2_.d_tiles,
			                                    int(st2_.tiles.size()), bases, 0, bxs, vout, 0, 1, 0,
			                                    stream_);
		} else {
			launch_cast_f64_to_f32(vin, d_x32_, stats_.n, stream_);
			const float* bases[BASE_COUNT] = { d_apack32_, d_bpack32_, d_x32_,
				                           static_cast<float*>(d_bx_) };
			launch_grouped_gemm<float, float>(st1_.d_tasks, st1_.d_segs, st1_.d_tiles,
			                                  int(st1_.tiles.size()), bases, 0, bxs,
			                                  static_cast<float*>(d_bx_), 0, 1, 0, stream_);
			DMK_HIP_CHECK(hipEventRecord(ev_[1], stream_));
			launch_grouped_gemm<float, double>(st2_.d_tasks, st2_.d_segs, st2_.d_tiles,
			                                   int(st2_.tiles.size()), bases, 0, bxs, vout, 0, 1, 0,
			                                   stream_);
		}
		DMK_HIP_CHECK(hipEventRecord(ev_[2], stream_));
		DMK_HIP_CHECK(hipEventSynchronize(ev_[2]));
		float a = 0, b = 0, c = 0;
		DMK_HIP_CHECK(hipEventElapsedTime(&a, ev_[0], ev_[1]));
		DMK_HIP_CHECK(hipEventElapsedTime(&b, ev_[1], ev_[2]));
		DMK_HIP_CHECK(hipEventElapsedTime(&c, ev_[0], ev_[2]));
		s1 += a;
		s2 += b;
		tot += c;
	}
	if (ms1)
		*ms1 = s1 / iters;
	if (ms2)
		*ms2 = s2 / iters;
	if (mstot)
		*mstot = tot / iters;
}

} // namespace dmk
