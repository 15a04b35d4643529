// The matvec facade of a superblock sector (MatrixVectorKron, Engine/MatrixVectorKron/
// MatrixVectorKron.h:113-166, as MatrixVector::create selects it, Engine/MatrixVector.hpp:
// 133-150): x += op(y), where op is H, or L^dagger L when H is given as its factor L (the
// LdaggerL option); with a superblock rank of at most max_rank_stored H is stored explicitly
// (CSR, one sparse product per call; the reference's matrixStored_ from fullHamiltonian),
// otherwise the Kron plugin computes it. The adjoint product L^dagger (.) of the Kron path, which
// the reference runs on the CPU Kron path (KronMatrix::matrixVectorProductAdjoint,
// KronMatrix.h:130-133; the batched plugin only serves the forward product, :142), runs here on
// a second plugin handle built from the adjoint blocks: block (J, I, k) of H^dagger is
// A_k(I,J)^H (x) B_k(I,J)^H. Every call is timed and accumulated (the reference's DeltaClock
// matrixVectorProduct, MatrixVectorKron.h:139-166).
#include "dmrgpp_amd.h"

#include "dmk_internal.h"
#include "kron_plan.h"
#include "pool.h"

#include <chrono>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace dmk {

void set_last_error(const std::string& msg);
void launch_csr_apply(const int64_t* rowptr, const int32_t* col, const double* val, int64_t rows,
                      int cplx, const double* y, double* x, hipStream_t s);
void launch_axpy_one(double* x, const double* y, int64_t n, hipStream_t s);

namespace {

constexpr int64_t kMaxStoredRank = 16384;

struct Csr {
	int64_t  rows    = 0;
	int64_t  nnz     = 0;
	int64_t* rowptr  = nullptr; // device
	int32_t* col     = nullptr;
	double*  val     = nullptr; // nnz doubles, or 2 nnz (interleaved complex)
	void     release()
	{
		pool_device_release(rowptr);
		pool_device_release(col);
		pool_device_release(val);
	}
};

using cplx = std::complex<double>;

// The superblock problem as the plugin constructor receives it (BatchedGemm.h:21-28).
struct Problem {
	size_t               nop = 0, np = 0;
	const size_t*        L   = nullptr;
	const size_t*        R   = nullptr;
	const double* const* A   = nullptr;
	const size_t*        ldA = nullptr;
	const double* const* B   = nullptr;
	const size_t*        ldB = nullptr;
	bool                 complex = false;

	size_t idx(size_t I, size_t J, size_t k) const { return I + J * np + k * np * np; }
	// Element (r, c) of a column-major block (complex: interleaved doubles).
	cplx at(const double* p, size_t ld, size_t r, size_t c) const
	{
		if (!complex)
			return cplx(p[r + c * ld], 0.0);
		return cplx(p[2 * (r + c * ld)], p[2 * (r + c * ld) + 1]);
	}
};

// H (or its adjoint) assembled row by row into CSR on the host, then uploaded. Row (I, r, l) of
// H is off_I + r + l R_I (the patch vector layout, InitKronHamiltonian.h:211-216).
Csr assemble(const Problem& p, bool adjoint, hipStream_t s)
{
	std::vector<int64_t> off(p.np + 1, 0);
	for (size_t I = 0; I < p.np; ++I)
		off[I + 1] = off[I] + int64_t(p.L[I] * p.R[I]);
	const int64_t                      n = off[p.np];
	std::vector<std::map<int32_t, cplx>> rows(static_cast<size_t>(n));
	for (size_t k = 0; k < p.nop; ++k)
		for (size_t J = 0; J < p.np; ++J)
			for (size_t I = 0; I < p.np; ++I) {
				const size_t  q = p.idx(I, J, k);
				const double* a = p.A[q];
				const double* b = p.B[q];
				if (!a || !b)
					continue;
				for (size_t l = 0; l < p.L[I]; ++l)
					for (size_t lp = 0; lp < p.L[J]; ++lp) {
						const cplx av = p.at(a, p.ldA[q], l, lp);
						if (av == cplx(0.0))
							continue;
						for (size_t r = 0; r < p.R[I]; ++r)
							for (size_t rp = 0; rp < p.R[J]; ++rp) {
								const cplx bv = p.at(b, p.ldB[q], r, rp);
								if (bv == cplx(0.0))
									continue;
								const int64_t row = off[I] + int64_t(r + l * p.R[I]);
								const int64_t col = off[J] + int64_t(rp + lp * p.R[J]);
								if (adjoint)
									rows[size_t(col)][int32_t(row)] += std::conj(av * bv);
								else
									rows[size_t(row)][int32_t(col)] += av * bv;
							}
					}
			}
	std::vector<int64_t> rp(size_t(n) + 1, 0);
	for (int64_t i = 0; i < n; ++i)
		rp[size_t(i) + 1] = rp[size_t(i)] + int64_t(rows[size_t(i)].size());
	const int64_t        nnz = rp[size_t(n)];
	std::vector<int32_t> col(size_t(std::max<int64_t>(nnz, 1)));
	std::vector<double>  val(size_t(std::max<int64_t>(nnz, 1)) * (p.complex ? 2 : 1));
	for (int64_t i = 0; i < n; ++i) {
		int64_t e = rp[size_t(i)];
		for (const auto& kv : rows[size_t(i)]) {
			col[size_t(e)] = kv.first;
			if (p.complex) {
				val[size_t(2 * e)]     = kv.second.real();
				val[size_t(2 * e + 1)] = kv.second.imag();
			} else {
				val[size_t(e)] = kv.second.real();
			}
			++e;
		}
	}
	Csr c;
	c.rows = n;
	c.nnz  = nnz;
	try {
		c.rowptr = static_cast<int64_t*>(pool_device_alloc(rp.size() * sizeof(int64_t)));
		c.col    = static_cast<int32_t*>(pool_device_alloc(col.size() * sizeof(int32_t)));
		c.val    = static_cast<double*>(pool_device_alloc(val.size() * sizeof(double)));
		DMK_HIP_CHECK(hipMemcpyAsync(c.rowptr, rp.data(), rp.size() * sizeof(int64_t),
		                             hipMemcpyHostToDevice, s));
		DMK_HIP_CHECK(hipMemcpyAsync(c.col, col.data(), col.size() * sizeof(int32_t),
		                             hipMemcpyHostToDevice, s));
		DMK_HIP_CHECK(hipMemcpyAsync(c.val, val.data(), val.size() * sizeof(double),
		                             hipMemcpyHostToDevice, s));
		DMK_HIP_CHECK(hipStreamSynchronize(s));
	} catch (...) {
		(void)hipStreamSynchronize(s);
		c.release();
		throw;
	}
	return c;
}

} // namespace
} // namespace dmk

struct dmk_facade {
	bool        complex   = false;
	bool        ldaggerl  = false;
	bool        stored    = false;
	int64_t     n         = 0; // vector length in elements of T
	hipStream_t stream    = nullptr;
	dmk_handle* fwd       = nullptr; // Kron path
	dmk_handle* adj       = nullptr; // Kron path, L^dagger (LdaggerL)
	dmk::Csr    h, ht;               // stored path: H and H^dagger (LdaggerL)
	double*     tmp       = nullptr; // intermediate L y / H y (doubles: n or 2 n)
	double*     tmp2      = nullptr; // complex Kron path: the second product
	int64_t     matvecs   = 0;
	double      ms_total  = 0;

	~dmk_facade()
	{
		if (stream)
			(void)hipStreamSynchronize(stream);
		dmk_destroy(fwd);
		dmk_destroy(adj);
		h.release();
		ht.release();
		dmk::pool_device_release(tmp);
		dmk::pool_device_release(tmp2);
		if (stream)
			(void)hipStreamDestroy(stream);
	}
};

namespace {

template <typename F> int facade_guarded(F&& f)
{
	try {
		return f();
	} catch (const std::invalid_argument& e) {
		dmk::set_last_error(e.what());
		return DMK_ERR_ARGUMENT;
	} catch (const std::exception& e) {
		dmk::set_last_error(e.what());
		return DMK_ERR_DEVICE;
	}
}

int create(bool cplx, size_t noperator, size_t npatches, const size_t* L, const size_t* R,
           const double* const* A, const size_t* ldA, const double* const* B, const size_t* ldB,
           int64_t max_rank_stored, int ldaggerl, dmk_facade** out)
{
	return facade_guarded([&] {
		if (!out)
			throw std::invalid_argument("dmk_facade_create: null output");
		*out = nullptr;
		if (!L || !R || !A || !ldA || !B || !ldB || npatches == 0)
			throw std::invalid_argument("dmk_facade_create: null or empty problem");
		if (max_rank_stored < 0)
			throw std::invalid_argument("dmk_facade_create: max_rank_stored must be >= 0");
		// The stored matrix and the adjoint blocks are assembled on the host.
		if (dmk::blocks_on_device(int64_t(noperator), int64_t(npatches), A, B))
			throw std::invalid_argument("dmk_facade_create: the blocks must be host memory");
		auto f      = std::make_unique<dmk_facade>();
		f->complex  = cplx;
		f->ldaggerl = ldaggerl != 0;
		for (size_t I = 0; I < npatches; ++I)
			f->n += int64_t(L[I] * R[I]);
		const size_t per = cplx ? 2 : 1; // doubles per element
		dmk::Problem p { noperator, npatches, L, R, A, ldA, B, ldB, cplx };
		// MatrixVectorKron.h:123-125: stored when the rank is at most MaxMatrixRankStored.
		f->stored = f->n <= max_rank_stored;
		if (f->stored) {
			if (f->n > dmk::kMaxStoredRank)
				throw std::invalid_argument("dmk_facade_create: a stored H is limited to rank "
				                            + std::to_string(dmk::kMaxStoredRank));
			DMK_HIP_CHECK(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
			f->h = dmk::assemble(p, false, f->stream);
			if (f->ldaggerl)
				f->ht = dmk::assemble(p, true, f->stream);
		} else {
			auto mk = cplx ? dmk_create_z : dmk_create_d;
			if (mk(noperator, npatches, L, R, A, ldA, B, ldB, &f->fwd) != DMK_OK)
				throw std::runtime_error(std::string("dmk_facade_create: ") + dmk_last_error());
			if (f->ldaggerl) {
				// The adjoint problem: block (J, I, k) = (A_k(I,J)^H, B_k(I,J)^H), dense copies.
				const size_t           np = npatches, nb = noperator * np * np;
				std::vector<std::vector<double>> store;
				std::vector<const double*>     aa(nb, nullptr), bb(nb, nullptr);
				std::vector<size_t>            la(nb, 1), lb(nb, 1);
				auto adjoint_of = [&](const double* src, size_t ld, size_t rows, size_t cols) {
					std::vector<double> d(rows * cols * per);
					for (size_t c = 0; c < cols; ++c)
						for (size_t r = 0; r < rows; ++r) {
							// (r, c) of the source becomes (c, r) of the adjoint (cols x rows).
							const size_t s = (r + c * ld) * per, t = (c + r * cols) * per;
							d[t] = src[s];
							if (cplx)
								d[t + 1] = -src[s + 1];
						}
					store.push_back(std::move(d));
					return static_cast<const double*>(store.back().data());
				};
				store.reserve(2 * nb);
				for (size_t k = 0; k < noperator; ++k)
					for (size_t J = 0; J < np; ++J)
						for (size_t I = 0; I < np; ++I) {
							const size_t q = p.idx(I, J, k), qa = p.idx(J, I, k);
							if (!A[q] || !B[q])
								continue;
							aa[qa] = adjoint_of(A[q], ldA[q], L[I], L[J]);
							la[qa] = L[J];
							bb[qa] = adjoint_of(B[q], ldB[q], R[I], R[J]);
							lb[qa] = R[J];
						}
				if (mk(noperator, npatches, L, R, aa.data(), la.data(), bb.data(), lb.data(),
				       &f->adj)
				    != DMK_OK)
					throw std::runtime_error(std::string("dmk_facade_create (adjoint): ")
					                         + dmk_last_error());
			}
		}
		// Intermediate vectors: L y of LdaggerL, and the complex Kron path's products (its
		// plugin calls overwrite their output).
		const size_t vbytes = size_t(f->n) * per * sizeof(double);
		if (f->ldaggerl || (cplx && !f->stored))
			f->tmp = static_cast<double*>(dmk::pool_device_alloc(vbytes));
		if (cplx && !f->stored && f->ldaggerl)
			f->tmp2 = static_cast<double*>(dmk::pool_device_alloc(vbytes));
		*out = f.release();
		return DMK_OK;
	});
}

} // namespace

extern "C" {

int dmk_facade_create_d(size_t noperator, size_t npatches, const size_t* left_patch_size,
                        const size_t* right_patch_size, const double* const* Amatrix,
                        const size_t* ld_Amatrix, const double* const* Bmatrix,
                        const size_t* ld_Bmatrix, int64_t max_rank_stored, int ldaggerl,
                        dmk_facade** facade)
{
	return create(false, noperator, npatches, left_patch_size, right_patch_size, Amatrix,
	              ld_Amatrix, Bmatrix, ld_Bmatrix, max_rank_stored, ldaggerl, facade);
}

int dmk_facade_create_z(size_t noperator, size_t npatches, const size_t* left_patch_size,
                        const size_t* right_patch_size, const double* const* Amatrix,
                        const size_t* ld_Amatrix, const double* const* Bmatrix,
                        const size_t* ld_Bmatrix, int64_t max_rank_stored, int ldaggerl,
                        dmk_facade** facade)
{
	return create(true, noperator, npatches, left_patch_size, right_patch_size, Amatrix,
	              ld_Amatrix, Bmatrix, ld_Bmatrix, max_rank_stored, ldaggerl, facade);
}

void dmk_facade_destroy(dmk_facade* facade) { delete facade; }

int dmk_facade_matvec(dmk_facade* f, double* x, const double* y)
{
	return facade_guarded([&] {
		if (!f || !x || !y)
			throw std::invalid_argument("dmk_facade_matvec: null argument");
		const auto t0  = std::chrono::steady_clock::now();
		const int    c     = f->complex ? 1 : 0;
		const size_t bytes = size_t(f->n) * (f->complex ? 2 : 1) * sizeof(double);
		if (f->stored) {
			hipStream_t s = f->stream;
			if (!f->ldaggerl) {
				dmk::launch_csr_apply(f->h.rowptr, f->h.col, f->h.val, f->h.rows, c, y, x, s);
			} else {
				// intermediate = L y (zeroed, then accumulated), x += L^dagger intermediate.
				DMK_HIP_CHECK(hipMemsetAsync(f->tmp, 0, bytes, s));
				dmk::launch_csr_apply(f->h.rowptr, f->h.col, f->h.val, f->h.rows, c, y, f->tmp, s);
				dmk::launch_csr_apply(f->ht.rowptr, f->ht.col, f->ht.val, f->ht.rows, c, f->tmp,
				                      x, s);
			}
			DMK_HIP_CHECK(hipStreamSynchronize(s));
		} else {
			auto check = [](int st) {
				if (st != DMK_OK)
					throw std::runtime_error(std::string("dmk_facade_matvec: ") + dmk_last_error());
			};
			const size_t n = size_t(f->n);
			if (!f->complex) {
				// The plugin accumulates with beta = 1 (x += H y).
				if (!f->ldaggerl) {
					check(dmk_apply_multi_d(f->fwd, 1, y, n, x, n, 1));
					check(dmk_sync(f->fwd));
				} else {
					check(dmk_apply_d(f->fwd, y, f->tmp));
					check(dmk_sync(f->fwd));
					check(dmk_apply_multi_d(f->adj, 1, f->tmp, n, x, n, 1));
					check(dmk_sync(f->adj));
				}
			} else {
				// Complex products overwrite their output: into tmp (and tmp2), then x += it.
				check(dmk_apply_z(f->fwd, y, f->tmp));
				check(dmk_sync(f->fwd));
				const double* last = f->tmp;
				if (f->ldaggerl) {
					check(dmk_apply_z(f->adj, f->tmp, f->tmp2));
					check(dmk_sync(f->adj));
					last = f->tmp2;
				}
				dmk::launch_axpy_one(x, last, int64_t(2 * n), nullptr);
				DMK_HIP_CHECK(hipStreamSynchronize(nullptr));
			}
		}
		f->matvecs++;
		f->ms_total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
		return DMK_OK;
	});
}

int dmk_facade_get_info(const dmk_facade* f, dmk_facade_info* info)
{
	return facade_guarded([&] {
		if (!f || !info)
			throw std::invalid_argument("dmk_facade_get_info: null argument");
		std::memset(info, 0, sizeof(*info));
		info->n        = f->n;
		info->stored   = f->stored ? 1 : 0;
		info->ldaggerl = f->ldaggerl ? 1 : 0;
		info->is_complex = f->complex ? 1 : 0;
		info->nnz      = f->stored ? f->h.nnz : 0;
		info->matvecs  = f->matvecs;
		info->ms_matvec_total = f->ms_total;
		return DMK_OK;
	});
}

} // extern "C"
