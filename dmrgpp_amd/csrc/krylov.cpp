// Batched Krylov time step (TimeVectorsKrylov::calcTimeVectors, TimeVectorsKrylov.h:165-322).
//
// The reference tridiagonalizes one vector phi per sector with the Lanczos recurrence
// (ParallelTriDiag.h:132-161), diagonalizes T = S diag(eps) S^T, and forms each target time as
//   r(tau) = V S diag(exp(-i (eps_k - E_g) tau d)) S^T e_0 |phi|
// (KrylovHelper::calcR with the default abridgement V^H phi = |phi| e_0, then two GEMVs). Here
// several independent vectors are propagated together: their recurrences run in lockstep and
// every Krylov step makes ONE batched matvec launch over all of them (each complex vector is two
// real columns of a real H, or one column of a complex H's real embedding). The Krylov vectors
// stay in HBM; per step only alpha and beta of each target come back to the host for the stop
// test. Each target keeps the Krylov dimension at which its own test passed, so its result is
// the one a separate run would give.
#include "krylov.h"

#include "pool.h"
#include "tridiag.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>

namespace dmk {

int  krylov_reduction_blocks();
int  krylov_max_times();
void launch_kry_dot(const double* x, const double* y, int64_t stride, int64_t len, int nvec,
                    double* partial, double* out, int ostride, int sqrt_inv, hipStream_t s);
void launch_kry_update_beta(double* w, const double* v1, const double* v0, const double* A,
                            double* B, double* HN, double* INV, int j, int cap, int64_t stride,
                            int64_t len, int nvec, double* partial, hipStream_t s);
void launch_kry_scale(double* x, const double* sc, int sstride, int64_t stride, int64_t len,
                      int nvec, hipStream_t s);
void launch_kry_combine(double* out, int64_t ldout, const double* const* vptr, int64_t vstride,
                        const double* coef, int m, int ntimes, int64_t n, int nvec, hipStream_t s);

void KrylovOperator::apply(const double* in, double* out, int nvec) const
{
	if (embedded)
		plan->apply(in, 2 * nc, out, 2 * nc, nvec, 0);
	else
		plan->apply(in, nc, out, nc, 2 * nvec, 0);
}

KrylovWorkspace::~KrylovWorkspace() { release(); }

void KrylovWorkspace::release()
{
	for (double* c : chunks_)
		pool_device_free(c);
	chunks_.clear();
	pool_device_free(partial_);
	pool_device_free(scal_);
	pool_device_free(coef_);
	pool_device_free(vptr_);
	pool_host_free(host_);
	partial_ = scal_ = coef_ = host_ = nullptr;
	vptr_                           = nullptr;
	nc_ = block_elems_ = 0;
	nvec_ = cols_ = ntimes_ = 0;
	host_elems_             = 0;
}

void KrylovWorkspace::reset(int64_t nc, int nvec, int columns, int ntimes)
{
	if (nc == nc_ && nvec == nvec_ && columns <= cols_ && ntimes <= ntimes_ && scal_)
		return;
	release();
	nc_          = nc;
	nvec_        = nvec;
	cols_        = columns;
	ntimes_      = ntimes;
	block_elems_ = 2 * nc * nvec;
	auto dev     = [](size_t elems) { return static_cast<double*>(pool_device_alloc(elems * sizeof(double))); };
	partial_     = dev(size_t(nvec) * size_t(krylov_reduction_blocks()));
	const size_t nscal = size_t(2 * nvec * columns + 4 * nvec);
	const size_t ncoef = size_t(2) * size_t(nvec) * size_t(ntimes) * size_t(columns);
	scal_        = dev(nscal);
	coef_        = dev(ncoef);
	vptr_        = static_cast<double**>(pool_device_alloc(size_t(columns) * sizeof(double*)));
	host_elems_  = nscal + ncoef + size_t(columns);
	host_        = static_cast<double*>(pool_host_alloc(host_elems_ * sizeof(double)));
}

double* KrylovWorkspace::block(int j)
{
	if (j < 0 || j >= cols_)
		throw std::runtime_error("krylov: basis column out of range");
	const size_t c = size_t(j / kChunk);
	while (chunks_.size() <= c)
		chunks_.push_back(static_cast<double*>(
		    pool_device_alloc(size_t(kChunk) * size_t(block_elems_) * sizeof(double))));
	return chunks_[c] + int64_t(j % kChunk) * block_elems_;
}

void KrylovWorkspace::upload_vptr(int m, hipStream_t s)
{
	const size_t nscal = size_t(2 * nvec_ * cols_ + 4 * nvec_);
	const size_t ncoef = size_t(2) * size_t(nvec_) * size_t(ntimes_) * size_t(cols_);
	double**     h     = reinterpret_cast<double**>(host_ + nscal + ncoef);
	for (int j = 0; j < m; ++j)
		h[j] = block(j);
	DMK_HIP_CHECK(hipMemcpyAsync(vptr_, h, size_t(m) * sizeof(double*), hipMemcpyHostToDevice, s));
}

namespace {

// Complex coefficients c_j = |phi| sum_k S(j,k) S(0,k) exp(-i (eps_k - E) tau d), j < m, written
// as (re, im) pairs (KrylovHelper.h calcR with abridgement, TimeVectorsKrylov.h:301-322).
void krylov_coefficients(const std::vector<double>& eps, const std::vector<double>& S, int m,
                         double norm, double tau, const KrylovParams& p, double* c)
{
	std::vector<double> wr(static_cast<size_t>(m)), wi(static_cast<size_t>(m));
	for (int k = 0; k < m; ++k) {
		const double th = (eps[size_t(k)] - p.energy_shift) * tau * p.time_direction;
		const double s0 = S[size_t(k) * size_t(m)]; // S(0, k)
		wr[size_t(k)]   = norm * s0 * std::cos(th);
		wi[size_t(k)]   = -norm * s0 * std::sin(th);
	}
	for (int j = 0; j < m; ++j) {
		double re = 0, im = 0;
		for (int k = 0; k < m; ++k) {
			const double sjk = S[size_t(j) + size_t(k) * size_t(m)];
			re += sjk * wr[size_t(k)];
			im += sjk * wi[size_t(k)];
		}
		c[2 * j]     = re;
		c[2 * j + 1] = im;
	}
}

struct TargetState {
	std::vector<double> a, b; // alpha_0.., beta (b[0] = 0)
	int                 m    = 0;     // Krylov dimension when done
	bool                done = false;
	double              eold = 0;
	KrylovReport        rep;
};

} // namespace

std::vector<KrylovReport> krylov_evolve(const KrylovOperator& op, KrylovWorkspace& ws, int nvec,
                                        const double* phi, int64_t ldphi, int ntimes,
                                        const double* times, double* out, int64_t ldout,
                                        const KrylovParams& p, double* ms_total, double* ms_matvec,
                                        int* matvec_launches)
{
	const auto t0 = std::chrono::steady_clock::now();
	if (!op.plan)
		throw std::invalid_argument("krylov: null plan");
	if (nvec < 1 || ntimes < 1 || ntimes > krylov_max_times())
		throw std::invalid_argument("krylov: nvec must be >= 1 and ntimes in [1, "
		                            + std::to_string(krylov_max_times()) + "]");
	if (!phi || !out || !times)
		throw std::invalid_argument("krylov: null argument");
	if (p.max_steps < 1 || p.min_steps < 0 || !(p.tolerance >= 0) || (p.stop != 0 && p.stop != 1))
		throw std::invalid_argument("krylov: bad parameters");
	if (p.precision != PREC_FP64 && p.precision != PREC_FP32 && p.precision != PREC_BF16X3)
		throw std::runtime_error("krylov: unsupported precision mode");
	const int64_t nc  = op.nc;
	const int64_t len = 2 * nc; // split complex length
	if (ldphi < nc || ldout < nc)
		throw std::invalid_argument("krylov: leading dimension smaller than the vector length");
	// At most nc steps: the complex Krylov space of phi has dimension <= nc (the recurrence in the
	// real embedding reproduces the complex one, alpha and beta being real for Hermitian H).
	int maxn = p.max_steps;
	if (int64_t(maxn) > nc)
		maxn = int(nc);
	const int cap = maxn + 1;
	ws.reset(nc, nvec, cap, ntimes);

	Plan&       plan = *op.plan;
	hipStream_t s    = plan.stream();
	double*     A    = ws.scal();
	double*     B    = A + int64_t(nvec) * cap;
	double*     HN   = B + int64_t(nvec) * cap;
	double*     INV  = HN + nvec;
	double*     NRM0 = INV + nvec; // (norm, 1/norm) per target
	double*     host = ws.host();
	const int   saved_prec = plan.precision();

	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	DMK_HIP_CHECK(hipEventCreate(&ev0));
	DMK_HIP_CHECK(hipEventCreate(&ev1));
	double ms_mv   = 0;
	int    nlaunch = 0;

	std::vector<TargetState> st(static_cast<size_t>(nvec));
	int                      j = 0;
	try {
		plan.set_precision(p.precision);
		DMK_HIP_CHECK(hipMemsetAsync(A, 0, size_t(2 * nvec * cap + 2 * nvec) * sizeof(double), s));
		// V_0 = phi / |phi| (LanczosCore.h:165-172), split layout.
		launch_kry_deinterleave(phi, ldphi, ws.block(0), len, nc, nvec, s);
		launch_kry_dot(ws.block(0), ws.block(0), len, len, nvec, ws.partial(), NRM0, 2, 1, s);
		launch_kry_scale(ws.block(0), NRM0 + 1, 2, len, len, nvec, s);
		DMK_HIP_CHECK(hipMemcpyAsync(host, NRM0, size_t(2 * nvec) * sizeof(double),
		                             hipMemcpyDeviceToHost, s));
		DMK_HIP_CHECK(hipStreamSynchronize(s));
		for (int t = 0; t < nvec; ++t) {
			st[size_t(t)].rep.norm = host[2 * t];
			if (!(host[2 * t] > 0)) {
				// A zero vector stays zero at every time.
				st[size_t(t)].done = true;
				st[size_t(t)].m    = 0;
			}
		}
		double taumax = 0;
		for (int q = 0; q < ntimes; ++q)
			taumax = std::max(taumax, std::fabs(times[q]));

		auto all_done = [&] {
			for (auto& x : st)
				if (!x.done)
					return false;
			return true;
		};
		for (j = 0; j < maxn && !all_done(); ++j) {
			double*       v1 = ws.block(j);
			const double* v0 = (j > 0) ? ws.block(j - 1) : v1;
			double*       w  = ws.block(j + 1);
			// One matvec launch for all targets (oneStepDecomposition, LanczosVectors.h:203-237).
			DMK_HIP_CHECK(hipEventRecord(ev0, s));
			op.apply(v1, w, nvec);
			DMK_HIP_CHECK(hipEventRecord(ev1, s));
			++nlaunch;
			launch_kry_dot(v1, w, len, len, nvec, ws.partial(), A + j, cap, 0, s);
			launch_kry_update_beta(w, v1, v0, A, B, HN, INV, j, cap, len, len, nvec, ws.partial(), s);
			launch_kry_scale(w, INV, 1, len, len, nvec, s);
			DMK_HIP_CHECK(hipMemcpyAsync(host, A, size_t(2 * nvec * cap) * sizeof(double),
			                             hipMemcpyDeviceToHost, s));
			DMK_HIP_CHECK(hipStreamSynchronize(s));
			float ms = 0;
			DMK_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
			ms_mv += ms;
			for (int t = 0; t < nvec; ++t) {
				TargetState& x = st[size_t(t)];
				if (x.done)
					continue;
				const double alpha = host[int64_t(t) * cap + j];
				const double beta  = host[int64_t(nvec) * cap + int64_t(t) * cap + j + 1];
				x.a.push_back(alpha);
				if (x.b.empty())
					x.b.push_back(0.0);
				x.b.push_back(beta); // b[j + 1]
				const int mm = j + 1;
				bool      stop = (j == maxn - 1);
				if (beta == 0.0) {
					x.rep.breakdown = 1;
					stop            = true;
				}
				if (!stop && j >= 1) {
					bool exitFlag = false;
					if (p.stop == 0) {
						// Lowest Ritz value settled (LanczosCore.h:220-240).
						const double e = tridiag_eigenvalue(x.a.data(), x.b.data(), mm, 0);
						exitFlag       = std::fabs(e - x.eold) < p.tolerance;
						x.eold         = e;
					} else {
						std::vector<double> eps(static_cast<size_t>(mm)), S(static_cast<size_t>(mm * mm)),
						    c(static_cast<size_t>(2 * mm));
						tridiag_eigen_all(x.a.data(), x.b.data(), mm, eps.data(), S.data());
						krylov_coefficients(eps, S, mm, 1.0, taumax, p, c.data());
						const double tail = std::hypot(c[size_t(2 * (mm - 1))], c[size_t(2 * mm - 1)]);
						exitFlag          = beta * tail < p.tolerance;
					}
					if (exitFlag && (nc <= 4 || j >= p.min_steps))
						stop = true;
				} else if (j == 0 && p.stop == 0) {
					x.eold = 0.0;
				}
				if (stop) {
					x.done = true;
					x.m    = mm;
				}
			}
		}
		// Targets still running when the step budget ran out keep all steps.
		for (auto& x : st)
			if (!x.done) {
				x.done = true;
				x.m    = int(x.a.size());
			}
		int mmax = 0;
		for (auto& x : st)
			mmax = std::max(mmax, x.m);
		if (mmax < 1)
			mmax = 1;

		// Coefficients of every (target, time), zero past each target's own dimension.
		const size_t nscal = size_t(2 * nvec * cap + 4 * nvec);
		double*      hc    = host + nscal;
		std::fill(hc, hc + size_t(2) * size_t(nvec) * size_t(ntimes) * size_t(mmax), 0.0);
		for (int t = 0; t < nvec; ++t) {
			TargetState& x = st[size_t(t)];
			x.rep.steps    = x.m;
			if (x.m == 0)
				continue;
			const int           m = x.m;
			std::vector<double> eps(static_cast<size_t>(m)), S(static_cast<size_t>(m) * static_cast<size_t>(m)),
			    c(static_cast<size_t>(2 * m));
			x.b.resize(size_t(m + 1), 0.0);
			tridiag_eigen_all(x.a.data(), x.b.data(), m, eps.data(), S.data());
			for (int q = 0; q < ntimes; ++q) {
				krylov_coefficients(eps, S, m, x.rep.norm, times[q], p, c.data());
				std::copy(c.begin(), c.end(), hc + size_t(2) * (size_t(t * ntimes + q) * size_t(mmax)));
				if (std::fabs(times[q]) == taumax)
					x.rep.error_estimate = x.b[size_t(m)] * std::hypot(c[size_t(2 * (m - 1))], c[size_t(2 * m - 1)]);
			}
		}
		DMK_HIP_CHECK(hipMemcpyAsync(ws.coef(), hc,
		                             size_t(2) * size_t(nvec) * size_t(ntimes) * size_t(mmax) * sizeof(double),
		                             hipMemcpyHostToDevice, s));
		ws.upload_vptr(mmax, s);
		launch_kry_combine(out, ldout, ws.vptr(), len, ws.coef(), mmax, ntimes, nc, nvec, s);
		DMK_HIP_CHECK(hipStreamSynchronize(s));
		plan.set_precision(saved_prec);
	} catch (...) {
		(void)hipStreamSynchronize(s);
		(void)hipEventDestroy(ev0);
		(void)hipEventDestroy(ev1);
		try {
			plan.set_precision(saved_prec);
		} catch (...) {
		}
		throw;
	}
	(void)hipEventDestroy(ev0);
	(void)hipEventDestroy(ev1);
	std::vector<KrylovReport> reps;
	for (auto& x : st)
		reps.push_back(x.rep);
	if (ms_matvec)
		*ms_matvec = ms_mv;
	if (matvec_launches)
		*matvec_launches = nlaunch;
	if (ms_total)
		*ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	return reps;
}

} // namespace dmk
