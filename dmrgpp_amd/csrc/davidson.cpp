// Davidson eigensolver for the lowest state, device resident.
//
// There is no Davidson in the reference (its solvers are Lanczos and Chebyshev, ParametersDmrg
// Solver.h:154-158); this is the learned-preconditioner variant requested beside the Lanczos
// path. Each iteration: Rayleigh-Ritz on the basis V (projected matrix G = V^T H V, kept on the
// host in FP64, diagonalized by Jacobi), Ritz vector u and residual r = H u - theta u in one
// fused pass, a correction t from r (diagonal: Olsen's M^-1 (eps u - r) with M = diag H - theta;
// learned: the MLP evaluated on per-element features), repeated classical Gram-Schmidt against
// V, and one matvec for the new basis vector. When the basis is full it restarts from the Ritz
// vector and the previous Ritz vector (with a fresh matvec for the latter). The stop test is the
// residual norm of the unit Ritz vector, and the reported residual is recomputed with a fresh
// FP64 matvec.
//
// Training pairs: with harvest set, up to harvest_iters iterates u_j are kept together with the
// features of harvest_samples sampled elements; once the eigenvector psi is known, the exact
// correction t* = psi / (psi . u_j) - u_j of each iterate, minus Olsen's correction, gives the
// targets: the network learns what the diagonal method misses.
#include "davidson.h"

#include "dmk_internal.h"
#include "pool.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <stdexcept>

namespace dmk {

int  davidson_reduction_blocks();
void launch_multidot(const double* V, int64_t ldv, const double* t, int64_t n, int k,
                     double* partial, double* out, hipStream_t s);
void launch_sub_combine(double* t, const double* V, int64_t ldv, const double* c, int k, int64_t n,
                        hipStream_t s);
void launch_ritz(const double* V, const double* W, int64_t ld, const double* sv, int k,
                 double theta, int64_t n, double* u, double* hu, double* r, double* partial,
                 double* out2, hipStream_t s);
void launch_precond_diag(const double* r, const double* D, double theta, double floor_, int64_t n,
                         double* t, hipStream_t s);
void launch_precond_mlp(const float* params, const double* r, const double* u, const double* D,
                        double theta, double floor_, const double* sums, double sD, double rms_r,
                        double rms_u, int64_t n, double* t, hipStream_t s);
void launch_olsen_sums(const double* r, const double* u, const double* D, double theta,
                       double floor_, int64_t n, double* partial, double* sums, hipStream_t s);
void launch_harvest_features(const double* r, const double* u, const double* D, double theta,
                             double floor_, const double* sums, double sD, double rms_r,
                             double rms_u, int64_t n, int64_t phase, int64_t stride, int64_t count,
                             float* x_out, float* tols_out, hipStream_t s);
void launch_harvest_targets(const double* psi, const double* u, const double* pu, const float* tols,
                            double unit_inv, int64_t n, int64_t phase, int64_t stride,
                            int64_t count, float* t_out, hipStream_t s);
void launch_dot(const double* x, const double* y, int64_t n, double* partial, double* out,
                int mode, hipStream_t s);
int  reduction_blocks();
void launch_scale_copy(double* y, const double* x, const double* scal, int64_t n, hipStream_t s);
void launch_scale(double* x, const double* scal, int64_t n, hipStream_t s);
void launch_combine(double* z, const double* V, int64_t ld, const double* c, int m, int64_t n,
                    hipStream_t s);
void launch_residual(const double* hx, const double* x, double e, int64_t n, double* partial,
                     double* out, hipStream_t s);
void launch_copy(double* y, const double* x, int64_t n, hipStream_t s);
void launch_fill(double* x, double v, int64_t n, hipStream_t s);
void launch_precond_olsen(const double* r, const double* u, const double* D, double theta,
                          double floor_, int64_t n, const double* sums, double* t, hipStream_t s);

double dense_lowest(const std::vector<double>& G, int k, std::vector<double>& vec)
{
	std::vector<double> A(size_t(k) * size_t(k)), Z(size_t(k) * size_t(k), 0.0);
	auto a = [&](int i, int j) -> double& { return A[size_t(i) + size_t(j) * size_t(k)]; };
	auto z = [&](int i, int j) -> double& { return Z[size_t(i) + size_t(j) * size_t(k)]; };
	double norm = 0;
	for (int j = 0; j < k; ++j)
		for (int i = 0; i < k; ++i) {
			a(i, j) = 0.5 * (G[size_t(i) + size_t(j) * size_t(k)] + G[size_t(j) + size_t(i) * size_t(k)]);
			norm += a(i, j) * a(i, j);
		}
	for (int i = 0; i < k; ++i)
		z(i, i) = 1.0;
	for (int sweep = 0; sweep < 100; ++sweep) {
		double off = 0;
		for (int j = 0; j < k; ++j)
			for (int i = 0; i < j; ++i)
				off += a(i, j) * a(i, j);
		if (off <= 1e-32 * norm || off == 0.0)
			break;
		for (int p = 0; p < k; ++p)
			for (int q = p + 1; q < k; ++q) {
				const double apq = a(p, q);
				if (apq == 0.0)
					continue;
				// Rotation zeroing A(p,q): tan of the angle from the classic Jacobi formula.
				const double th = (a(q, q) - a(p, p)) / (2.0 * apq);
				const double t  = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
				const double c  = 1.0 / std::sqrt(t * t + 1.0);
				const double s  = t * c;
				for (int r = 0; r < k; ++r) {
					const double arp = a(r, p), arq = a(r, q);
					a(r, p)          = c * arp - s * arq;
					a(r, q)          = s * arp + c * arq;
				}
				for (int r = 0; r < k; ++r) {
					const double apr = a(p, r), aqr = a(q, r);
					a(p, r)          = c * apr - s * aqr;
					a(q, r)          = s * apr + c * aqr;
				}
				for (int r = 0; r < k; ++r) {
					const double zrp = z(r, p), zrq = z(r, q);
					z(r, p)          = c * zrp - s * zrq;
					z(r, q)          = s * zrp + c * zrq;
				}
			}
	}
	int best = 0;
	for (int i = 1; i < k; ++i)
		if (a(i, i) < a(best, best))
			best = i;
	vec.resize(size_t(k));
	for (int i = 0; i < k; ++i)
		vec[size_t(i)] = z(i, best);
	return a(best, best);
}

namespace {

struct Buffers {
	std::vector<void*> owned;
	double*            alloc(size_t elems)
	{
		void* p = pool_device_alloc(std::max<size_t>(elems, 1) * sizeof(double));
		owned.push_back(p);
		return static_cast<double*>(p);
	}
	~Buffers()
	{
		for (void* p : owned)
			pool_device_free(p);
	}
};

struct HarvestMeta {
	double  theta, sD, rms_r;
	int64_t phase;
};

} // namespace

DavidsonReport davidson_ground(Plan& plan, const double* init, double* zout,
                               const DavidsonParams& p, Mlp* mlp)
{
	const auto t_start = std::chrono::steady_clock::now();
	if (!init || !zout)
		throw std::invalid_argument("davidson: null vector");
	if (p.max_basis < 3 || p.max_basis > 128 || !(p.tolerance > 0) || p.max_matvecs < 2)
		throw std::invalid_argument("davidson: bad parameters (max_basis in [3, 128], tolerance > 0)");
	if (p.precond < PRECOND_NONE || p.precond > PRECOND_MLP)
		throw std::invalid_argument("davidson: unknown preconditioner");
	if ((p.precond == PRECOND_MLP || p.harvest) && !mlp)
		throw std::invalid_argument("davidson: the MLP preconditioner and harvest need an MLP");
	if (p.harvest && (p.harvest_samples < 1 || p.harvest_iters < 1))
		throw std::invalid_argument("davidson: bad harvest sizes");

	const int64_t n  = plan.n();
	hipStream_t   s  = plan.stream();
	const int     MB = p.max_basis;
	const int     saved_prec = plan.precision();
	plan.set_precision(PREC_FP64);

	Buffers B;
	double* V    = B.alloc(size_t(MB) * size_t(n));
	double* W    = B.alloc(size_t(MB) * size_t(n));
	double* u    = B.alloc(size_t(n));
	double* hu   = B.alloc(size_t(n));
	double* r    = B.alloc(size_t(n));
	double* t    = B.alloc(size_t(n));
	double* up   = B.alloc(size_t(n));
	double* hup  = B.alloc(size_t(n));
	double* D    = B.alloc(size_t(n));
	double* part = B.alloc(size_t(std::max(2 * davidson_reduction_blocks() * std::max(MB, 2),
	                                       reduction_blocks())));
	double* scal = B.alloc(size_t(4 * MB + 64));
	double* d_sv = scal;             // Ritz coefficients (MB)
	double* d_c  = scal + MB;        // projections (MB)
	double* d_sp = scal + 2 * MB;    // previous Ritz coefficients (MB)
	double* d_g  = scal + 3 * MB;    // new column of G (MB)
	double* NRM  = scal + 4 * MB;    // norm, inverse
	double* OUT2 = scal + 4 * MB + 4;
	double* MISC = scal + 4 * MB + 8;
	std::vector<double> host(size_t(4 * MB + 64));

	const int     HS     = p.harvest ? std::min<int64_t>(p.harvest_samples, n) : 0;
	const int     HI     = p.harvest ? p.harvest_iters : 0;
	const int64_t stride = HS ? std::max<int64_t>(1, n / HS) : 1;
	double*       Ubuf   = HI ? B.alloc(size_t(HI) * size_t(n)) : nullptr;
	float*        Xst    = HI ? reinterpret_cast<float*>(B.alloc(size_t(HI) * size_t(HS) * kMlpIn / 2 + 1))
	                          : nullptr;
	float*        Tst    = HI ? reinterpret_cast<float*>(B.alloc(size_t(HI) * size_t(HS) / 2 + 1))
	                          : nullptr; // Olsen's correction at the samples, in units
	double*       d_pu   = HI ? B.alloc(size_t(HI)) : nullptr;
	std::vector<HarvestMeta> meta;

	hipEvent_t e0 = nullptr, e1 = nullptr;
	DMK_HIP_CHECK(hipEventCreate(&e0));
	DMK_HIP_CHECK(hipEventCreate(&e1));
	DavidsonReport rep;
	auto fetch = [&](const double* dsrc, size_t count, double* dst) {
		DMK_HIP_CHECK(hipMemcpyAsync(dst, dsrc, count * sizeof(double), hipMemcpyDeviceToHost, s));
		DMK_HIP_CHECK(hipStreamSynchronize(s));
	};
	auto put = [&](double* ddst, const double* src, size_t count) {
		DMK_HIP_CHECK(hipMemcpyAsync(ddst, src, count * sizeof(double), hipMemcpyHostToDevice, s));
		DMK_HIP_CHECK(hipStreamSynchronize(s));
	};
	auto matvec = [&](const double* x, double* y) {
		DMK_HIP_CHECK(hipEventRecord(e0, s));
		plan.apply(x, n, y, n, 1, 0);
		DMK_HIP_CHECK(hipEventRecord(e1, s));
		DMK_HIP_CHECK(hipEventSynchronize(e1));
		float ms = 0;
		DMK_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
		rep.ms_matvec += ms;
		rep.matvecs++;
	};
	auto Vc = [&](int j) { return V + int64_t(j) * n; };
	auto Wc = [&](int j) { return W + int64_t(j) * n; };

	try {
		// diag(H) and its first two moments (for the feature scale s_D = rms(D - theta)).
		plan.diagonal(D);
		launch_fill(r, 1.0, n, s);
		launch_dot(D, r, n, part, NRM, 0, s);
		launch_dot(D, D, n, part, NRM + 1, 0, s);
		fetch(NRM, 2, host.data());
		const double meanD  = host[0] / double(n);
		const double meanD2 = host[1] / double(n);

		// V_0 = init / |init|, W_0 = H V_0, G = [V_0 . W_0].
		launch_dot(init, init, n, part, NRM, 2, s);
		launch_scale_copy(Vc(0), init, NRM + 1, n, s);
		matvec(Vc(0), Wc(0));
		std::vector<double> G(size_t(MB) * size_t(MB), 0.0);
		launch_multidot(V, n, Wc(0), n, 1, part, d_g, s);
		fetch(d_g, 1, host.data());
		G[0]       = host[0];
		int  k     = 1;
		bool have_prev = false;
		std::vector<double> sv, sprev;
		double theta = G[0];
		double rnorm = 0, unorm = 1;
		while (true) {
			std::vector<double> Gk(size_t(k) * size_t(k));
			for (int j = 0; j < k; ++j)
				for (int i = 0; i < k; ++i)
					Gk[size_t(i) + size_t(j) * size_t(k)] = G[size_t(i) + size_t(j) * size_t(MB)];
			theta = dense_lowest(Gk, k, sv);
			put(d_sv, sv.data(), size_t(k));
			launch_ritz(V, W, n, d_sv, k, theta, n, u, hu, r, part, OUT2, s);
			fetch(OUT2, 2, host.data());
			rnorm = host[0];
			unorm = host[1] > 0 ? host[1] : 1.0;
			rep.iterations++;
			const double res   = rnorm / unorm;
			const double rms_r = rnorm / std::sqrt(double(n));
			const double rms_u = unorm / std::sqrt(double(n));
			const double sD    = std::max(1e-12, std::sqrt(std::max(0.0, meanD2 - 2.0 * theta * meanD + theta * theta)));
			const double floor_ = 1e-8 * std::max(1.0, std::fabs(theta));
			double*      SUMS   = MISC + 2; // Olsen's u . M^-1 r and u . M^-1 u
			if (p.precond != PRECOND_NONE || HI)
				launch_olsen_sums(r, u, D, theta, floor_, n, part, SUMS, s);
			if (HI && int(meta.size()) < HI && rnorm > 0) {
				const int     j     = int(meta.size());
				const int64_t phase = (int64_t(j) * 7919 + 13) % stride;
				launch_copy(Ubuf + int64_t(j) * n, u, n, s);
				launch_harvest_features(r, u, D, theta, floor_, SUMS, sD, rms_r, rms_u, n, phase,
				                        stride, HS, Xst + int64_t(j) * HS * kMlpIn,
				                        Tst + int64_t(j) * HS, s);
				meta.push_back(HarvestMeta{ theta, sD, rms_r, phase });
			}
			if (res < p.tolerance) {
				rep.converged = 1;
				break;
			}
			if (rep.matvecs >= p.max_matvecs)
				break;
			// Correction.
			DMK_HIP_CHECK(hipEventRecord(e0, s));
			if (p.precond == PRECOND_NONE)
				launch_copy(t, r, n, s);
			else if (p.precond == PRECOND_DIAG)
				launch_precond_olsen(r, u, D, theta, floor_, n, SUMS, t, s);
			else
				launch_precond_mlp(mlp->params(), r, u, D, theta, floor_, SUMS, sD, rms_r, rms_u, n,
				                   t, s);
			DMK_HIP_CHECK(hipEventRecord(e1, s));
			DMK_HIP_CHECK(hipEventSynchronize(e1));
			{
				float ms = 0;
				DMK_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
				rep.ms_precond += ms;
			}
			if (k == MB) {
				// Thick restart: [u, previous Ritz vector orthogonalized against u].
				std::vector<double> inv(2, 1.0 / unorm);
				if (have_prev) {
					std::vector<double> spad(size_t(k), 0.0);
					std::copy(sprev.begin(), sprev.end(), spad.begin());
					put(d_sp, spad.data(), size_t(k));
					launch_combine(up, V, n, d_sp, k, n, s);
					launch_combine(hup, W, n, d_sp, k, n, s);
				}
				put(MISC, inv.data(), 1);
				launch_scale_copy(Vc(0), u, MISC, n, s);
				launch_scale_copy(Wc(0), hu, MISC, n, s);
				k = 1;
				if (have_prev) {
					// The previous Ritz vector is nearly parallel to u near convergence: its W
					// column is recomputed by a matvec rather than combined, since combining
					// amplifies the roundoff of the cancellation by 1 / |up - (u.up) u|.
					launch_multidot(V, n, up, n, 1, part, d_c, s);
					launch_sub_combine(up, V, n, d_c, 1, n, s);
					launch_dot(up, up, n, part, NRM, 2, s);
					fetch(NRM, 2, host.data());
					if (host[0] > 1e-6) {
						launch_scale_copy(Vc(1), up, NRM + 1, n, s);
						matvec(Vc(1), Wc(1));
						k = 2;
					}
				}
				std::fill(G.begin(), G.end(), 0.0);
				for (int j = 0; j < k; ++j) {
					launch_multidot(V, n, Wc(j), n, k, part, d_g, s);
					fetch(d_g, size_t(k), host.data());
					for (int i = 0; i < k; ++i)
						G[size_t(i) + size_t(j) * size_t(MB)] = host[size_t(i)];
				}
				have_prev = false;
				rep.restarts++;
			}
			// Classical Gram-Schmidt against V_0..V_{k-1} on the normalized correction, repeated
			// (at most three rounds) while a round removes more than half of the norm: near
			// convergence most of t lies in span V and one round leaves roundoff of relative
			// size eps / |t_new| in the new direction.
			double tn = 0;
			for (int attempt = 0; attempt < 2; ++attempt) {
				launch_dot(t, t, n, part, NRM, 2, s);
				launch_scale(t, NRM + 1, n, s);
				double before = 1.0;
				for (int round = 0; round < 3; ++round) {
					launch_multidot(V, n, t, n, k, part, d_c, s);
					launch_sub_combine(t, V, n, d_c, k, n, s);
					launch_dot(t, t, n, part, NRM, 2, s);
					fetch(NRM, 2, host.data());
					tn = host[0];
					if (tn > 0.5 * before)
						break;
					before = tn;
				}
				if (tn > 1e-12 || attempt == 1)
					break;
				launch_copy(t, r, n, s); // correction lies in span V: fall back to the residual
			}
			if (!(tn > 0))
				break;
			launch_scale_copy(Vc(k), t, NRM + 1, n, s);
			matvec(Vc(k), Wc(k));
			launch_multidot(V, n, Wc(k), n, k + 1, part, d_g, s);
			fetch(d_g, size_t(k + 1), host.data());
			for (int i = 0; i <= k; ++i) {
				G[size_t(i) + size_t(k) * size_t(MB)] = host[size_t(i)];
				G[size_t(k) + size_t(i) * size_t(MB)] = host[size_t(i)];
			}
			sprev     = sv;
			have_prev = true;
			++k;
		}
		// z = u / |u|; FP64 residual with a fresh matvec.
		std::vector<double> inv(1, 1.0 / unorm);
		put(MISC, inv.data(), 1);
		launch_scale_copy(zout, u, MISC, n, s);
		rep.energy = theta;
		matvec(zout, t);
		launch_residual(t, zout, theta, n, part, NRM, s);
		fetch(NRM, 1, host.data());
		rep.residual = host[0];

		// Training pairs from the recorded iterates (only once the eigenvector is known).
		if (HI && rep.converged && !meta.empty()) {
			const int h = int(meta.size());
			launch_multidot(Ubuf, n, zout, n, h, part, d_pu, s);
			std::vector<double> pu(static_cast<size_t>(h));
			fetch(d_pu, size_t(h), pu.data());
			for (int j = 0; j < h; ++j) {
				const double uu = 1.0; // iterates are unit up to roundoff (V orthonormal)
				if (std::fabs(pu[size_t(j)]) < 0.1 * uu)
					continue;
				const int64_t first = mlp->append(HS);
				const int64_t got   = mlp->samples() - first;
				if (got <= 0)
					break;
				DMK_HIP_CHECK(hipMemcpyAsync(mlp->samples_x() + first * kMlpIn,
				                             Xst + int64_t(j) * HS * kMlpIn,
				                             size_t(got) * kMlpIn * sizeof(float),
				                             hipMemcpyDeviceToDevice, s));
				launch_harvest_targets(zout, Ubuf + int64_t(j) * n, d_pu + j, Tst + int64_t(j) * HS,
				                       meta[size_t(j)].sD / meta[size_t(j)].rms_r, n,
				                       meta[size_t(j)].phase, stride, got,
				                       mlp->samples_t() + first, s);
				rep.samples += got;
			}
		}
		DMK_HIP_CHECK(hipStreamSynchronize(s));
		plan.set_precision(saved_prec);
	} catch (...) {
		(void)hipStreamSynchronize(s);
		(void)hipEventDestroy(e0);
		(void)hipEventDestroy(e1);
		try {
			plan.set_precision(saved_prec);
		} catch (...) {
		}
		throw;
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	rep.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
	return rep;
}

} // namespace dmk
