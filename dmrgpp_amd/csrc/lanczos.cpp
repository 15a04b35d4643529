// Device-resident Lanczos recurrence (PsimagLite LanczosCore::decomposition,
// LanczosCore.h:147-272, and LanczosSolver::computeAllStatesBelow, LanczosSolver.h:68-88).
//
// The Lanczos vectors never leave HBM: V_j is column j of the basis buffer, the matvec writes
// V_{j+1} in place, and the alpha/beta scalars are produced and consumed on the device. Per step
// the host copies back (alpha_j, beta_{j+1}) only, to run the reference's stop test on the
// lowest eigenvalue of T_{j+1}. The ground vector z = sum_j c_j V_j is one pass over the basis.
//
// Mixed precision: the first run may use the FP32 matvec (FP64 recurrence); with refine set it
// is followed by an FP64 restart from the first run's Ritz vector, which converges in a few
// FP64 matvecs because that vector is already close to the FP64 eigenvector.
#include "lanczos.h"

#include "davidson.h"

#include "pool.h"
#include "tridiag.h"

#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>

namespace dmk {

int  reduction_blocks();
void launch_dot(const double* x, const double* y, int64_t n, double* partial, double* out,
                int mode, hipStream_t s);
void launch_lanczos_update(double* v2, const double* v1, const double* v0, const double* alpha,
                           const double* beta, int64_t n, double* partial, double* norm_out,
                           hipStream_t s);
void launch_scale(double* x, const double* scal, int64_t n, hipStream_t s);
void launch_scale_copy(double* y, const double* x, const double* scal, int64_t n, hipStream_t s);
void launch_axpy_neg(double* y, const double* x, const double* coef, int64_t n, hipStream_t s);
void launch_combine(double* z, const double* V, int64_t ld, const double* c, int m, int64_t n,
                    hipStream_t s);
void launch_residual(const double* hx, const double* x, double e, int64_t n, double* partial,
                     double* out, hipStream_t s);
void launch_dot_at(const double* x, const double* y, int64_t n, double* partial, double* base,
                   const int* jdev, int offset, hipStream_t s);
void launch_lanczos_update_at(double* v2, const double* v1, const double* v0, const double* A,
                              const double* B, const int* jdev, int64_t n, double* partial,
                              double* norm_out, hipStream_t s);
void launch_finish_step(double* v2, double* basis, int64_t n, const double* nrm, const int* jdev,
                        hipStream_t s);
void launch_step_scalars(const double* A, double* B, const double* nrm, int* jdev, int maxn,
                         double* report, hipStream_t s);
void launch_copy(double* y, const double* x, int64_t n, hipStream_t s);

LanczosWorkspace::~LanczosWorkspace()
{
	pool_device_free(basis_);
	pool_device_free(partial_);
	pool_device_free(scal_);
	pool_device_free(tmp_);
	pool_device_free(tmp2_);
	pool_device_free(ring_);
	pool_device_free(jdev_);
	pool_host_free(host_);
}

void LanczosWorkspace::reserve(int64_t n, int columns)
{
	if (n == n_ && columns <= cols_ && basis_)
		return;
	pool_device_free(basis_);
	pool_device_free(partial_);
	pool_device_free(scal_);
	pool_device_free(tmp_);
	pool_device_free(tmp2_);
	pool_device_free(ring_);
	pool_device_free(jdev_);
	pool_host_free(host_);
	basis_ = partial_ = scal_ = tmp_ = tmp2_ = ring_ = host_ = nullptr;
	jdev_  = nullptr;
	n_    = 0;
	cols_ = 0;
	auto dev = [](size_t bytes) { return static_cast<double*>(pool_device_alloc(bytes)); };
	basis_   = dev(size_t(n) * size_t(columns) * sizeof(double));
	partial_ = dev(size_t(reduction_blocks()) * sizeof(double));
	scal_    = dev(size_t(3 * columns + 8) * sizeof(double));
	tmp_     = dev(size_t(n) * sizeof(double));
	tmp2_    = dev(size_t(n) * sizeof(double));
	ring_    = dev(size_t(3) * size_t(n) * sizeof(double));
	jdev_    = static_cast<int*>(pool_device_alloc(sizeof(int)));
	host_    = static_cast<double*>(pool_host_alloc(size_t(3 * columns + 8) * sizeof(double)));
	n_    = n;
	cols_ = columns;
}

namespace {

struct Decomp {
	std::vector<double> a, b; // a[0..m), b[0..m) with b[0] = 0
	int                 m       = 0;
	int                 matvecs = 0;
	double              energy  = 0;
};

// One Lanczos decomposition from init (any norm) with the plan's current precision. Leaves the
// kept vectors in ws.basis(0..m).
// The three captured phase graphs of one decomposition (ring roles rotate with period 3).
struct StepGraphs {
	hipGraph_t     g[3] = { nullptr, nullptr, nullptr };
	hipGraphExec_t e[3] = { nullptr, nullptr, nullptr };
	~StepGraphs()
	{
		for (int i = 0; i < 3; ++i) {
			if (e[i])
				(void)hipGraphExecDestroy(e[i]);
			if (g[i])
				(void)hipGraphDestroy(g[i]);
		}
	}
};

// noise: relative matvec noise of a reduced-precision run that an FP64 restart follows (0 for
// a plain run); see noise_reached below.
Decomp decompose(Plan& plan, LanczosWorkspace& ws, const double* init, int max_steps, int min_steps,
                 double tol, double noise, double stop_res, int reortho, bool use_graph,
                 double& ms_matvec, hipEvent_t ev0, hipEvent_t ev1)
{
	const int64_t n      = plan.n();
	hipStream_t   s      = plan.stream();
	int           maxn   = max_steps;
	if (int64_t(maxn) > n)
		maxn = int(n);
	if (maxn < 1)
		maxn = 1;
	const int cap = ws.columns() - 1;
	if (maxn > cap)
		throw std::runtime_error("lanczos: workspace too small");
	double* A    = ws.scal();           // alpha[0..cap)
	double* B    = ws.scal() + cap;     // beta[0..cap)
	double* NRM  = ws.scal() + 2 * cap; // norm, 1/norm
	double* DOT  = NRM + 2;             // scratch scalar pair
	double* host = ws.host();
	double* part = ws.partial();

	auto matvec = [&](const double* x, double* y) {
		DMK_HIP_CHECK(hipEventRecord(ev0, s));
		plan.apply(x, n, y, n, 1, 0);
		DMK_HIP_CHECK(hipEventRecord(ev1, s));
		DMK_HIP_CHECK(hipEventSynchronize(ev1));
		float ms = 0;
		DMK_HIP_CHECK(hipEventElapsedTime(&ms, ev0, ev1));
		ms_matvec += ms;
	};

	Decomp d;
	d.a.assign(size_t(maxn), 0.0);
	d.b.assign(size_t(maxn), 0.0);

	// V0 = init / |init| (LanczosCore.h:165-172); beta_0 = 0.
	DMK_HIP_CHECK(hipMemsetAsync(B, 0, sizeof(double), s));
	launch_dot(init, init, n, part, NRM, 2, s);
	launch_scale_copy(ws.basis(0), init, NRM + 1, n, s);

	// First step (LanczosCore.h:181-208): V1 = H V0, alpha_0, V1 -= alpha_0 V0, beta_1, V1 /= beta_1.
	matvec(ws.basis(0), ws.basis(1));
	d.matvecs++;
	launch_dot(ws.basis(0), ws.basis(1), n, part, A, 0, s);
	launch_lanczos_update(ws.basis(1), ws.basis(0), ws.basis(0), A, B, n, part, NRM, s);
	launch_scale(ws.basis(1), NRM + 1, n, s);
	DMK_HIP_CHECK(hipMemcpyAsync(host, A, sizeof(double), hipMemcpyDeviceToHost, s));
	DMK_HIP_CHECK(hipMemcpyAsync(host + 1, NRM, sizeof(double), hipMemcpyDeviceToHost, s));
	DMK_HIP_CHECK(hipStreamSynchronize(s));
	d.a[0] = host[0];
	if (maxn > 1) {
		d.b[1] = host[1];
		DMK_HIP_CHECK(hipMemcpyAsync(B + 1, NRM, sizeof(double), hipMemcpyDeviceToDevice, s));
	}

	double eold  = 0;
	int    j     = 1;
	bool   broke = false;
	// A reduced-precision run (noise > 0) stops when the residual estimate of its lowest Ritz
	// pair, |beta_{j+1} s_j| with s the unit eigenvector of T_{j+1}, drops below
	// noise * ||H|| (||H|| estimated by the extreme Ritz values): past that point its matvec noise
	// dominates, and the FP64 restart takes over.
	// With stop_res > 0 the same estimate is also compared with that absolute threshold.
	auto noise_reached = [&](int jj, double e0) {
		if ((noise <= 0 && stop_res <= 0) || jj + 1 >= maxn)
			return false;
		const double bn = d.b[size_t(jj + 1)];
		std::vector<double> c = tridiag_eigenvector(d.a.data(), d.b.data(), jj + 1, e0);
		const double est = std::fabs(bn * c[size_t(jj)]);
		if (stop_res > 0 && est < stop_res)
			return true;
		if (noise <= 0)
			return false;
		const double emax  = tridiag_eigenvalue(d.a.data(), d.b.data(), jj + 1, jj);
		const double hnorm = std::max(std::fabs(e0), std::fabs(emax));
		return est < noise * hnorm;
	};
	if (use_graph && !reortho && maxn > 1) {
		// Ring of three fixed vectors: V_{j-1}, V_j, V_{j+1} live in ring[(j-1)%3], ring[j%3],
		// ring[(j+1)%3]; each captured step also writes V_{j+1} to basis column j+1.
		double* report = ws.scal() + 3 * cap + 4;
		launch_copy(ws.ring(0), ws.basis(0), n, s);
		launch_copy(ws.ring(1), ws.basis(1), n, s);
		const int one = 1;
		DMK_HIP_CHECK(hipMemcpyAsync(ws.jdev(), &one, sizeof(int), hipMemcpyHostToDevice, s));
		DMK_HIP_CHECK(hipStreamSynchronize(s));
		StepGraphs sg;
		for (int p = 0; p < 3; ++p) {
			double* v0 = ws.ring((p + 2) % 3);
			double* v1 = ws.ring(p);
			double* v2 = ws.ring((p + 1) % 3);
			DMK_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
			// Every call recorded into the capture is checked, and the capture is ended on every
			// path: a failure inside it (a launch error, a copy the runtime refuses, an exception
			// of the matvec) must not leave the stream capturing or a partial graph behind.
			hipError_t ec = hipSuccess;
			try {
				plan.apply(v1, n, v2, n, 1, 0);
				launch_dot_at(v1, v2, n, part, A, ws.jdev(), 0, s);
				launch_lanczos_update_at(v2, v1, v0, A, B, ws.jdev(), n, part, NRM, s);
				launch_finish_step(v2, ws.basis(0), n, NRM, ws.jdev(), s);
				launch_step_scalars(A, B, NRM, ws.jdev(), maxn, report, s);
				ec = hipGetLastError();
				if (ec == hipSuccess)
					ec = hipMemcpyAsync(host, report, 2 * sizeof(double), hipMemcpyDeviceToHost, s);
			} catch (...) {
				hipGraph_t partial = nullptr;
				(void)hipStreamEndCapture(s, &partial);
				if (partial)
					(void)hipGraphDestroy(partial);
				throw;
			}
			const hipError_t ee = hipStreamEndCapture(s, &sg.g[p]);
			DMK_HIP_CHECK(ec);
			DMK_HIP_CHECK(ee);
			if (!sg.g[p])
				throw std::runtime_error("lanczos: stream capture produced no graph");
			DMK_HIP_CHECK(hipGraphInstantiate(&sg.e[p], sg.g[p], nullptr, nullptr, 0));
		}
		const auto t0 = std::chrono::steady_clock::now();
		for (j = 1; j < maxn; ++j) {
			DMK_HIP_CHECK(hipGraphLaunch(sg.e[j % 3], s));
			DMK_HIP_CHECK(hipStreamSynchronize(s));
			d.matvecs++;
			d.a[size_t(j)] = host[0];
			if (j + 1 < maxn)
				d.b[size_t(j + 1)] = host[1];
			const double e     = tridiag_eigenvalue(d.a.data(), d.b.data(), j + 1, 0);
			const double delta = std::fabs(e - eold);
			eold               = e;
			bool exitFlag = (delta < tol) || (j == maxn - 1) || noise_reached(j, e);
			if (exitFlag && (n <= 4 || j >= min_steps)) {
				broke = true;
				break;
			}
		}
		// Device time is not split out per matvec inside the graphs; charge the loop.
		ms_matvec += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	} else
	for (j = 1; j < maxn; ++j) {
		double* v0 = ws.basis(j - 1);
		double* v1 = ws.basis(j);
		double* v2 = ws.basis(j + 1);
		// oneStepDecomposition (LanczosVectors.h:203-237).
		matvec(v1, v2);
		d.matvecs++;
		launch_dot(v1, v2, n, part, A + j, 0, s);
		launch_lanczos_update(v2, v1, v0, A + j, B + j, n, part, NRM, s);
		launch_scale(v2, NRM + 1, n, s);
		DMK_HIP_CHECK(hipMemcpyAsync(host, A + j, sizeof(double), hipMemcpyDeviceToHost, s));
		DMK_HIP_CHECK(hipMemcpyAsync(host + 1, NRM, sizeof(double), hipMemcpyDeviceToHost, s));
		if (j + 1 < maxn)
			DMK_HIP_CHECK(hipMemcpyAsync(B + j + 1, NRM, sizeof(double), hipMemcpyDeviceToDevice, s));
		if (reortho) {
			// reorthoIfNecessary (LanczosVectors.h:243-271): V2 -= <V_i|V2> V_i for i = 0..j in turn,
			// then renormalize.
			for (int i = 0; i <= j; ++i) {
				launch_dot(ws.basis(i), v2, n, part, DOT, 0, s);
				launch_axpy_neg(v2, ws.basis(i), DOT, n, s);
			}
			launch_dot(v2, v2, n, part, DOT, 2, s);
			launch_scale(v2, DOT + 1, n, s);
		}
		DMK_HIP_CHECK(hipStreamSynchronize(s));
		d.a[size_t(j)] = host[0];
		if (j + 1 < maxn)
			d.b[size_t(j + 1)] = host[1];
		// ab.diag(tmpEigs, j+1) and the stop test (LanczosCore.h:220-240).
		const double e     = tridiag_eigenvalue(d.a.data(), d.b.data(), j + 1, 0);
		const double delta = std::fabs(e - eold);
		eold               = e;
		bool exitFlag      = (delta < tol) || (j == maxn - 1) || noise_reached(j, e);
		if (exitFlag && n <= 4) {
			broke = true;
			break;
		}
		if (exitFlag && j >= min_steps) {
			broke = true;
			break;
		}
	}
	d.m = (broke || j < maxn) ? j + 1 : maxn;
	if (d.m > maxn)
		d.m = maxn;
	d.a.resize(size_t(d.m));
	d.b.resize(size_t(d.m));
	d.energy = tridiag_eigenvalue(d.a.data(), d.b.data(), d.m, 0);
	return d;
}

// z = sum_j c_j V_j with c the unit lowest eigenvector of T_m.
void ritz_vector(Plan& plan, LanczosWorkspace& ws, const Decomp& d, double* z)
{
	std::vector<double> c    = tridiag_eigenvector(d.a.data(), d.b.data(), d.m, d.energy);
	const int           cap  = ws.columns() - 1;
	double*             coef = ws.scal() + 2 * cap + 4;
	DMK_HIP_CHECK(hipMemcpyAsync(coef, c.data(), size_t(d.m) * sizeof(double),
	                             hipMemcpyHostToDevice, plan.stream()));
	launch_combine(z, ws.basis(0), plan.n(), coef, d.m, plan.n(), plan.stream());
	DMK_HIP_CHECK(hipStreamSynchronize(plan.stream()));
}

} // namespace

LanczosReport lanczos_ground(Plan& plan, LanczosWorkspace& ws, const double* init, double* z,
                             const LanczosParams& p)
{
	if (!init || !z)
		throw std::invalid_argument("lanczos: null vector");
	if (p.max_steps < 1 || p.min_steps < 0)
		throw std::invalid_argument("lanczos: bad step limits");
	if (p.precision != PREC_FP64 && p.precision != PREC_FP32 && p.precision != PREC_BF16X3)
		throw std::runtime_error("lanczos: unsupported precision mode");
	const auto    t0 = std::chrono::steady_clock::now();
	const int64_t n  = plan.n();
	ws.reserve(n, std::max(p.max_steps, 2) + 1);
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
	DMK_HIP_CHECK(hipEventCreate(&ev0));
	DMK_HIP_CHECK(hipEventCreate(&ev1));
	const int     saved_prec = plan.precision();
	LanczosReport r;
	try {
		plan.set_precision(p.precision);
		// Hand-over level of the reduced-precision first run, relative to ||H||: near the
		// matvec noise (about 16 units of roundoff of the FP32 accumulation; the (hi, lo) bf16
		// pair representation for BF16x3) before an FP64 Lanczos restart, which starts from
		// scratch; earlier (3e-5 for FP32) before the FP64 Davidson refinement, which gains
		// the remaining digits in a few matvecs (ladder superblock at m = 2048: FP32 70 -> 66
		// ms, the restart slower with any larger value, BF16x3 flat between 3e-5 and 1e-3).
		const double noise = p.precision == PREC_FP64 ? 0.0
		                     : p.switch_noise > 0     ? p.switch_noise
		                     : p.precision == PREC_FP32 ? (p.refine == 2 ? 3e-5 : 1e-6)
		                                                : 1e-4;
		const bool mixed_first = (p.precision != PREC_FP64 && p.refine);
		Decomp d = decompose(plan, ws, init, p.max_steps, p.min_steps, p.tolerance,
		                     (p.refine ? noise : 0.0), mixed_first ? 0.0 : p.stop_residual,
		                     p.reortho, p.graph != 0, r.ms_matvec, ev0, ev1);
		ritz_vector(plan, ws, d, z);
		r.energy_first    = d.energy;
		r.steps_first     = d.m;
		r.matvecs_first   = d.matvecs;
		r.precision_first = p.precision;
		r.energy          = d.energy;
		r.steps           = d.m;
		r.matvecs         = d.matvecs;
		bool residual_done = false;
		if (p.precision != PREC_FP64 && p.refine == 2) {
			// FP64 Davidson (Olsen's diagonal correction) from the mixed-precision Ritz vector,
			// to a residual of sqrt(refine_tolerance): the Rayleigh quotient's error is of the
			// order of the squared residual over the gap, the size of the Lanczos |dE| test.
			DavidsonParams dp;
			dp.tolerance   = std::sqrt(p.refine_tolerance);
			dp.max_basis   = 12;
			dp.max_matvecs = std::max(p.max_steps, 8);
			dp.precond     = PRECOND_DIAG;
			DMK_HIP_CHECK(hipMemcpyAsync(ws.tmp2(), z, size_t(n) * sizeof(double),
			                             hipMemcpyDeviceToDevice, plan.stream()));
			const DavidsonReport dr = davidson_ground(plan, ws.tmp2(), z, dp, nullptr);
			r.energy   = dr.energy;
			r.steps    = dr.iterations;
			r.matvecs += dr.matvecs;
			r.residual = dr.residual;
			residual_done = true;
		} else if (p.precision != PREC_FP64 && p.refine) {
			plan.set_precision(PREC_FP64);
			// Restart from the mixed-precision Ritz vector (copied: z is overwritten below).
			DMK_HIP_CHECK(hipMemcpyAsync(ws.tmp2(), z, size_t(n) * sizeof(double),
			                             hipMemcpyDeviceToDevice, plan.stream()));
			Decomp d2 = decompose(plan, ws, ws.tmp2(), p.max_steps, p.refine_min_steps,
			                      p.refine_tolerance, 0.0, p.stop_residual, p.reortho,
			                      p.graph != 0, r.ms_matvec, ev0, ev1);
			ritz_vector(plan, ws, d2, z);
			r.energy = d2.energy;
			r.steps  = d2.m;
			r.matvecs += d2.matvecs;
		}
		// FP64 residual ||H z - E z|| / ||z|| (the Davidson refinement reports its own).
		plan.set_precision(PREC_FP64);
		if (!residual_done) {
			double* NRM = ws.scal() + 2 * (ws.columns() - 1);
			plan.apply(z, n, ws.tmp(), n, 1, 0);
			launch_residual(ws.tmp(), z, r.energy, n, ws.partial(), NRM, plan.stream());
			launch_dot(z, z, n, ws.partial(), NRM + 1, 1, plan.stream());
			DMK_HIP_CHECK(hipMemcpyAsync(ws.host(), NRM, 2 * sizeof(double),
			                             hipMemcpyDeviceToHost, plan.stream()));
			DMK_HIP_CHECK(hipStreamSynchronize(plan.stream()));
			r.residual = ws.host()[0] / (ws.host()[1] > 0 ? ws.host()[1] : 1.0);
		}
		plan.set_precision(saved_prec);
	} catch (...) {
		(void)hipEventDestroy(ev0);
		(void)hipEventDestroy(ev1);
		throw;
	}
	(void)hipEventDestroy(ev0);
	(void)hipEventDestroy(ev1);
	r.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	return r;
}

} // namespace dmk
