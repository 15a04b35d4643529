// Device-resident Lanczos ground-state solver over a Plan's matvec.
#pragma once

#include "kron_plan.h"

#include <vector>

namespace dmk {

struct LanczosParams {
	int    max_steps        = 200;   // ParametersForSolver.h:97 (LanczosSteps)
	int    min_steps        = 4;     // ParametersForSolver.h:98
	double tolerance        = 1e-12; // ParametersForSolver.h:99
	int    reortho          = 0;     // full re-orthogonalization (LanczosVectors.h:243-271)
	int    precision        = PREC_FP64;
	int    refine           = 1;     // FP64 restart from the mixed-precision Ritz vector
	int    refine_min_steps = 4;
	double refine_tolerance = 1e-12;
	int    graph            = 1;     // capture each step in a HIP graph (not with reortho)
	double stop_residual    = 0;     // > 0: stop once |beta_{j+1} s_j| < stop_residual
	double switch_noise     = 0;     // > 0: hand-over level of a refined mixed first run
};

struct LanczosReport {
	double energy          = 0;
	double residual        = 0; // ||H z - E z|| / ||z|| with the FP64 matvec
	int    steps           = 0;
	int    matvecs         = 0;
	double energy_first    = 0; // energy of the first (possibly mixed-precision) run
	int    steps_first     = 0;
	int    matvecs_first   = 0;
	int    precision_first = PREC_FP64;
	double ms_total        = 0;
	double ms_matvec       = 0;
};

// Reusable device workspace of the solver (basis vectors and scalars).
class LanczosWorkspace {
public:
	~LanczosWorkspace();
	void    reserve(int64_t n, int columns);
	double* basis(int j) { return basis_ + int64_t(j) * n_; }
	double* partial() { return partial_; }
	double* scal() { return scal_; } // device scalars: a[cap], b[cap], norm/inverse pairs, coefs
	double* tmp() { return tmp_; }
	double* tmp2() { return tmp2_; }
	double* ring(int i) { return ring_ + int64_t(i) * n_; } // three work vectors (graph mode)
	int*    jdev() { return jdev_; }
	double* host() { return host_; }
	int     columns() const { return cols_; }

private:
	int64_t n_       = 0;
	int     cols_    = 0;
	double* basis_   = nullptr;
	double* partial_ = nullptr;
	double* scal_    = nullptr;
	double* tmp_     = nullptr;
	double* tmp2_    = nullptr;
	double* ring_    = nullptr;
	int*    jdev_    = nullptr;
	double* host_    = nullptr; // pinned
};

// Ground state of the plan's H from init (device, n). z (device, n) receives the Ritz vector.
LanczosReport lanczos_ground(Plan& plan, LanczosWorkspace& ws, const double* init, double* z,
                             const LanczosParams& p);

} // namespace dmk
