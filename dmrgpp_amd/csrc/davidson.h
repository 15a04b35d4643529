// Davidson ground-state solver over a Plan's matvec, with a diagonal or a learned (MLP)
// correction, and the harvest of training pairs for the MLP.
#pragma once

#include "kron_plan.h"
#include "mlp.h"

#include <vector>

namespace dmk {

enum Precond : int
{
	PRECOND_NONE = 0, // t = r
	PRECOND_DIAG = 1, // Olsen: t = M^-1 (eps u - r), M = diag(H) - theta, t orthogonal to u
	PRECOND_MLP  = 2, // t = MLP(features) * rms(r) / s_D
};

struct DavidsonParams {
	int    max_matvecs     = 400;
	double tolerance       = 1e-9; // ||H u - theta u|| (u unit) at which to stop
	int    max_basis       = 24;   // restart size (thick restart keeps 2 vectors)
	int    precond         = PRECOND_DIAG;
	int    harvest         = 0;     // record iterates and append training pairs to the MLP set
	int    harvest_samples = 8192;  // sampled elements per recorded iterate
	int    harvest_iters   = 32;    // at most this many iterates recorded
};

struct DavidsonReport {
	double  energy      = 0;
	double  residual    = 0; // ||H z - E z|| / ||z|| with a fresh FP64 matvec
	int     iterations  = 0;
	int     matvecs     = 0; // including the residual check
	int     restarts    = 0;
	int     converged   = 0;
	int64_t samples     = 0; // training pairs appended
	double  ms_total    = 0;
	double  ms_matvec   = 0;
	double  ms_precond  = 0;
};

// Ground state of the plan's H (FP64 matvec) from init (device, n); z (device, n) receives the
// unit Ritz vector. mlp is required for PRECOND_MLP and for harvest.
DavidsonReport davidson_ground(Plan& plan, const double* init, double* z,
                               const DavidsonParams& p, Mlp* mlp);

// Lowest eigenpair of the dense symmetric k x k matrix G (column major) by cyclic Jacobi.
double dense_lowest(const std::vector<double>& G, int k, std::vector<double>& vec);

} // namespace dmk
