// Batched Krylov propagator exp(-i (H - E_shift) tau) phi over a Plan's matvec.
#pragma once

#include "kron_plan.h"

#include <vector>

namespace dmk {

struct KrylovParams {
	int    max_steps      = 200;   // ParametersForSolver.h:98 defaults of the "Tridiag" solver
	int    min_steps      = 4;
	double tolerance      = 1e-12;
	int    stop           = 0; // 0: lowest Ritz value settled (LanczosCore.h:220-240);
	                           // 1: error estimate of the propagated vector below tolerance*|phi|
	int    precision      = PREC_FP64; // matvec precision; the recurrence is FP64
	double time_direction = 1.0;       // TimeVectorsKrylov.h:287 (timeDirection)
	double energy_shift   = 0.0;       // E_g subtracted from the Ritz values (TimeVectorsKrylov.h:291)
};

struct KrylovReport {
	int    steps          = 0; // Krylov dimension of this target
	double norm           = 0; // |phi|
	double error_estimate = 0; // beta_m |[exp(-i T tau_max) e_0]_{m-1}| |phi| at the largest |tau|
	int    breakdown      = 0; // an invariant subspace was reached
};

// How a plan's real matvec acts on split complex vectors (real parts, then imaginary parts).
struct KrylovOperator {
	Plan*   plan     = nullptr;
	bool    embedded = false; // the plan is the real embedding of a complex H (size 2 nc)
	int64_t nc       = 0;     // complex length of one target vector

	// out_t = H in_t for nvec split complex vectors at stride 2 nc.
	void apply(const double* in, double* out, int nvec) const;
};

class KrylovWorkspace {
public:
	~KrylovWorkspace();
	// Room for `columns` Krylov vectors of nvec targets of complex length nc (allocated lazily,
	// in chunks, as the recurrence advances).
	void    reset(int64_t nc, int nvec, int columns, int ntimes);
	double* block(int j); // step j: nvec split vectors at stride 2 nc
	double* partial() { return partial_; }
	double* scal() { return scal_; }
	double* coef() { return coef_; }
	double* host() { return host_; }
	const double* const* vptr() { return vptr_; }
	void    upload_vptr(int m, hipStream_t s);

private:
	static constexpr int kChunk = 8;
	int64_t              nc_ = 0, block_elems_ = 0;
	int                  nvec_ = 0, cols_ = 0, ntimes_ = 0;
	std::vector<double*> chunks_;
	double*              partial_ = nullptr;
	double*              scal_    = nullptr;
	double*              coef_    = nullptr;
	double**             vptr_    = nullptr; // device array of block pointers
	double*              host_    = nullptr; // pinned: scalars, then coefficients, then pointers
	size_t               host_elems_ = 0;
	void                 release();
};

// out_{t,s} = exp(-i (H - E) tau_s d) phi_t for nvec targets and ntimes times, the reference's
// Krylov time step (TimeVectorsKrylov.h:165-322, ParallelTriDiag.h:132-161, KrylovHelper.h),
// with all targets advancing through one batched matvec per Krylov step. phi: interleaved
// complex device vectors at ldphi (complex elements); out: interleaved at ldout, target t,
// time s at out + 2 ldout (t ntimes + s).
std::vector<KrylovReport> krylov_evolve(const KrylovOperator& op, KrylovWorkspace& ws, int nvec,
                                        const double* phi, int64_t ldphi, int ntimes,
                                        const double* times, double* out, int64_t ldout,
                                        const KrylovParams& p, double* ms_total, double* ms_matvec,
                                        int* matvec_launches);

void launch_kry_deinterleave(const double* z, int64_t ldz, double* out, int64_t ldout, int64_t n,
                             int nvec, hipStream_t s);
void launch_kry_interleave(const double* in, int64_t ldin, double* z, int64_t ldz, int64_t n,
                           int nvec, hipStream_t s);

} // namespace dmk
