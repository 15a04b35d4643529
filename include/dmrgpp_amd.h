/*
 * dmrgpp_amd — MI355X (gfx950) superblock Hamiltonian matvec for DMRG++.
 *
 * C ABI of the drop-in replacement for DMRG++'s dmrg/GPUPlugin batched-GEMM plugin
 * (selected in the reference by -DDMRG_BATCHED_GEMM_TYPE=Plugin, dmrg/CMakeLists.txt:8-11, and
 * at run time by SolverOptions=BatchedGemm, BatchedGemmPluginSc.h:137).
 *
 * Conventions shared with the reference plugin:
 *  - a patch I has left size L_I and right size R_I; its vector segment is an R_I x L_I
 *    column-major matrix; segments are concatenated in patch order (InitKronHamiltonian.h:211-216);
 *  - block pointers are indexed I + J*npatches + k*npatches*npatches; A_k(I,J) is L_I x L_J,
 *    B_k(I,J) is R_I x R_J, both column major with the given leading dimensions, NULL meaning a
 *    zero block (BatchedGemmPluginSc.h:96-112, setup_nC.cpp:67-101);
 *  - the product is y_I = sum_{J,k} B_k(I,J) x_J A_k(I,J)^T, i.e. y = sum_k (A_k (x) B_k) x,
 *    with the plain transpose also for complex scalars (apply_Htarget_sparse_not_ready.cpp:384);
 *  - dmk_apply overwrites its output (beta = 0); the caller accumulates (KronMatrix.h:142-151).
 *
 * All functions return 0 on success and a negative status on failure; dmk_last_error() then
 * describes the failure (the reference asserts instead, dmrg_malloc_hip.cpp:17-30).
 * Vectors passed to dmk_apply*, dmk_lanczos_* are device pointers unless a name says _host.
 * A handle owns one HIP stream; calls on one handle are not thread safe.
 */
#ifndef DMRGPP_AMD_H
#define DMRGPP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dmk_handle dmk_handle;

enum dmk_status
{
	DMK_OK             = 0,
	DMK_ERR_ARGUMENT   = -1,
	DMK_ERR_DEVICE     = -2,
	DMK_ERR_UNSUPPORTED = -3,
	DMK_ERR_NUMERICAL  = -4,
};

enum dmk_precision
{
	DMK_PREC_FP64   = 0, /* FP64 MFMA, FP64 accumulate */
	DMK_PREC_FP32   = 1, /* FP32 MFMA on FP32 copies of factors and input, FP32 accumulate */
	DMK_PREC_BF16X3 = 2, /* operands as bf16 (hi, lo) pairs, three BF16 MFMAs, FP32 accumulate */
};

typedef struct dmk_info {
	int64_t n;                 /* superblock dimension */
	int64_t npatches;
	int64_t noperator;
	int64_t nterms;            /* nonzero (I,J,k) terms */
	int64_t nterms_identity;   /* terms whose A or B factor is an identity */
	int64_t stage1_tasks, stage2_tasks, stage1_tiles, stage2_tiles;
	int64_t max_patch_left, max_patch_right;
	int64_t apack_elems, bpack_elems, bx_elems;
	double  flops_algorithmic; /* sum_terms min(method1, method2); identity terms as one GEMM */
	double  flops_plugin;      /* reference plugin count gflops1+gflops2 (all terms dense) */
	double  flops_executed;    /* flops the two stages execute */
	double  flops_stage1;      /* of which the stage-1 launch (B_k X_J products) */
	double  flops_stage2;      /* of which the stage-2 launch (panel and identity products) */
	double  bytes_compulsory;  /* FP64 bytes: non-identity factors + x read + y write */
	double  bytes_bx;          /* FP64 bytes of the stage-1 scratch written and read */
	int32_t precision;
	int32_t pad;
	int64_t split_tasks;       /* stage-2 tasks cut into split-k pieces under the current tuning */
	int64_t ws_elems;          /* split-k partial workspace (elements per vector) */
	int32_t is_complex;        /* 1: created by dmk_create_z (n above is 2 n_complex) */
	int32_t device_blocks;     /* 1: the blocks were device (or managed) memory, scanned and packed
	                              in place on the device (no staging, no upload) */
	int64_t n_complex;         /* complex vector length (equals n for a real handle) */
	double  ms_scan;           /* handle creation, host wall ms: block scan (nonzero, identity;
	                              on the device for device-resident blocks) */
	double  ms_stage;          /*   copy of the uploaded blocks into pinned staging (0 for
	                              device-resident blocks) */
	double  ms_upload;         /*   upload and panel packing on the device */
	double  raw_bytes;         /*   bytes uploaded (0 for device-resident blocks) */
} dmk_info;

/* Launch shape of a handle: output tile (64 or 128 each way; FP64 and FP32 also 96 x 96,
 * 96 x 64, 64 x 96 and the 32-wide tiles 32 x 32, 32 x 64, 64 x 32, 32 x 128, 128 x 32) of the
 * stage-1 and stage-2 grouped-GEMM launches, the split-k
 * granularity of stage 2 (target contraction length per piece, 0 = never split; pieces are
 * summed in a fixed order, so results stay deterministic),
 * the workgroup order (0 = by decreasing task cost, 1 = XCD-aware regions), the MFMA block form
 * and the load pipeline depth (1 or 2 K chunks in flight from global memory to LDS; 0 means 1). */
typedef struct dmk_tuning {
	int32_t stage1_tile_m, stage1_tile_n;
	int32_t stage2_tile_m, stage2_tile_n;
	int32_t splitk;
	int32_t xcd_order;
	int32_t mfma32;            /* 1 = 32x32 MFMA blocks (FP32: 32x32x2, BF16x3: 32x32x16), 0 = 16x16 */
	int32_t prefetch;          /* K chunks in flight during the MFMAs of a chunk, in a ring of
	                              prefetch + 1 LDS stages: 1 or 2 (0 = 1) */
	int32_t kchunk;            /* K chunk depth: 1 (0 = 1) = 16 (FP64) / 32 (FP32, BF16x3) elements
	                              per LDS stage, 2 = twice that (FP64 / FP32, tiles with
	                              tile_m + tile_n <= 128 in both stages) */
	int32_t persist;           /* k = 1..8: stage 2 as one balanced persistent launch of k
	                              workgroups per CU (at most as many as stay resident), each given
	                              an equal share of the stage's K chunks (tiles cut between
	                              workgroups are summed in a fixed order afterwards; splitk is
	                              then ignored); 0: one tile per workgroup */
} dmk_tuning;

/* Replaces BatchedGemm<double>::BatchedGemm(noperator, npatches, left_patch_size,
 * right_patch_size, Amatrix, ld_Amatrix, Bmatrix, ld_Bmatrix) — dmrg/GPUPlugin/BatchedGemm.h:21-28,
 * BatchedGemm.cpp:8-70. Blocks are host pointers (scanned on the host, staged, uploaded and
 * packed on the device) or, all of them, device / managed memory pointers, as from a caller that
 * keeps its block operators in HBM (scanned and packed on the device, read in place, nothing
 * uploaded; dmk_info.device_blocks = 1). Mixing host and device blocks is an error. Either way
 * the blocks may be freed after the call (as setup_sparse_batch.cpp:281-343 copies them). */
int dmk_create_d(size_t noperator, size_t npatches, const size_t* left_patch_size,
                 const size_t* right_patch_size, const double* const* Amatrix,
                 const size_t* ld_Amatrix, const double* const* Bmatrix, const size_t* ld_Bmatrix,
                 dmk_handle** handle);

/* Replaces BatchedGemm<std::complex<double>>::BatchedGemm (the same constructor,
 * BatchedGemm.h:21-28, instantiated at BatchedGemm.cpp:119). Blocks are std::complex<double>
 * column-major host arrays passed as interleaved (re, im) doubles, leading dimensions in complex
 * elements. The handle holds the real embedding of H on 2n reals (real parts of all patches,
 * then imaginary parts); blocks with zero imaginary parts cost nothing extra. The embedding is
 * built on the host: device-resident blocks are rejected. */
int dmk_create_z(size_t noperator, size_t npatches, const size_t* left_patch_size,
                 const size_t* right_patch_size, const double* const* Amatrix,
                 const size_t* ld_Amatrix, const double* const* Bmatrix, const size_t* ld_Bmatrix,
                 dmk_handle** handle);

/* Replaces BatchedGemm<T>::~BatchedGemm — BatchedGemm.h:33, BatchedGemm.cpp:72-79. */
void dmk_destroy(dmk_handle* handle);

/* Replaces BatchedGemm<double>::apply_Htarget(vin, vout) — BatchedGemm.h:30-31,
 * BatchedGemm.cpp:80-113 -> apply_Htarget_sparse. vout = H vin, device pointers, n elements.
 * Asynchronous on the handle's stream; dmk_sync waits for it. */
int dmk_apply_d(dmk_handle* handle, const double* vin, double* vout);

/* Same on host vectors (the BatchedGemmPluginSc::matrixVector round trip,
 * BatchedGemmPluginSc.h:139-161): copies vin in, applies, copies vout out, synchronous. */
int dmk_apply_host_d(dmk_handle* handle, const double* vin, double* vout);

/* Replaces BatchedGemm<std::complex<double>>::apply_Htarget (BatchedGemm.h:30-31): vout = H vin
 * on interleaved complex device vectors of n_complex elements. Also valid on a real handle
 * (real H times a complex vector: real and imaginary parts go through one two-column launch). */
int dmk_apply_z(dmk_handle* handle, const double* vin, double* vout);

/* The apply_Htarget(Kokkos::View<const KokkosScalar*, Kokkos::SharedSpace> vin,
 * Kokkos::View<KokkosScalar*, Kokkos::SharedSpace> vout) of BatchedGemm.h:30-31 on the views'
 * data pointers: vout = H vin with vin and vout in any memory — device, managed
 * (Kokkos::SharedSpace is hipMallocManaged memory on ROCm) or host. Device and managed vectors
 * are read and written in place; host vectors go through device copies. Real or complex by the
 * handle (n or 2 n_complex doubles). Synchronous: returns after vout is written. */
int dmk_apply_shared(dmk_handle* handle, const void* vin, void* vout);

/* vout = H vin like dmk_apply_d (one real vector, device memory, beta = 0, the handle's current
 * precision), allocating nothing: no hipMalloc / hipFree and no allocating library call, and no
 * hipMemsetAsync. No counterpart in the reference; it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid allocation). The matvec's
 * scratch (stage-1 products, split-k partials, input panels) lives in a static device array of
 * DMK_NOALLOC_SCRATCH_BYTES per GPU, held from the library's first use on that GPU until the
 * process ends; a handle whose matvec needs more in its current precision is rejected with
 * DMK_ERR_ARGUMENT before any work is enqueued (dmk_noalloc_scratch_bytes tells). All calls on
 * one GPU share that array: calls that overlap in time on one GPU (several streams or threads)
 * race and must be serialised by the caller. Enqueued on the handle's stream; returns before
 * completion (dmk_sync waits). The handle must have been created, and its precision set, before
 * (both allocate). */
#define DMK_NOALLOC_SCRATCH_BYTES (64ull << 20)
int dmk_apply_noalloc_d(dmk_handle* handle, const double* vin, double* vout);

/* Scratch bytes one matvec of the handle needs in its current precision (what
 * dmk_apply_noalloc_d compares with DMK_NOALLOC_SCRATCH_BYTES). Host-only. */
int dmk_noalloc_scratch_bytes(const dmk_handle* handle, size_t* bytes);

/* Sector ordering of the caller's vectors: map[ip] (host, n entries, a permutation of 0..n-1)
 * is the sector index of patch-order element ip, i.e. permutationInverse[i + j*nl] - offset for
 * the ip = vstart[patch] + iright + ileft*sizeRight of InitKronHamiltonian::copyIn
 * (InitKronHamiltonian.h:158-221). Uploaded once per handle. */
int dmk_set_sector_map(dmk_handle* handle, const int64_t* map);

/* vout += H vin on sector-ordered device vectors (KronMatrix::matrixVectorProductImpl with the
 * plugin enabled, KronMatrix.h:137-151: copyIn, apply, xout += tmp, copyOut), with the gather
 * and scatter on the device instead of the host. Needs dmk_set_sector_map. Asynchronous. */
int dmk_matvec_sector_d(dmk_handle* handle, const double* vin, double* vout);

/* nvec right-hand sides in one launch pair: vout[v*ldout..] = H vin[v*ldin..] (beta = 0) or
 * += (beta = 1). No reference counterpart (ParallelTriDiag.h:132-161 tridiagonalizes one vector
 * at a time); used by the batched Krylov time evolution. */
int dmk_apply_multi_d(dmk_handle* handle, int nvec, const double* vin, size_t ldin, double* vout,
                      size_t ldout, int beta);

/* Precision of the A_k (x) B_k products (enum dmk_precision). No reference counterpart: the
 * plugin is FP64/complex-FP64 only (GPUPluginConfig.h.in). */
int dmk_set_precision(dmk_handle* handle, int precision);

int dmk_get_info(const dmk_handle* handle, dmk_info* info);

/* Tuning of the launches. No reference counterpart (the Kokkos path hard-codes its shape,
 * BatchedGemmKokkos.h:110-130). dmk_set_tuning rebuilds the launch tables. A tuning belongs to
 * the handle's current precision mode: dmk_set_precision switches to the tuning last set or
 * autotuned for the selected mode (else the cached or the default one). */
int dmk_get_tuning(const dmk_handle* handle, dmk_tuning* tuning);
int dmk_set_tuning(dmk_handle* handle, const dmk_tuning* tuning);

/* Tunes the launch shape on this handle's problem (iters matvecs per candidate, in the current
 * precision, on the device vectors vin/vout): coordinate descent from the mode's default over the
 * MFMA form, load depth, stage-2 tile, split-k, tile order and stage-1 tile, two passes, keeping
 * the fastest. The choice is cached in the
 * process under the block-size histogram of the problem, and handles created later for a
 * problem with the same histogram start from it. chosen may be NULL. */
int dmk_autotune_d(dmk_handle* handle, const double* vin, double* vout, int iters,
                   dmk_tuning* chosen);

/* Waits for all work queued on the handle's stream. */
int dmk_sync(dmk_handle* handle);

/* Average device time per call (ms) of stage 1, stage 2 and the whole matvec, over iters calls
 * timed with HIP events on the handle's stream (the plugin prints GFLOP/s with idebug=1,
 * apply_Htarget_sparse_not_ready.cpp:451-463). */
int dmk_profile_apply_d(dmk_handle* handle, const double* vin, double* vout, int iters,
                        double* ms_stage1, double* ms_stage2, double* ms_total);

/* Parameters of the device-resident Lanczos ground-state solver. Defaults (from
 * dmk_lanczos_default_params) are those of PsimagLite ParametersForSolver
 * (ParametersForSolver.h:97-128): 200 steps, at least 4, tolerance 1e-12, no re-orthogonalization. */
typedef struct dmk_lanczos_params {
	int32_t max_steps;
	int32_t min_steps;
	double  tolerance;        /* stop when the lowest Ritz value moves less than this */
	int32_t reortho;          /* 1: full Gram-Schmidt each step (LanczosVectors.h:243-271) */
	int32_t precision;        /* matvec precision of the first run (enum dmk_precision) */
	int32_t refine;           /* after a non-FP64 run: 1 = restart Lanczos in FP64 from its Ritz
	                             vector; 2 = FP64 Davidson (diagonal correction) from it to a
	                             residual of sqrt(refine_tolerance) */
	int32_t refine_min_steps;
	double  refine_tolerance;
	int32_t graph;            /* 1: each step is one replay of a captured HIP graph (default) */
	int32_t pad;
	double  stop_residual;    /* > 0: also stop once the Ritz residual estimate |beta_{j+1} s_j|
	                             of the unit lowest Ritz vector drops below it (0 = off) */
	double  switch_noise;     /* a refined non-FP64 first run hands over to the FP64 refinement
	                             once its residual estimate drops below switch_noise * ||H||
	                             (0 = the mode's default: FP32 3e-5 before the Davidson
	                             refinement, 1e-6 before the Lanczos restart; BF16x3 1e-4) */
} dmk_lanczos_params;

typedef struct dmk_lanczos_result {
	double  energy;           /* lowest eigenvalue (after refinement when it ran) */
	double  residual;         /* ||H z - E z|| / ||z|| evaluated with the FP64 matvec */
	int32_t steps;            /* Lanczos steps of the last run */
	int32_t matvecs;          /* all matvecs of the solve (both runs) */
	double  energy_first;     /* energy of the first run, in its own precision */
	int32_t steps_first;
	int32_t matvecs_first;
	int32_t precision_first;
	int32_t pad;
	double  ms_total;         /* wall time of the solve */
	double  ms_matvec;        /* device time inside the matvec launches */
} dmk_lanczos_result;

int dmk_lanczos_default_params(dmk_lanczos_params* params);

/* Lowest eigenpair of the handle's H: LanczosSolver::computeAllStatesBelow for the ground state
 * (PsimagLite/LanczosSolver.h:68-88) over LanczosCore::decomposition (LanczosCore.h:147-272) and
 * LanczosVectors::oneStepDecomposition / excitedVector (LanczosVectors.h:192-237), with the
 * Lanczos vectors resident on the device. init and z are device vectors of n elements;
 * z receives the (unnormalized Ritz combination of unit Lanczos vectors) ground vector. */
int dmk_lanczos_ground_d(dmk_handle* handle, const double* init, double* z,
                         const dmk_lanczos_params* params, dmk_lanczos_result* result);

/* Lowest eigenvalue and unit eigenvector of the m x m symmetric tridiagonal matrix with diagonal
 * a[0..m) and off-diagonal b[1..m) (T(i,i+1) = b[i+1]), as TridiagonalMatrix::buildDenseMatrix
 * lays it out (TridiagonalMatrix.h:149-163); replaces its LAPACK dsyev call (Matrix.cpp:15-45)
 * for the ground state. evec may be NULL. Host arrays. */
int dmk_tridiag_lowest(int m, const double* a, const double* b, double* eval, double* evec);

/* All eigenpairs of the same tridiagonal matrix, ascending: evals[0..m), evecs column major
 * m x m (column k belongs to evals[k]). Replaces PsimagLite::diag(T, eigs, 'V') of the Krylov
 * step (TimeVectorsKrylov.h:224-225). Host arrays. */
int dmk_tridiag_eigen(int m, const double* a, const double* b, double* evals, double* evecs);

/* Krylov time step. Defaults (dmk_krylov_default_params) are the "Tridiag" solver defaults of
 * ParametersForSolver (ParametersForSolver.h:97-128) used by ParallelTriDiag.h:146-147. */
enum dmk_krylov_stop
{
	DMK_KRYLOV_STOP_ENERGY = 0, /* lowest Ritz value moved less than tolerance (the reference) */
	DMK_KRYLOV_STOP_EXPM   = 1, /* beta_m |[exp(-i T tau) e_0]_{m-1}| < tolerance (relative) */
};

typedef struct dmk_krylov_params {
	int32_t max_steps;
	int32_t min_steps;
	double  tolerance;
	int32_t stop;            /* enum dmk_krylov_stop */
	int32_t precision;       /* matvec precision (enum dmk_precision); recurrence in FP64 */
	double  time_direction;  /* d in exp(-i (H - E) tau d), TimeVectorsKrylov.h:287 */
	double  energy_shift;    /* E (the reference passes E_g), TimeVectorsKrylov.h:291 */
} dmk_krylov_params;

typedef struct dmk_krylov_result {
	int32_t steps;           /* Krylov dimension of this target */
	int32_t breakdown;       /* 1: an invariant subspace was reached */
	double  norm;            /* |phi| */
	double  error_estimate;  /* beta_m |[exp(-i T tau_max) e_0]_{m-1}| |phi| */
} dmk_krylov_result;

typedef struct dmk_krylov_stats {
	double  ms_total;        /* wall time of the call */
	double  ms_matvec;       /* device time inside the batched matvec launches */
	int32_t matvec_launches; /* batched matvecs (one per Krylov step, all targets together) */
	int32_t pad;
} dmk_krylov_stats;

int dmk_krylov_default_params(dmk_krylov_params* params);

/* out_{t,s} = exp(-i (H - E) tau_s d) phi_t for nvec target vectors phi_t and ntimes (<= 8)
 * times tau_s: the Krylov time step of TimeVectorsKrylov::calcTimeVectors
 * (TimeVectorsKrylov.h:165-322) over ParallelTriDiag::triDiag (ParallelTriDiag.h:132-161) and
 * KrylovHelper::calcR (KrylovHelper.h, default abridgement). The reference tridiagonalizes one
 * vector per call; here all nvec targets share one batched matvec launch per Krylov step, each
 * keeping the Krylov dimension its own stop test gives. phi: interleaved complex device vectors
 * at ldphi complex elements; out: target t, time s at out + 2*ldout*(t*ntimes + s); times on the
 * host. results (nvec entries) and stats may be NULL. Works on real and complex handles. */
int dmk_krylov_evolve_z(dmk_handle* handle, int nvec, const double* phi, size_t ldphi, int ntimes,
                        const double* times, double* out, size_t ldout,
                        const dmk_krylov_params* params, dmk_krylov_result* results,
                        dmk_krylov_stats* stats);

/* Eigen-decomposition of a block-diagonal symmetric matrix on the GPU: replaces
 * DiagBlockDiagMatrix::diagonalise (DiagBlockDiagMatrix.h:73-88), i.e. PsimagLite::diag(block,
 * eigs, 'V') on every symmetry block (BlockDiagonalMatrix.h:317-322), which the truncation calls
 * on the reduced density matrix (DensityMatrixLocal.h:201-204). Block b is n[b] x n[b], column
 * major, in host memory at blocks[b] with leading dimension n[b] (0 <= n[b] <= 32768); only its
 * upper triangle is read. It is overwritten by its orthonormal eigenvectors (column j belongs to
 * eigenvalue j), and its eigenvalues, ascending, go to eigs[o_b .. o_b + n[b]) with
 * o_b = n[0] + ... + n[b-1]. rocSOLVER dsyevd on the device: blocks of similar size (down to
 * half the largest of their class, at most 16) as one strided batch, each padded to the class's
 * size with a diagonal below its spectrum; classes and single blocks up to four at once on
 * their own streams and host threads (largest first); synchronous. The reference's
 * EnforcePhase (EnforcePhase.h:23-39) multiplies each eigenvector by the sign of a norm, which is
 * +1 for real entries, so no sign convention is imposed here either. */
int dmk_density_diag_d(int nblocks, const int64_t* n, double* const* blocks, double* eigs);

/* The same on blocks already in device memory, as the density matrices of a truncation that
 * forms them on the device from the device-resident ground state (DensityMatrixLocal.h:201-204
 * builds them on the host): block b is n[b] x n[b], compact column major at blocks[b] (device;
 * read, not modified; only its upper triangle is used), its eigenvectors go to the host array
 * vecs[b] (n[b] x n[b], column major) and its eigenvalues, ascending, to eigs as above.
 * Synchronous. */
int dmk_density_diag_dev_d(int nblocks, const int64_t* n, const double* const* blocks,
                           double* const* vecs, double* eigs);

/* The SVD form of the same step: replaces DensityMatrixSvd::diag (DensityMatrixSvd.h:517-554),
 * whose ParallelSvd::doTask (DensityMatrixSvd.h:382-410) calls PsimagLite::Svd with jobz 'A' on
 * every group's wave-function matrix M_b (the group's states x its partner states, weighted by
 * the square roots of the target weights); the eigenvectors of the density matrix M_b M_b^T are
 * the left singular vectors and its eigenvalues the squared singular values. Block b is
 * rows[b] x cols[b], column major, at blocks[b] with leading dimension rows[b] (both sizes
 * 0..32768); it is not modified. u[b] receives all rows[b] x rows[b] left singular vectors
 * (column j belongs to singular value j; the columns past min(rows, cols) span the null space,
 * and a block with no columns gets the identity), s[o_b .. o_b + min(rows[b], cols[b])) the
 * singular values in descending order, o_b = sum of min(rows, cols) over the earlier blocks.
 * rocSOLVER dgesvd on the device, up to four blocks at once; synchronous. */
int dmk_density_svd_d(int nblocks, const int64_t* rows, const int64_t* cols,
                      const double* const* blocks, double* const* u, double* s);

/* Variable-size batched GEMM: replaces dmrg_Xgemm_vbatch<T> (GPUPlugin/dmrg_vbatch.h:118-133,
 * dmrg_vbatch.cpp:129-652), the free function the reference plugin and its test_vbatch driver
 * call. For every entry e of every group g: C_e = alpha op(A_e) op(B_e) + beta C_e, column major,
 * op from transa/transb ('N', 'T', or 'C' = conjugate transpose as in BLAS). m/n/k/lda/ldb/ldc
 * and group_size have group_count entries; a/b/c hold one device pointer per entry, in group
 * order (host arrays). C is not read when beta == 0. Runs on the MFMA grouped kernel (FP64);
 * synchronous. Used by the sweep driver for the wave-function transformation GEMMs
 * (WftAccelBlocks.h:162-322). */
int dmk_gemm_vbatch_d(char transa, char transb, const int* m_array, const int* n_array,
                      const int* k_array, double alpha, const double* const* a_array,
                      const int* lda_array, const double* const* b_array, const int* ldb_array,
                      double beta, double* const* c_array, const int* ldc_array,
                      size_t group_count, const int* group_size);

/* Complex instantiation (dmrg_vbatch.cpp template for std::complex<double>): matrices are
 * interleaved (re, im), leading dimensions in complex elements; alpha and beta point to
 * (re, im) pairs. */
int dmk_gemm_vbatch_z(char transa, char transb, const int* m_array, const int* n_array,
                      const int* k_array, const double* alpha, const double* const* a_array,
                      const int* lda_array, const double* const* b_array, const int* ldb_array,
                      const double* beta, double* const* c_array, const int* ldc_array,
                      size_t group_count, const int* group_size);

/* diag(H) into diag (device, n elements): sum over the diagonal-patch terms (I, I, k) of
 * diag(A_k(I,I)) (x) diag(B_k(I,I)). No reference counterpart (used by the Davidson
 * corrections below). Synchronous. */
int dmk_diagonal_d(dmk_handle* handle, double* diag);

/* Learned Davidson preconditioner: a per-element MLP (4 features -> 64 -> 64 -> 1, tanh, FP32)
 * with its training set on the device. No reference counterpart (the reference has no
 * Davidson, ParametersDmrgSolver.h:154-158). Parameters are DMK_MLP_PARAMS floats: W1 (64x4
 * row major), b1 (64), W2 (64x64 row major), b2 (64), W3 (64), b3 (1). */
#define DMK_MLP_PARAMS 4545
#define DMK_MLP_FEATURES 4
typedef struct dmk_mlp dmk_mlp;
int  dmk_mlp_create(uint64_t seed, int64_t sample_capacity, dmk_mlp** mlp);
void dmk_mlp_destroy(dmk_mlp* mlp);
int  dmk_mlp_get_params(const dmk_mlp* mlp, float* params);
int  dmk_mlp_set_params(dmk_mlp* mlp, const float* params);
/* y[s] = MLP(x[s]) for count samples; x (count x 4, row major) and y are device arrays. */
int  dmk_mlp_forward(dmk_mlp* mlp, const float* x, int64_t count, float* y);
/* Gradient (host, DMK_MLP_PARAMS floats) of the mean squared error (loss, host) over the
 * device batch (x, t). */
int  dmk_mlp_gradient(dmk_mlp* mlp, const float* x, const float* t, int64_t count, float* grad,
                      double* loss);
/* Appends count device samples (x, t) to the training set (truncated at its capacity). */
int  dmk_mlp_add_samples(dmk_mlp* mlp, const float* x, const float* t, int64_t count);
int  dmk_mlp_samples(const dmk_mlp* mlp, int64_t* count);
int  dmk_mlp_clear_samples(dmk_mlp* mlp);
/* iters Adam steps (lr) over minibatches of batch consecutive training samples; loss_last is
 * the mean squared error of the last minibatch, loss_all of the whole set after training (either
 * may be NULL). */
int  dmk_mlp_train(dmk_mlp* mlp, int iters, int64_t batch, float lr, double* loss_last,
                   double* loss_all);

enum dmk_precond
{
	DMK_PRECOND_NONE = 0, /* t = r */
	DMK_PRECOND_DIAG = 1, /* Olsen: t = M^-1 (eps u - r), M = diag H - theta, t orthogonal to u */
	DMK_PRECOND_MLP  = 2, /* t = MLP(features) in units of rms(r) / rms(diag H - theta) */
};

typedef struct dmk_davidson_params {
	int32_t max_matvecs;
	int32_t max_basis;       /* restart size (3..128) */
	double  tolerance;       /* stop when ||H u - theta u|| < tolerance for the unit Ritz vector */
	int32_t precond;         /* enum dmk_precond */
	int32_t harvest;         /* 1: append training pairs from this run to the MLP's set */
	int32_t harvest_samples; /* sampled elements per recorded iterate */
	int32_t harvest_iters;   /* recorded iterates at most */
} dmk_davidson_params;

typedef struct dmk_davidson_result {
	double  energy;
	double  residual;        /* ||H z - E z|| with a fresh FP64 matvec (z unit) */
	int32_t iterations;
	int32_t matvecs;         /* including the residual check */
	int32_t restarts;
	int32_t converged;
	int64_t samples;         /* training pairs appended */
	double  ms_total;
	double  ms_matvec;
	double  ms_precond;
} dmk_davidson_result;

int dmk_davidson_default_params(dmk_davidson_params* params);

/* Lowest eigenpair by Davidson with the chosen correction (FP64 matvecs, device-resident basis);
 * mlp is needed for DMK_PRECOND_MLP and for harvest, else may be NULL. z receives the unit
 * Ritz vector. */
int dmk_davidson_ground_d(dmk_handle* handle, const double* init, double* z,
                          const dmk_davidson_params* params, dmk_mlp* mlp,
                          dmk_davidson_result* result);

/* ---- The matvec facade of a superblock sector: MatrixVectorKron
 * (Engine/MatrixVectorKron/MatrixVectorKron.h:113-166) as MatrixVector::create selects it
 * (Engine/MatrixVector.hpp:133-150). x += op(y), op = H, or L^dagger L with ldaggerl (the
 * LdaggerL option, MatrixVectorKron.h:151-160) where the blocks describe L. A superblock whose
 * rank is at most max_rank_stored (MaxMatrixRankStored, ParametersDmrgSolver.h:326; the
 * reference's default 0 never stores) is assembled once into CSR on the device and every call
 * is one sparse product (two with ldaggerl; stored ranks are limited to 16384); otherwise the
 * Kron plugin computes the product, and L^dagger comes from a second plugin handle built from
 * the adjoint blocks (the reference runs that adjoint on its CPU Kron path). Each call is timed
 * on the host and accumulated (DeltaClock matrixVectorProduct, MatrixVectorKron.h:139-166).
 * The blocks are host memory (the stored matrix and the adjoint blocks are assembled on the
 * host). x and y are device vectors of n elements (interleaved complex for _z); synchronous. */
typedef struct dmk_facade dmk_facade;

typedef struct dmk_facade_info {
	int64_t n;               /* rank (vector length in elements of T) */
	int64_t nnz;             /* stored path: nonzeros of H (0 on the Kron path) */
	int64_t matvecs;         /* calls so far */
	double  ms_matvec_total; /* accumulated wall time of the calls (DeltaClock) */
	int32_t stored;          /* 1: stored CSR path, 0: Kron plugin */
	int32_t ldaggerl;
	int32_t is_complex;
	int32_t pad;
} dmk_facade_info;

int  dmk_facade_create_d(size_t noperator, size_t npatches, const size_t* left_patch_size,
                         const size_t* right_patch_size, const double* const* Amatrix,
                         const size_t* ld_Amatrix, const double* const* Bmatrix,
                         const size_t* ld_Bmatrix, int64_t max_rank_stored, int ldaggerl,
                         dmk_facade** facade);
int  dmk_facade_create_z(size_t noperator, size_t npatches, const size_t* left_patch_size,
                         const size_t* right_patch_size, const double* const* Amatrix,
                         const size_t* ld_Amatrix, const double* const* Bmatrix,
                         const size_t* ld_Bmatrix, int64_t max_rank_stored, int ldaggerl,
                         dmk_facade** facade);
int  dmk_facade_matvec(dmk_facade* facade, double* x, const double* y);
int  dmk_facade_get_info(const dmk_facade* facade, dmk_facade_info* info);
void dmk_facade_destroy(dmk_facade* facade);

/* Device memory helpers for callers without their own allocator. */
int dmk_device_alloc(size_t bytes, void** ptr);
int dmk_device_free(void* ptr);
int dmk_memcpy_h2d(void* dst, const void* src, size_t bytes);
int dmk_memcpy_d2h(void* dst, const void* src, size_t bytes);
int dmk_device_count(int* count);

/* Books of the library's allocation caches (no reference counterpart; diagnostics): for the
 * device pool (host = 0) or the pinned host pool (host = 1), out[0..6) = bytes in use, bytes
 * cached, allocator calls, buffers returned to the allocator, requests served from the cache,
 * and releases of pointers the cache did not hold as in use (a stale pointer in the library:
 * 0 in a correct run). Host-only bookkeeping: no GPU call. */
int dmk_pool_counters(int host, uint64_t* out);

/* dst[i] = src[i], i < n, by a plain grid-stride copy kernel on the default stream (synchronous).
 * No reference counterpart: it calibrates HBM byte counters (FETCH_SIZE / WRITE_SIZE) against a
 * kernel of known traffic. */
int dmk_device_copy_d(double* dst, const double* src, size_t n);

const char* dmk_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* DMRGPP_AMD_H */
