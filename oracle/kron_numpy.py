"""CPU baseline and full-size FP64 oracle: the reference's CPU Kron matvec with host BLAS GEMMs.

Test infrastructure only (tests/, bench.py's cpu_baseline leg and its FP64 oracle energy).
Restates the reference's CPU path of H y:

* KronMatrix::matrixVectorProductImpl -> KronConnections::doTask over every out-patch I, summing
  over in-patches J and connections k (dmrg/Engine/MatrixVectorKron/KronConnections.h:124-181);
* every factor block is stored dense when nnz >= threshold * rows * cols and as CSR otherwise
  (ArrayOfMatStruct.h:300-367, threshold = DenseSparseThreshold, default 0.2,
  ParametersDmrgSolver.h:246);
* kronMult dispatches on (dense, sparse) (KronUtil/MatrixDenseOrSparse.h:133-200):
  - dense x dense: den_kron_mult method 1 ((B Y) A^T) or 2 (B (Y A^T)), the cheaper by
    estimate_kron_cost (KronUtil/estimate_kron_cost.cpp:17-118);
  - a CSR identity factor makes the term one GEMM: X += B Y when A = I
    (csr_den_kron_mult.cpp:431-460), X += Y A^T when B = I (den_csr_kron_mult.cpp:492-520);
  - other CSR factors: the same two association orders with sparse x dense products
    (den_csr_kron_mult.cpp, csr_den_kron_mult.cpp, csr_kron_mult.cpp), the order again by
    estimate_kron_cost's nnz-based counts.
The dense GEMMs run in the BLAS that numpy links (blas_library() names it, with its thread
count); the CSR products in scipy.sparse. The reference calls BLAS dgemm through GemmR ->
kokkos_gemm (PsimagLite/GemmR.h:36-54).

lanczos() restates LanczosCore::decomposition (PsimagLite/LanczosCore.h:147-272) with this
matvec: it gives the FP64 oracle energy of superblocks too large for the C++ restatement's plain
loops. tests/test_oracle.py pins apply() against the C++ oracle's plugin_apply and kron_form, and
lanczos() against the C++ oracle's lanczos.
"""
from __future__ import annotations

import numpy as np

DENSE_SPARSE_THRESHOLD = 0.2  # ParametersDmrgSolver.h:246


def blas_library() -> str:
    """Name, version and thread count of the BLAS numpy's GEMMs run in."""
    try:
        from threadpoolctl import threadpool_info

        for d in threadpool_info():
            if d.get("user_api") == "blas":
                return (f"{d.get('internal_api', 'blas')} {d.get('version', '')} "
                        f"({d.get('num_threads', '?')} threads)").strip()
    except Exception:
        pass
    return "numpy BLAS (unknown)"


def blas_threads() -> int:
    try:
        from threadpoolctl import threadpool_info

        return max([d.get("num_threads", 1) for d in threadpool_info()
                    if d.get("user_api") == "blas"] or [1])
    except Exception:
        return 1


def _is_identity(m: np.ndarray) -> bool:
    """csr_is_eye / den_is_eye: square with ones on the diagonal and zeros elsewhere."""
    if m.shape[0] != m.shape[1]:
        return False
    d = np.diagonal(m)
    if not np.all(d == 1):
        return False
    return np.count_nonzero(m) == m.shape[0]


class _Factor:
    """One factor block as the reference stores it: dense, CSR or CSR identity."""

    __slots__ = ("kind", "dense", "csr", "nnz", "rows", "cols")

    def __init__(self, m: np.ndarray, threshold: float):
        self.rows, self.cols = m.shape
        self.nnz = int(np.count_nonzero(m))
        self.dense = None
        self.csr = None
        if self.nnz >= threshold * self.rows * self.cols:
            self.kind = "dense"
            self.dense = m
        elif _is_identity(m):
            self.kind = "eye"
        else:
            from scipy import sparse

            self.kind = "csr"
            self.csr = sparse.csr_matrix(m)


def _method(fa: _Factor, fb: _Factor, discount: float = 1.0) -> int:
    """estimate_kron_cost's method 1 or 2 for X += B Y A^T (A: L_I x L_J, B: R_I x R_J).
    Method 3 (explicit Kronecker of the nonzeros) is evaluated as method 1 or 2: the sum is
    the same, only the order of the additions differs."""
    dense_a = fa.nnz == fa.rows * fa.cols
    dense_b = fb.nnz == fb.rows * fb.cols
    da = discount if dense_a else 1.0
    db = discount if dense_b else 1.0
    # estimate_kron_cost.cpp:44-72 with nrow_A = L_I, ncol_A = L_J, nrow_B = R_I, ncol_B = R_J.
    m1 = db * 2.0 * fb.nnz * fa.cols + da * 2.0 * fa.nnz * fb.rows
    m2 = da * 2.0 * fa.nnz * fb.cols + db * 2.0 * fb.nnz * fa.rows
    return 1 if m1 < m2 else 2 if m2 < m1 else 1


def prepare(prob, threshold: float = DENSE_SPARSE_THRESHOLD):
    """Classifies every factor block (dense / CSR / identity) and picks each term's order.
    Terms are grouped per out-patch in (J, k) order, as KronConnections walks them."""
    per_out = {}
    for (I, J, k), (a, b) in sorted(prob.blocks.items(), key=lambda t: (t[0][0], t[0][1], t[0][2])):
        fa, fb = _Factor(a, threshold), _Factor(b, threshold)
        if fa.nnz == 0 or fb.nnz == 0:
            continue
        per_out.setdefault(I, []).append((J, fa, fb, _method(fa, fb)))
    return per_out


def _left(f: _Factor, m: np.ndarray) -> np.ndarray:
    """f @ m for a dense or CSR factor."""
    return f.dense @ m if f.kind == "dense" else f.csr @ m


def _right_t(m: np.ndarray, f: _Factor) -> np.ndarray:
    """m @ f^T for a dense or CSR factor."""
    return m @ f.dense.T if f.kind == "dense" else (f.csr @ m.T).T


def apply(prob, prepared, y: np.ndarray) -> np.ndarray:
    """x = H y (KronMatrix::matrixVectorProductImpl with xout = 0)."""
    off = prob.offsets
    x = np.zeros(prob.n, dtype=np.result_type(prob.dtype, y.dtype))
    for I, terms in prepared.items():
        XI = x[off[I]:off[I + 1]].reshape((prob.R[I], prob.L[I]), order="F")
        for J, fa, fb, method in terms:
            YJ = y[off[J]:off[J + 1]].reshape((prob.R[J], prob.L[J]), order="F")
            if fa.kind == "eye" and fb.kind == "eye":
                XI += YJ
            elif fa.kind == "eye":
                XI += _left(fb, YJ)
            elif fb.kind == "eye":
                XI += _right_t(YJ, fa)
            elif method == 1:
                XI += _right_t(_left(fb, YJ), fa)
            else:
                XI += _left(fb, _right_t(YJ, fa))
    return x


def lanczos(prob, prepared, init: np.ndarray, max_steps: int = 200, min_steps: int = 4,
            tol: float = 1e-12):
    """Lowest eigenvalue by LanczosCore::decomposition (LanczosCore.h:147-272) with apply():
    V0 = init/|init|; each step V_{j+1} = H V_j - a_j V_j - b_j V_{j-1}, b_{j+1} = |V_{j+1}|;
    stop once the lowest eigenvalue of T_{j+1} moved by less than tol after at least min_steps
    steps (LanczosCore.h:218-249), or after max_steps. Returns (energy, steps, matvecs)."""
    from scipy.linalg import eigvalsh_tridiagonal

    n = prob.n
    maxn = max(1, min(max_steps, n))
    v0 = np.asarray(init, dtype=np.float64)
    v0 = v0 / np.sqrt(np.dot(v0, v0))
    # First step (LanczosCore.h:165-208): no stop test.
    v1 = apply(prob, prepared, v0)
    a = [float(np.dot(v0, v1))]
    v1 -= a[0] * v0
    b = [0.0, float(np.sqrt(np.dot(v1, v1)))]
    if b[1] > 0:
        v1 /= b[1]
    matvecs, eold, steps = 1, 0.0, maxn
    for j in range(1, maxn):
        # oneStepDecomposition (LanczosVectors.h:203-237).
        v2 = apply(prob, prepared, v1)
        matvecs += 1
        a.append(float(np.dot(v1, v2)))
        v2 -= a[j] * v1
        v2 -= b[j] * v0
        beta = float(np.sqrt(np.dot(v2, v2)))
        if beta > 0:
            v2 /= beta
        e = float(eigvalsh_tridiagonal(np.array(a), np.array(b[1:j + 1]),
                                       select="i", select_range=(0, 0))[0])
        delta = abs(e - eold)
        eold = e
        if (delta < tol or j == maxn - 1) and (n <= 4 or j >= min_steps):
            steps = j + 1
            break
        b.append(beta)
        v0, v1 = v1, v2
    a = np.array(a[:steps])
    off = np.array(b[1:steps])
    energy = float(eigvalsh_tridiagonal(a, off, select="i", select_range=(0, 0))[0])
    return energy, steps, matvecs
