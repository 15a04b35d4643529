"""CPU baseline: KronConnections::doTask over all out-patches with BLAS GEMMs through numpy.

Test infrastructure / baseline only (bench.py's cpu_baseline leg). Restates
dmrg/Engine/MatrixVectorKron/KronConnections.h:124-181 -> kronMult (dense x dense,
KronUtil/MatrixDenseOrSparse.h:149-160) -> den_kron_mult method 1 or 2 chosen by
estimate_kron_cost (KronUtil/den_kron_mult.cpp, estimate_kron_cost.cpp), with the GEMMs in the
host BLAS that numpy links (the reference calls BLAS dgemm through GemmR -> kokkos_gemm,
PsimagLite/GemmR.h:36-54). Threads: the BLAS library's own threads (the reference's default is
Threads=1 over out-patches plus a threaded BLAS).
"""
from __future__ import annotations

import numpy as np


def prepare(prob):
    """Groups the blocks per out-patch and picks method 1 or 2 per block (dense blocks:
    method 3 is never cheapest for dense factors)."""
    per_out = {}
    for (I, J, k), (a, b) in sorted(prob.blocks.items(), key=lambda t: (t[0][0], t[0][1], t[0][2])):
        LI, LJ, RI, RJ = prob.L[I], prob.L[J], prob.R[I], prob.R[J]
        m1 = 2.0 * RI * RJ * LJ + 2.0 * LI * LJ * RI
        m2 = 2.0 * LI * LJ * RJ + 2.0 * RI * RJ * LI
        method = 2 if m2 <= m1 else 1
        per_out.setdefault(I, []).append((J, a, b, method))
    return per_out


def apply(prob, prepared, y: np.ndarray) -> np.ndarray:
    """x = H y."""
    off = prob.offsets
    x = np.zeros_like(y)
    for I, terms in prepared.items():
        XI = x[off[I]:off[I + 1]].reshape((prob.R[I], prob.L[I]), order="F")
        for J, a, b, method in terms:
            YJ = y[off[J]:off[J + 1]].reshape((prob.R[J], prob.L[J]), order="F")
            if method == 1:
                XI += (b @ YJ) @ a.T
            else:
                XI += b @ (YJ @ a.T)
    return x
