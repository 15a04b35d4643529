"""The matvec facade (MatrixVectorKron, Engine/MatrixVectorKron/MatrixVectorKron.h:113-166):
x += op(y) with op = H or L^dagger L (the LdaggerL option), through the stored CSR path when the
rank is at most max_rank_stored and through the Kron plugin otherwise, with the accumulated
matvec clock (DeltaClock). Each path and mode against the oracle's explicit Kron form of the
same problem (kron_form, pinned in test_oracle.py) at 1e-12."""
import numpy as np
import pytest

import dmrgpp_amd as dm
from dmrgpp_amd import problems
from oracle import oracle


def _problem(ldaggerl: bool, cplx: bool):
    dtype = np.complex128 if cplx else np.float64
    if ldaggerl:  # a non-Hermitian factor L: random ragged blocks
        return problems.random_problem([3, 5, 4], [4, 2, 6], nC=3, density=0.6, seed=22,
                                       dtype=dtype)
    return problems.hermitian_problem([3, 5, 4], [4, 2, 6], nC=2, density=0.6, seed=21,
                                      dtype=dtype)


@pytest.mark.gpu
@pytest.mark.parametrize("stored", [False, True])
@pytest.mark.parametrize("ldaggerl", [False, True])
@pytest.mark.parametrize("cplx", [False, True])
def test_gpu_facade_matches_the_kron_form(stored, ldaggerl, cplx):
    prob = _problem(ldaggerl, cplx)
    H = oracle.kron_form(prob)
    n = prob.n
    rng = np.random.default_rng(5)
    y = rng.uniform(-1, 1, n)
    x0 = rng.uniform(-1, 1, n)
    if cplx:
        y = y + 1j * rng.uniform(-1, 1, n)
        x0 = x0 + 1j * rng.uniform(-1, 1, n)
    op = H.conj().T @ (H @ y) if ldaggerl else H @ y
    want = x0 + op
    f = dm.Facade(prob, max_rank_stored=n if stored else 0, ldaggerl=ldaggerl)
    info = f.info()
    assert info["stored"] == int(stored) and info["ldaggerl"] == int(ldaggerl)
    assert info["n"] == n
    got = f.matvec(x0, y)
    got2 = f.matvec(x0, y)  # the same call again: the clock and the count accumulate
    assert np.array_equal(got, got2)
    err = np.linalg.norm(got - want) / np.linalg.norm(op)
    assert err < 1e-12, err
    info = f.info()
    assert info["matvecs"] == 2 and info["ms_matvec_total"] > 0
    if stored:
        assert info["nnz"] == np.count_nonzero(H)
    f.close()


@pytest.mark.gpu
def test_gpu_facade_rejects_a_negative_rank_limit():
    with pytest.raises(dm.DmkError):
        dm.Facade(_problem(False, False), max_rank_stored=-1)
