"""Pins the CPU oracle (oracle/) against the reference's own fixtures and against mathematics
that does not depend on the oracle's code paths (explicit Kronecker products, closed forms,
LAPACK via numpy)."""
import itertools

import numpy as np
import pytest

from oracle import oracle
from tests import fixtures


def test_batchedgemm_fixture():
    # dmrg/Engine/tests/test_BatchedGemm.cpp:234-258, Approx(...).epsilon(1e-12).
    prob, vin, expected, tol = fixtures.batchedgemm_fixture()
    for got in (oracle.plugin_apply(prob, vin), oracle.kron_connections(prob, vin)):
        np.testing.assert_allclose(got, expected, rtol=tol, atol=0)


def test_kron_connections_fixture():
    # dmrg/Engine/tests/test_KronConnections.cpp:132-178, hard-coded expected values.
    prob = fixtures.kron_connections_fixture()
    inp = np.array([1 + 1j, 2 - 1j])
    x0 = np.array([3 + 1j, -1 + 2j])
    fwd = oracle.kron_connections(prob, inp, x0=x0)
    np.testing.assert_allclose(fwd, [10 + 15j, 1 + 6j], atol=1e-12)
    matrix = np.array([[3 + 1j, 5j], [5 - 5j, -4]])
    np.testing.assert_allclose(oracle.kron_form(prob), matrix, atol=1e-12)
    np.testing.assert_allclose(matrix @ inp, [7 + 14j, 2 + 4j], atol=1e-12)
    adj = oracle.kron_connections(prob, inp, x0=x0, apply_adjoint=True)
    np.testing.assert_allclose(adj, x0 + matrix.conj().T @ inp, atol=1e-12)
    # The adjoint is not the transpose for this non-Hermitian complex factor.
    assert np.all(np.abs(matrix.conj().T @ inp - matrix.T @ inp) > 1e-12)
    # A^H A as forward followed by adjoint.
    hy = oracle.kron_connections(prob, inp)
    fact = oracle.kron_connections(prob, hy, x0=x0, apply_adjoint=True)
    np.testing.assert_allclose(fact, x0 + matrix.conj().T @ (matrix @ inp), atol=1e-12)
    # Plugin layout computes the plain product (no conjugation, BatchedGemm semantics).
    np.testing.assert_allclose(oracle.plugin_apply(prob, inp), matrix @ inp, atol=1e-12)


def _np_op(m, t):
    return {"n": m, "t": m.T, "c": m.conj().T}[t]


@pytest.mark.parametrize("cplx", [False, True])
def test_den_kron_mult_methods_vs_explicit_kron(cplx):
    # KronUtil/test1.cpp:10-60 sweep: sizes {1,4,7} x {1,4,7,10}, op in {n,t,c}, methods 1-3,
    # against the explicit Kronecker product (den_kron_form), tol 1e-9 there, 1e-12 here.
    rng = np.random.default_rng(7)
    sizes_a = [1, 4, 7]
    sizes_b = [1, 4, 7, 10]
    for nra, nca, nrb, ncb in itertools.product(sizes_a, sizes_a, sizes_b, sizes_b):
        if rng.random() > 0.35:
            continue
        a = rng.uniform(-1, 1, (nra, nca))
        b = rng.uniform(-1, 1, (nrb, ncb))
        if cplx:
            a = a + 1j * rng.uniform(-1, 1, a.shape)
            b = b + 1j * rng.uniform(-1, 1, b.shape)
        for ta, tb in itertools.product("ntc" if cplx else "nt", repeat=2):
            if (ta == "c") != (tb == "c"):
                continue  # KronConnections always passes the same op for A and B
            K = np.kron(_np_op(a, ta), _np_op(b, tb))
            y = rng.uniform(-1, 1, K.shape[1])
            x0 = rng.uniform(-1, 1, K.shape[0])
            if cplx:
                y = y + 1j * rng.uniform(-1, 1, y.shape)
            want = x0 + K @ y
            for m in (1, 2, 3):
                got = oracle.den_kron_mult(m, ta, tb, a, b, y, x0)
                np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * max(1, np.abs(want).max()))


def test_estimate_kron_cost_choices():
    # estimate_kron_cost.cpp: dense square blocks prefer method 1 when B x Y is cheaper, method 2
    # when Y x A^T is, and very sparse blocks go to method 3.
    assert oracle.estimate_kron_method(10, 10, 100, 10, 10, 100) == 2  # tie -> method 2 wins last
    # A 40x10 dense, B 4x4 dense: method 1 = 2*16*10 + 2*400*4 = 3520 < method 2 = 4480.
    assert oracle.estimate_kron_method(40, 10, 400, 4, 4, 16) == 1
    # A 10x40: method 1 = 2*16*40 + 2*400*4 = 4480 > method 2 = 2*400*4 + 2*16*10 = 3520.
    assert oracle.estimate_kron_method(10, 40, 400, 4, 4, 16) == 2
    assert oracle.estimate_kron_method(100, 100, 3, 100, 100, 2) == 3


def test_kron31_fixture_products_agree():
    full, lower, y, ents = fixtures.kron31_problem()
    assert full.n == len(y) == 70
    H = oracle.kron_form(full)
    want = H @ y
    scale = np.abs(want).max()
    np.testing.assert_allclose(oracle.plugin_apply(full, y), want, atol=1e-13 * scale)
    # The reference's own call pattern: useLowerPart with 't' on the mirrored blocks.
    got = oracle.kron_connections(lower, y, use_lower_part=True)
    np.testing.assert_allclose(got, want, atol=1e-13 * scale)
    # Offsets of the fixture match the patch layout.
    off = full.offsets
    for e in ents:
        assert off[e["out"]] == e["offsetX"] and off[e["in"]] == e["offsetY"]
    # Hermiticity invariant of the assembled superblock Hamiltonian.
    assert np.abs(H - H.T).max() < 1e-12


def test_kron31_lanczos_matches_exact_diagonalization():
    # input31 is the Heisenberg chain L=8; the dumped superblock is the full Sz=0 sector
    # (70 = C(8,4) states), so its lowest eigenvalue is the ED value -3.374932599
    # (dmrg/CMakeLists.txt:214-221 checks -3.37493 at 2e-5).
    full, _, y, _ = fixtures.kron31_problem()
    H = oracle.kron_form(full)
    e_exact = np.linalg.eigvalsh(0.5 * (H + H.T))[0]
    assert abs(e_exact - (-3.374932598687897)) < 1e-9
    init = np.random.default_rng(1234).uniform(-1, 1, full.n)
    e, z, steps, a, b = oracle.lanczos(full, init, max_steps=200, min_steps=4, tol=1e-12)
    assert abs(e - e_exact) < 1e-10 * abs(e_exact)
    zn = z / np.linalg.norm(z)
    assert np.linalg.norm(H @ zn - e * zn) < 1e-5


def test_tridiag_eigen_vs_lapack():
    rng = np.random.default_rng(3)
    for n in (1, 2, 3, 10, 57, 200):
        d = rng.uniform(-2, 2, n)
        e = rng.uniform(-1, 1, max(n - 1, 0))
        T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
        w, v = np.linalg.eigh(T)
        got, z = oracle.tridiag_eigen(d, e)
        np.testing.assert_allclose(got, w, atol=1e-12 * max(1, np.abs(w).max()))
        # Eigenvectors up to sign.
        for j in range(n):
            s = np.sign(z[:, j] @ v[:, j])
            np.testing.assert_allclose(s * z[:, j], v[:, j], atol=1e-8)
    # Degenerate spectrum (identity) and a zero off-diagonal split.
    got, _ = oracle.tridiag_eigen(np.ones(5), np.zeros(4))
    np.testing.assert_allclose(got, np.ones(5))


def test_lanczos_free_fermion_spectrum():
    # A diagonal operator with a known spectrum written as a Kron problem: H = D_L (x) I + I (x) D_R
    # has eigenvalues dL_i + dR_j; the lowest is min(dL) + min(dR).
    rng = np.random.default_rng(11)
    dl = np.sort(rng.uniform(-1, 1, 6))
    dr = np.sort(rng.uniform(-1, 1, 5))
    from dmrgpp_amd.problem import KronProblem
    # One patch holding everything; dense rotated factors so that nothing is trivially diagonal.
    ql, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    qr, _ = np.linalg.qr(rng.standard_normal((5, 5)))
    HL = ql @ np.diag(dl) @ ql.T
    HR = qr @ np.diag(dr) @ qr.T
    prob = KronProblem(L=[6], R=[5], nC=2, blocks={(0, 0, 0): (HL, np.eye(5)),
                                                   (0, 0, 1): (np.eye(6), HR)})
    e, z, steps, _, _ = oracle.lanczos(prob, rng.uniform(-1, 1, 30), tol=1e-14, max_steps=30)
    assert abs(e - (dl[0] + dr[0])) < 1e-12


def test_plugin_matches_connections_on_random_problems():
    from dmrgpp_amd import problems
    for seed, cplx in [(1, False), (2, True)]:
        prob = problems.random_problem([3, 5, 2, 7], [4, 1, 6, 3], nC=3, density=0.6, seed=seed,
                                       dtype=np.complex128 if cplx else np.float64)
        x = np.random.default_rng(seed).uniform(-1, 1, prob.n).astype(prob.dtype)
        want = oracle.kron_form(prob) @ x
        np.testing.assert_allclose(oracle.plugin_apply(prob, x), want, atol=1e-12)
        np.testing.assert_allclose(oracle.kron_connections(prob, x), want, atol=1e-12)
        np.testing.assert_allclose(prob.apply_numpy(x), want, atol=1e-12)


def _expm_apply(H, phi, tau, Eg=0.0, d=1.0):
    w, U = np.linalg.eigh(H)
    return U @ (np.exp(-1j * (w - Eg) * tau * d) * (U.conj().T @ phi))


@pytest.mark.parametrize("cplx", [False, True])
def test_krylov_oracle_full_space_is_exact(cplx):
    # With tolerance 0 the recurrence runs to the full space (n steps) and the Krylov step is
    # the exact exponential of the assembled H.
    from dmrgpp_amd import problems
    prob = problems.hermitian_problem([2, 3, 2], [3, 2, 2], nC=2, density=0.7, seed=5,
                                      dtype=np.complex128 if cplx else np.float64)
    H = oracle.kron_form(prob)
    assert np.abs(H - H.conj().T).max() < 1e-12
    rng = np.random.default_rng(9)
    phi = rng.uniform(-1, 1, prob.n) + 1j * rng.uniform(-1, 1, prob.n)
    times = [0.0, 0.05, 0.2, -0.3]
    out, steps = oracle.krylov(prob, phi, times, Eg=0.7, time_direction=1.0, max_steps=200,
                               min_steps=4, tol=0.0)
    assert steps == prob.n
    for q, t in enumerate(times):
        want = _expm_apply(H, phi, t, Eg=0.7)
        assert np.linalg.norm(out[q] - want) < 1e-10 * np.linalg.norm(phi)


def test_krylov_oracle_unitary_and_time_reversal():
    full, _, _, _ = fixtures.kron31_problem()
    H = oracle.kron_form(full)
    H = 0.5 * (H + H.T)
    rng = np.random.default_rng(2)
    phi = rng.uniform(-1, 1, full.n) + 1j * rng.uniform(-1, 1, full.n)
    out, steps = oracle.krylov(full, phi, [0.1], Eg=-3.0)
    assert 4 <= steps <= full.n
    # The reference stop test (lowest Ritz value settled to 1e-12) leaves a Krylov space large
    # enough that the short-time propagator is exact to roundoff here.
    want = _expm_apply(H, phi, 0.1, Eg=-3.0)
    assert np.linalg.norm(out[0] - want) < 1e-10 * np.linalg.norm(phi)
    assert abs(np.linalg.norm(out[0]) - np.linalg.norm(phi)) < 1e-10 * np.linalg.norm(phi)
    back, _ = oracle.krylov(full, out[0], [0.1], Eg=-3.0, time_direction=-1.0)
    assert np.linalg.norm(back[0] - phi) < 1e-9 * np.linalg.norm(phi)


def test_comb_patch_sizes_follow_state_counting():
    # GPUPlugin/gen_patches_comb.cpp for the test_vbatch defaults (8+8 sites, 8 up 8 down,
    # keep 8000 / 2000): every (n_up, n_down) split is valid; sizes are the rounded shares of
    # C(8, n_up) C(8, n_down) / 2^16, so the largest left patch is round(70*70/65536*8000).
    from dmrgpp_amd import problems
    L, R = problems.comb_patch_sizes()
    assert max(L) == round(70 * 70 / 65536 * 8000) and max(R) == round(70 * 70 / 65536 * 2000)
    assert abs(sum(L) - 8000) < 100 and abs(sum(R) - 2000) < 60
    assert len(L) == len(R) and all(v >= 1 for v in L + R)


def _sweep_superblocks(model, lat, m, count, target=(0,)):
    """The first `count` superblocks the sweep driver hands to the eigensolver (identity
    factors, CSR-sparse link blocks and dense block Hamiltonians, as DMRG++ produces them)."""
    from dmrgpp_amd import sweep

    got = []

    class Capture:
        def __call__(self, prob):
            got.append(prob)
            init = np.random.default_rng(1234).uniform(-1, 1, prob.n)
            e, z, s, _, _ = oracle.lanczos(prob, init)
            return e, z, {"steps": s, "matvecs": s}

    sweep.run(model, lat, lambda N: target, m, [], Capture(), log=lambda s: None,
              stop_after=count)
    return got


def test_kron_numpy_matches_oracle_with_identity_and_sparse_factors():
    # The timed CPU baseline (oracle/kron_numpy.py) against the C++ restatement of the plugin and
    # of KronConnections, and against the explicit Kronecker form, on sweep-driver superblocks:
    # connection 0 and 1 carry identity factors, the link operators of the enlarged blocks fall
    # under the 0.2 density threshold and go through the CSR branches.
    from dmrgpp_amd import models
    from oracle import kron_numpy

    probs = _sweep_superblocks(models.heisenberg(), models.ladder(4, 2), 24, 3)
    probs += _sweep_superblocks(models.hubbard(4.0), models.chain(6), 32, 2, target=(3, 3))
    seen = {"eye": 0, "csr": 0, "dense": 0}
    for prob in probs:
        prep = kron_numpy.prepare(prob)
        for terms in prep.values():
            for _, fa, fb, _ in terms:
                seen[fa.kind] += 1
                seen[fb.kind] += 1
        x = np.random.default_rng(5).uniform(-1, 1, prob.n)
        want = oracle.plugin_apply(prob, x)
        got = kron_numpy.apply(prob, prep, x)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
        np.testing.assert_allclose(got, oracle.kron_connections(prob, x),
                                   atol=1e-12 * np.abs(want).max())
        if prob.n <= 400:
            np.testing.assert_allclose(got, oracle.kron_form(prob) @ x,
                                       atol=1e-12 * np.abs(want).max())
    assert min(seen.values()) > 0, seen


def test_kron_numpy_factor_kinds_follow_the_threshold():
    # ArrayOfMatStruct.h:307: dense iff nnz >= 0.2 rows cols (so a 5x5 identity is dense and a
    # 10x10 one CSR); a CSR identity is one GEMM.
    from oracle import kron_numpy
    from dmrgpp_amd.problem import KronProblem

    rng = np.random.default_rng(4)
    a_dense = rng.uniform(-1, 1, (10, 10))
    a_sparse = np.zeros((10, 10))
    a_sparse[1, 3] = 2.0
    a_sparse[4, 0] = -1.0  # 2 of 100 < 0.2 * 100
    eye = np.eye(10)
    b = rng.uniform(-1, 1, (3, 3))
    blocks = {(0, 0, 0): (a_dense, b), (0, 0, 1): (a_sparse, b), (0, 0, 2): (eye, b),
              (0, 0, 3): (a_dense, np.eye(3))}  # a 3x3 identity is dense (3 >= 1.8)
    prob = KronProblem(L=[10], R=[3], nC=4, blocks=blocks)
    prep = kron_numpy.prepare(prob)
    kinds = [(fa.kind, fb.kind) for _, fa, fb, _ in prep[0]]
    assert kinds == [("dense", "dense"), ("csr", "dense"), ("eye", "dense"), ("dense", "dense")]
    small = kron_numpy.prepare(KronProblem(L=[5], R=[3], nC=1,
                                           blocks={(0, 0, 0): (np.eye(5), b)}))
    assert small[0][0][1].kind == "dense"
    x = rng.uniform(-1, 1, prob.n)
    want = sum(np.kron(a, bb) for a, bb in blocks.values()) @ x
    np.testing.assert_allclose(kron_numpy.apply(prob, prep, x), want, atol=1e-13)
    # A planted mistake (the identity term dropped) must fail the same comparison.
    bad = {k: v for k, v in blocks.items() if k != (0, 0, 2)}
    bprob = KronProblem(L=[10], R=[3], nC=4, blocks=bad)
    wrong = kron_numpy.apply(bprob, kron_numpy.prepare(bprob), x)
    assert np.abs(wrong - want).max() > 1e-3


def test_kron_numpy_lanczos_matches_oracle_lanczos():
    # The full-size FP64 oracle energy (numpy matvec + LanczosCore recurrence) against the C++
    # restatement's Lanczos and against ED on the kron31 dump (Heisenberg L=8, -3.374932599).
    from oracle import kron_numpy

    full, _, _, _ = fixtures.kron31_problem()
    init = np.random.default_rng(1234).uniform(-1, 1, full.n)
    e_np, steps_np, mv = kron_numpy.lanczos(full, kron_numpy.prepare(full), init)
    e_c, _, steps_c, _, _ = oracle.lanczos(full, init)
    assert abs(e_np - e_c) < 1e-12 * abs(e_c)
    assert steps_np == steps_c and mv == steps_np
    assert abs(e_np - (-3.374932598687897)) < 1e-10
    from dmrgpp_amd import models

    for prob in _sweep_superblocks(models.heisenberg(), models.ladder(4, 2), 24, 3)[1:]:
        init = np.random.default_rng(7).uniform(-1, 1, prob.n)
        e_np, s_np, _ = kron_numpy.lanczos(prob, kron_numpy.prepare(prob), init)
        e_c, _, s_c, _, _ = oracle.lanczos(prob, init)
        assert abs(e_np - e_c) < 1e-12 * abs(e_c) and s_np == s_c
