"""Complex matvec (dmk_create_z / dmk_apply_z) and the batched Krylov time step
(dmk_krylov_evolve_z) against the oracle and the exact exponential."""
import numpy as np
import pytest

import dmrgpp_amd as dm
from dmrgpp_amd import problems
from dmrgpp_amd.problem import KronProblem
from oracle import oracle
from tests import fixtures


def rel_err(got, want):
    return np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300)


def to_device_complex(vecs):
    """Interleaved complex device array holding the given complex vectors back to back."""
    flat = np.ascontiguousarray(np.concatenate([np.asarray(v, dtype=np.complex128) for v in vecs]))
    return dm.DeviceArray.from_host(flat.view(np.float64))


def from_device_complex(d, count, n):
    return d.download(2 * count * n).view(np.complex128).reshape(count, n)


def expm_apply(H, phi, tau, Eg=0.0, d=1.0):
    w, U = np.linalg.eigh(H)
    return U @ (np.exp(-1j * (w - Eg) * tau * d) * (U.conj().T @ phi))


def as_complex_problem(prob):
    return KronProblem(L=prob.L, R=prob.R, nC=prob.nC,
                       blocks={k: (a.astype(np.complex128), b.astype(np.complex128))
                               for k, (a, b) in prob.blocks.items()},
                       dtype=np.complex128, name=prob.name + "-z")


# ----------------------------------------------------------------------------- CPU tests


def test_tridiag_eigen_all_pairs_vs_lapack():
    rng = np.random.default_rng(4)
    for m in (1, 2, 5, 40, 200):
        a = rng.uniform(-2, 2, m)
        b = np.r_[0.0, rng.uniform(-1, 1, m - 1)]
        T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
        w = np.linalg.eigvalsh(T)
        ev, S = dm.tridiag_eigen(a, b)
        np.testing.assert_allclose(ev, w, atol=1e-12 * max(1.0, np.abs(w).max()))
        assert np.abs(T @ S - S * ev).max() < 1e-12 * max(1.0, np.abs(w).max())
        assert np.abs(S.T @ S - np.eye(m)).max() < 1e-12
    # Lanczos-like matrix with a near-degenerate pair (a converged Ritz value and its ghost).
    a = np.array([1.0, 1.0 + 1e-13, 3.0])
    b = np.array([0.0, 1e-9, 0.5])
    ev, S = dm.tridiag_eigen(a, b)
    assert np.abs(S.T @ S - np.eye(3)).max() < 1e-12


def test_krylov_rejects_bad_arguments():
    p = dm.krylov_params()
    assert p.max_steps == 200 and p.min_steps == 4 and p.tolerance == 1e-12
    rc = dm.lib().dmk_krylov_evolve_z(None, 1, None, 1, 1, None, None, 1, None, None, None)
    assert rc == -1


# ----------------------------------------------------------------------------- GPU tests


@pytest.mark.gpu
def test_gpu_complex_fixture_matches_oracle():
    # test_KronConnections.cpp:132-178: H = [[3+i, 5i], [5-5i, -4]], H (1+i, 2-i) = (7+14i, 2+4i).
    prob = fixtures.kron_connections_fixture()
    x = np.array([1 + 1j, 2 - 1j])
    want = np.array([7 + 14j, 2 + 4j])
    op = dm.BatchedGemm(prob)
    info = op.info()
    assert info["is_complex"] == 1 and info["n_complex"] == 2 and info["n"] == 4
    dx, dy = to_device_complex([x]), to_device_complex([np.zeros(2)])
    op.apply_z(dx, dy)
    got = from_device_complex(dy, 1, 2)[0]
    assert rel_err(got, want) < 1e-15


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gpu_complex_random_problems_match_oracle(seed):
    prob = problems.random_problem([7, 33, 65, 1, 20], [12, 70, 3, 40, 9], nC=3, density=0.6,
                                   seed=seed, dtype=np.complex128)
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, prob.n) + 1j * rng.uniform(-1, 1, prob.n)
    want = oracle.plugin_apply(prob, x)
    op = dm.BatchedGemm(prob)
    dx, dy = to_device_complex([x]), to_device_complex([np.zeros(prob.n)])
    op.apply_z(dx, dy)
    assert rel_err(from_device_complex(dy, 1, prob.n)[0], want) < 1e-13


@pytest.mark.gpu
def test_gpu_real_handle_complex_vector_matches_complex_handle():
    # A real H applied to a complex vector: the real handle runs re/im as two columns, the
    # complex handle runs the real embedding; both must agree with the oracle.
    prob = problems.hermitian_problem([10, 30, 17], [25, 8, 40], nC=2, density=0.8, seed=7,
                                      dtype=np.float64)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, prob.n) + 1j * rng.uniform(-1, 1, prob.n)
    want = oracle.plugin_apply(as_complex_problem(prob), x)
    for p in (prob, as_complex_problem(prob)):
        op = dm.BatchedGemm(p)
        dx, dy = to_device_complex([x]), to_device_complex([np.zeros(prob.n)])
        op.apply_z(dx, dy)
        assert rel_err(from_device_complex(dy, 1, prob.n)[0], want) < 1e-13
        # The embedding of a real H drops the zero imaginary blocks: same term count as 2 copies.
        if p is not prob:
            assert op.info()["nterms"] == 2 * len(prob.blocks)


@pytest.mark.gpu
@pytest.mark.parametrize("cplx", [False, True])
def test_gpu_krylov_matches_oracle_and_exact(cplx):
    prob = problems.hermitian_problem([6, 14, 9, 3], [11, 4, 13, 7], nC=2, density=0.7, seed=11,
                                      dtype=np.complex128 if cplx else np.float64)
    H = oracle.kron_form(prob)
    H = 0.5 * (H + H.conj().T)
    rng = np.random.default_rng(8)
    phi = rng.uniform(-1, 1, prob.n) + 1j * rng.uniform(-1, 1, prob.n)
    times = [0.0, 0.02, 0.05, -0.05]
    Eg = -1.5
    want_o, steps_o = oracle.krylov(prob, phi, times, Eg=Eg)
    op = dm.BatchedGemm(prob)
    dphi = to_device_complex([phi])
    dout = to_device_complex([np.zeros(prob.n)] * len(times))
    res, st = op.krylov_evolve(dphi, dout, 1, times, energy_shift=Eg)
    got = from_device_complex(dout, len(times), prob.n)
    assert st["matvec_launches"] >= res[0]["steps"]
    # The stop test compares Ritz values at 1e-12, so a different summation order can move the
    # stop by one step; the truncation error is far below the tolerance either way.
    assert abs(res[0]["steps"] - steps_o) <= 1, (res[0]["steps"], steps_o)
    assert abs(res[0]["norm"] - np.linalg.norm(phi)) < 1e-12 * np.linalg.norm(phi)
    for q, t in enumerate(times):
        assert rel_err(got[q], want_o[q]) < 1e-10, (q, rel_err(got[q], want_o[q]))
        assert rel_err(got[q], expm_apply(H, phi, t, Eg=Eg)) < 1e-10
    assert np.linalg.norm(got[0] - phi) < 1e-12 * np.linalg.norm(phi)


@pytest.mark.gpu
def test_gpu_krylov_batched_targets_equal_single_runs():
    # Four targets in one batched launch per step give bitwise the results of four separate
    # runs: each target keeps its own Krylov dimension and its own reductions.
    prob = problems.hermitian_problem([20, 45, 33], [30, 12, 50], nC=2, density=0.8, seed=21,
                                      dtype=np.float64)
    rng = np.random.default_rng(5)
    n = prob.n
    phis = [rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n) for _ in range(4)]
    phis[2] = phis[2] * 1e-3  # different norms and convergence
    times = [0.01, 0.03]
    op = dm.BatchedGemm(prob)
    dphi = to_device_complex(phis)
    dout = to_device_complex([np.zeros(n)] * (4 * len(times)))
    res, st = op.krylov_evolve(dphi, dout, 4, times, energy_shift=0.0)
    got = from_device_complex(dout, 4 * len(times), n).reshape(4, len(times), n)
    assert st["matvec_launches"] == max(r["steps"] for r in res)
    for t in range(4):
        d1 = to_device_complex([phis[t]])
        o1 = to_device_complex([np.zeros(n)] * len(times))
        r1, _ = op.krylov_evolve(d1, o1, 1, times, energy_shift=0.0)
        single = from_device_complex(o1, len(times), n)
        assert r1[0]["steps"] == res[t]["steps"]
        np.testing.assert_array_equal(got[t], single)


@pytest.mark.gpu
def test_gpu_krylov_conserves_norm_and_energy_over_steps():
    # Twenty steps of dt = 0.05 on a 47570-dimensional real H with two targets.
    left = problems.gaussian_histogram(500, 1.5, range(-4, 5))
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(3), seed=3)
    n = prob.n
    op = dm.BatchedGemm(prob)
    rng = np.random.default_rng(1)
    psi = [rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n) for _ in range(2)]
    d_in = to_device_complex(psi)
    d_out = to_device_complex([np.zeros(n)] * 2)
    x = dm.DeviceArray(2 * n)
    y = dm.DeviceArray(2 * n)

    def energy(v):
        x.upload(v.view(np.float64))
        op.apply_z(x, y)
        hv = y.download().view(np.complex128)
        return np.vdot(v, hv)

    e0 = [energy(v) for v in psi]
    n0 = [np.linalg.norm(v) for v in psi]
    for _ in range(20):
        res, _ = op.krylov_evolve(d_in, d_out, 2, [0.05], stop="expm", tolerance=1e-13)
        d_in, d_out = d_out, d_in
    final = from_device_complex(d_in, 2, n)
    for t in range(2):
        assert abs(np.linalg.norm(final[t]) - n0[t]) < 1e-10 * n0[t]
        et = energy(final[t])
        assert abs(et.imag) < 1e-9 * abs(e0[t])
        assert abs(et.real - e0[t].real) < 1e-9 * abs(e0[t])


@pytest.mark.gpu
def test_gpu_krylov_fp32_matvec_within_bound():
    prob = problems.hermitian_problem([40, 60], [50, 30], nC=2, density=1.0, seed=2,
                                      dtype=np.float64)
    n = prob.n
    rng = np.random.default_rng(2)
    phi = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    op = dm.BatchedGemm(prob)
    outs = {}
    for prec in ("fp64", "fp32"):
        dphi = to_device_complex([phi])
        dout = to_device_complex([np.zeros(n)])
        op.krylov_evolve(dphi, dout, 1, [0.02], stop="expm", tolerance=1e-12, precision=prec)
        outs[prec] = from_device_complex(dout, 1, n)[0]
    assert rel_err(outs["fp32"], outs["fp64"]) < 1e-5
    assert op.info()["precision"] == dm.PREC_FP64  # restored after the call


@pytest.mark.gpu
def test_gpu_timeevo_workload_small_ladder():
    # The multi-vector time-evolution workload on a 4x2 Hubbard ladder: four targets (the
    # ground state and three random vectors) evolved together for ten steps of dt = 0.05.
    from dmrgpp_amd import timeevo

    prob, z, e = timeevo.superblock("hubbard-ladder-4x2", log=lambda s: None)
    out = timeevo.evolve(prob, z, targets=4, dt=0.05, steps=10, log=lambda s: None)
    assert out["norm_drift"] < 1e-10
    assert out["energy_drift_rel"] < 1e-9
