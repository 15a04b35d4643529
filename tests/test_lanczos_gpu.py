"""The device-resident Lanczos solver against the oracle's restatement of LanczosCore, the ED
value of the reference's dumped Kron problem, and its own FP64 path for the mixed modes."""
import numpy as np
import pytest

import dmrgpp_amd as dm
from dmrgpp_amd import problems
from oracle import oracle
from tests import fixtures


def test_tridiag_lowest_vs_lapack():
    rng = np.random.default_rng(0)
    for m in (1, 2, 3, 8, 64, 200):
        a = rng.uniform(-2, 2, m)
        b = np.concatenate([[0.0], rng.uniform(-1, 1, m - 1)])
        T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
        w, v = np.linalg.eigh(T)
        e, vec = dm.tridiag_lowest(a, b)
        assert abs(e - w[0]) < 1e-13 * max(1.0, abs(w[0]))
        assert min(np.linalg.norm(vec - v[:, 0]), np.linalg.norm(vec + v[:, 0])) < 1e-10
    # Reducible matrix (a zero off-diagonal) and a repeated diagonal.
    a = np.array([1.0, -3.0, 2.0, -3.0])
    b = np.array([0.0, 0.5, 0.0, 0.25])
    T = np.diag(a) + np.diag(b[1:], 1) + np.diag(b[1:], -1)
    e, vec = dm.tridiag_lowest(a, b)
    assert abs(e - np.linalg.eigvalsh(T)[0]) < 1e-13
    assert np.linalg.norm(T @ vec - e * vec) < 1e-12


def gpu_lanczos(prob, init, **kw):
    op = dm.BatchedGemm(prob)
    dinit = dm.DeviceArray.from_host(init)
    dz = dm.DeviceArray(prob.n)
    r = op.lanczos_ground(dinit, dz, **kw)
    z = dz.download()
    op.close()
    return r, z


def _qn(nb, left, seed=3):
    return problems.qn_problem(left, left, 0, problems.heisenberg_links(nb), seed=seed)


@pytest.mark.gpu
def test_gpu_lanczos_kron31_matches_ed_and_oracle():
    full, _, _, _ = fixtures.kron31_problem()
    init = np.random.default_rng(1234).uniform(-1, 1, full.n)
    e_orc, z_orc, steps_orc, _, _ = oracle.lanczos(full, init)
    r, z = gpu_lanczos(full, init)
    assert abs(r["energy"] - e_orc) < 1e-10 * abs(e_orc)
    assert abs(r["energy"] - (-3.374932598687897)) < 1e-9
    assert abs(r["steps"] - steps_orc) <= 2
    assert r["residual"] < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("nb,left", [(2, {-2: 30, 0: 50, 2: 30}),
                                     (3, {-4: 10, -2: 60, 0: 100, 2: 60, 4: 10})])
def test_gpu_lanczos_fp64_matches_oracle(nb, left):
    prob = _qn(nb, left)
    init = np.random.default_rng(1234).uniform(-1, 1, prob.n)
    e_orc, z_orc, steps_orc, a_orc, _ = oracle.lanczos(prob, init)
    r, z = gpu_lanczos(prob, init)
    assert abs(r["energy"] - e_orc) < 1e-10 * abs(e_orc), (r, e_orc)
    assert abs(r["steps"] - steps_orc) <= 3
    # Ritz vectors agree up to sign and the convergence level of the recurrence.
    zn = z / np.linalg.norm(z)
    zo = z_orc / np.linalg.norm(z_orc)
    assert min(np.linalg.norm(zn - zo), np.linalg.norm(zn + zo)) < 1e-4
    assert r["residual"] < 1e-4


@pytest.mark.gpu
def test_gpu_lanczos_fp32_with_fp64_refinement():
    prob = _qn(3, {-4: 10, -2: 60, 0: 100, 2: 60, 4: 10})
    init = np.random.default_rng(1234).uniform(-1, 1, prob.n)
    r64, _ = gpu_lanczos(prob, init)
    r32, z = gpu_lanczos(prob, init, precision="fp32", refine=1)
    assert r32["precision_first"] == dm.PREC_FP32
    # The FP64 restart recovers the FP64 energy; the FP32 run alone does not reach 1e-10.
    assert abs(r32["energy"] - r64["energy"]) < 1e-10 * abs(r64["energy"]), (r32, r64)
    assert r32["matvecs"] - r32["matvecs_first"] <= 20
    assert r32["residual"] < 1e-4
    # Without refinement the result carries the FP32 error.
    r32n, _ = gpu_lanczos(prob, init, precision="fp32", refine=0)
    assert r32n["matvecs"] == r32n["matvecs_first"]
    assert abs(r32n["energy"] - r64["energy"]) < 1e-5 * abs(r64["energy"])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_gpu_lanczos_mixed_with_davidson_refinement(precision):
    # refine = 2: the FP64 refinement is a Davidson run (diagonal correction) from the mixed Ritz
    # vector to a residual of sqrt(refine_tolerance) = 1e-6; its energy is the FP64 Lanczos
    # energy within 1e-10 relative, and the residual it reports is below the target.
    prob = _qn(3, {-4: 10, -2: 60, 0: 100, 2: 60, 4: 10})
    init = np.random.default_rng(1234).uniform(-1, 1, prob.n)
    r64, _ = gpu_lanczos(prob, init)
    rd, z = gpu_lanczos(prob, init, precision=precision, refine=2)
    assert abs(rd["energy"] - r64["energy"]) < 1e-10 * abs(r64["energy"]), (rd, r64)
    assert rd["residual"] <= 1e-6 * (1 + 1e-9)
    assert rd["matvecs"] > rd["matvecs_first"]
    # The returned vector is the unit Ritz vector of that energy.
    zn = z / np.linalg.norm(z)
    hz = oracle.plugin_apply(prob, zn)
    assert abs(zn @ hz - rd["energy"]) < 1e-10 * abs(rd["energy"])


@pytest.mark.gpu
def test_gpu_lanczos_switch_noise_moves_the_hand_over():
    # A larger hand-over level stops the FP32 first run earlier (no more first-run matvecs) and
    # the FP64 Davidson refinement still reaches the FP64 energy; a negative level is rejected.
    prob = _qn(3, {-4: 10, -2: 60, 0: 100, 2: 60, 4: 10})
    init = np.random.default_rng(1234).uniform(-1, 1, prob.n)
    r64, _ = gpu_lanczos(prob, init)
    lo, _ = gpu_lanczos(prob, init, precision="fp32", refine=2, switch_noise=1e-7)
    hi, _ = gpu_lanczos(prob, init, precision="fp32", refine=2, switch_noise=1e-4)
    assert hi["matvecs_first"] <= lo["matvecs_first"]
    for r in (lo, hi):
        assert abs(r["energy"] - r64["energy"]) < 1e-10 * abs(r64["energy"]), (r, r64)
    with pytest.raises(dm.DmkError):
        gpu_lanczos(prob, init, precision="fp32", refine=2, switch_noise=-1.0)


@pytest.mark.gpu
def test_gpu_lanczos_reortho():
    prob = _qn(2, {-2: 30, 0: 50, 2: 30})
    init = np.random.default_rng(1234).uniform(-1, 1, prob.n)
    e_orc, _, _, _, _ = oracle.lanczos(prob, init, reortho=True)
    r, _ = gpu_lanczos(prob, init, reortho=1)
    assert abs(r["energy"] - e_orc) < 1e-10 * abs(e_orc)


@pytest.mark.gpu
def test_gpu_lanczos_bf16x3_with_fp64_refinement():
    prob = _qn(3, {-4: 10, -2: 60, 0: 100, 2: 60, 4: 10})
    init = np.random.default_rng(1234).uniform(-1, 1, prob.n)
    r64, _ = gpu_lanczos(prob, init)
    r16, _ = gpu_lanczos(prob, init, precision="bf16x3", refine=1)
    assert r16["precision_first"] == dm.PREC_BF16X3
    assert abs(r16["energy"] - r64["energy"]) < 1e-10 * abs(r64["energy"]), (r16, r64)
    assert r16["residual"] < 1e-4


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
def test_gpu_lanczos_graph_replays_match_eager(precision):
    # Each step is a replay of one of three captured graphs, with host synchronization and the
    # eager stop test between replays; the captured kernels are the eager ones, so the energy,
    # the step count and the Ritz vector agree bitwise with the eager path.
    prob = _qn(3, {-4: 10, -2: 60, 0: 100, 2: 60, 4: 10})
    init = np.random.default_rng(1234).uniform(-1, 1, prob.n)
    rg, zg = gpu_lanczos(prob, init, precision=precision, graph=1)
    re, ze = gpu_lanczos(prob, init, precision=precision, graph=0)
    assert rg["energy"] == re["energy"] and rg["steps"] == re["steps"], (rg, re)
    assert rg["matvecs"] == re["matvecs"]
    assert np.array_equal(zg, ze)
    # Two graph solves in a row (fresh captures) give the same result again.
    rg2, zg2 = gpu_lanczos(prob, init, precision=precision, graph=1)
    assert rg2["energy"] == rg["energy"] and np.array_equal(zg2, zg)
