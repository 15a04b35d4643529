"""Davidson eigensolver with the diagonal and the learned (MLP) correction: the MLP oracle
pinned by finite differences, the GPU gradient kernels against it, diag(H) against the
assembled matrix, and Davidson energies against exact diagonalization and Lanczos."""
import numpy as np
import pytest

import dmrgpp_amd as dm
from dmrgpp_amd import problems
from oracle import mlp_oracle, oracle
from tests import fixtures


def _params(seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.4, 0.4, mlp_oracle.NPARAMS)


# ----------------------------------------------------------------------------- CPU tests


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params()
    x = rng.normal(size=(37, 4))
    t = rng.normal(size=37)
    g = mlp_oracle.gradient(p, x, t)
    idx = list(range(0, mlp_oracle.NPARAMS, 97)) + [mlp_oracle.OFF_B3, mlp_oracle.OFF_W3 + 5,
                                                   mlp_oracle.OFF_B1 + 7, mlp_oracle.OFF_B2 + 11]
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp_oracle.loss(p + e, x, t) - mlp_oracle.loss(p - e, x, t)) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(fd)), (i, fd, g[i])


def test_davidson_and_mlp_params_reject_bad_arguments():
    p = dm.davidson_params()
    assert p.tolerance == 1e-9 and p.precond == 1
    assert dm.lib().dmk_davidson_ground_d(None, None, None, None, None, None) == -1
    assert dm.lib().dmk_mlp_get_params(None, None) == -1


# ----------------------------------------------------------------------------- GPU tests


@pytest.mark.gpu
def test_gpu_mlp_forward_and_gradient_match_oracle():
    rng = np.random.default_rng(2)
    net = dm.Mlp(seed=5, capacity=1 << 12)
    p = _params(4).astype(np.float32)
    net.set_params(p)
    np.testing.assert_array_equal(net.params(), p)
    for count in (1, 63, 64, 1000, 20011):
        x = rng.normal(size=(count, 4)).astype(np.float32)
        t = rng.normal(size=count).astype(np.float32)
        y = net.forward(x)
        y_ref, _, _ = mlp_oracle.forward(p.astype(np.float64), x.astype(np.float64))
        assert np.abs(y - y_ref).max() < 1e-5 * max(1.0, np.abs(y_ref).max())
        g, loss = net.gradient(x, t)
        g_ref = mlp_oracle.gradient(p.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
        assert np.linalg.norm(g - g_ref) < 1e-5 * np.linalg.norm(g_ref), count
        assert abs(loss - mlp_oracle.loss(p, x, t)) < 1e-5 * mlp_oracle.loss(p, x, t)
    # Deterministic: the same batch gives bitwise the same gradient.
    g1, _ = net.gradient(x, t)
    g2, _ = net.gradient(x, t)
    np.testing.assert_array_equal(g1, g2)


@pytest.mark.gpu
def test_gpu_mlp_training_reduces_loss():
    # Fit the clipped diagonal correction f3 = -f1 / f0 (a known function of the features).
    rng = np.random.default_rng(3)
    x = rng.normal(size=(32768, 4)).astype(np.float32)
    d = np.where(np.abs(x[:, 0]) < 0.05, np.sign(x[:, 0]) * 0.05 + (x[:, 0] == 0) * 0.05, x[:, 0])
    t = np.clip(-x[:, 1] / d, -10, 10).astype(np.float32)
    x[:, 3] = t
    net = dm.Mlp(seed=7, capacity=1 << 16)
    net.add_samples(x, t)
    assert net.samples() == x.shape[0]
    _, loss0 = net.gradient(x, t)
    last, after = net.train(iters=1500, batch=4096, lr=2e-3)
    assert after < 0.05 * loss0, (loss0, after)


def _dense_diag(prob):
    return np.diag(oracle.kron_form(prob))


@pytest.mark.gpu
def test_gpu_diagonal_matches_assembled_matrix():
    full, _, _, _ = fixtures.kron31_problem()
    prob = problems.random_problem([3, 5, 2, 7], [4, 1, 6, 3], nC=3, density=0.7, seed=9)
    for p in (full, prob):
        op = dm.BatchedGemm(p)
        d = dm.DeviceArray(p.n)
        op.diagonal(d)
        np.testing.assert_allclose(d.download(), _dense_diag(p), atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("precond", ["none", "diag"])
def test_gpu_davidson_kron31_matches_ed(precond):
    full, _, _, _ = fixtures.kron31_problem()
    op = dm.BatchedGemm(full)
    init = dm.DeviceArray.from_host(np.random.default_rng(1234).uniform(-1, 1, full.n))
    z = dm.DeviceArray(full.n)
    r = op.davidson_ground(init, z, precond=precond, tolerance=1e-10, max_basis=12)
    assert r["converged"] == 1, r
    assert abs(r["energy"] - (-3.374932598687897)) < 1e-10
    assert r["residual"] < 1e-9
    zz = z.download()
    assert abs(np.linalg.norm(zz) - 1.0) < 1e-12


@pytest.mark.gpu
def test_gpu_davidson_matches_lanczos_on_qn_problem():
    left = problems.gaussian_histogram(160, 1.5, range(-3, 4))
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(3), seed=4)
    op = dm.BatchedGemm(prob)
    init = dm.DeviceArray.from_host(np.random.default_rng(1234).uniform(-1, 1, prob.n))
    z = dm.DeviceArray(prob.n)
    lz = op.lanczos_ground(init, z, tolerance=0.0, stop_residual=1e-9, max_steps=400)
    dv = op.davidson_ground(init, z, precond="diag", tolerance=1e-9, max_basis=20)
    assert dv["converged"] == 1 and dv["residual"] < 2e-9
    assert lz["residual"] < 2e-9
    assert abs(dv["energy"] - lz["energy"]) < 1e-10 * abs(lz["energy"])


@pytest.mark.gpu
def test_gpu_learned_preconditioner_comparison():
    # Train on the first sweep of a Hubbard chain, then run the next sweep with plain Lanczos,
    # Davidson with the diagonal correction and Davidson with the network's own correction (no
    # fallback), all at FP64 residual 1e-9, each step from the same transformed wave function.
    # Equal energies; the network's matvec count is compared with the diagonal one directly.
    # The measured outcome is that the network does not gain: it needs 1.15x the diagonal
    # count here (1.14x on an 8x2 ladder and 1.22x on a 16x4 cylinder at m = 256), so the
    # assertion pins that it converges without costing more than 25 % extra matvecs, and that the
    # validation step, which decides from its own counts, keeps the diagonal method whenever
    # the network is not at least 5 % better on the validation superblocks.
    from dmrgpp_amd import learned

    out = learned.compare("hubbard-chain-12", train_sweeps=1, test_sweeps=1, train_iters=1500,
                          log=print)
    e = out["lanczos"]["energy"]
    for key in ("davidson-diag", "davidson-mlp", "davidson-learned"):
        assert abs(out[key]["energy"] - e) < 1e-9 * abs(e)
        assert out[key]["max_residual"] < 1.1e-9
    assert out["samples"] > 0 and out["loss"] < out["loss_untrained"]
    mlp, diag = out["davidson-mlp"]["matvecs"], out["davidson-diag"]["matvecs"]
    print(f"network / diagonal matvecs: {mlp} / {diag} = {mlp / diag:.3f}")
    assert mlp <= 1.25 * diag
    val = out["validation_matvecs"]
    assert out["chosen"] == ("mlp" if val["mlp"] < 0.95 * val["diag"] else "diag")
    assert out["lanczos"]["matvecs"] > diag


@pytest.mark.gpu
def test_gpu_superblock_compare_small():
    # The one-superblock comparison on a small Hubbard ladder superblock: Lanczos and Davidson
    # reach the FP64 residual at the same energy, and the mixed-precision ground states
    # (FP32 / BF16x3 with FP64 refinement) match FP64 Lanczos to 1e-10.
    from dmrgpp_amd import learned

    out = learned.superblock_compare("hubbard-ladder-8x2", at=5, log=lambda s: None)
    assert out["n"] > 10000 and out["npatches"] > 1
    e = out["lanczos"]["energy"]
    assert out["lanczos"]["residual"] < 1.1e-9 and out["davidson-diag"]["residual"] < 1.1e-9
    assert abs(out["davidson-diag"]["energy"] - e) < 1e-10 * abs(e)
    for prec in ("fp64", "fp32", "bf16x3"):
        assert out[prec]["ms_per_matvec"] > 0
        assert abs(out["ground_" + prec]["energy"] - e) < 1e-10 * abs(e)
