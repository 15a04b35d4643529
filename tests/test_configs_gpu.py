"""The five benchmark configurations at their full sizes on the GPU (config 3, the m = 2048
Heisenberg ladder, is the bench workload and has its full-size tests in test_matvec_gpu.py).

* Config 1 (Heisenberg chain L = 32, m = 128, 3 finite sweeps, FP64): every step's superblock
  is solved by the GPU Lanczos and by the oracle's Lanczos (the LanczosCore restatement over
  the plugin-layout matvec) from the same seed-1234 start vector; the energies agree to 1e-10
  relative at every step, and the final energy meets the reference's own pin (input3100,
  -13.9973 at 2e-5, dmrg/CMakeLists.txt:343-344).
* Config 2 (Hubbard chain L = 48, U = 8, m = 1024): the last superblock of the infinite growth
  (the full chain, n = 1.47e6 over 36 patches); the FP64 matvec against the oracle's CPU Kron matvec (KronConnections with
  host BLAS) at 1e-12, the FP32 and BF16x3 matvec errors against FP64 within their bounds, and
  the mixed-precision ground states with the FP64 refinement against FP64 Lanczos at 1e-10.
* Config 4 (Hubbard 24 x 2 ladder, U = 4, m = 2048, complex): four targets evolved together by
  the Krylov step, dt = 0.05, 40 steps; batched equals single bitwise on the first steps; norm
  and energy drift bounds in FP64; the FP32 matvec's drift and distance from FP64.
* Config 5 (Hubbard 16 x 4 cylinder, U = 8, m = 4096): a superblock of the growth at m = 4096
  (both blocks at m states); Davidson and Lanczos to the same FP64 residual 1e-9 give the same
  energy, and the mixed-precision ground states with FP64 refinement match FP64 at 1e-10.

Each test also checks that the library's allocation cache saw no release of a pointer it did
not hold during the run."""
import numpy as np
import pytest

import dmrgpp_amd as dm
from dmrgpp_amd import configs, learned, models, sweep, timeevo

pytestmark = pytest.mark.gpu


def _log(line):
    # Progress of the long runs (one line per DMRG step), visible with pytest -s.
    print(line, flush=True)


def _no_stale_release():
    for host in (False, True):
        assert dm.pool_counters(host)["bad_releases"] == 0


class _Keep:
    """sweep.run solver that keeps the last superblock and its ground state."""

    def __init__(self, inner):
        self.inner, self.prob, self.z, self.e = inner, None, None, None

    def __call__(self, prob, init=None):
        e, z, r = self.inner(prob) if init is None else self.inner(prob, init)
        self.prob, self.z, self.e = prob, z, e
        return e, z, r


class _GpuAndOracle:
    """Solves each superblock on the GPU and with the oracle from the same start vector and
    records both energies; the sweep continues with the GPU's ground state."""

    def __init__(self):
        self.gpu = sweep.GpuSolver(precision="fp64")
        self.pairs = []

    def __call__(self, prob, init=None):
        from oracle import oracle

        e, z, r = self.gpu(prob)
        x0 = np.random.default_rng(1234).uniform(-1.0, 1.0, prob.n)
        eo, _, steps, _, _ = oracle.lanczos(prob, x0)
        self.pairs.append((prob.n, e, eo))
        return e, z, r


def test_config1_chain32_three_sweeps_against_the_oracle_every_step():
    sol = _GpuAndOracle()
    e, recs, _ = configs.run_config("heisenberg-chain-32", solver=sol, log=_log)
    c = configs.CONFIGS["heisenberg-chain-32"]
    assert c["sweeps"] == 3 and c["m"] == 128
    assert len(sol.pairs) == len(recs) > 3 * 2 * (32 - 3)
    worst = max(abs(eg - eo) / abs(eo) for _, eg, eo in sol.pairs)
    assert worst < 1e-10, worst
    assert abs(e - (-13.9973)) < 2e-5  # input3100 (dmrg/CMakeLists.txt:343-344)
    _no_stale_release()


def test_config2_hubbard_chain48_m1024_precision_study():
    from oracle import kron_numpy

    c = configs.CONFIGS["hubbard-chain-48"]
    keep = _Keep(sweep.GpuSolver(precision="fp64"))
    sweep.run(c["model"](), c["lattice"](), c["target"], c["m"], [], keep, log=_log,
              eig=sweep.GpuDensityEig(512))
    prob = keep.prob
    # The full chain's middle superblock: n = 1 467 818 over 36 patches of up to 688 x 688.
    assert prob.n > 10**6, prob.n
    assert prob.npatches >= 30, prob.npatches
    x = np.random.default_rng(1234).uniform(-1.0, 1.0, prob.n)
    y_cpu = kron_numpy.apply(prob, kron_numpy.prepare(prob), x)
    op = dm.BatchedGemm(prob, "fp64")
    dx, dy = dm.DeviceArray.from_host(x), dm.DeviceArray(prob.n)
    ys = {}
    for prec in ("fp64", "fp32", "bf16x3"):
        op.set_precision(prec)
        op.apply_Htarget(dx, dy)
        op.sync()
        ys[prec] = dy.download()
    nrm = np.linalg.norm(y_cpu)
    assert np.linalg.norm(ys["fp64"] - y_cpu) / nrm < 1e-12
    # Mixed-precision matvec error bounds (FP32 accumulation; BF16x3 keeps ~16 mantissa bits
    # of each operand).
    assert np.linalg.norm(ys["fp32"] - ys["fp64"]) / nrm < 1e-5
    assert np.linalg.norm(ys["bf16x3"] - ys["fp64"]) / nrm < 1e-4
    # Ground states from the same start vector. The spectrum of this superblock is dense at the
    # bottom: the FP64 Lanczos needs more than the default 200 steps to meet its |dE| < 1e-12
    # stop rule, so the step limit is raised and checked to be unreached.
    e = {}
    dz = dm.DeviceArray(prob.n)
    op.set_precision("fp64")
    r = op.lanczos_ground(dx, dz, max_steps=800)
    assert r["steps"] < 800
    e["fp64"] = r["energy"]
    for prec in ("fp32", "bf16x3"):
        for refine in (1, 2):
            r = op.lanczos_ground(dx, dz, precision=prec, refine=refine, max_steps=800)
            assert r["steps"] < 800
            e[(prec, refine)] = r["energy"]
    op.close()
    for a in (dx, dy, dz):
        a.free()
    for k, v in e.items():
        assert abs(v - e["fp64"]) / abs(e["fp64"]) < 1e-10, (k, v, e["fp64"])
    _no_stale_release()


def test_config4_hubbard_ladder24x2_krylov_four_targets_40_steps():
    prob, z, e0 = timeevo.superblock("hubbard-ladder-24x2", log=_log)
    assert prob.n > 10**6
    n, targets, dt = prob.n, 4, 0.05
    rng = np.random.default_rng(99)
    psi = [z / np.linalg.norm(z) + 0j]
    while len(psi) < targets:
        v = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        psi.append(v / np.linalg.norm(v))
    op = dm.BatchedGemm(prob, "fp64")
    flat = np.concatenate(psi).view(np.float64)
    # Batched equals single, bitwise, over the first two steps.
    d_a, d_b = dm.DeviceArray.from_host(flat), dm.DeviceArray(flat.size)
    for _ in range(2):
        op.krylov_evolve(d_a, d_b, targets, [dt], stop="expm", tolerance=1e-12)
        d_a, d_b = d_b, d_a
    batched = d_a.download()
    for t in range(targets):
        s_a = dm.DeviceArray.from_host(psi[t].view(np.float64).copy())
        s_b = dm.DeviceArray(2 * n)
        for _ in range(2):
            op.krylov_evolve(s_a, s_b, 1, [dt], stop="expm", tolerance=1e-12)
            s_a, s_b = s_b, s_a
        assert np.array_equal(s_a.download(), batched[2 * n * t:2 * n * (t + 1)]), t
        s_a.free()
        s_b.free()
    d_a.free()
    d_b.free()
    op.close()
    # 40 steps in FP64 and with the FP32 matvec (FP64 Krylov recurrence).
    r64 = timeevo.evolve(prob, z, targets, dt, 40, "fp64", log=_log)
    r32 = timeevo.evolve(prob, z, targets, dt, 40, "fp32", log=_log)
    assert r64["steps"] == 40 and r64["targets"] == 4
    assert r64["norm_drift"] < 1e-12
    assert r64["energy_drift_rel"] < 1e-11
    assert r32["norm_drift"] < 1e-5
    assert r32["energy_drift_rel"] < 1e-5
    _no_stale_release()


def test_config5_hubbard_cylinder16x4_m4096_davidson_vs_lanczos():
    out = learned.superblock_compare("hubbard-cylinder-16x4", at=8, tol=1e-9, max_matvecs=400,
                                     log=_log)
    assert out["m"] == 4096 and out["n"] > 10**6
    el, ed = out["lanczos"]["energy"], out["davidson-diag"]["energy"]
    assert abs(el - ed) / abs(el) < 1e-10
    assert out["davidson-diag"]["residual"] < 1e-9
    for prec in ("fp32", "bf16x3"):
        assert out["ground_" + prec]["rel_dE_vs_fp64"] < 1e-10
    _no_stale_release()
