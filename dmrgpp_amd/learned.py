"""The learned-preconditioner workload: Davidson with an MLP correction trained on pairs harvested
from the solver's own converged runs, compared with plain Lanczos and with Davidson's diagonal
correction at the same FP64 residual.

  python -m dmrgpp_amd.learned --config hubbard-ladder-8x2 [--m M] [--train-sweeps 1]

Phase 1 runs the first sweeps with Davidson (diagonal correction) and harvests training pairs
from every converged step; the MLP is then trained on the GPU. Validation: the last three
training superblocks are solved again with the diagonal and with the learned correction; the
learned one is kept only if it needs at least 5% fewer matvecs over them (otherwise the
diagonal method is the fallback).
Phase 2 runs the remaining sweeps from the same block stores with plain Lanczos, Davidson +
diagonal, Davidson + MLP and the validated choice, each step started from the same transformed
wave function, and counts matvecs and time.

  python -m dmrgpp_amd.learned --config hubbard-cylinder-16x4 --superblock-at 8

runs the workload at its full m on one superblock instead (sweeps at m = 4096 keep the host
side of the driver busy for longer than a measurement should take): the infinite growth stops
at the given step, once both blocks hold m states, and that superblock's matvec is timed in the
three precisions and solved by Lanczos and by Davidson (diagonal correction) to the same FP64
residual from the same start vector.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import models, sweep


class DavidsonSolver:
    """sweep.run solver: Davidson with the given correction; harvests into mlp when asked and
    keeps the last `keep` problems it solved (for validation)."""

    def __init__(self, precond="diag", mlp=None, harvest=False, tol=1e-9, max_basis=24,
                 max_matvecs=600, seed=1234, keep=3):
        self.precond, self.mlp, self.harvest = precond, mlp, harvest
        self.tol, self.max_basis, self.max_matvecs, self.seed = tol, max_basis, max_matvecs, seed
        self.keep, self.recent = keep, []

    @property
    def last(self):
        return self.recent[-1] if self.recent else None

    def __call__(self, prob, init=None):
        self.recent = (self.recent + [prob])[-self.keep:]
        from . import BatchedGemm, DeviceArray

        own = not isinstance(init, DeviceArray)
        if init is None:
            init = np.random.default_rng(self.seed).uniform(-1.0, 1.0, prob.n)
        t0 = time.perf_counter()
        op = BatchedGemm(prob, "fp64")
        t1 = time.perf_counter()
        dinit = DeviceArray.from_host(init) if own else init
        dz = DeviceArray(prob.n)
        r = op.davidson_ground(dinit, dz, mlp=self.mlp, precond=self.precond, tolerance=self.tol,
                               max_basis=self.max_basis, max_matvecs=self.max_matvecs,
                               harvest=1 if self.harvest else 0)
        z = dz.download()
        t2 = time.perf_counter()
        op.close()
        if own:
            dinit.free()
        r["steps"] = r["iterations"]
        r["z_dev"] = dz
        r["timing"] = {"plan_ms": 1e3 * (t1 - t0), "lanczos_ms": 1e3 * (t2 - t1)}
        return r["energy"], z, r


class LanczosResidualSolver(sweep.GpuSolver):
    """Plain Lanczos stopped at the same FP64 residual as the Davidson runs."""

    def __init__(self, tol=1e-9, max_steps=600):
        super().__init__(max_steps=max_steps, tol=0.0)
        self.stop_residual = tol

    def __call__(self, prob, init=None):
        from . import BatchedGemm, DeviceArray

        own = not isinstance(init, DeviceArray)
        if init is None:
            init = np.random.default_rng(self.seed).uniform(-1.0, 1.0, prob.n)
        t0 = time.perf_counter()
        op = BatchedGemm(prob, "fp64")
        t1 = time.perf_counter()
        dinit = DeviceArray.from_host(init) if own else init
        dz = DeviceArray(prob.n)
        r = op.lanczos_ground(dinit, dz, max_steps=self.max_steps, min_steps=self.min_steps,
                              tolerance=0.0, stop_residual=self.stop_residual, graph=1)
        z = dz.download()
        t2 = time.perf_counter()
        op.close()
        if own:
            dinit.free()
        r["z_dev"] = dz
        r["timing"] = {"plan_ms": 1e3 * (t1 - t0), "lanczos_ms": 1e3 * (t2 - t1)}
        return r["energy"], z, r


CONFIGS = {
    # Hubbard 4-leg cylinder 16x4, U=8, t=1, half filling, m=4096 (the learned-preconditioner
    # workload); smaller stand-ins of the same model for tests.
    "hubbard-cylinder-16x4": dict(model=lambda: models.hubbard(8.0),
                                  lattice=lambda: models.ladder(16, 4, periodic_y=True),
                                  target=lambda N: (N // 2, N // 2), m=4096),
    "hubbard-ladder-8x2": dict(model=lambda: models.hubbard(8.0), lattice=lambda: models.ladder(8, 2),
                               target=lambda N: (N // 2, N // 2), m=256),
    "hubbard-chain-12": dict(model=lambda: models.hubbard(8.0), lattice=lambda: models.chain(12),
                             target=lambda N: (N // 2, N // 2), m=128),
}


def compare(name: str, m=None, train_sweeps=1, test_sweeps=1, tol=1e-9, train_iters=3000,
            log=print, mlp_seed=11):
    """Returns a dict with per-method totals over the test sweeps."""
    from . import Mlp

    c = CONFIGS[name]
    model, lat, target = c["model"](), c["lattice"](), c["target"]
    m = m or c["m"]
    mlp = Mlp(seed=mlp_seed, capacity=1 << 22)
    state = {}
    t0 = time.perf_counter()
    harvest = DavidsonSolver("diag", mlp=mlp, harvest=True, tol=tol)
    sweep.run(model, lat, target, m, [m] * train_sweeps, harvest, log=log, state=state, wft=True)
    t1 = time.perf_counter()
    nsamp = mlp.samples()
    _, loss0 = mlp.train(0, batch=8192, lr=2e-3)  # loss of the untrained (zero) correction
    loss_last, loss_all = mlp.train(train_iters, batch=8192, lr=2e-3)
    t2 = time.perf_counter()
    log(f"harvested {nsamp} pairs in {t1 - t0:.1f} s; trained {train_iters} Adam steps in "
        f"{t2 - t1:.2f} s, loss {loss0:.4g} -> {loss_all:.4g}")
    # Validation on the last training superblocks: keep the learned correction only if it wins
    # clearly.
    val = {"diag": 0, "mlp": 0}
    for prob in harvest.recent:
        for pc in ("diag", "mlp"):
            e, _, r = DavidsonSolver(pc, mlp=mlp, tol=tol)(prob)
            r["z_dev"].free()
            val[pc] += r["matvecs"]
    chosen = "mlp" if val["mlp"] < 0.95 * val["diag"] else "diag"
    log(f"validation matvecs: diag {val['diag']}, mlp {val['mlp']} -> {chosen}")
    out = {"config": name, "m": m, "samples": nsamp, "train_s": t2 - t1, "loss_untrained": loss0,
           "loss": loss_all, "validation_matvecs": val, "chosen": chosen}
    solvers = {"lanczos": LanczosResidualSolver(tol=tol),
               "davidson-diag": DavidsonSolver("diag", tol=tol),
               "davidson-mlp": DavidsonSolver("mlp", mlp=mlp, tol=tol),
               "davidson-learned": DavidsonSolver(chosen, mlp=mlp, tol=tol)}
    for key, solver in solvers.items():
        st = sweep.clone_state(state)
        ts = time.perf_counter()
        e, recs = sweep.run(model, lat, target, m, [m] * test_sweeps, solver, log=lambda s: None,
                            state=st, resume=True, wft=True)
        te = time.perf_counter()
        mv = sum(r.matvecs for r in recs)
        res = max(r.info.get("residual", 0.0) for r in recs)
        out[key] = {"energy": e, "matvecs": mv, "wall_s": te - ts, "max_residual": res,
                    "steps": len(recs)}
        log(f"{key:16s} E={e:.12f} matvecs={mv:6d} wall={te - ts:7.2f}s max_res={res:.2e}")
    out["gain_matvecs_vs_diag"] = (out["davidson-diag"]["matvecs"]
                                   / max(1, out["davidson-learned"]["matvecs"]))
    out["gain_matvecs_vs_lanczos"] = (out["lanczos"]["matvecs"]
                                      / max(1, out["davidson-learned"]["matvecs"]))
    return out


class _KeepLast:
    """sweep.run solver that keeps the last superblock it solved."""

    def __init__(self, inner):
        self.inner, self.prob = inner, None

    def __call__(self, prob):
        self.prob = prob
        return self.inner(prob)


def superblock_compare(name: str, m=None, at=8, tol=1e-9, max_matvecs=300, log=print):
    """One superblock of the workload at its full m (see the module docstring)."""
    from . import BatchedGemm, DeviceArray

    c = CONFIGS[name]
    m = m or c["m"]
    rec = _KeepLast(sweep.GpuSolver(precision="fp64"))
    t0 = time.perf_counter()
    sweep.run(c["model"](), c["lattice"](), c["target"], m, [], rec, log=log, stop_after=at,
              eig=sweep.GpuDensityEig(512))
    prob = rec.prob
    out = {"config": name, "m": m, "growth_steps": at, "growth_s": time.perf_counter() - t0,
           "n": prob.n, "npatches": prob.npatches, "nterms": len(prob.blocks),
           "histogram": prob.histogram()}
    log(f"superblock: n={prob.n} patches={prob.npatches} terms={len(prob.blocks)}")
    x = np.random.default_rng(1234).uniform(-1.0, 1.0, prob.n)
    dx, dy = DeviceArray.from_host(x), DeviceArray(prob.n)
    op = BatchedGemm(prob, "fp64")
    flops = op.info()["flops_algorithmic"]
    for prec in ("fp64", "fp32", "bf16x3"):
        op.set_precision(prec)
        op.autotune(dx, dy, 3)
        for _ in range(3):
            op.apply_Htarget(dx, dy)
        op.sync()
        t = time.perf_counter()
        for _ in range(10):
            op.apply_Htarget(dx, dy)
        op.sync()
        ms = 1e3 * (time.perf_counter() - t) / 10
        out[prec] = {"ms_per_matvec": ms, "tflops": flops / (ms * 1e-3) / 1e12,
                     "tuning": op.tuning()}
        log(f"{prec}: {ms:.3f} ms/matvec, {flops / (ms * 1e-3) / 1e12:.1f} TF/s")
    op.close()
    dx.free()
    dy.free()
    for key, solver in (("lanczos", LanczosResidualSolver(tol=tol, max_steps=max_matvecs)),
                        ("davidson-diag", DavidsonSolver("diag", tol=tol,
                                                         max_matvecs=max_matvecs))):
        t = time.perf_counter()
        e, _, r = solver(prob)
        dt = time.perf_counter() - t
        r["z_dev"].free()
        out[key] = {"energy": e, "matvecs": r["matvecs"], "wall_s": dt,
                    "residual": r.get("residual", None)}
        log(f"{key:14s} E={e:.12f} matvecs={r['matvecs']} wall={dt:.2f}s "
            f"residual={r.get('residual', float('nan')):.2e}")
    out["matvec_ratio_lanczos_vs_davidson"] = (out["lanczos"]["matvecs"]
                                               / max(1, out["davidson-diag"]["matvecs"]))
    # The mixed-precision ground state (FP32 / BF16x3 Lanczos, FP64 Davidson refinement) against
    # FP64 Lanczos, all stopped at the energy criterion of the sweeps.
    for prec in ("fp64", "fp32", "bf16x3"):
        t = time.perf_counter()
        e, _, r = sweep.GpuSolver(precision=prec, refine=2)(prob)
        dt = time.perf_counter() - t
        if hasattr(r.get("z_dev"), "free"):
            r["z_dev"].free()
        out["ground_" + prec] = {"energy": e, "matvecs": r.get("matvecs"), "wall_s": dt,
                                 "residual": r.get("residual")}
        log(f"ground {prec:7s} E={e:.12f} matvecs={r.get('matvecs')} wall={dt:.2f}s")
    e64 = out["ground_fp64"]["energy"]
    for prec in ("fp32", "bf16x3"):
        out["ground_" + prec]["rel_dE_vs_fp64"] = abs(out["ground_" + prec]["energy"] - e64) / abs(e64)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="hubbard-ladder-8x2", choices=sorted(CONFIGS))
    ap.add_argument("--m", type=int, default=None)
    ap.add_argument("--train-sweeps", type=int, default=1)
    ap.add_argument("--test-sweeps", type=int, default=1)
    ap.add_argument("--train-iters", type=int, default=3000)
    ap.add_argument("--superblock-at", type=int, default=0, metavar="STEP",
                    help="time one superblock at full m, reached after STEP growth steps, "
                         "instead of running sweeps")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    if a.superblock_at > 0:
        out = superblock_compare(a.config, a.m, a.superblock_at,
                                 log=lambda s: print(s, flush=True))
        print(json.dumps(out), flush=True)
        if a.json:
            with open(a.json, "w") as fh:
                json.dump(out, fh)
        return 0
    out = compare(a.config, a.m, a.train_sweeps, a.test_sweeps, train_iters=a.train_iters,
                  log=lambda s: print(s, flush=True))
    print(json.dumps(out), flush=True)
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(out, fh)


if __name__ == "__main__":
    sys.exit(main())
