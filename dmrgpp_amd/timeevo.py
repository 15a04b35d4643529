"""The multi-vector time-evolution workload: several target vectors of one superblock evolved
together by the Krylov time step, all targets in one grouped Kron launch per Krylov step.

  python -m dmrgpp_amd.timeevo --config hubbard-ladder-24x2 [--m 2048] [--targets 4]
                               [--dt 0.05] [--steps 40] [--precision fp64[,fp32,bf16x3]]
                               [--save-problem F.npz] (with --steps 0: save and stop)

The superblock is the middle of the lattice after infinite-size growth (sweep driver, GPU
Lanczos). The targets are the ground state and seeded random vectors of the same sector
(complex), normalized. Each time step calls dmk_krylov_evolve_z once for all targets; the run
reports the time per step, Krylov dimensions, and the drift of each target's norm and energy
(both conserved by exact evolution). Several precisions (comma separated) evolve the same
superblock and start vectors one after the other.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import models, sweep

CONFIGS = {
    # Hubbard two-leg ladder 24x2, U=4, t_leg = t_rung = 1, half filling, m=2048.
    "hubbard-ladder-24x2": dict(model=lambda: models.hubbard(4.0), lattice=lambda: models.ladder(24, 2),
                                target=lambda N: (N // 2, N // 2), m=2048),
    # Small stand-in of the same model for tests.
    "hubbard-ladder-4x2": dict(model=lambda: models.hubbard(4.0), lattice=lambda: models.ladder(4, 2),
                               target=lambda N: (N // 2, N // 2), m=64),
}


class _Keep:
    def __init__(self, inner):
        self.inner, self.prob, self.z = inner, None, None

    def __call__(self, prob):
        e, z, r = self.inner(prob)
        self.prob, self.z, self.e = prob, z, e
        return e, z, r


def superblock(name, m=None, log=print):
    c = CONFIGS[name]
    keep = _Keep(sweep.GpuSolver(precision="fp64"))
    sweep.run(c["model"](), c["lattice"](), c["target"], m or c["m"], [], keep, log=log)
    return keep.prob, keep.z, keep.e


def evolve(prob, ground, targets=4, dt=0.05, steps=40, precision="fp64", seed=99, log=print,
           stop="expm", tol=1e-12):
    from . import BatchedGemm, DeviceArray

    n = prob.n
    rng = np.random.default_rng(seed)
    psi = [ground / np.linalg.norm(ground) + 0j]
    while len(psi) < targets:
        v = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
        psi.append(v / np.linalg.norm(v))
    op = BatchedGemm(prob, "fp64")
    flat = np.concatenate(psi).view(np.float64)
    d_a = DeviceArray.from_host(flat)
    d_b = DeviceArray(flat.size)
    x, y = DeviceArray(2 * n), DeviceArray(2 * n)

    def energies(dvec):
        host = dvec.download().view(np.complex128).reshape(targets, n)
        out = []
        for t in range(targets):
            x.upload(host[t].view(np.float64))
            op.apply_z(x, y)
            hv = y.download().view(np.complex128)
            out.append((float(np.linalg.norm(host[t])), complex(np.vdot(host[t], hv))))
        return out

    e0 = energies(d_a)
    per, dims, mv_ms = [], [], 0.0
    t_all = time.perf_counter()
    for s in range(steps):
        t0 = time.perf_counter()
        res, st = op.krylov_evolve(d_a, d_b, targets, [dt], stop=stop, tolerance=tol,
                                   precision=precision)
        per.append(time.perf_counter() - t0)
        dims.append([r["steps"] for r in res])
        mv_ms += st["ms_matvec"]
        d_a, d_b = d_b, d_a
        if s % 10 == 0 or s == steps - 1:
            log(f"step {s:3d}: {1e3 * per[-1]:8.2f} ms, Krylov dims {dims[-1]}, "
                f"launches {st['matvec_launches']}")
    wall = time.perf_counter() - t_all
    e1 = energies(d_a)
    out = {"n": n, "targets": targets, "dt": dt, "steps": steps, "precision": precision,
           "wall_s": wall, "ms_per_step": 1e3 * wall / steps,
           "ms_matvec_total": mv_ms, "krylov_dims_mean": float(np.mean(dims)),
           "norm_drift": max(abs(a[0] - b[0]) for a, b in zip(e0, e1)),
           "energy_drift_rel": max(abs(a[1].real - b[1].real) / max(abs(a[1].real), 1e-300)
                                   for a, b in zip(e0, e1))}
    op.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="hubbard-ladder-24x2", choices=sorted(CONFIGS))
    ap.add_argument("--m", type=int, default=None)
    ap.add_argument("--targets", type=int, default=4)
    ap.add_argument("--dt", type=float, default=0.05)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--json", default=None)
    ap.add_argument("--save-problem", default=None,
                    help="write the superblock to this .npz (for microbench --replay)")
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    prob, z, e = superblock(a.config, a.m, log=log)
    log(f"superblock n={prob.n} patches={prob.npatches} terms={len(prob.blocks)} E0={e:.12f} "
        f"({time.perf_counter() - t0:.1f} s)")
    if a.save_problem:
        from .problem import save_npz

        save_npz(prob, a.save_problem)
        if a.steps == 0:
            return 0
    out = {"config": a.config, "E0": e, "npatches": prob.npatches, "nterms": len(prob.blocks),
           "runs": {}}
    for prec in a.precision.split(","):
        log(f"-- {prec}")
        out["runs"][prec] = evolve(prob, z, a.targets, a.dt, a.steps, prec, log=log)
        log(json.dumps(out["runs"][prec]))
    print(json.dumps(out), flush=True)
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(out, fh)


if __name__ == "__main__":
    sys.exit(main())
