"""The benchmark workloads as sweep-driver settings, and a command-line runner.

  python -m dmrgpp_amd.configs --config heisenberg-chain-32 [--precision fp64|fp32|bf16x3]
                               [--refine 1|2] [--sweeps N] [--stop-after K] [--wft]
                               [--json out.json]

Each run prints one line per DMRG step and, with --json, writes the step records and the total
sweep wall time.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

from . import models, sweep

CONFIGS = {
    # Spin-1/2 Heisenberg chain, L=32, OBC, J=1, Sz=0; infinite growth then 3 finite sweeps at
    # m=128; Lanczos tol 1e-12, <= 200 steps, initial vectors U[-1,1) from seed 1234. The
    # reference's own pin of this model is input3100 (-13.9973 at 2e-5, dmrg/CMakeLists.txt:343).
    "heisenberg-chain-32": dict(model=lambda: models.heisenberg(1.0), lattice=lambda: models.chain(32),
                                target=lambda N: (0,), m=128, sweeps=3),
    # Fermi-Hubbard chain L=48, U=8, t=1, half filling, m=1024, 4 sweeps.
    "hubbard-chain-48": dict(model=lambda: models.hubbard(8.0), lattice=lambda: models.chain(48),
                             target=lambda N: (N // 2, N // 2), m=1024, sweeps=4),
    # Heisenberg two-leg ladder 32x2, J_leg = J_rung = 1, Sz=0, m=2048.
    "heisenberg-ladder-32x2": dict(model=lambda: models.heisenberg(1.0),
                                   lattice=lambda: models.ladder(32, 2), target=lambda N: (0,),
                                   m=2048, sweeps=2),
    # Fermi-Hubbard two-leg ladder 24x2, U=4, t_leg = t_rung = 1, half filling, m=2048 (the
    # time-evolution workload's lattice), one finite sweep.
    "hubbard-ladder-24x2": dict(model=lambda: models.hubbard(4.0),
                                lattice=lambda: models.ladder(24, 2),
                                target=lambda N: (N // 2, N // 2), m=2048, sweeps=1),
    # Small cases used by the tests.
    "heisenberg-chain-8": dict(model=lambda: models.heisenberg(1.0), lattice=lambda: models.chain(8),
                               target=lambda N: (0,), m=16, sweeps=1),
    "heisenberg-ladder-4x2": dict(model=lambda: models.heisenberg(1.0),
                                  lattice=lambda: models.ladder(4, 2), target=lambda N: (0,),
                                  m=16, sweeps=1),
    "hubbard-chain-8-U0": dict(model=lambda: models.hubbard(0.0), lattice=lambda: models.chain(8),
                               target=lambda N: (N // 2, N // 2), m=256, sweeps=1),
}


def run_config(name: str, precision="fp64", m=None, sweeps=None, stop_after=None, log=print,
               solver=None, wft=None, refine=2, gpu_trunc_min=0, gpu_basis=False, svd=False):
    """gpu_trunc_min > 0: the truncation diagonalises density-matrix blocks of at least that
    many rows on the GPU; gpu_basis: it also changes the basis of the block operators there;
    svd: the truncation takes the SVD of the wave-function blocks instead, on the host (faster
    than rocSOLVER's dgesvd at these sizes; sweep.SvdDensityEig(device=True) runs it there)."""
    c = CONFIGS[name]
    m = m or c["m"]
    nsw = c["sweeps"] if sweeps is None else sweeps
    solver = solver or sweep.GpuSolver(precision=precision, refine=refine)
    if svd:
        eig = sweep.SvdDensityEig()
    else:
        eig = sweep.GpuDensityEig(gpu_trunc_min) if gpu_trunc_min > 0 else None
    t0 = time.perf_counter()
    e, recs = sweep.run(c["model"](), c["lattice"](), c["target"], m, [m] * nsw, solver, log=log,
                        stop_after=stop_after, wft=wft, eig=eig,
                        rotate=sweep.GpuBasisChange() if gpu_basis else None)
    wall = time.perf_counter() - t0
    return e, recs, wall


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", required=True, choices=sorted(CONFIGS))
    ap.add_argument("--precision", default="fp64")
    ap.add_argument("--m", type=int, default=None)
    ap.add_argument("--sweeps", type=int, default=None)
    ap.add_argument("--stop-after", type=int, default=None)
    ap.add_argument("--wft", action="store_true",
                    help="start finite-sweep Lanczos runs from the transformed previous state")
    ap.add_argument("--refine", type=int, default=2,
                    help="FP64 refinement of a mixed-precision solve: 1 = Lanczos restart, "
                         "2 = Davidson (default)")
    ap.add_argument("--gpu-trunc-min", type=int, default=512,
                    help="diagonalise density-matrix blocks of at least this many rows on the "
                         "GPU (0: all on the host)")
    ap.add_argument("--gpu-basis", action="store_true",
                    help="do the truncation's change of basis on the GPU")
    ap.add_argument("--svd", action="store_true",
                    help="truncate by the SVD of the wave-function blocks (DensityMatrixSvd)")
    ap.add_argument("--json", default=None)
    a = ap.parse_args(argv)
    e, recs, wall = run_config(a.config, a.precision, a.m, a.sweeps, a.stop_after,
                               log=lambda s: print(s, flush=True), wft=True if a.wft else None,
                               refine=a.refine, gpu_trunc_min=a.gpu_trunc_min,
                               gpu_basis=a.gpu_basis, svd=a.svd)
    sweep_steps = [r for r in recs if r.phase != "infinite"]
    out = {"config": a.config, "precision": a.precision, "energy": e, "wall_s": wall,
           "steps": len(recs), "matvecs": sum(r.matvecs for r in recs),
           "ms_matvec": sum(r.ms_matvec for r in recs),
           "ms_lanczos": sum(r.ms_lanczos for r in recs),
           "sweep_wall_s_est": sum(r.ms_lanczos for r in sweep_steps) / 1e3,
           "sweep_steps": len(sweep_steps),
           "sweep_step_s": (sum(sum(v for k, v in r.info.get("timing", {}).items()
                                    if k in ("grow_ms", "kron_ms", "wft_ms", "truncate_ms"))
                                + r.ms_lanczos for r in sweep_steps)
                            / 1e3 / max(1, len(sweep_steps))),
           "sweep_eigensolve_ms_per_step": (sum(r.ms_lanczos for r in sweep_steps)
                                            / max(1, len(sweep_steps))),
           "records": [r.__dict__ for r in recs]}
    print(json.dumps({k: v for k, v in out.items() if k != "records"}), flush=True)
    if a.json:
        with open(a.json, "w") as fh:
            json.dump(out, fh)


if __name__ == "__main__":
    sys.exit(main())
