"""Benchmark of the superblock Kron matvec on one MI355X.

Workload: the Heisenberg two-leg ladder 32x2 (J_leg = J_rung = 1, Sz = 0) at m = 2048 kept
states. The sweep driver grows the system by the infinite-size algorithm with the eigensolve on
the GPU; the superblock of its last step (the middle of the 64-site ladder: left block S+s and
right block e+E of 2m states each) is the matvec workload. Untimed setup.

A step is one matvec y = H x with x and y resident in HBM. value = matvecs per second of the
FP32-MFMA matvec (FP32 accumulation), the mode the mixed-precision eigensolver uses, provided
the ground state of the same superblock by FP32 Lanczos + FP64 refinement matches the FP64
oracle energy (the CPU Lanczos recurrence over the CPU Kron matvec, same start vector) to
|dE| < 1e-10, absolute and relative (the gate is reported; without it the line says so).
Beside it: the FP64 and BF16x3 matvec rates with their errors against FP64; the ground state in
all modes against the oracle energy; one whole finite sweep (every step started from the
transformed previous wave function) in FP64 and in the mixed modes, timed; the roofline of the
dominant grouped-GEMM launch of each mode from HIP events on its stream; and the CPU baseline
(the reference's CPU Kron path restated with host BLAS and CSR factors) on a bounded number of
matvecs, with its y against the GPU FP64 y.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--sweep-steps -1] [--no-cpu]
                       [--save-problem F.npz | --load-problem F.npz]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X dense peaks: FP64 MFMA, FP32 MFMA, bf16 MFMA. A BF16x3 launch runs three bf16 MFMAs per
# product (hi*hi + hi*lo + lo*hi), so its roofline counts 3x the algorithmic flops of the
# product against the bf16 peak.
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0}
MFMA_WORK = {"fp64": 1, "fp32": 1, "bf16x3": 3}
M = 2048
LX = 32
# HBM bytes per launch of the dominant kernel (stage 2) from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE on this workload (profiles/panel_gemm_ladder_m2048.md), FETCH scaled by the
# copy-kernel calibration factor 2.0, WRITE by 1.0. Keyed by precision and the stage-2 launch
# shape (tile m, tile n, splitk, xcd order) the counters were taken with; the MFMA form and the
# load pipeline depth change the loads' timing and the instruction mix, not the bytes (the two
# bf16x3 forms on 128x64 tiles measured within 3 %). The bytes belong to one superblock,
# TRAFFIC_SUPERBLOCK (n, patches, terms): the grown superblock can differ from run to run (the
# order of degenerate density-matrix eigenvalues moves with roundoff), and a run on another
# superblock, or whose autotuner picks another shape, reports None.
TRAFFIC_SUPERBLOCK = (2858152, 15, 116)
TRAFFIC = {
    ("fp64", (32, 32, 4096, 1)): 2.4442e9 + 0.0258e9,
    ("fp64", (32, 32, 2048, 1)): 2.0231e9 + 0.0653e9,
    ("fp64", (64, 32, 2048, 1)): 1.7916e9 + 0.0650e9,
    ("fp32", (64, 64, 2048, 1)): 0.5997e9 + 0.0420e9,
    ("fp32", (64, 32, 2048, 1)): 0.7202e9 + 0.0345e9,
    ("fp32", (32, 32, 4096, 1)): 0.7558e9 + 0.0241e9,
    ("bf16x3", (128, 64, 2048, 1)): 0.6202e9 + 0.0331e9,
    ("bf16x3", (64, 128, 0, 1)): 0.4910e9 + 0.0233e9,
    ("bf16x3", (128, 64, 0, 1)): 0.4953e9 + 0.0232e9,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def shape_key(t):
    """The stage-2 launch shape of a tuning (the key of TRAFFIC); a balanced persistent stage 2
    has no measured entry."""
    if t.get("persist", 0):
        return None
    return (t["stage2_tile_m"], t["stage2_tile_n"], t["splitk"], t["xcd_order"])


def host_cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(prob, x, y_gpu64, budget_s=12.0, max_calls=200):
    """The reference's CPU Kron path (oracle/kron_numpy: KronConnections with dense / CSR /
    identity factors as ArrayOfMatStruct stores them) timed on this host on a bounded number
    of matvecs of the bench superblock; its y is compared with the GPU FP64 matvec's."""
    from oracle import kron_numpy

    threads = kron_numpy.blas_threads()
    prep = kron_numpy.prepare(prob)
    y_cpu = kron_numpy.apply(prob, prep, x)  # warm-up, and the parity check of the GPU y
    err = float(np.linalg.norm(y_gpu64 - y_cpu) / np.linalg.norm(y_cpu))
    t0 = time.perf_counter()
    calls = 0
    while calls < max_calls:
        kron_numpy.apply(prob, prep, x)
        calls += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": round(calls / dt, 4), "unit": "matvecs/s", "cores": int(threads),
            "kind": "port", "host": f"{host_cpu_model()}, {os.cpu_count()} logical CPUs",
            "blas": kron_numpy.blas_library(),
            "gpu_fp64_rel_err_vs_cpu": err,
            "sample": f"{calls} matvecs of the same m=2048 ladder superblock (KronConnections "
                      f"restatement: dense blocks in BLAS GEMMs, sparse and identity factors as "
                      f"CSR, {threads} BLAS threads), {dt:.1f} s"}, prep


def save_problem(prob, path):
    from dmrgpp_amd.problem import save_npz

    save_npz(prob, path)


def load_problem(path):
    from dmrgpp_amd.problem import load_npz

    return load_npz(path, name="heisenberg-ladder-32x2-m2048")


class Recorder:
    """GPU solver that keeps the last superblock it solved."""

    def __init__(self, inner):
        self.inner = inner
        self.prob = None
        self.energy = None

    def __call__(self, prob):
        self.prob = prob
        e, z, r = self.inner(prob)
        self.energy = e
        return e, z, r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--sweep-steps", type=int, default=-1,
                    help="finite-sweep DMRG steps timed per precision mode (-1 = one whole "
                         "finite sweep, 0 = skip)")
    ap.add_argument("--no-oracle-energy", action="store_true",
                    help="skip the FP64 oracle Lanczos of the bench superblock on the CPU")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--save-problem", default=None, help="write the superblock to this .npz")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch shapes instead of autotuning each mode")
    ap.add_argument("--save-tuning", default=None, help="write the chosen tunings (JSON)")
    ap.add_argument("--load-tuning", default=None,
                    help="apply tunings from this JSON instead of autotuning (profiled runs)")
    ap.add_argument("--load-problem", default=None,
                    help="read the superblock from this .npz instead of growing it (no sweep)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="N",
                    help="only run the plain copy kernel on N doubles 5 times (HBM counter "
                         "calibration for rocprofv3 --pmc) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: only --gpus 1 is supported (asked for {args.gpus}); this build runs the "
              "matvec on one GPU and does not shard it", file=sys.stderr)
        sys.exit(2)

    import dmrgpp_amd as dm
    from dmrgpp_amd import models, sweep

    if dm.device_count() < 1:
        raise SystemExit("bench.py: no GPU visible")
    if args.copy_calibration:
        import ctypes

        n = args.copy_calibration
        src = dm.DeviceArray.from_host(np.ones(n))
        dst = dm.DeviceArray(n)
        for _ in range(5):
            dm._check(dm.lib().dmk_device_copy_d(dst.ptr, src.ptr, ctypes.c_size_t(n)),
                      "dmk_device_copy_d")
        log(f"copy calibration: 5 copies of {n} doubles ({8 * n} bytes read + written each)")
        return

    # ---- setup: infinite-size growth of the 32x2 ladder at m=2048 (untimed)
    t0 = time.perf_counter()
    model, lattice = models.heisenberg(1.0), models.ladder(LX, 2)
    state = {}
    if args.load_problem:
        prob = load_problem(args.load_problem)
        recs_inf = []
    else:
        rec = Recorder(sweep.GpuSolver(precision="fp64"))
        e_inf, recs_inf = sweep.run(model, lattice, lambda N: (0,), M, [], rec, log=log,
                                    state=state)
        prob = rec.prob
        if args.save_problem:
            save_problem(prob, args.save_problem)
    t_setup = time.perf_counter() - t0
    log(f"setup: infinite growth {len(recs_inf)} steps in {t_setup:.1f} s; superblock "
        f"{prob.histogram()}")
    x = np.random.default_rng(1234).uniform(-1.0, 1.0, prob.n)

    op = dm.BatchedGemm(prob, "fp64")
    info = op.info()
    log("plan:", json.dumps({k: info[k] for k in ("nterms", "nterms_identity", "stage1_tasks",
                                                   "stage2_tasks", "stage1_tiles", "stage2_tiles",
                                                   "flops_algorithmic", "flops_executed",
                                                   "flops_plugin", "bytes_compulsory")}))
    dx = dm.DeviceArray.from_host(x)
    dy = dm.DeviceArray(prob.n)

    # ---- timed matvecs (each mode with the autotuner's launch shapes)
    results = {}
    tunings = {}
    loaded = json.load(open(args.load_tuning)) if args.load_tuning else None
    for prec in ("fp64", "fp32", "bf16x3"):
        op.set_precision(prec)
        if loaded:
            op.set_tuning(**loaded[prec])
        elif not args.no_autotune:
            t_at = time.perf_counter()
            op.autotune(dx, dy, 5)
            log(f"{prec}: autotuned in {time.perf_counter() - t_at:.2f} s -> {op.tuning()}")
        tunings[prec] = op.tuning()
        for _ in range(args.warmup):
            op.apply_Htarget(dx, dy)
        op.sync()
        t = time.perf_counter()
        for _ in range(args.steps):
            op.apply_Htarget(dx, dy)
        op.sync()
        dt = time.perf_counter() - t
        y = dy.download()
        ms1, ms2, mst = op.profile(dx, dy, max(5, min(args.steps, 20)))
        results[prec] = {"ms_per_step": 1e3 * dt / args.steps, "y": y, "ms_stage1": ms1,
                         "ms_stage2": ms2, "ms_total_events": mst}
        log(f"{prec}: {1e3 * dt / args.steps:.3f} ms/matvec (events: stage1 {ms1:.3f} ms, "
            f"stage2 {ms2:.3f} ms, total {mst:.3f} ms)")
    if args.save_tuning:
        json.dump(tunings, open(args.save_tuning, "w"))
    y64 = results["fp64"]["y"]
    for prec in ("fp32", "bf16x3"):
        results[prec]["rel_err"] = float(np.linalg.norm(results[prec]["y"] - y64)
                                         / np.linalg.norm(y64))

    # ---- ground state of the same superblock: FP64 vs FP32 + FP64 refinement
    dz = dm.DeviceArray(prob.n)
    op.set_precision("fp64")
    lz = {}
    modes = (("fp64", {}), ("fp32+refine", {"precision": "fp32", "refine": 1}),
             ("bf16x3+refine", {"precision": "bf16x3", "refine": 1}),
             ("fp32+davidson", {"precision": "fp32", "refine": 2}),
             ("bf16x3+davidson", {"precision": "bf16x3", "refine": 2}))
    for mode, kw in modes:
        r = op.lanczos_ground(dx, dz, **kw)
        lz[mode] = r
        log(f"lanczos {mode}: E={r['energy']:.14f} steps={r['steps']} matvecs={r['matvecs']} "
            f"first={r['matvecs_first']} residual={r['residual']:.2e} {r['ms_total']:.1f} ms")
    e64 = lz["fp64"]["energy"]
    op.close()

    # ---- one finite sweep per mode (or a bounded segment), continuing from the grown blocks
    sweep_out = None
    if args.sweep_steps != 0 and state:
        sweep_out = {}
        stop = None if args.sweep_steps < 0 else args.sweep_steps
        # Every step of a finite sweep starts from the transformed wave function; the mixed
        # modes run those solves in their own precision (mixed_warm) with the FP64 Davidson
        # refinement. The truncation diagonalises density-matrix blocks of 512 rows or more on
        # the GPU while the host LAPACK does the smaller ones.
        mk = lambda p, r: sweep.GpuSolver(precision=p, refine=r, mixed_warm=True)  # noqa: E731
        gpu_eig = sweep.GpuDensityEig(512)
        # rocSOLVER loads its kernels on the first call of each size class, and each of the
        # solver's four streams sizes its workspace on its first block: not timed.
        w = np.random.default_rng(0).standard_normal((1280, 1280))
        dm.density_diag([w + w.T] * 4 + [np.eye(64) + 0.5])
        for mode, solver in (("fp64", sweep.GpuSolver(precision="fp64")),
                             ("fp32+davidson", mk("fp32", 2)),
                             ("bf16x3+davidson", mk("bf16x3", 2))):
            st = sweep.clone_state(state)
            t = time.perf_counter()
            e, recs = sweep.run(model, lattice, lambda N: (0,), M, [M], solver, log=log,
                                state=st, resume=True, stop_after=stop, wft=True, eig=gpu_eig)
            dt = time.perf_counter() - t
            per = dt / max(1, len(recs))
            lz_ms = sum(r.ms_lanczos for r in recs) / max(1, len(recs))
            tr_ms = sum(r.info["timing"]["truncate_ms"] for r in recs) / max(1, len(recs))
            parts = {k: round(sum(r.info["timing"].get("truncate_" + k, 0.0) for r in recs)
                              / max(1, len(recs)), 1) for k in ("rho_ms", "eig_ms", "basis_ms")}
            sweep_out[mode] = {"steps": len(recs), "whole_sweep": stop is None,
                               "sweep_s": round(dt, 2), "s_per_step": round(per, 4),
                               "eigensolve_ms_per_step": round(lz_ms, 1),
                               "eigensolve_s": round(sum(r.ms_lanczos for r in recs) / 1e3, 2),
                               "truncate_ms_per_step": round(tr_ms, 1),
                               "truncate_parts_ms_per_step": parts,
                               "lanczos_matvecs": int(sum(r.matvecs for r in recs)),
                               "energy_last": e,
                               "energy_min": min(r.energy for r in recs)}
            log(f"sweep {mode}: {len(recs)} steps in {dt:.2f} s, {per:.3f} s/step "
                f"(eigensolve {lz_ms:.1f} ms/step, truncation {tr_ms:.1f} ms/step {parts})")
        e_ref = sweep_out["fp64"]["energy_min"]
        for mode in ("fp32+davidson", "bf16x3+davidson"):
            so = sweep_out[mode]
            so["sweep_speedup_vs_fp64"] = round(sweep_out["fp64"]["sweep_s"]
                                                / max(1e-9, so["sweep_s"]), 3)
            so["eigensolve_speedup_vs_fp64"] = round(
                sweep_out["fp64"]["eigensolve_ms_per_step"]
                / max(1e-9, so["eigensolve_ms_per_step"]), 3)
            so["rel_dE_min_vs_fp64"] = abs(so["energy_min"] - e_ref) / abs(e_ref)

    def roofline_of(prec):
        r = results[prec]
        if r["ms_stage2"] >= r["ms_stage1"]:
            kname, kms, kflops = "stage2", r["ms_stage2"], info["flops_stage2"]
        else:
            kname, kms, kflops = "stage1", r["ms_stage1"], info["flops_stage1"]
        achieved = MFMA_WORK[prec] * kflops / (kms * 1e-3) / 1e12
        kern = {"fp64": "panel_gemm_kernel<double,OpF64>",
                "fp32": "panel_gemm_kernel<double,OpF32|OpF32x32>",
                "bf16x3": "panel_bf16x3_kernel"}[prec]
        out = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_TFLOPS[prec],
               "unit": "TFLOP/s", "frac": round(achieved / PEAK_TFLOPS[prec], 4),
               "traffic": (TRAFFIC.get((prec, shape_key(tunings[prec])))
                           if (prob.n, prob.npatches, len(prob.blocks)) == TRAFFIC_SUPERBLOCK
                           else None),
               "kernel": f"{kern} ({kname})",
               "ms_per_launch": round(kms, 4), "flops_per_launch": kflops}
        if prec == "bf16x3":
            out["note"] = ("achieved counts the 3 bf16 MFMA products per multiply-add against "
                           "the bf16 dense peak; FP32-equivalent rate "
                           f"{kflops / (kms * 1e-3) / 1e12:.1f} TFLOP/s")
        return out

    r = results["fp32"]
    value = 1e3 / r["ms_per_step"]
    roofline = roofline_of("fp32")

    cpu, prep = None, None
    if not args.no_cpu:
        cpu, prep = cpu_baseline(prob, x, y64)
        log("cpu baseline:", cpu)
    # FP64 oracle energy of the bench superblock: the CPU Lanczos recurrence (LanczosCore) over
    # the CPU Kron matvec from the same start vector; every GPU mode is gated against it.
    oracle_e = None
    if not args.no_cpu and not args.no_oracle_energy:
        from oracle import kron_numpy

        t_or = time.perf_counter()
        eo, steps_o, mv_o = kron_numpy.lanczos(prob, prep, x)
        oracle_e = {"energy": eo, "steps": steps_o, "matvecs": mv_o,
                    "s": round(time.perf_counter() - t_or, 1)}
        log(f"oracle lanczos (CPU, FP64): E={eo:.14f} steps={steps_o} in {oracle_e['s']} s")

    e_ref = oracle_e["energy"] if oracle_e else e64
    e_rel = {m: abs(lz[m]["energy"] - e_ref) / abs(e_ref) for m in lz}
    e_abs = {m: abs(lz[m]["energy"] - e_ref) for m in lz}
    line = {
        "metric": "superblock matvecs/s + sweep time at m=2048, mixed vs FP64 GPU at |\u0394E|<1e-10",
        "value": round(value, 3),
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic model run: the superblock of the last infinite-growth step of the "
                "Heisenberg 32x2 ladder at m=2048 (GPU Lanczos, seed-1234 U[-1,1) start vectors)",
        "config": {"workload": "heisenberg-ladder-32x2-m2048-superblock-matvec",
                   "precision": "fp32-mfma matvec, fp64 recurrence + fp64 refinement",
                   "n": prob.n, "npatches": prob.npatches, "nC": prob.nC,
                   "nterms": len(prob.blocks), "flops_per_matvec": info["flops_algorithmic"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "gate": {"mode": "fp32+refine",
                 "against": "fp64 oracle (CPU Lanczos)" if oracle_e else "fp64 GPU Lanczos",
                 "rel_dE": e_rel["fp32+refine"], "abs_dE": e_abs["fp32+refine"],
                 "passed": bool(e_rel["fp32+refine"] < 1e-10 and e_abs["fp32+refine"] < 1e-10)},
        "oracle_energy": oracle_e,
        "modes": {prec: {"matvecs_per_s": round(1e3 / results[prec]["ms_per_step"], 3),
                         "ms_per_step": round(results[prec]["ms_per_step"], 4),
                         "rel_err_vs_fp64": results[prec].get("rel_err", 0.0),
                         "tflops_per_matvec": round(info["flops_algorithmic"]
                                                    / (results[prec]["ms_total_events"] * 1e-3)
                                                    / 1e12, 3),
                         "roofline": roofline_of(prec), "tuning": tunings[prec]}
                  for prec in ("fp64", "fp32", "bf16x3")},
        "ground_state": {m: {"energy": lz[m]["energy"], "rel_dE_vs_oracle": e_rel[m],
                             "abs_dE_vs_oracle": e_abs[m],
                             "gate_1e-10": bool(e_rel[m] < 1e-10),
                             "ms": round(lz[m]["ms_total"], 2), "matvecs": lz[m]["matvecs"],
                             "matvecs_first": lz[m]["matvecs_first"],
                             "residual": lz[m]["residual"]} for m in lz},
        "sweep": sweep_out,
        "setup_s": round(t_setup, 1),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
