"""Benchmark of the superblock Kron matvec on one MI355X.

Workload: the superblock of the Heisenberg 32x2 two-leg ladder at m = 2048 kept states,
Sz_total = 0 (the m = 2048 ladder configuration): 17 patches, 11 connections (H_L (x) I,
I (x) H_R, three Heisenberg bonds across the cut), block sizes from the symmetry-sector
histogram, values U(-1,1) from seed 1234 (synthetic; no DMRG run produces it here).

A step is one matvec y = H x with x and y resident in HBM. value = matvecs per second in the
FP64 path; the FP32-MFMA rate and its error against FP64 are reported beside it. The roofline
entry is the dominant grouped-GEMM launch: algorithmic flops of that launch / its average
device time from HIP events on the launch stream. cpu_baseline times the KronConnections
restatement with numpy BLAS GEMMs on a bounded number of matvecs.

Usage: python bench.py [--gpus 1] [--steps 50] [--warmup 5]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(prob, x, budget_s=15.0, max_calls=20):
    from oracle import kron_numpy

    try:
        from threadpoolctl import threadpool_info

        threads = max([d.get("num_threads", 1) for d in threadpool_info()] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    prep = kron_numpy.prepare(prob)
    kron_numpy.apply(prob, prep, x)  # warm-up
    t0 = time.perf_counter()
    calls = 0
    while calls < max_calls:
        kron_numpy.apply(prob, prep, x)
        calls += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    return {"value": calls / dt, "unit": "matvecs/s", "cores": int(threads), "kind": "port",
            "sample": f"{calls} matvecs of the same m=2048 ladder superblock (KronConnections "
                      f"restatement, numpy BLAS GEMMs, {threads} BLAS threads), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="N",
                    help="only run the plain copy kernel on N doubles 5 times (HBM counter "
                         "calibration for rocprofv3 --pmc) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: only --gpus 1 is supported (asked for {args.gpus}); this build runs the "
              "matvec on one GPU and does not shard it", file=sys.stderr)
        sys.exit(2)

    import dmrgpp_amd as dm
    from dmrgpp_amd import problems

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

    t0 = time.perf_counter()
    prob = problems.ladder_m2048_problem(seed=1234)
    log(f"problem {prob.name}: {prob.histogram()} ({time.perf_counter() - t0:.1f} s)")
    x = np.random.default_rng(1234).uniform(-1.0, 1.0, prob.n)

    op = dm.BatchedGemm(prob, "fp64")
    info = op.info()
    log("plan:", json.dumps({k: info[k] for k in ("nterms", "nterms_identity", "stage1_tasks",
                                                   "stage2_tasks", "stage1_tiles", "stage2_tiles",
                                                   "flops_algorithmic", "flops_executed",
                                                   "flops_plugin", "bytes_compulsory")}))
    dx = dm.DeviceArray.from_host(x)
    dy = dm.DeviceArray(prob.n)

    results = {}
    for prec in ("fp64", "fp32"):
        op.set_precision(prec)
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
    y64 = results["fp64"]["y"]
    fp32_err = float(np.linalg.norm(results["fp32"]["y"] - y64) / np.linalg.norm(y64))
    op.close()

    r = results["fp64"]
    value = 1e3 / r["ms_per_step"]
    # Dominant kernel: the longer of the two grouped-GEMM launches.
    if r["ms_stage2"] >= r["ms_stage1"]:
        kname, kms, kflops = "stage2", r["ms_stage2"], info["flops_stage2"]
    else:
        kname, kms, kflops = "stage1", r["ms_stage1"], info["flops_stage1"]
    achieved = kflops / (kms * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_TFLOPS["fp64"],
                "unit": "TFLOP/s", "frac": round(achieved / PEAK_TFLOPS["fp64"], 4),
                "traffic": None, "kernel": f"grouped_gemm_kernel<double,double> ({kname})"}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(prob, x)
        log("cpu baseline:", cpu)

    r32 = results["fp32"]
    line = {
        "metric": "superblock matvecs/s (m=2048 Heisenberg 32x2 ladder superblock)",
        "value": round(value, 3),
        "unit": "matvecs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (symmetry-sector histogram of the m=2048 ladder superblock, "
                "U(-1,1) block values, seed 1234)",
        "config": {"workload": "heisenberg-ladder-32x2-m2048-superblock-matvec", "n": prob.n,
                   "npatches": prob.npatches, "nC": prob.nC, "nterms": len(prob.blocks),
                   "flops_per_matvec": info["flops_algorithmic"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "fp32_mfma": {"matvecs_per_s": round(1e3 / r32["ms_per_step"], 3),
                      "ms_per_step": round(r32["ms_per_step"], 4),
                      "rel_err_vs_fp64": fp32_err,
                      "tflops": round(info["flops_algorithmic"] / (r32["ms_total_events"] * 1e-3)
                                      / 1e12, 3)},
        "fp64_tflops_per_matvec": round(info["flops_algorithmic"] / (r["ms_total_events"] * 1e-3)
                                        / 1e12, 3),
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
