"""Driver-free matvec microbenchmarks on one GPU:

  (i)   the plugin benchmark's shape (GPUPlugin bench_BatchedGemm.cpp:195-245): 40 patches,
        34 of size 60 and 6 of size 900, one operator, block diagonal, U(-1,1) values;
  (ii)  the plugin test driver's shape (GPUPlugin/test_vbatch.cpp): comb patch sizes for 8+8
        sites at 8 up / 8 down, keep 8000 / 2000 states, two operators, every block dense,
        X[i] = (i + 1) / n;
  (iii) a replay of a saved superblock (bench.py --save-problem), e.g. the m=2048 ladder.

  python -m dmrgpp_amd.microbench [--replay F.npz [--load-tuning T.json]] [--iters 30] [--autotune]
  python -m dmrgpp_amd.microbench --replay F.npz --multi 1,2,4,8 [--iters 12]
  python -m dmrgpp_amd.microbench --density 1093,986,654,538 [--iters 5]
  python -m dmrgpp_amd.microbench --vbatch [--iters 30]
  python -m dmrgpp_amd.microbench --replay F.npz --plan [--iters 10]


With --plan, the replayed superblock's handle is created from host blocks and from the same
blocks already in device memory, and the creation times are reported (scan, staging, upload).
Each case reports the average matvec time from HIP events (one warm-up, then iters calls) and
the algorithmic TFLOP/s in FP64, FP32 and BF16x3. With --multi, the replayed superblock is
multiplied into several vectors per launch (dmk_apply_multi_d, the batched form of the Krylov
time step) for each listed count, in FP64 and FP32, timed by wall clock around iters calls.
With --density, the truncation's eigensolver (dmk_density_diag_d) diagonalises random density
matrices of the listed sizes (the largest symmetry blocks of an m=2048 ladder truncation by
default), all in one call and one at a time, timed by wall clock.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import problems


def run_case(name, prob, x, iters, log, tunings=None, autotune=False):
    from . import BatchedGemm, DeviceArray

    op = BatchedGemm(prob, "fp64")
    info = op.info()
    dx, dy = DeviceArray.from_host(x), DeviceArray(prob.n)
    out = {"case": name, "n": prob.n, "npatches": prob.npatches, "nterms": info["nterms"],
           "flops": info["flops_algorithmic"]}
    for prec in ("fp64", "fp32", "bf16x3"):
        op.set_precision(prec)
        if tunings and prec in tunings:
            op.set_tuning(**tunings[prec])
        elif autotune:
            op.autotune(dx, dy, 5)
        op.profile(dx, dy, 1)
        _, _, ms = op.profile(dx, dy, iters)
        out[prec] = {"ms": round(ms, 4), "tflops": round(info["flops_algorithmic"] / (ms * 1e-3)
                                                         / 1e12, 2)}
        out[prec]["tuning"] = op.tuning()
        log(f"{name:28s} {prec:7s} {ms:8.3f} ms  {out[prec]['tflops']:7.2f} TF/s  {op.tuning()}")
    op.close()
    return out


def run_multi(prob, counts, iters, log):
    import time

    from . import BatchedGemm, DeviceArray

    out = []
    for prec in ("fp64", "fp32"):
        op = BatchedGemm(prob, prec)
        for nvec in counts:
            X = np.random.default_rng(3).uniform(-1, 1, nvec * prob.n)
            dX, dY = DeviceArray.from_host(X), DeviceArray(nvec * prob.n)
            op.apply_multi(dX, dY, nvec)
            op.sync()
            t = time.perf_counter()
            for _ in range(iters):
                op.apply_multi(dX, dY, nvec)
            op.sync()
            ms = (time.perf_counter() - t) / iters * 1e3
            out.append({"prec": prec, "nvec": nvec, "ms": round(ms, 3),
                        "ms_per_vec": round(ms / nvec, 3)})
            log(f"multi {prec:5s} nvec={nvec}: {ms:8.3f} ms, {ms / nvec:7.3f} ms per vector")
            dX.free()
            dY.free()
        op.close()
    return out


def run_vbatch(iters, log):
    """The batched GEMM (dmk_gemm_vbatch_d) on the shapes of the wave-function transformation
    of an m = 2048 ladder step: one (N, N) and one (T, N) product per symmetry block, blocks of
    the sizes below (m = n = k = s), timed by wall clock around each synchronous call."""
    from . import DeviceArray, gemm_vbatch

    sizes = [1040, 1038, 986, 654, 538, 245, 134, 60, 17, 4]
    rng = np.random.default_rng(5)
    keep, groups = [], []
    flops = 0.0
    for s in sizes:
        a, b, c = (DeviceArray.from_host(rng.uniform(-1, 1, s * s)) for _ in range(3))
        keep += [a, b, c]
        groups.append((s, s, s, s, s, s, [(a.ptr, b.ptr, c.ptr)]))
        flops += 2.0 * s ** 3
    out = {"sizes": sizes, "flops": flops}
    for ta in ("N", "T"):
        gemm_vbatch(ta, "N", groups)  # warm-up
        t = time.perf_counter()
        for _ in range(iters):
            gemm_vbatch(ta, "N", groups)
        ms = 1e3 * (time.perf_counter() - t) / iters
        out[ta + "N"] = {"ms": round(ms, 3), "tflops": round(flops / (ms * 1e-3) / 1e12, 2)}
        log(f"gemm_vbatch {ta}N {len(sizes)} blocks: {ms:.3f} ms, "
            f"{flops / (ms * 1e-3) / 1e12:.2f} TF/s")
    for a in keep:
        a.free()
    return out


def run_plan(prob, iters, log):
    """Handle creation (dmk_create_d: scan, staging, upload, packing) from host blocks and from
    the same blocks already in device memory (scan and packing on the device, nothing staged or
    uploaded): wall clock of the create call and the handle's own breakdown, medians over iters
    creations after one warm-up each. The copy of the blocks into device memory is the
    caller's and is not timed."""
    import ctypes

    from . import DmkInfo, _check, _device_plugin_args, lib

    out = {}
    for mode in ("host", "device"):
        args = prob.plugin_args()
        dev = None
        if mode == "device":
            args, dev = _device_plugin_args(prob, args)
        rec = []
        for it in range(iters + 1):
            h = ctypes.c_void_p()
            t = time.perf_counter()
            _check(lib().dmk_create_d(args.noperator, args.npatches, ctypes.addressof(args.L),
                                      ctypes.addressof(args.R), ctypes.addressof(args.A),
                                      ctypes.addressof(args.ldA), ctypes.addressof(args.B),
                                      ctypes.addressof(args.ldB), ctypes.byref(h)),
                   "dmk_create_d")
            ms = 1e3 * (time.perf_counter() - t)
            inf = DmkInfo()
            _check(lib().dmk_get_info(h.value, ctypes.byref(inf)), "dmk_get_info")
            info = inf.as_dict()
            lib().dmk_destroy(h.value)
            if it:
                rec.append((ms, info["ms_scan"], info["ms_stage"], info["ms_upload"]))
        if dev is not None:
            dev.free()
        r = np.median(np.array(rec), axis=0)
        out[mode] = {"ms_create": round(float(r[0]), 2), "ms_scan": round(float(r[1]), 2),
                     "ms_stage": round(float(r[2]), 2), "ms_upload": round(float(r[3]), 2),
                     "raw_mib": round(info["raw_bytes"] / 2**20, 1),
                     "device_blocks": info["device_blocks"]}
        log(f"plan from {mode} blocks: {out[mode]}")
    return out


def run_density(sizes, iters, log):
    from . import density_diag

    rng = np.random.default_rng(3)
    mats = []
    for n in sizes:
        x = rng.standard_normal((max(1, n // 2), n))
        rho = x.T @ x
        mats.append(rho / np.trace(rho))
    density_diag(mats)  # start-up of the solver's streams and workspaces: not timed
    out = {"sizes": list(sizes)}
    t = time.perf_counter()
    for _ in range(iters):
        density_diag(mats)
    out["ms_together"] = 1e3 * (time.perf_counter() - t) / iters
    single = []
    for a in mats:
        t = time.perf_counter()
        for _ in range(iters):
            density_diag([a])
        single.append(1e3 * (time.perf_counter() - t) / iters)
    out["ms_each"] = single
    out["ms_one_at_a_time"] = sum(single)
    log(f"density_diag {list(sizes)}: {out['ms_together']:.1f} ms together, "
        f"{out['ms_one_at_a_time']:.1f} ms one at a time ({', '.join(f'{v:.1f}' for v in single)})")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--replay", default=None, help="superblock .npz written by bench.py")
    ap.add_argument("--iters", type=int, default=30)
    ap.add_argument("--load-tuning", default=None,
                    help="per-precision tunings (JSON written by bench.py --save-tuning) for the "
                         "replayed superblock")
    ap.add_argument("--multi", default=None,
                    help="comma-separated vector counts: time dmk_apply_multi_d on the replayed "
                         "superblock instead of the single-vector cases")
    ap.add_argument("--vbatch", action="store_true",
                    help="time the batched GEMM on wave-function-transformation shapes instead")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune each case and precision (else the default launch shapes)")
    ap.add_argument("--density", default=None, nargs="?", const="1093,986,654,538",
                    help="comma-separated block sizes: time the truncation eigensolver instead")
    ap.add_argument("--plan", action="store_true",
                    help="time handle creation on the replayed superblock from host blocks and "
                         "from device-resident blocks instead")
    a = ap.parse_args(argv)
    tunings = json.load(open(a.load_tuning)) if a.load_tuning else None
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    if a.vbatch:
        print(json.dumps(run_vbatch(a.iters, log)), flush=True)
        return 0
    if a.density:
        sizes = [int(c) for c in a.density.split(",")]
        print(json.dumps(run_density(sizes, a.iters, log)), flush=True)
        return 0
    if a.plan:
        from .problem import load_npz

        if not a.replay:
            ap.error("--plan needs --replay")
        print(json.dumps(run_plan(load_npz(a.replay, name="superblock"), a.iters, log)),
              flush=True)
        return 0
    if a.multi:
        from .problem import load_npz

        if not a.replay:
            ap.error("--multi needs --replay")
        counts = [int(c) for c in a.multi.split(",")]
        print(json.dumps(run_multi(load_npz(a.replay, name="superblock"), counts, a.iters, log)),
              flush=True)
        return 0
    res = []
    p1 = problems.block_diagonal_problem()
    res.append(run_case("bench_BatchedGemm-40-patches", p1,
                        np.random.default_rng(42).uniform(-1, 1, p1.n), a.iters, log,
                        autotune=a.autotune))
    p2 = problems.comb_problem()
    res.append(run_case("test_vbatch-comb-8+8", p2, (np.arange(p2.n) + 1.0) / p2.n, a.iters, log,
                        autotune=a.autotune))
    if a.replay:
        from .problem import load_npz

        p3 = load_npz(a.replay, name="superblock")
        res.append(run_case("replay:" + p3.name, p3,
                            np.random.default_rng(1234).uniform(-1, 1, p3.n), a.iters, log,
                            tunings, autotune=a.autotune))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    sys.exit(main())
