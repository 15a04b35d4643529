"""Parity of the HIP matvec (through the C ABI) with the CPU oracle, plus the properties that
hold at full size. GPU tests are marked ``gpu``; the C-ABI load/export and argument-error tests
run on the CPU."""
import ctypes
import re

import numpy as np
import pytest

import dmrgpp_amd as dm
from dmrgpp_amd import problems
from oracle import oracle
from tests import fixtures

HEADER = "include/dmrgpp_amd.h"


def rel_err(got, want):
    return np.linalg.norm(got - want) / max(np.linalg.norm(want), 1e-300)


# ----------------------------------------------------------------------------- CPU tests


def test_library_exports_every_declared_symbol():
    import os
    src = open(os.path.join(os.path.dirname(os.path.dirname(__file__)), HEADER)).read()
    declared = set(re.findall(r"\b(dmk_[a-z0-9_]+)\s*\(", src))
    assert declared, "no declarations parsed"
    lib = ctypes.CDLL(dm.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(dm.SIGNATURES), declared ^ set(dm.SIGNATURES)


def test_create_rejects_bad_arguments():
    L = (ctypes.c_size_t * 2)(3, 0)
    R = (ctypes.c_size_t * 2)(2, 2)
    A = (ctypes.c_void_p * 4)()
    ld = (ctypes.c_size_t * 4)(3, 3, 3, 3)
    h = ctypes.c_void_p()
    rc = dm.lib().dmk_create_d(1, 2, ctypes.addressof(L), ctypes.addressof(R),
                               ctypes.addressof(A), ctypes.addressof(ld), ctypes.addressof(A),
                               ctypes.addressof(ld), ctypes.byref(h))
    assert rc == -1
    assert b"patch sizes" in dm.lib().dmk_last_error()
    rc = dm.lib().dmk_create_d(0, 2, ctypes.addressof(L), ctypes.addressof(R),
                               ctypes.addressof(A), ctypes.addressof(ld), ctypes.addressof(A),
                               ctypes.addressof(ld), ctypes.byref(h))
    assert rc == -1
    assert h.value is None
    assert dm.lib().dmk_apply_d(None, None, None) == -1


# ----------------------------------------------------------------------------- GPU tests


def gpu_apply(prob, x, precision="fp64"):
    op = dm.BatchedGemm(prob, precision)
    dx = dm.DeviceArray.from_host(x)
    dy = dm.DeviceArray(prob.n)
    op.apply_Htarget(dx, dy)
    op.sync()
    y = dy.download()
    op.close()
    return y


@pytest.mark.gpu
def test_gpu_batchedgemm_fixture():
    prob, vin, expected, tol = fixtures.batchedgemm_fixture()
    got = gpu_apply(prob, vin)
    np.testing.assert_allclose(got, expected, rtol=tol, atol=0)


@pytest.mark.gpu
def test_gpu_kron31_fixture():
    full, _, y, _ = fixtures.kron31_problem()
    want = oracle.plugin_apply(full, y)
    got = gpu_apply(full, y)
    assert rel_err(got, want) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("seed,L,R,nC,density", [
    (1, [1, 1, 1], [1, 1, 1], 2, 1.0),
    (2, [3, 70, 5, 130], [65, 2, 64, 9], 3, 0.7),
    (3, [64, 128, 63, 1], [64, 1, 129, 200], 4, 0.5),
    (4, [17] * 9, [33] * 9, 5, 0.3),
    (5, [200, 7], [3, 150], 2, 1.0),
])
def test_gpu_random_problems_match_oracle(seed, L, R, nC, density):
    prob = problems.random_problem(L, R, nC, density=density, seed=seed)
    x = np.random.default_rng(seed).uniform(-1, 1, prob.n)
    want = oracle.plugin_apply(prob, x)
    got = gpu_apply(prob, x)
    assert rel_err(got, want) < 1e-12


@pytest.mark.gpu
def test_gpu_patch_without_terms_is_zeroed():
    # Patch 1 has no (1, J, k) block at all: its output segment must be overwritten with zeros.
    rng = np.random.default_rng(0)
    from dmrgpp_amd.problem import KronProblem
    blocks = {(0, 0, 0): (rng.uniform(-1, 1, (5, 5)), rng.uniform(-1, 1, (4, 4))),
              (0, 1, 0): (rng.uniform(-1, 1, (5, 3)), rng.uniform(-1, 1, (4, 6)))}
    prob = KronProblem(L=[5, 3], R=[4, 6], nC=1, blocks=blocks)
    x = rng.uniform(-1, 1, prob.n)
    op = dm.BatchedGemm(prob)
    dx = dm.DeviceArray.from_host(x)
    dy = dm.DeviceArray.from_host(np.full(prob.n, 7.0))
    op.apply_Htarget(dx, dy)
    op.sync()
    got = dy.download()
    assert np.all(got[20:] == 0.0)
    assert rel_err(got, oracle.plugin_apply(prob, x)) < 1e-12


@pytest.mark.gpu
def test_gpu_identity_and_general_terms_qn_problem():
    # QN-structured problem with identity factors (H_L (x) I, I (x) H_R) and link terms.
    left = {-3: 7, -1: 40, 1: 70, 3: 20}
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(2), seed=5)
    x = np.random.default_rng(2).uniform(-1, 1, prob.n)
    want = oracle.plugin_apply(prob, x)
    op = dm.BatchedGemm(prob)
    info = op.info()
    assert info["nterms_identity"] == 2 * prob.npatches
    op.close()
    got = gpu_apply(prob, x)
    assert rel_err(got, want) < 1e-12


@pytest.mark.gpu
def test_gpu_ladder_full_size_properties():
    # Full-size m=2048 ladder shape: linearity, symmetry <x|Hy> = <Hx|y>, bitwise determinism,
    # and agreement with per-block BLAS products.
    prob = problems.ladder_m2048_problem()
    rng = np.random.default_rng(1234)
    x = rng.uniform(-1, 1, prob.n)
    y = rng.uniform(-1, 1, prob.n)
    op = dm.BatchedGemm(prob)
    dx, dy, dz = dm.DeviceArray.from_host(x), dm.DeviceArray.from_host(y), dm.DeviceArray(prob.n)
    hx_d, hy_d = dm.DeviceArray(prob.n), dm.DeviceArray(prob.n)
    op.apply_Htarget(dx, hx_d)
    op.apply_Htarget(dy, hy_d)
    a, b = 0.75, -1.25
    dz.upload(a * x + b * y)
    hz_d = dm.DeviceArray(prob.n)
    op.apply_Htarget(dz, hz_d)
    op.sync()
    hx, hy, hz = hx_d.download(), hy_d.download(), hz_d.download()
    assert rel_err(hz, a * hx + b * hy) < 1e-12
    assert abs(x @ hy - hx @ y) < 1e-10 * abs(x @ hy)
    op.apply_Htarget(dx, hz_d)
    op.sync()
    assert np.array_equal(hz_d.download(), hx)
    want = prob.apply_numpy(x)
    assert rel_err(hx, want) < 1e-12
    op.close()


@pytest.mark.gpu
def test_gpu_fp32_mode_error_bound():
    left = {-2: 90, 0: 150, 2: 90}
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(3), seed=9)
    x = np.random.default_rng(3).uniform(-1, 1, prob.n)
    want = oracle.plugin_apply(prob, x)
    got = gpu_apply(prob, x, "fp32")
    err = rel_err(got, want)
    # FP32 products with FP32 accumulation: error ~ eps32 * sqrt(K), far below 1e-4.
    assert 1e-10 < err < 1e-5, err


@pytest.mark.gpu
def test_gpu_bf16x3_mode_error_bound():
    # (hi, lo) bf16 pairs carry ~2^-17 relative representation error per operand; with FP32
    # accumulation the matvec error sits between the FP32 mode's and 1e-4.
    left = {-2: 90, 0: 150, 2: 90}
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(3), seed=9)
    x = np.random.default_rng(3).uniform(-1, 1, prob.n)
    want = oracle.plugin_apply(prob, x)
    err16 = rel_err(gpu_apply(prob, x, "bf16x3"), want)
    err32 = rel_err(gpu_apply(prob, x, "fp32"), want)
    assert err32 < err16 < 1e-4, (err32, err16)


@pytest.mark.gpu
def test_gpu_multi_vector_and_accumulate():
    prob = problems.random_problem([30, 70, 12], [50, 9, 66], 3, density=0.8, seed=4)
    nvec = 4
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (nvec, prob.n))
    op = dm.BatchedGemm(prob)
    dX = dm.DeviceArray.from_host(X.reshape(-1))
    dY = dm.DeviceArray(nvec * prob.n)
    op.apply_multi(dX, dY, nvec)
    op.sync()
    Y = dY.download().reshape(nvec, prob.n)
    for v in range(nvec):
        assert np.array_equal(Y[v], gpu_apply(prob, X[v]))
    # beta = 1 accumulates on top of the previous result.
    op.apply_multi(dX, dY, nvec, beta=1)
    op.sync()
    Y2 = dY.download().reshape(nvec, prob.n)
    np.testing.assert_allclose(Y2, 2 * Y, rtol=1e-15, atol=0)
    # Host round trip gives the same numbers.
    assert np.array_equal(op.apply_host(X[0]), Y[0])
    op.close()


TUNINGS = [
    dict(stage1_tile_m=64, stage1_tile_n=64, stage2_tile_m=64, stage2_tile_n=64, splitk=0,
         xcd_order=0, mfma32=0, prefetch=1),
    dict(stage1_tile_m=128, stage1_tile_n=64, stage2_tile_m=64, stage2_tile_n=128, splitk=16,
         xcd_order=0, mfma32=1, prefetch=2),
    dict(stage1_tile_m=64, stage1_tile_n=128, stage2_tile_m=128, stage2_tile_n=128, splitk=32,
         xcd_order=1, mfma32=1, prefetch=1),
    dict(stage1_tile_m=128, stage1_tile_n=128, stage2_tile_m=128, stage2_tile_n=64, splitk=48,
         xcd_order=1, mfma32=0, prefetch=2),
    dict(stage1_tile_m=64, stage1_tile_n=64, stage2_tile_m=64, stage2_tile_n=64, splitk=64,
         xcd_order=0, mfma32=1, prefetch=2),
    # 96-wide tiles (FP64 / FP32; BF16x3 rejects them), in both MFMA settings.
    dict(stage1_tile_m=96, stage1_tile_n=64, stage2_tile_m=96, stage2_tile_n=96, splitk=32,
         xcd_order=1, mfma32=1, prefetch=2),
    dict(stage1_tile_m=64, stage1_tile_n=96, stage2_tile_m=64, stage2_tile_n=96, splitk=0,
         xcd_order=0, mfma32=0, prefetch=1),
    dict(stage1_tile_m=96, stage1_tile_n=96, stage2_tile_m=96, stage2_tile_n=64, splitk=48,
         xcd_order=1, mfma32=0, prefetch=2),
    # 32-wide tiles (FP64 / FP32; BF16x3 rejects them).
    dict(stage1_tile_m=32, stage1_tile_n=64, stage2_tile_m=64, stage2_tile_n=32, splitk=16,
         xcd_order=1, mfma32=1, prefetch=2),
    dict(stage1_tile_m=32, stage1_tile_n=32, stage2_tile_m=128, stage2_tile_n=32, splitk=0,
         xcd_order=0, mfma32=0, prefetch=1),
    dict(stage1_tile_m=32, stage1_tile_n=128, stage2_tile_m=32, stage2_tile_n=32, splitk=32,
         xcd_order=1, mfma32=0, prefetch=2),
    # Double-depth K chunks (FP64 / FP32; BF16x3 rejects them), both MFMA forms, forced split-k.
    dict(stage1_tile_m=64, stage1_tile_n=64, stage2_tile_m=64, stage2_tile_n=64, splitk=48,
         xcd_order=1, mfma32=1, prefetch=2, kchunk=2),
    dict(stage1_tile_m=32, stage1_tile_n=64, stage2_tile_m=64, stage2_tile_n=32, splitk=0,
         xcd_order=0, mfma32=0, prefetch=1, kchunk=2),
    # The balanced persistent stage 2 (on these small problems most tiles are cut between
    # workgroups, so nearly every output goes through the partials and the fix-up).
    dict(stage1_tile_m=64, stage1_tile_n=64, stage2_tile_m=64, stage2_tile_n=64, splitk=0,
         xcd_order=1, mfma32=1, prefetch=2, persist=8),
    dict(stage1_tile_m=64, stage1_tile_n=64, stage2_tile_m=128, stage2_tile_n=128, splitk=0,
         xcd_order=0, mfma32=0, prefetch=1, persist=1),
    dict(stage1_tile_m=64, stage1_tile_n=32, stage2_tile_m=128, stage2_tile_n=64, splitk=0,
         xcd_order=1, mfma32=1, prefetch=1, persist=2),
    dict(stage1_tile_m=64, stage1_tile_n=64, stage2_tile_m=96, stage2_tile_n=64, splitk=0,
         xcd_order=1, mfma32=0, prefetch=2, persist=3),
    dict(stage1_tile_m=32, stage1_tile_n=32, stage2_tile_m=32, stage2_tile_n=64, splitk=0,
         xcd_order=0, mfma32=0, prefetch=2, kchunk=2, persist=5),
]


@pytest.mark.gpu
@pytest.mark.parametrize("tuning", TUNINGS)
@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16x3", 1e-4)])
def test_gpu_tunings_match_oracle(tuning, prec, tol):
    # Every tile shape (64 / 128, 96-wide, 32-wide), forced split-k (splitk 16-64: cuts on
    # multiples of 64 along K, so some pieces are empty and the last is ragged), both
    # workgroup orders and both FP32 MFMA forms give the oracle's result; the same tuning twice
    # is bitwise identical.
    shapes = (tuning["stage1_tile_m"], tuning["stage1_tile_n"], tuning["stage2_tile_m"],
              tuning["stage2_tile_n"])
    if prec == "bf16x3" and (96 in shapes or 32 in shapes or tuning.get("kchunk", 1) == 2):
        op = dm.BatchedGemm(problems.random_problem([5, 7], [6, 3], 2, seed=1), prec)
        with pytest.raises(dm.DmkError):
            op.set_tuning(**tuning)
        op.close()
        return
    left = {-2: 70, 0: 130, 2: 70}
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(2), seed=21)
    rand = problems.random_problem([3, 70, 5, 130], [65, 2, 64, 9], 3, density=0.7, seed=2)
    for p in (prob, rand):
        x = np.random.default_rng(7).uniform(-1, 1, p.n)
        want = oracle.plugin_apply(p, x)
        op = dm.BatchedGemm(p, prec)
        op.set_tuning(**tuning)
        assert op.tuning() == dict({"kchunk": 1, "persist": 0}, **tuning)
        if tuning["splitk"]:
            assert op.info()["split_tasks"] > 0
        if tuning.get("persist"):
            assert op.info()["split_tasks"] == 0 and op.info()["ws_elems"] > 0
        dx, dy = dm.DeviceArray.from_host(x), dm.DeviceArray(p.n)
        op.apply_Htarget(dx, dy)
        op.sync()
        y1 = dy.download()
        op.apply_Htarget(dx, dy)
        op.sync()
        assert np.array_equal(dy.download(), y1)
        assert rel_err(y1, want) < tol
        # beta = 1 through the split-k reduction as well.
        op.apply_multi(dx, dy, 1, beta=1)
        op.sync()
        assert rel_err(dy.download(), 2 * want) < tol
        op.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_gpu_balanced_launch_multi_vector(prec):
    # Several vectors in one balanced persistent launch (each vector's partials in its own part
    # of the workspace) equal the vectors one at a time, bitwise, with beta = 0 and 1.
    left = {-2: 70, 0: 130, 2: 70}
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(2), seed=21)
    op = dm.BatchedGemm(prob, prec)
    op.set_tuning(stage1_tile_m=64, stage1_tile_n=64, stage2_tile_m=128, stage2_tile_n=64,
                  splitk=0, xcd_order=1, mfma32=0, prefetch=1, persist=8)
    nvec, n = 3, prob.n
    X = np.random.default_rng(4).uniform(-1, 1, nvec * n)
    Y0 = np.random.default_rng(5).uniform(-1, 1, nvec * n)
    dX, dY = dm.DeviceArray.from_host(X), dm.DeviceArray.from_host(Y0)
    op.apply_multi(dX, dY, nvec, beta=1)
    op.sync()
    multi = dY.download()
    for v in range(nvec):
        dx = dm.DeviceArray.from_host(X[v * n:(v + 1) * n])
        dy = dm.DeviceArray.from_host(Y0[v * n:(v + 1) * n])
        op.apply_multi(dx, dy, 1, beta=1)
        op.sync()
        assert np.array_equal(dy.download(), multi[v * n:(v + 1) * n]), v
        want = Y0[v * n:(v + 1) * n] + oracle.plugin_apply(prob, X[v * n:(v + 1) * n])
        tol = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 1e-4}[prec]
        assert rel_err(multi[v * n:(v + 1) * n], want) < tol
        dx.free()
        dy.free()
    op.close()


@pytest.mark.gpu
def test_gpu_autotune_keeps_results():
    left = {-2: 90, 0: 150, 2: 90}
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(3), seed=9)
    x = np.random.default_rng(3).uniform(-1, 1, prob.n)
    want = oracle.plugin_apply(prob, x)
    op = dm.BatchedGemm(prob)
    dx, dy = dm.DeviceArray.from_host(x), dm.DeviceArray(prob.n)
    chosen = op.autotune(dx, dy, iters=2)
    assert chosen == op.tuning()
    op.apply_Htarget(dx, dy)
    op.sync()
    assert rel_err(dy.download(), want) < 1e-12
    op.close()
    # A new handle for the same histogram starts from the cached choice.
    op2 = dm.BatchedGemm(prob)
    assert op2.tuning() == chosen
    op2.close()


@pytest.mark.gpu
def test_gpu_tuning_belongs_to_its_precision():
    # A tuning set in one precision mode is kept for that mode only: switching modes selects
    # each mode's own tuning, and the results stay the oracle's.
    left = {-2: 40, 0: 70, 2: 40}
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(2), seed=5)
    x = np.random.default_rng(4).uniform(-1, 1, prob.n)
    want = oracle.plugin_apply(prob, x)
    op = dm.BatchedGemm(prob, "fp64")
    t64 = op.tuning()
    big = dict(t64, stage1_tile_m=128, stage1_tile_n=128, stage2_tile_m=128, stage2_tile_n=128,
               splitk=0, mfma32=1, kchunk=1)
    op.set_precision("fp32")
    op.set_tuning(**big)
    op.set_precision("fp64")
    assert op.tuning() == t64
    dx, dy = dm.DeviceArray.from_host(x), dm.DeviceArray(prob.n)
    op.apply_Htarget(dx, dy)
    op.sync()
    assert rel_err(dy.download(), want) < 1e-12
    op.set_precision("fp32")
    assert op.tuning() == big
    op.apply_Htarget(dx, dy)
    op.sync()
    assert rel_err(dy.download(), want) < 1e-5
    op.close()


def test_set_tuning_rejects_bad_shapes():
    t = dm.Tuning(stage1_tile_m=96, stage1_tile_n=64, stage2_tile_m=64, stage2_tile_n=64,
                  splitk=0, xcd_order=0)
    assert dm.lib().dmk_set_tuning(None, ctypes.byref(t)) == -1


def test_sector_map_rejects_non_permutations():
    assert dm.lib().dmk_set_sector_map(None, None) == -1


@pytest.mark.gpu
def test_gpu_sector_ordered_matvec_accumulates():
    # KronMatrix semantics: vout += P^T H P vin with P the copyIn gather.
    prob = problems.random_problem([5, 9, 2, 13], [7, 3, 11, 4], nC=3, density=0.7, seed=12)
    rng = np.random.default_rng(12)
    perm = rng.permutation(prob.n)
    vin = rng.uniform(-1, 1, prob.n)
    vout0 = rng.uniform(-1, 1, prob.n)
    want = vout0.copy()
    want[perm] += oracle.plugin_apply(prob, vin[perm])
    op = dm.BatchedGemm(prob)
    op.set_sector_map(perm)
    din, dout = dm.DeviceArray.from_host(vin), dm.DeviceArray.from_host(vout0)
    op.matvec_sector(din, dout)
    op.sync()
    assert rel_err(dout.download(), want) < 1e-13
    bad = perm.copy()
    bad[0] = bad[1]
    with pytest.raises(dm.DmkError):
        op.set_sector_map(bad)


@pytest.mark.gpu
def test_gpu_shared_blocks_uploaded_once_and_packed_on_device():
    # Blocks shared by several terms are scanned and uploaded once; identity factors are not
    # uploaded at all (they are contracted directly). The device-packed panels give the
    # oracle's result, and the handle reports its construction timings.
    from dmrgpp_amd.problem import KronProblem

    rng = np.random.default_rng(17)
    f = lambda r, c: np.asfortranarray(rng.uniform(-1, 1, (r, c)))  # noqa: E731
    a_shared, b_shared = f(30, 30), f(20, 20)
    blocks = {(0, 0, 0): (a_shared, f(20, 20)), (1, 1, 0): (a_shared, f(20, 20)),
              (0, 1, 1): (f(30, 30), b_shared), (1, 0, 1): (f(30, 30), b_shared),
              (0, 0, 1): (f(30, 30), np.asfortranarray(np.eye(20)))}
    prob = KronProblem(L=[30, 30], R=[20, 20], nC=2, blocks=blocks)
    x = rng.uniform(-1, 1, prob.n)
    op = dm.BatchedGemm(prob)
    info = op.info()
    assert info["nterms"] == 5 and info["nterms_identity"] == 1
    assert info["raw_bytes"] == 8 * (4 * 30 * 30 + 3 * 20 * 20)
    assert info["ms_scan"] >= 0 and info["ms_upload"] > 0
    dx, dy = dm.DeviceArray.from_host(x), dm.DeviceArray(prob.n)
    op.apply_Htarget(dx, dy)
    op.sync()
    assert rel_err(dy.download(), oracle.plugin_apply(prob, x)) < 1e-12
    op.close()
