"""The allocation-free matvec dmk_apply_noalloc_d against dmk_apply_d, through the C ABI.

Same kernels and launch tables on a static device scratch instead of the plan's pool buffers:
the products must be bitwise equal, in every precision mode, on consecutive calls of
different superblocks (the second smaller; the static scratch keeps the last call's contents,
so a stale stage-1 padding would show), with no allocation on the calls; a superblock whose
scratch exceeds the static array is refused before any work."""
import ctypes

import numpy as np
import pytest

import dmrgpp_amd as dm
from dmrgpp_amd import problems

pytestmark = pytest.mark.gpu


def _problem(scale: int, seed: int):
    left = {-2: 20 * scale, 0: 36 * scale, 2: 20 * scale}
    return problems.qn_problem(left, left, 0, problems.heisenberg_links(2), seed=seed)


def _noalloc(op, dx, dy):
    st = dm.lib().dmk_apply_noalloc_d(op.handle, dx.ptr, dy.ptr)
    if st != 0:
        raise dm.DmkError(dm.lib().dmk_last_error().decode())
    op.sync()


def _scratch_bytes(op) -> int:
    b = ctypes.c_size_t(0)
    dm._check(dm.lib().dmk_noalloc_scratch_bytes(op.handle, ctypes.byref(b)),
              "dmk_noalloc_scratch_bytes")
    return b.value


def test_noalloc_matches_apply_on_consecutive_calls():
    big, small = _problem(6, 11), _problem(2, 12)
    ops = {}
    for name, prob in (("big", big), ("small", small)):
        ops[name] = (dm.BatchedGemm(prob, "fp64"), prob)
    limit = 64 << 20
    for prec in ("fp64", "fp32", "bf16x3"):
        # The big superblock first, then the smaller one, then the big one again: each call
        # finds the previous call's data in the shared static scratch.
        for name in ("big", "small", "big"):
            op, prob = ops[name]
            op.set_precision(prec)
            assert _scratch_bytes(op) <= limit
            x = np.random.default_rng(hash((prec, name)) % 2**32).uniform(-1, 1, prob.n)
            dx = dm.DeviceArray.from_host(x)
            want = dm.DeviceArray(prob.n)
            got = dm.DeviceArray.from_host(np.full(prob.n, np.nan))
            op.apply_Htarget(dx, want)
            op.sync()
            before = dm.pool_counters()
            _noalloc(op, dx, got)
            _noalloc(op, dx, got)  # twice in a row on the same handle
            after = dm.pool_counters()
            assert after["allocs"] == before["allocs"], "the allocation-free call allocated"
            assert after["reuses"] == before["reuses"], "the allocation-free call used the pool"
            w, g = want.download(), got.download()
            assert np.array_equal(w, g), (prec, name, float(np.max(np.abs(w - g))))
            for a in (dx, want, got):
                a.free()
    for op, _ in ops.values():
        op.close()
    assert dm.pool_counters()["bad_releases"] == 0


def test_noalloc_refuses_a_superblock_beyond_the_static_scratch():
    # n = 2 * 900^2 + 1300^2 = 3.31e6: the FP64 input panels take 2 n doubles (53 MB) and the
    # stage-1 products about 37 MB, together over the 64 MiB of the static array.
    left = {-2: 900, 0: 1300, 2: 900}
    prob = problems.qn_problem(left, left, 0, problems.heisenberg_links(1), seed=5)
    op = dm.BatchedGemm(prob, "fp64")
    need = _scratch_bytes(op)
    assert need > (64 << 20)
    dx = dm.DeviceArray.from_host(np.ones(prob.n))
    dy = dm.DeviceArray.from_host(np.zeros(prob.n))
    st = dm.lib().dmk_apply_noalloc_d(op.handle, dx.ptr, dy.ptr)
    assert st != 0
    msg = dm.lib().dmk_last_error().decode()
    assert "static scratch" in msg and str(need) in msg
    op.sync()
    assert not np.any(dy.download())  # refused before any work was enqueued
    for a in (dx, dy):
        a.free()
    op.close()
