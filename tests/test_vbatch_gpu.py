"""Variable-size batched GEMM (dmk_gemm_vbatch_{d,z}, the dmrg_Xgemm_vbatch boundary) against
numpy on the same inputs."""
import numpy as np
import pytest

import dmrgpp_amd as dm


def _op(x, t):
    return {"N": x, "T": x.T, "C": x.conj().T}[t]


def _run(transa, transb, cplx, alpha, beta, seed):
    rng = np.random.default_rng(seed)
    dt = np.complex128 if cplx else np.float64
    # Mixed shapes: tiny, ragged, tile-crossing, k = 0 and an empty entry.
    shapes = [(1, 1, 1), (7, 5, 3), (64, 64, 64), (65, 130, 17), (200, 33, 129), (3, 90, 0),
              (0, 4, 5), (150, 1, 70)]
    groups, keep, want, dev = [], [], [], []
    for (m, n, k) in shapes:
        ra, ca = (k, m) if transa != "N" else (m, k)
        rb, cb = (n, k) if transb != "N" else (k, n)
        lda, ldb, ldc = ra + 3, rb + 1, m + 2

        def rnd(r, c, ld):
            x = rng.uniform(-1, 1, (ld, max(c, 1)))
            if cplx:
                x = x + 1j * rng.uniform(-1, 1, (ld, max(c, 1)))
            return np.asfortranarray(x.astype(dt))
        entries = []
        for _ in range(2):
            A, B, C = rnd(ra, ca, max(lda, 1)), rnd(rb, cb, max(ldb, 1)), rnd(m, n, max(ldc, 1))
            if beta == 0:
                C[:] = np.nan  # must not be read
            Av, Bv = A[:ra, :ca], B[:rb, :cb]
            Cw = C.copy()
            if m and n:
                prod = _op(Av, transa) @ _op(Bv, transb) if k else np.zeros((m, n), dt)
                Cw[:m, :n] = alpha * prod + (beta * C[:m, :n] if beta != 0 else 0)
            dA = dm.DeviceArray.from_host(A.ravel(order="F").view(np.float64))
            dB = dm.DeviceArray.from_host(B.ravel(order="F").view(np.float64))
            dC = dm.DeviceArray.from_host(C.ravel(order="F").view(np.float64))
            dev += [dA, dB, dC]
            want.append((dC, Cw, m, n))
            entries.append((dA.ptr, dB.ptr, dC.ptr))
        groups.append((m, n, k, max(lda, 1), max(ldb, 1), max(ldc, 1), entries))
    dm.gemm_vbatch(transa, transb, groups, alpha=alpha, beta=beta, cplx=cplx)
    for dC, Cw, m, n in want:
        got = dC.download().view(dt).reshape(Cw.shape, order="F")
        if m and n:
            ref = Cw[:m, :n]
            err = np.abs(got[:m, :n] - ref).max() / max(1.0, np.abs(ref).max())
            assert err < 1e-13, (m, n, err)


def test_vbatch_rejects_bad_arguments():
    one = (dm.ctypes.c_int * 1)(4)
    ptr = (dm.ctypes.c_void_p * 1)(0)
    L = dm.lib()
    a = dm.ctypes.addressof
    assert L.dmk_gemm_vbatch_d(b"X", b"N", a(one), a(one), a(one), 1.0, a(ptr), a(one), a(ptr),
                               a(one), 0.0, a(ptr), a(one), 1, a(one)) == -1
    assert b"transa" in L.dmk_last_error()
    small = (dm.ctypes.c_int * 1)(2)
    assert L.dmk_gemm_vbatch_d(b"N", b"N", a(one), a(one), a(one), 1.0, a(ptr), a(small),
                               a(ptr), a(one), 0.0, a(ptr), a(one), 1, a(one)) == -1
    assert b"leading dimension" in L.dmk_last_error()


@pytest.mark.gpu
@pytest.mark.parametrize("transa,transb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (-0.5, 1.0), (2.0, 0.25)])
def test_gpu_vbatch_real(transa, transb, alpha, beta):
    _run(transa, transb, False, alpha, beta, seed=ord(transa) * 7 + ord(transb) + int(10 * alpha))


@pytest.mark.gpu
@pytest.mark.parametrize("transa,transb", [("N", "N"), ("C", "N"), ("T", "C")])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.5 - 1j, 0.3 + 0.2j)])
def test_gpu_vbatch_complex(transa, transb, alpha, beta):
    _run(transa, transb, True, alpha, beta, seed=7)


@pytest.mark.gpu
def test_gpu_vbatch_alpha_zero_scales_c_only():
    # alpha = 0: C = beta C, A and B (NaN here) are not used.
    m = n = k = 20
    A = np.full((m, k), np.nan)
    C = np.arange(m * n, dtype=float).reshape(m, n, order="F")
    dA = dm.DeviceArray.from_host(A.ravel(order="F"))
    dC = dm.DeviceArray.from_host(C.ravel(order="F"))
    dm.gemm_vbatch("N", "N", [(m, n, k, m, k, m, [(dA.ptr, dA.ptr, dC.ptr)])], alpha=0.0,
                   beta=3.0)
    np.testing.assert_array_equal(dC.download().reshape(m, n, order="F"), 3.0 * C)
