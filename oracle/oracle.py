"""ctypes binding of the CPU oracle (oracle/liboracle.so). Test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module. See
kron_oracle.cpp for the reference functions each entry point restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_VP = ctypes.c_void_p
_ARGS = [ctypes.c_int64, ctypes.c_int64, _VP, _VP, _VP, _VP, _VP, _VP]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for suf in ("d", "z"):
            f = getattr(L, f"orc_kron_connections_{suf}")
            f.restype = ctypes.c_int
            f.argtypes = _ARGS + [_VP, _VP, ctypes.c_int, ctypes.c_int, ctypes.c_double]
            f = getattr(L, f"orc_plugin_apply_{suf}")
            f.restype = ctypes.c_int
            f.argtypes = _ARGS + [_VP, _VP]
            f = getattr(L, f"orc_kron_form_{suf}")
            f.restype = ctypes.c_int
            f.argtypes = _ARGS + [_VP]
            f = getattr(L, f"orc_lanczos_{suf}")
            f.restype = ctypes.c_int
            f.argtypes = _ARGS + [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  _VP, _VP, _VP, _VP, _VP]
        for suf in ("d", "z"):
            f = getattr(L, f"orc_den_kron_mult_{suf}")
            f.restype = ctypes.c_int
            f.argtypes = [ctypes.c_int, ctypes.c_char, ctypes.c_char, _VP, ctypes.c_int64,
                          ctypes.c_int64, _VP, ctypes.c_int64, ctypes.c_int64, _VP, _VP]
        L.orc_estimate_kron_method.restype = ctypes.c_int
        L.orc_estimate_kron_method.argtypes = [ctypes.c_int64] * 6 + [ctypes.c_double]
        L.orc_krylov_z.restype = ctypes.c_int
        L.orc_krylov_z.argtypes = _ARGS + [_VP, ctypes.c_int, _VP, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_double, _VP, _VP]
        L.orc_tridiag_eigen.restype = ctypes.c_int
        L.orc_tridiag_eigen.argtypes = [ctypes.c_int, _VP, _VP, _VP]
        _lib = L
    return _lib


def _suffix(prob):
    return "z" if np.issubdtype(prob.dtype, np.complexfloating) else "d"


def _pargs(args):
    return (args.noperator, args.npatches, ctypes.addressof(args.L), ctypes.addressof(args.R),
            ctypes.addressof(args.A), ctypes.addressof(args.ldA), ctypes.addressof(args.B),
            ctypes.addressof(args.ldB))


def plugin_apply(prob, x: np.ndarray) -> np.ndarray:
    """vout = H vin as the reference plugin's two-stage evaluation."""
    args = prob.plugin_args()
    x = np.ascontiguousarray(x, dtype=prob.dtype)
    y = np.zeros(prob.n, dtype=prob.dtype)
    rc = getattr(lib(), f"orc_plugin_apply_{_suffix(prob)}")(*_pargs(args), x.ctypes.data,
                                                            y.ctypes.data)
    assert rc == 0
    return y


def kron_connections(prob, y: np.ndarray, x0: np.ndarray | None = None, apply_adjoint=False,
                     use_lower_part=False, dense_flop_discount=1.0) -> np.ndarray:
    """x += H y as KronConnections::doTask over all out-patches (x starts at x0 or 0)."""
    args = prob.plugin_args()
    y = np.ascontiguousarray(y, dtype=prob.dtype)
    x = np.zeros(prob.n, dtype=prob.dtype) if x0 is None else np.array(x0, dtype=prob.dtype)
    rc = getattr(lib(), f"orc_kron_connections_{_suffix(prob)}")(
        *_pargs(args), y.ctypes.data, x.ctypes.data, int(apply_adjoint), int(use_lower_part),
        float(dense_flop_discount))
    assert rc == 0
    return x


def kron_form(prob) -> np.ndarray:
    """Explicit dense H (n x n)."""
    args = prob.plugin_args()
    H = np.zeros((prob.n, prob.n), dtype=prob.dtype, order="F")
    rc = getattr(lib(), f"orc_kron_form_{_suffix(prob)}")(*_pargs(args), H.ctypes.data)
    assert rc == 0
    return H


def den_kron_mult(imethod: int, transA: str, transB: str, a: np.ndarray, b: np.ndarray,
                  y: np.ndarray, x: np.ndarray) -> np.ndarray:
    """x += kron(op(A), op(B)) y with an explicit method; y and x are column-major flattened."""
    cplx = np.iscomplexobj(a) or np.iscomplexobj(b) or np.iscomplexobj(y) or np.iscomplexobj(x)
    dt = np.complex128 if cplx else np.float64
    a = np.asfortranarray(a, dtype=dt)
    b = np.asfortranarray(b, dtype=dt)
    y = np.ascontiguousarray(y, dtype=dt)
    x = np.array(x, dtype=dt)
    f = getattr(lib(), "orc_den_kron_mult_z" if cplx else "orc_den_kron_mult_d")
    rc = f(int(imethod), transA.encode(), transB.encode(), a.ctypes.data, a.shape[0], a.shape[1],
           b.ctypes.data, b.shape[0], b.shape[1], y.ctypes.data, x.ctypes.data)
    assert rc == 0
    return x


def estimate_kron_method(nrow_A, ncol_A, nnz_A, nrow_B, ncol_B, nnz_B, discount=1.0) -> int:
    return lib().orc_estimate_kron_method(nrow_A, ncol_A, nnz_A, nrow_B, ncol_B, nnz_B, discount)


def tridiag_eigen(d: np.ndarray, e: np.ndarray, vectors=True):
    d = np.array(d, dtype=np.float64)
    n = d.size
    e = np.ascontiguousarray(np.concatenate([np.asarray(e, np.float64), [0.0]])[:max(n, 1)])
    z = np.zeros((n, n), dtype=np.float64, order="F") if vectors else None
    rc = lib().orc_tridiag_eigen(n, d.ctypes.data, e.ctypes.data,
                                 z.ctypes.data if vectors else None)
    if rc != 0:
        raise RuntimeError("tridiagonal QL did not converge")
    return (d, z) if vectors else d


def lanczos(prob, init: np.ndarray, max_steps=200, min_steps=4, tol=1e-12, reortho=False):
    """Lowest eigenpair by the LanczosCore recurrence with the plugin matvec."""
    args = prob.plugin_args()
    init = np.ascontiguousarray(init, dtype=prob.dtype)
    z = np.zeros(prob.n, dtype=prob.dtype)
    energy = ctypes.c_double()
    steps = ctypes.c_int()
    ms = max(1, min(max_steps, prob.n))
    alpha = np.zeros(ms)
    beta = np.zeros(ms)
    rc = getattr(lib(), f"orc_lanczos_{_suffix(prob)}")(
        *_pargs(args), init.ctypes.data, int(max_steps), int(min_steps), float(tol),
        int(reortho), ctypes.addressof(energy), z.ctypes.data, alpha.ctypes.data,
        beta.ctypes.data, ctypes.addressof(steps))
    if rc != 0:
        raise RuntimeError(f"oracle lanczos failed ({rc})")
    s = steps.value
    return energy.value, z, s, alpha[:s], beta[:s]


def krylov(prob, phi: np.ndarray, times, Eg=0.0, time_direction=1.0, max_steps=200, min_steps=4,
           tol=1e-12):
    """Krylov time step of one vector (TimeVectorsKrylov::calcTimeVectors restated): returns
    (out[ntimes, n] complex, steps). A real problem is promoted to complex blocks."""
    if not np.issubdtype(prob.dtype, np.complexfloating):
        from dmrgpp_amd.problem import KronProblem
        prob = KronProblem(L=prob.L, R=prob.R, nC=prob.nC,
                           blocks={k: (a.astype(np.complex128), b.astype(np.complex128))
                                   for k, (a, b) in prob.blocks.items()},
                           dtype=np.complex128, name=prob.name)
    args = prob.plugin_args()
    phi = np.ascontiguousarray(phi, dtype=np.complex128)
    times = np.ascontiguousarray(times, dtype=np.float64)
    out = np.zeros((times.size, prob.n), dtype=np.complex128)
    steps = ctypes.c_int()
    rc = lib().orc_krylov_z(*_pargs(args), phi.ctypes.data, int(times.size), times.ctypes.data,
                            float(Eg), float(time_direction), int(max_steps), int(min_steps),
                            float(tol), out.ctypes.data, ctypes.addressof(steps))
    if rc != 0:
        raise RuntimeError(f"oracle krylov failed ({rc})")
    return out, steps.value
