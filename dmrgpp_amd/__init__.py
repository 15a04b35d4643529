"""dmrgpp_amd: MI355X-native superblock Kron matvec for DMRG++ (Python host mirror).

The product is the C-ABI library ``libdmrgpp_amd.so`` built from ``dmrgpp_amd/csrc`` (HIP for
gfx950). This module loads it with ctypes and mirrors the reference plugin interface
``BatchedGemm<T>`` (dmrg/GPUPlugin/BatchedGemm.h:11-56): construct from the patch sizes and the
A/B block pointers, then ``apply_Htarget(vin, vout)`` computes vout = H vin.

There is no CPU fallback: if the library is missing or no GPU is visible the calls raise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .problem import KronProblem, PluginArgs  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdmrgpp_amd.so")

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16X3 = 2
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}


class DmkError(RuntimeError):
    pass


class DmkInfo(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64), ("npatches", ctypes.c_int64), ("noperator", ctypes.c_int64),
        ("nterms", ctypes.c_int64), ("nterms_identity", ctypes.c_int64),
        ("stage1_tasks", ctypes.c_int64), ("stage2_tasks", ctypes.c_int64),
        ("stage1_tiles", ctypes.c_int64), ("stage2_tiles", ctypes.c_int64),
        ("max_patch_left", ctypes.c_int64), ("max_patch_right", ctypes.c_int64),
        ("apack_elems", ctypes.c_int64), ("bpack_elems", ctypes.c_int64),
        ("bx_elems", ctypes.c_int64),
        ("flops_algorithmic", ctypes.c_double), ("flops_plugin", ctypes.c_double),
        ("flops_executed", ctypes.c_double), ("flops_stage1", ctypes.c_double),
        ("flops_stage2", ctypes.c_double), ("bytes_compulsory", ctypes.c_double),
        ("bytes_bx", ctypes.c_double), ("precision", ctypes.c_int32), ("pad", ctypes.c_int32),
        ("split_tasks", ctypes.c_int64), ("ws_elems", ctypes.c_int64),
        ("is_complex", ctypes.c_int32), ("device_blocks", ctypes.c_int32), ("n_complex", ctypes.c_int64),
        ("ms_scan", ctypes.c_double), ("ms_stage", ctypes.c_double),
        ("ms_upload", ctypes.c_double), ("raw_bytes", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_
                if name != "pad"}


class Tuning(ctypes.Structure):
    _fields_ = [("stage1_tile_m", ctypes.c_int32), ("stage1_tile_n", ctypes.c_int32),
                ("stage2_tile_m", ctypes.c_int32), ("stage2_tile_n", ctypes.c_int32),
                ("splitk", ctypes.c_int32), ("xcd_order", ctypes.c_int32),
                ("mfma32", ctypes.c_int32), ("prefetch", ctypes.c_int32),
                ("kchunk", ctypes.c_int32), ("persist", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class LanczosParams(ctypes.Structure):
    _fields_ = [("max_steps", ctypes.c_int32), ("min_steps", ctypes.c_int32),
                ("tolerance", ctypes.c_double), ("reortho", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("refine", ctypes.c_int32),
                ("refine_min_steps", ctypes.c_int32), ("refine_tolerance", ctypes.c_double),
                ("graph", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("stop_residual", ctypes.c_double), ("switch_noise", ctypes.c_double)]


class LanczosResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("residual", ctypes.c_double),
                ("steps", ctypes.c_int32), ("matvecs", ctypes.c_int32),
                ("energy_first", ctypes.c_double), ("steps_first", ctypes.c_int32),
                ("matvecs_first", ctypes.c_int32), ("precision_first", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("ms_total", ctypes.c_double),
                ("ms_matvec", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


KRYLOV_STOP = {"energy": 0, "expm": 1}


class KrylovParams(ctypes.Structure):
    _fields_ = [("max_steps", ctypes.c_int32), ("min_steps", ctypes.c_int32),
                ("tolerance", ctypes.c_double), ("stop", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("time_direction", ctypes.c_double),
                ("energy_shift", ctypes.c_double)]


class KrylovResult(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int32), ("breakdown", ctypes.c_int32),
                ("norm", ctypes.c_double), ("error_estimate", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class KrylovStats(ctypes.Structure):
    _fields_ = [("ms_total", ctypes.c_double), ("ms_matvec", ctypes.c_double),
                ("matvec_launches", ctypes.c_int32), ("pad", ctypes.c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad"}


PRECONDS = {"none": 0, "diag": 1, "mlp": 2}
MLP_PARAMS = 4545
MLP_FEATURES = 4


class DavidsonParams(ctypes.Structure):
    _fields_ = [("max_matvecs", ctypes.c_int32), ("max_basis", ctypes.c_int32),
                ("tolerance", ctypes.c_double), ("precond", ctypes.c_int32),
                ("harvest", ctypes.c_int32), ("harvest_samples", ctypes.c_int32),
                ("harvest_iters", ctypes.c_int32)]


class DavidsonResult(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("residual", ctypes.c_double),
                ("iterations", ctypes.c_int32), ("matvecs", ctypes.c_int32),
                ("restarts", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("samples", ctypes.c_int64), ("ms_total", ctypes.c_double),
                ("ms_matvec", ctypes.c_double), ("ms_precond", ctypes.c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None

# Every symbol include/dmrgpp_amd.h declares, with its ctypes signature.
_VP = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I = ctypes.c_int
_D = ctypes.c_double
SIGNATURES = {
    "dmk_create_d": (_I, [_SZ, _SZ, _VP, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(_VP)]),
    "dmk_create_z": (_I, [_SZ, _SZ, _VP, _VP, _VP, _VP, _VP, _VP, ctypes.POINTER(_VP)]),
    "dmk_destroy": (None, [_VP]),
    "dmk_apply_z": (_I, [_VP, _VP, _VP]),
    "dmk_apply_d": (_I, [_VP, _VP, _VP]),
    "dmk_apply_host_d": (_I, [_VP, _VP, _VP]),
    "dmk_apply_shared": (_I, [_VP, _VP, _VP]),
    "dmk_apply_noalloc_d": (_I, [_VP, _VP, _VP]),
    "dmk_noalloc_scratch_bytes": (_I, [_VP, ctypes.POINTER(_SZ)]),
    "dmk_apply_multi_d": (_I, [_VP, _I, _VP, _SZ, _VP, _SZ, _I]),
    "dmk_set_sector_map": (_I, [_VP, _VP]),
    "dmk_matvec_sector_d": (_I, [_VP, _VP, _VP]),
    "dmk_set_precision": (_I, [_VP, _I]),
    "dmk_get_tuning": (_I, [_VP, ctypes.POINTER(Tuning)]),
    "dmk_set_tuning": (_I, [_VP, ctypes.POINTER(Tuning)]),
    "dmk_autotune_d": (_I, [_VP, _VP, _VP, _I, ctypes.POINTER(Tuning)]),
    "dmk_get_info": (_I, [_VP, ctypes.POINTER(DmkInfo)]),
    "dmk_sync": (_I, [_VP]),
    "dmk_profile_apply_d": (_I, [_VP, _VP, _VP, _I, ctypes.POINTER(_D), ctypes.POINTER(_D),
                                 ctypes.POINTER(_D)]),
    "dmk_lanczos_default_params": (_I, [ctypes.POINTER(LanczosParams)]),
    "dmk_lanczos_ground_d": (_I, [_VP, _VP, _VP, ctypes.POINTER(LanczosParams),
                                  ctypes.POINTER(LanczosResult)]),
    "dmk_tridiag_lowest": (_I, [_I, _VP, _VP, ctypes.POINTER(_D), _VP]),
    "dmk_tridiag_eigen": (_I, [_I, _VP, _VP, _VP, _VP]),
    "dmk_gemm_vbatch_d": (_I, [ctypes.c_char, ctypes.c_char, _VP, _VP, _VP, _D, _VP, _VP, _VP,
                               _VP, _D, _VP, _VP, _SZ, _VP]),
    "dmk_gemm_vbatch_z": (_I, [ctypes.c_char, ctypes.c_char, _VP, _VP, _VP, _VP, _VP, _VP, _VP,
                               _VP, _VP, _VP, _VP, _SZ, _VP]),
    "dmk_krylov_default_params": (_I, [ctypes.POINTER(KrylovParams)]),
    "dmk_krylov_evolve_z": (_I, [_VP, _I, _VP, _SZ, _I, _VP, _VP, _SZ,
                                 ctypes.POINTER(KrylovParams), _VP, ctypes.POINTER(KrylovStats)]),
    "dmk_diagonal_d": (_I, [_VP, _VP]),
    "dmk_density_diag_d": (_I, [_I, _VP, _VP, _VP]),
    "dmk_density_diag_dev_d": (_I, [_I, _VP, _VP, _VP, _VP]),
    "dmk_density_svd_d": (_I, [_I, _VP, _VP, _VP, _VP, _VP]),
    "dmk_mlp_create": (_I, [ctypes.c_uint64, ctypes.c_int64, ctypes.POINTER(_VP)]),
    "dmk_mlp_destroy": (None, [_VP]),
    "dmk_mlp_get_params": (_I, [_VP, _VP]),
    "dmk_mlp_set_params": (_I, [_VP, _VP]),
    "dmk_mlp_forward": (_I, [_VP, _VP, ctypes.c_int64, _VP]),
    "dmk_mlp_gradient": (_I, [_VP, _VP, _VP, ctypes.c_int64, _VP, ctypes.POINTER(_D)]),
    "dmk_mlp_add_samples": (_I, [_VP, _VP, _VP, ctypes.c_int64]),
    "dmk_mlp_samples": (_I, [_VP, ctypes.POINTER(ctypes.c_int64)]),
    "dmk_mlp_clear_samples": (_I, [_VP]),
    "dmk_mlp_train": (_I, [_VP, _I, ctypes.c_int64, ctypes.c_float, ctypes.POINTER(_D),
                           ctypes.POINTER(_D)]),
    "dmk_davidson_default_params": (_I, [ctypes.POINTER(DavidsonParams)]),
    "dmk_davidson_ground_d": (_I, [_VP, _VP, _VP, ctypes.POINTER(DavidsonParams), _VP,
                                   ctypes.POINTER(DavidsonResult)]),
    "dmk_device_alloc": (_I, [_SZ, ctypes.POINTER(_VP)]),
    "dmk_device_free": (_I, [_VP]),
    "dmk_memcpy_h2d": (_I, [_VP, _VP, _SZ]),
    "dmk_memcpy_d2h": (_I, [_VP, _VP, _SZ]),
    "dmk_device_count": (_I, [ctypes.POINTER(_I)]),
    "dmk_facade_create_d": (_I, [_SZ, _SZ, _VP, _VP, _VP, _VP, _VP, _VP, ctypes.c_int64, _I,
                                 ctypes.POINTER(_VP)]),
    "dmk_facade_create_z": (_I, [_SZ, _SZ, _VP, _VP, _VP, _VP, _VP, _VP, ctypes.c_int64, _I,
                                 ctypes.POINTER(_VP)]),
    "dmk_facade_matvec": (_I, [_VP, _VP, _VP]),
    "dmk_facade_get_info": (_I, [_VP, _VP]),
    "dmk_facade_destroy": (None, [_VP]),
    "dmk_pool_counters": (_I, [_I, _VP]),
    "dmk_device_copy_d": (_I, [_VP, _VP, _SZ]),
    "dmk_last_error": (ctypes.c_char_p, []),
}


def lib():
    """Loads libdmrgpp_amd.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DmkError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "or `make -C dmrgpp_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(status: int, what: str):
    if status != 0:
        msg = lib().dmk_last_error()
        raise DmkError(f"{what} failed ({status}): {msg.decode() if msg else ''}")


def device_count() -> int:
    c = ctypes.c_int(0)
    _check(lib().dmk_device_count(ctypes.byref(c)), "dmk_device_count")
    return c.value


def pool_counters(host: bool = False) -> dict:
    """Books of the device (or pinned host) allocation cache: bytes in use / cached, allocator
    calls, buffers given back, cache hits, and releases of pointers the cache did not hold
    (0 in a correct run)."""
    out = np.zeros(6, dtype=np.uint64)
    _check(lib().dmk_pool_counters(1 if host else 0, out.ctypes.data), "dmk_pool_counters")
    keys = ("used_bytes", "cached_bytes", "allocs", "frees", "reuses", "bad_releases")
    return {k: int(v) for k, v in zip(keys, out)}


def tridiag_lowest(a, b):
    """Lowest eigenpair of the Lanczos tridiagonal matrix (diagonal a, off-diagonal b[1:])."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m = a.size
    assert b.size >= m
    ev = ctypes.c_double()
    vec = np.zeros(m)
    _check(lib().dmk_tridiag_lowest(m, a.ctypes.data, b.ctypes.data, ctypes.byref(ev),
                                    vec.ctypes.data), "dmk_tridiag_lowest")
    return ev.value, vec


def tridiag_eigen(a, b):
    """All eigenpairs (ascending) of the Lanczos tridiagonal matrix (diagonal a, off-diagonal
    b[1:]); eigenvectors are the columns of the returned matrix."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    m = a.size
    assert b.size >= m
    ev = np.zeros(m)
    vec = np.zeros(m * m)
    _check(lib().dmk_tridiag_eigen(m, a.ctypes.data, b.ctypes.data, ev.ctypes.data,
                                   vec.ctypes.data), "dmk_tridiag_eigen")
    return ev, vec.reshape((m, m), order="F")


def density_diag(blocks):
    """Eigenpairs of symmetric blocks on the GPU (dmk_density_diag_d, the truncation's
    DiagBlockDiagMatrix::diagonalise): returns [(w ascending, v with eigenvectors in columns)]."""
    mats = [np.array(b, dtype=np.float64, order="F", copy=True) for b in blocks]
    for a in mats:
        assert a.ndim == 2 and a.shape[0] == a.shape[1]
    nb = len(mats)
    sizes = (ctypes.c_int64 * max(nb, 1))(*[a.shape[0] for a in mats])
    ptrs = (ctypes.c_void_p * max(nb, 1))(*[a.ctypes.data for a in mats])
    eigs = np.zeros(max(1, sum(a.shape[0] for a in mats)))
    _check(lib().dmk_density_diag_d(nb, ctypes.addressof(sizes), ctypes.addressof(ptrs),
                                    eigs.ctypes.data), "dmk_density_diag_d")
    out, off = [], 0
    for a in mats:
        k = a.shape[0]
        out.append((eigs[off:off + k].copy(), a))
        off += k
    return out


def density_diag_device(blocks):
    """Eigenpairs of symmetric blocks that are already in device memory (dmk_density_diag_dev_d):
    blocks is a list of (device address, n) of compact column-major n x n matrices; returns
    [(w ascending, v with eigenvectors in columns)] on the host."""
    nb = len(blocks)
    vecs = [np.empty((n, n), order="F") for _, n in blocks]
    sizes = (ctypes.c_int64 * max(nb, 1))(*[n for _, n in blocks])
    ptrs = (ctypes.c_void_p * max(nb, 1))(*[p for p, _ in blocks])
    vptr = (ctypes.c_void_p * max(nb, 1))(*[v.ctypes.data for v in vecs])
    eigs = np.zeros(max(1, sum(n for _, n in blocks)))
    _check(lib().dmk_density_diag_dev_d(nb, ctypes.addressof(sizes), ctypes.addressof(ptrs),
                                        ctypes.addressof(vptr), eigs.ctypes.data),
           "dmk_density_diag_dev_d")
    out, off = [], 0
    for v in vecs:
        k = v.shape[0]
        out.append((eigs[off:off + k].copy(), v))
        off += k
    return out


def density_matrices_device(z_ptr: int, offsets, R, L, side: str, scale: float = 1.0):
    """The reduced density-matrix blocks of a device-resident superblock vector, formed on the
    device: patch I's segment X_I (R_I x L_I, column major at z_ptr + 8 offsets[I]) gives
    rho_I = scale X_I^T X_I (L_I x L_I) for the left block and scale X_I X_I^T (R_I x R_I) for
    the right block (DensityMatrixLocal.h:201-204 forms them on the host), all in one batched
    GEMM (dmk_gemm_vbatch_d). Returns (DeviceArray holding the blocks, [(address, n)])."""
    R = [int(r) for r in R]
    L = [int(c) for c in L]
    ns = L if side == "left" else R
    at, offs = 0, []
    for n in ns:
        offs.append(at)
        at += n * n
    dev = DeviceArray(max(at, 1))
    groups = []
    for I, n in enumerate(ns):
        x = z_ptr + 8 * int(offsets[I])
        c = dev.ptr + 8 * offs[I]
        if side == "left":   # C = X^T X: m = n = L_I, k = R_I
            groups.append((n, n, R[I], R[I], R[I], n, [(x, x, c)]))
        else:                # C = X X^T: m = n = R_I, k = L_I
            groups.append((n, n, L[I], R[I], R[I], n, [(x, x, c)]))
    try:
        gemm_vbatch("T" if side == "left" else "N", "N" if side == "left" else "T", groups,
                    alpha=scale)
    except BaseException:
        dev.free()
        raise
    return dev, [(dev.ptr + 8 * o, n) for o, n in zip(offs, ns)]


def density_svd(blocks):
    """Singular values and all left singular vectors of wave-function blocks on the GPU
    (dmk_density_svd_d, the truncation's DensityMatrixSvd::diag): returns [(s descending, u)]
    with u square (rows x rows)."""
    mats = [np.asfortranarray(b, dtype=np.float64) for b in blocks]
    for a in mats:
        assert a.ndim == 2
    nb = len(mats)
    rows = (ctypes.c_int64 * max(nb, 1))(*[a.shape[0] for a in mats])
    cols = (ctypes.c_int64 * max(nb, 1))(*[a.shape[1] for a in mats])
    us = [np.empty((a.shape[0], a.shape[0]), order="F") for a in mats]
    aps = (ctypes.c_void_p * max(nb, 1))(*[a.ctypes.data for a in mats])
    ups = (ctypes.c_void_p * max(nb, 1))(*[u.ctypes.data for u in us])
    s = np.zeros(max(1, sum(min(a.shape) for a in mats)))
    _check(lib().dmk_density_svd_d(nb, ctypes.addressof(rows), ctypes.addressof(cols),
                                   ctypes.addressof(aps), ctypes.addressof(ups), s.ctypes.data),
           "dmk_density_svd_d")
    out, off = [], 0
    for a, u in zip(mats, us):
        k = min(a.shape)
        out.append((s[off:off + k].copy(), u))
        off += k
    return out


def gemm_vbatch(transa: str, transb: str, groups, alpha=1.0, beta=0.0, cplx: bool = False):
    """Variable-size batched GEMM on device matrices (dmrg_Xgemm_vbatch semantics):
    groups is a list of (m, n, k, lda, ldb, ldc, [(a_ptr, b_ptr, c_ptr), ...]); pointers are
    device addresses (ints). Complex matrices are interleaved; alpha/beta may be complex."""
    G = len(groups)
    ia = lambda vals: (ctypes.c_int * max(G, 1))(*vals)
    m = ia([g[0] for g in groups])
    n = ia([g[1] for g in groups])
    k = ia([g[2] for g in groups])
    lda = ia([g[3] for g in groups])
    ldb = ia([g[4] for g in groups])
    ldc = ia([g[5] for g in groups])
    gs = ia([len(g[6]) for g in groups])
    ents = [e for g in groups for e in g[6]]
    nE = max(len(ents), 1)
    A = (ctypes.c_void_p * nE)(*[e[0] for e in ents])
    B = (ctypes.c_void_p * nE)(*[e[1] for e in ents])
    C = (ctypes.c_void_p * nE)(*[e[2] for e in ents])
    args_tail = [ctypes.addressof(ldc), G, ctypes.addressof(gs)]
    if not cplx:
        _check(lib().dmk_gemm_vbatch_d(transa.encode(), transb.encode(), ctypes.addressof(m),
                                       ctypes.addressof(n), ctypes.addressof(k), float(alpha),
                                       ctypes.addressof(A), ctypes.addressof(lda),
                                       ctypes.addressof(B), ctypes.addressof(ldb), float(beta),
                                       ctypes.addressof(C), *args_tail), "dmk_gemm_vbatch_d")
    else:
        al = (ctypes.c_double * 2)(complex(alpha).real, complex(alpha).imag)
        be = (ctypes.c_double * 2)(complex(beta).real, complex(beta).imag)
        _check(lib().dmk_gemm_vbatch_z(transa.encode(), transb.encode(), ctypes.addressof(m),
                                       ctypes.addressof(n), ctypes.addressof(k),
                                       ctypes.addressof(al), ctypes.addressof(A),
                                       ctypes.addressof(lda), ctypes.addressof(B),
                                       ctypes.addressof(ldb), ctypes.addressof(be),
                                       ctypes.addressof(C), *args_tail), "dmk_gemm_vbatch_z")


def krylov_params(**kw) -> KrylovParams:
    p = KrylovParams()
    _check(lib().dmk_krylov_default_params(ctypes.byref(p)), "dmk_krylov_default_params")
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        if k == "stop" and isinstance(v, str):
            v = KRYLOV_STOP[v]
        setattr(p, k, v)
    return p


def davidson_params(**kw) -> DavidsonParams:
    p = DavidsonParams()
    _check(lib().dmk_davidson_default_params(ctypes.byref(p)), "dmk_davidson_default_params")
    for k, v in kw.items():
        if k == "precond" and isinstance(v, str):
            v = PRECONDS[v]
        setattr(p, k, v)
    return p


class Mlp:
    """The Davidson MLP preconditioner (dmk_mlp_*): parameters, training set, training."""

    def __init__(self, seed: int = 7, capacity: int = 1 << 20):
        h = ctypes.c_void_p()
        _check(lib().dmk_mlp_create(seed, capacity, ctypes.byref(h)), "dmk_mlp_create")
        self.handle = h.value

    def params(self) -> np.ndarray:
        p = np.zeros(MLP_PARAMS, dtype=np.float32)
        _check(lib().dmk_mlp_get_params(self.handle, p.ctypes.data), "dmk_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == MLP_PARAMS
        _check(lib().dmk_mlp_set_params(self.handle, p.ctypes.data), "dmk_mlp_set_params")

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float32)
        cnt = x.shape[0]
        dx = DeviceArray.from_host(x.ravel())
        dy = DeviceArray(max(cnt, 1), np.float32)
        _check(lib().dmk_mlp_forward(self.handle, dx.ptr, cnt, dy.ptr), "dmk_mlp_forward")
        return dy.download(cnt)

    def gradient(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        dx, dt = DeviceArray.from_host(x.ravel()), DeviceArray.from_host(t)
        g = np.zeros(MLP_PARAMS, dtype=np.float32)
        loss = ctypes.c_double()
        _check(lib().dmk_mlp_gradient(self.handle, dx.ptr, dt.ptr, t.size, g.ctypes.data,
                                      ctypes.byref(loss)), "dmk_mlp_gradient")
        return g, loss.value

    def add_samples(self, x: np.ndarray, t: np.ndarray):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        dx, dt = DeviceArray.from_host(x.ravel()), DeviceArray.from_host(t)
        _check(lib().dmk_mlp_add_samples(self.handle, dx.ptr, dt.ptr, t.size),
               "dmk_mlp_add_samples")

    def samples(self) -> int:
        c = ctypes.c_int64()
        _check(lib().dmk_mlp_samples(self.handle, ctypes.byref(c)), "dmk_mlp_samples")
        return c.value

    def clear_samples(self):
        _check(lib().dmk_mlp_clear_samples(self.handle), "dmk_mlp_clear_samples")

    def train(self, iters: int, batch: int = 4096, lr: float = 1e-3):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().dmk_mlp_train(self.handle, iters, batch, lr, ctypes.byref(a), ctypes.byref(b)),
               "dmk_mlp_train")
        return a.value, b.value

    def close(self):
        if getattr(self, "handle", None):
            lib().dmk_mlp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lanczos_params(**kw) -> LanczosParams:
    p = LanczosParams()
    _check(lib().dmk_lanczos_default_params(ctypes.byref(p)), "dmk_lanczos_default_params")
    for k, v in kw.items():
        if k == "precision" and isinstance(v, str):
            v = PRECISIONS[v]
        setattr(p, k, v)
    return p


class DeviceArray:
    """A device allocation of ``size`` elements of ``dtype`` owned through the C ABI."""

    def __init__(self, size: int, dtype=np.float64):
        self.size = int(size)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.size * self.dtype.itemsize
        p = ctypes.c_void_p()
        _check(lib().dmk_device_alloc(self.nbytes, ctypes.byref(p)), "dmk_device_alloc")
        self.ptr = p.value

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.size
        _check(lib().dmk_memcpy_h2d(self.ptr, a.ctypes.data, a.nbytes), "dmk_memcpy_h2d")

    def download(self, count: int | None = None) -> np.ndarray:
        count = self.size if count is None else int(count)
        out = np.empty(count, dtype=self.dtype)
        _check(lib().dmk_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes), "dmk_memcpy_d2h")
        return out

    def free(self):
        if self.ptr:
            _check(lib().dmk_device_free(self.ptr), "dmk_device_free")
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _device_plugin_args(prob: KronProblem, args):
    """The plugin arguments of a real problem with every distinct block copied, column major,
    into one device allocation: the same argument arrays with device pointers (and leading
    dimension = rows). Returns (args, the DeviceArray holding the blocks)."""
    if prob.dtype != np.float64:
        raise DmkError("device-resident blocks: real (float64) problems only")
    offs, host = {}, []
    at = 0
    for a, b in prob.blocks.values():
        for m in (a, b):
            if id(m) not in offs:
                offs[id(m)] = at
                host.append(np.asarray(m, dtype=np.float64).ravel(order="F"))
                at += m.size
    dev = DeviceArray(max(at, 1))
    if host:
        dev.upload(np.concatenate(host))
    pa, pb = np.array(args._arrays[2]), np.array(args._arrays[3])
    lda, ldb = np.array(args._arrays[0]), np.array(args._arrays[1])
    for (I, J, k), (a, b) in prob.blocks.items():
        i = prob.index(I, J, k)
        pa[i] = dev.ptr + 8 * offs[id(a)]
        pb[i] = dev.ptr + 8 * offs[id(b)]
        lda[i] = a.shape[0]
        ldb[i] = b.shape[0]

    class _Args:
        pass

    d = _Args()
    d.noperator, d.npatches, d.L, d.R = args.noperator, args.npatches, args.L, args.R
    total = pa.size
    d._arrays = (lda, ldb, pa, pb)
    d.A = (ctypes.c_void_p * total).from_buffer(pa)
    d.B = (ctypes.c_void_p * total).from_buffer(pb)
    d.ldA = (ctypes.c_size_t * total).from_buffer(lda)
    d.ldB = (ctypes.c_size_t * total).from_buffer(ldb)
    return d, dev


class BatchedGemm:
    """Mirror of ``BatchedGemm<T>`` (dmrg/GPUPlugin/BatchedGemm.h:11-56) over the C ABI, for
    T = double (dmk_create_d) and T = std::complex<double> (dmk_create_z, chosen by the
    problem's dtype)."""

    def __init__(self, prob: KronProblem, precision: str | int = "fp64",
                 device_blocks: bool = False):
        """device_blocks: copy the distinct blocks into one device allocation first and hand the
        plugin device pointers (the blocks of a caller that keeps its block operators in HBM:
        the handle scans and packs them on the device, with no staging or upload). Real
        problems only."""
        if prob.dtype == np.float64:
            create, self.is_complex = lib().dmk_create_d, False
        elif prob.dtype == np.complex128:
            create, self.is_complex = lib().dmk_create_z, True
        else:
            raise DmkError(f"unsupported scalar type {prob.dtype}")
        args = prob.plugin_args()
        self._blocks_dev = None
        if device_blocks:
            args, self._blocks_dev = _device_plugin_args(prob, args)
        h = ctypes.c_void_p()
        _check(create(args.noperator, args.npatches, ctypes.addressof(args.L),
                      ctypes.addressof(args.R), ctypes.addressof(args.A),
                      ctypes.addressof(args.ldA), ctypes.addressof(args.B),
                      ctypes.addressof(args.ldB), ctypes.byref(h)),
               "dmk_create_z" if self.is_complex else "dmk_create_d")
        self.handle = h.value
        self.n = prob.n
        self.set_precision(precision)

    def set_precision(self, precision: str | int):
        p = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        _check(lib().dmk_set_precision(self.handle, p), "dmk_set_precision")
        self.precision = p

    def apply_Htarget(self, vin: DeviceArray, vout: DeviceArray):
        """vout = H vin on device arrays (BatchedGemm.h:30-31)."""
        _check(lib().dmk_apply_d(self.handle, vin.ptr, vout.ptr), "dmk_apply_d")

    def set_sector_map(self, perm):
        """perm[ip] = sector index of patch-order element ip (InitKronHamiltonian::copyIn)."""
        perm = np.ascontiguousarray(perm, dtype=np.int64)
        assert perm.size == self.n
        _check(lib().dmk_set_sector_map(self.handle, perm.ctypes.data), "dmk_set_sector_map")

    def matvec_sector(self, vin: DeviceArray, vout: DeviceArray):
        """vout += H vin in sector order (KronMatrix::matrixVectorProductImpl)."""
        _check(lib().dmk_matvec_sector_d(self.handle, vin.ptr, vout.ptr), "dmk_matvec_sector_d")

    def apply_z(self, vin: DeviceArray, vout: DeviceArray):
        """vout = H vin on interleaved complex device vectors (2n doubles each); waits for the
        result (the C call itself is asynchronous on the handle's stream)."""
        _check(lib().dmk_apply_z(self.handle, vin.ptr, vout.ptr), "dmk_apply_z")
        self.sync()

    def krylov_evolve(self, phi: DeviceArray, out: DeviceArray, nvec: int, times,
                      ld: int | None = None, **kw):
        """out[t, s] = exp(-i (H - E) tau_s d) phi[t] for nvec interleaved complex vectors
        (TimeVectorsKrylov::calcTimeVectors); returns per-target results and call stats."""
        times = np.ascontiguousarray(times, dtype=np.float64)
        ld = self.n if ld is None else int(ld)
        p = krylov_params(**kw)
        res = (KrylovResult * nvec)()
        st = KrylovStats()
        _check(lib().dmk_krylov_evolve_z(self.handle, nvec, phi.ptr, ld, times.size,
                                         times.ctypes.data, out.ptr, ld, ctypes.byref(p),
                                         ctypes.addressof(res), ctypes.byref(st)),
               "dmk_krylov_evolve_z")
        return [r.as_dict() for r in res], st.as_dict()

    def apply_multi(self, vin: DeviceArray, vout: DeviceArray, nvec: int, beta: int = 0):
        _check(lib().dmk_apply_multi_d(self.handle, nvec, vin.ptr, self.n, vout.ptr, self.n, beta),
               "dmk_apply_multi_d")

    def apply_host(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.size == self.n
        y = np.empty_like(x)
        _check(lib().dmk_apply_host_d(self.handle, x.ctypes.data, y.ctypes.data),
               "dmk_apply_host_d")
        return y

    def sync(self):
        _check(lib().dmk_sync(self.handle), "dmk_sync")

    def tuning(self) -> dict:
        t = Tuning()
        _check(lib().dmk_get_tuning(self.handle, ctypes.byref(t)), "dmk_get_tuning")
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning()
        _check(lib().dmk_get_tuning(self.handle, ctypes.byref(t)), "dmk_get_tuning")
        for k, v in kw.items():
            setattr(t, k, v)
        _check(lib().dmk_set_tuning(self.handle, ctypes.byref(t)), "dmk_set_tuning")

    def autotune(self, vin: DeviceArray, vout: DeviceArray, iters: int = 3) -> dict:
        t = Tuning()
        _check(lib().dmk_autotune_d(self.handle, vin.ptr, vout.ptr, iters, ctypes.byref(t)),
               "dmk_autotune_d")
        return t.as_dict()

    def info(self) -> dict:
        inf = DmkInfo()
        _check(lib().dmk_get_info(self.handle, ctypes.byref(inf)), "dmk_get_info")
        return inf.as_dict()

    def lanczos_ground(self, init: DeviceArray, z: DeviceArray, **kw) -> dict:
        """Device-resident Lanczos ground state (LanczosSolver::computeAllStatesBelow)."""
        p = lanczos_params(**kw)
        r = LanczosResult()
        _check(lib().dmk_lanczos_ground_d(self.handle, init.ptr, z.ptr, ctypes.byref(p),
                                          ctypes.byref(r)), "dmk_lanczos_ground_d")
        return r.as_dict()

    def diagonal(self, out: DeviceArray):
        """diag(H) into a device array of n elements."""
        _check(lib().dmk_diagonal_d(self.handle, out.ptr), "dmk_diagonal_d")

    def davidson_ground(self, init: DeviceArray, z: DeviceArray, mlp: "Mlp | None" = None,
                        **kw) -> dict:
        """Davidson lowest state with the diagonal or MLP correction (dmk_davidson_ground_d)."""
        p = davidson_params(**kw)
        r = DavidsonResult()
        _check(lib().dmk_davidson_ground_d(self.handle, init.ptr, z.ptr, ctypes.byref(p),
                                           mlp.handle if mlp is not None else None,
                                           ctypes.byref(r)), "dmk_davidson_ground_d")
        return r.as_dict()

    def profile(self, vin: DeviceArray, vout: DeviceArray, iters: int) -> tuple:
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().dmk_profile_apply_d(self.handle, vin.ptr, vout.ptr, iters, ctypes.byref(a),
                                         ctypes.byref(b), ctypes.byref(c)), "dmk_profile_apply_d")
        return a.value, b.value, c.value

    def close(self):
        if getattr(self, "handle", None):
            lib().dmk_destroy(self.handle)
            self.handle = None
        if getattr(self, "_blocks_dev", None) is not None:
            self._blocks_dev.free()
            self._blocks_dev = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FacadeInfo(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nnz", ctypes.c_int64), ("matvecs", ctypes.c_int64),
                ("ms_matvec_total", ctypes.c_double), ("stored", ctypes.c_int32),
                ("ldaggerl", ctypes.c_int32), ("is_complex", ctypes.c_int32),
                ("pad", ctypes.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class Facade:
    """Mirror of MatrixVectorKron (Engine/MatrixVectorKron/MatrixVectorKron.h:113-166) over the
    C ABI: matvec(x, y) returns x + op(y) with op = H, or L^dagger L with ldaggerl; H is stored
    as CSR when the rank is at most max_rank_stored (MaxMatrixRankStored), otherwise the Kron
    plugin computes it. Host vectors in and out (device copies inside)."""

    def __init__(self, prob: KronProblem, max_rank_stored: int = 0, ldaggerl: bool = False):
        if prob.dtype == np.float64:
            create, self.is_complex = lib().dmk_facade_create_d, False
        elif prob.dtype == np.complex128:
            create, self.is_complex = lib().dmk_facade_create_z, True
        else:
            raise DmkError(f"unsupported scalar type {prob.dtype}")
        args = prob.plugin_args()
        h = ctypes.c_void_p()
        _check(create(args.noperator, args.npatches, ctypes.addressof(args.L),
                      ctypes.addressof(args.R), ctypes.addressof(args.A),
                      ctypes.addressof(args.ldA), ctypes.addressof(args.B),
                      ctypes.addressof(args.ldB), ctypes.c_int64(int(max_rank_stored)),
                      1 if ldaggerl else 0, ctypes.byref(h)), "dmk_facade_create")
        self.handle = h.value
        self.n = prob.n

    def matvec(self, x: np.ndarray, y: np.ndarray) -> np.ndarray:
        dt = np.complex128 if self.is_complex else np.float64
        xv = np.ascontiguousarray(x, dtype=dt)
        yv = np.ascontiguousarray(y, dtype=dt)
        dx = DeviceArray.from_host(xv.view(np.float64))
        dy = DeviceArray.from_host(yv.view(np.float64))
        try:
            _check(lib().dmk_facade_matvec(self.handle, dx.ptr, dy.ptr), "dmk_facade_matvec")
            return dx.download().view(dt)
        finally:
            dx.free()
            dy.free()

    def info(self) -> dict:
        i = FacadeInfo()
        _check(lib().dmk_facade_get_info(self.handle, ctypes.byref(i)), "dmk_facade_get_info")
        return i.as_dict()

    def close(self):
        if self.handle:
            lib().dmk_facade_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
