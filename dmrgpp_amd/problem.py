"""The superblock Kron problem as the plugin interface sees it.

A ``KronProblem`` holds what ``BatchedGemmPluginSc::setup_`` hands to ``BatchedGemm<T>``
(dmrg/Engine/MatrixVectorKron/BatchedGemmPluginSc.h:50-125): the patch sizes and, for every
(out-patch I, in-patch J, connection k), the dense blocks A_k(I,J) (L_I x L_J) and B_k(I,J)
(R_I x R_J). Blocks are stored column major (Fortran order), as PsimagLite::Matrix stores them.
The vector segment of patch I is an R_I x L_I column-major matrix and segments are
concatenated in patch order (InitKronHamiltonian.h:211-216).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np


@dataclass
class KronProblem:
    L: np.ndarray                      # left patch sizes (np,)
    R: np.ndarray                      # right patch sizes (np,)
    nC: int                            # number of connections
    blocks: dict = field(default_factory=dict)   # (I, J, k) -> (A, B), Fortran-ordered
    dtype: type = np.float64
    name: str = ""

    def __post_init__(self):
        self.L = np.asarray(self.L, dtype=np.int64)
        self.R = np.asarray(self.R, dtype=np.int64)
        assert self.L.shape == self.R.shape
        assert np.all(self.L >= 1) and np.all(self.R >= 1)
        for key, (a, b) in list(self.blocks.items()):
            I, J, k = key
            if not (a.flags.f_contiguous and a.dtype == self.dtype):
                a = np.asfortranarray(a, dtype=self.dtype)
            if not (b.flags.f_contiguous and b.dtype == self.dtype):
                b = np.asfortranarray(b, dtype=self.dtype)
            assert a.shape == (self.L[I], self.L[J]), (key, a.shape)
            assert b.shape == (self.R[I], self.R[J]), (key, b.shape)
            assert 0 <= k < self.nC
            self.blocks[key] = (a, b)

    @property
    def npatches(self) -> int:
        return int(self.L.shape[0])

    @property
    def offsets(self) -> np.ndarray:
        off = np.zeros(self.npatches + 1, dtype=np.int64)
        np.cumsum(self.L * self.R, out=off[1:])
        return off

    @property
    def n(self) -> int:
        return int(np.sum(self.L * self.R))

    def index(self, I: int, J: int, k: int) -> int:
        """Plugin block index (BatchedGemmPluginSc.h:96-101)."""
        npch = self.npatches
        return I + J * npch + k * npch * npch

    def plugin_args(self) -> "PluginArgs":
        return PluginArgs(self)

    def histogram(self) -> dict:
        """Patch-size and term statistics (the BatchedGemmKokkos.h:745-753 style log)."""
        sizes = (self.L * self.R)
        return {
            "npatches": self.npatches,
            "nC": self.nC,
            "n": self.n,
            "nterms": len(self.blocks),
            "L_min": int(self.L.min()), "L_max": int(self.L.max()),
            "R_min": int(self.R.min()), "R_max": int(self.R.max()),
            "patch_elems_max": int(sizes.max()),
        }

    def apply_numpy(self, x: np.ndarray) -> np.ndarray:
        """y = H x by per-block numpy products (small problems, test use)."""
        off = self.offsets
        y = np.zeros(self.n, dtype=np.result_type(self.dtype, x.dtype))
        for (I, J, k), (a, b) in self.blocks.items():
            XJ = x[off[J]:off[J + 1]].reshape((self.R[J], self.L[J]), order="F")
            YI = b @ XJ @ a.T
            y[off[I]:off[I + 1]] += YI.reshape(-1, order="F")
        return y


def save_npz(prob: KronProblem, path: str):
    """Writes the problem (patch sizes, nC, every block) to an .npz file."""
    keys = sorted(prob.blocks)
    arrays = {"L": prob.L, "R": prob.R, "nC": np.array([prob.nC]),
              "keys": np.array(keys, dtype=np.int64).reshape(-1, 3)}
    for i, k in enumerate(keys):
        a, b = prob.blocks[k]
        arrays[f"a{i}"] = a
        arrays[f"b{i}"] = b
    np.savez(path, **arrays)


def load_npz(path: str, name: str = "") -> KronProblem:
    """Reads a problem written by save_npz (no pickled objects)."""
    d = np.load(path, allow_pickle=False)
    keys = [tuple(int(v) for v in k) for k in d["keys"]]
    blocks = {k: (np.asfortranarray(d[f"a{i}"]), np.asfortranarray(d[f"b{i}"]))
              for i, k in enumerate(keys)}
    dt = np.complex128 if any(np.iscomplexobj(a) for a, _ in blocks.values()) else np.float64
    return KronProblem(L=d["L"], R=d["R"], nC=int(d["nC"][0]), blocks=blocks, dtype=dt,
                       name=name or path)


class PluginArgs:
    """ctypes views of a KronProblem in the BatchedGemm argument layout. Keeps the arrays alive."""

    def __init__(self, prob: KronProblem):
        npch, nC = prob.npatches, prob.nC
        total = npch * npch * nC
        self.noperator = nC
        self.npatches = npch
        self.L = (ctypes.c_size_t * npch)(*[int(v) for v in prob.L])
        self.R = (ctypes.c_size_t * npch)(*[int(v) for v in prob.R])
        # numpy arrays behind the ctypes views (O(blocks) Python work, not O(np^2 nC)).
        lda = np.tile(np.asarray(prob.L, dtype=np.uint64), npch * nC)
        ldb = np.tile(np.asarray(prob.R, dtype=np.uint64), npch * nC)
        pa = np.zeros(total, dtype=np.uint64)
        pb = np.zeros(total, dtype=np.uint64)
        self._keep = []
        for (I, J, k), (a, b) in prob.blocks.items():
            i = prob.index(I, J, k)
            self._keep += [a, b]
            pa[i] = a.ctypes.data
            pb[i] = b.ctypes.data
            lda[i] = a.shape[0]
            ldb[i] = b.shape[0]
        self._arrays = (lda, ldb, pa, pb)
        self.A = (ctypes.c_void_p * total).from_buffer(pa)
        self.B = (ctypes.c_void_p * total).from_buffer(pb)
        self.ldA = (ctypes.c_size_t * total).from_buffer(lda)
        self.ldB = (ctypes.c_size_t * total).from_buffer(ldb)
