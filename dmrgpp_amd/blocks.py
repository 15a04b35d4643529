"""Block-sparse operators on quantum-number-graded bases, the algebra a two-site DMRG step needs
to produce the superblock Kron problem.

A ``Basis`` is a list of groups (quantum number, size) sorted by quantum number; a state is
(group, index). A ``BSOp`` stores its nonzero (out group, in group) dense blocks. Combining two
bases A (x) B (A's modes first in the Jordan-Wigner order) groups the product states by total
quantum number; inside a group the (gA, gB) sub-blocks follow each other and a product state
(a, b) sits at a + b * dim(gA) of its sub-block (left index fastest, as DMRG++ orders
superblock states i + j * nl, InitKronHamiltonian.h:193-199).

Fermionic operators are Jordan-Wigner dressed inside their block: a fermionic operator of the
right factor acts as P_A (x) O_B on A (x) B, with P_A the fermion parity of A's states.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Basis:
    qns: list            # quantum number tuple per group, sorted
    sizes: list          # group sizes

    def __post_init__(self):
        order = sorted(range(len(self.qns)), key=lambda i: self.qns[i])
        self.qns = [tuple(self.qns[i]) for i in order]
        self.sizes = [int(self.sizes[i]) for i in order]
        self.index = {q: i for i, q in enumerate(self.qns)}
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)

    @property
    def dim(self) -> int:
        return int(self.offsets[-1])

    @property
    def ngroups(self) -> int:
        return len(self.qns)


@dataclass
class BSOp:
    """Block-sparse operator: blocks[(go, gi)] is a dense (size_out[go] x size_in[gi]) array,
    kept column major (Fortran order), the layout the Kron plugin takes, so that building a
    superblock problem copies nothing."""
    bout: Basis
    bin: Basis
    blocks: dict = field(default_factory=dict)
    fermionic: bool = False

    def transpose(self) -> "BSOp":
        return BSOp(self.bin, self.bout,
                    {(gi, go): np.asfortranarray(b.T) for (go, gi), b in self.blocks.items()},
                    self.fermionic)

    def scaled(self, c: float) -> "BSOp":
        return BSOp(self.bout, self.bin, {k: c * b for k, b in self.blocks.items()}, self.fermionic)

    def add_(self, other: "BSOp", c: float = 1.0) -> "BSOp":
        for k, b in other.blocks.items():
            if k in self.blocks:
                self.blocks[k] = self.blocks[k] + c * b
            else:
                self.blocks[k] = c * b
        return self

    def times_parity_in(self, parity: np.ndarray) -> "BSOp":
        """O @ P with P = diag(parity per in-group) (+1/-1)."""
        return BSOp(self.bout, self.bin,
                    {(go, gi): (b if parity[gi] > 0 else -b) for (go, gi), b in self.blocks.items()},
                    self.fermionic)

    def dense(self) -> np.ndarray:
        M = np.zeros((self.bout.dim, self.bin.dim))
        for (go, gi), b in self.blocks.items():
            M[self.bout.offsets[go]:self.bout.offsets[go + 1],
              self.bin.offsets[gi]:self.bin.offsets[gi + 1]] = b
        return M

    def transform(self, wout: dict, bout: Basis, win: dict, bin_: Basis) -> "BSOp":
        """W_out^T O W_in with per-group truncation matrices (old group -> (old size x kept))."""
        out = {}
        for (go, gi), b in self.blocks.items():
            if go not in wout or gi not in win:
                continue
            wo, (ngo, _) = wout[go]
            wi, (ngi, _) = win[gi]
            # (wi^T b^T wo)^T: the C-ordered product's transpose is the Fortran-ordered result.
            out[(ngo, ngi)] = ((wi.T @ b.T) @ wo).T
        return BSOp(bout, bin_, out, self.fermionic)


def identity(b: Basis) -> BSOp:
    return BSOp(b, b, {(g, g): np.eye(s, order="F") for g, s in enumerate(b.sizes)})


def add_qn(a: tuple, b: tuple) -> tuple:
    return tuple(x + y for x, y in zip(a, b))


def sub_qn(a: tuple, b: tuple) -> tuple:
    return tuple(x - y for x, y in zip(a, b))


class Product:
    """The graded product basis A (x) B and the embedding of (gA, gB) sub-blocks."""

    def __init__(self, A: Basis, B: Basis):
        self.A, self.B = A, B
        groups = {}
        for ga, qa in enumerate(A.qns):
            for gb, qb in enumerate(B.qns):
                groups.setdefault(add_qn(qa, qb), []).append((ga, gb))
        qns = sorted(groups)
        sizes = []
        self.place = {}  # (ga, gb) -> (G, offset within G)
        for G, q in enumerate(qns):
            off = 0
            for ga, gb in sorted(groups[q], key=lambda t: (t[1], t[0])):
                self.place[(ga, gb)] = (G, off)
                off += A.sizes[ga] * B.sizes[gb]
            sizes.append(off)
        self.basis = Basis(qns, sizes)
        # Basis sorts by qn; qns are already sorted so group indices are unchanged.
        assert self.basis.qns == qns

    def kron(self, opA: BSOp | None, opB: BSOp | None, coef: float = 1.0,
             signA: np.ndarray | None = None, acc: BSOp | None = None) -> BSOp:
        """coef * (opA (x) opB) in the product basis. None stands for the identity; with signA,
        a None opA is diag(signA[g]) on each group g of A (the fermion parity). With acc, the
        product is added into acc's blocks (which acc must own) and acc is returned.

        Element (i na + p, j na' + q) of the (gA, gB) sub-block is b[i, j] a[p, q]. The sub-block
        is addressed as a 4-index strided view (p, i, q, j), so an identity factor touches only
        its diagonal and a dense pair loops over the nonzeros of the smaller factor: no dense
        Kronecker product of an identity is ever formed."""
        from numpy.lib.stride_tricks import as_strided

        A, B = self.A, self.B
        blocksA = opA.blocks if opA is not None else {(g, g): None for g in range(A.ngroups)}
        blocksB = opB.blocks if opB is not None else {(g, g): None for g in range(B.ngroups)}
        out = acc.blocks if acc is not None else {}
        for (gao, gai), a in blocksA.items():
            sa = coef * (float(signA[gao]) if (a is None and signA is not None) else 1.0)
            na_o, na_i = A.sizes[gao], A.sizes[gai]
            for (gbo, gbi), b in blocksB.items():
                Go, oo = self.place[(gao, gbo)]
                Gi, oi = self.place[(gai, gbi)]
                nb_o, nb_i = B.sizes[gbo], B.sizes[gbi]
                blk = out.get((Go, Gi))
                if blk is None:
                    blk = np.zeros((self.basis.sizes[Go], self.basis.sizes[Gi]), order="F")
                    out[(Go, Gi)] = blk
                assert blk.flags.f_contiguous
                sub = blk[oo:oo + na_o * nb_o, oi:oi + na_i * nb_i]
                r, c = sub.strides
                if a is None and b is None:      # (p, i, p, i)
                    d = as_strided(sub, (na_o, nb_o), (r + c, (r + c) * na_o))
                    d += sa
                elif a is None:                  # (p, i, p, j): b on every diagonal p
                    d = as_strided(sub, (na_o, nb_o, nb_i), (r + c, r * na_o, c * na_i))
                    d += sa * b[None, :, :]
                elif b is None:                  # (p, i, q, i): a on every diagonal i
                    d = as_strided(sub, (na_o, na_i, nb_o), (r, c, r * na_o + c * na_i))
                    d += sa * a[:, :, None]
                else:
                    v = as_strided(sub, (na_o, nb_o, na_i, nb_i), (r, r * na_o, c, c * na_i))
                    if b.size <= a.size:
                        for i, j in zip(*np.nonzero(b)):
                            v[:, i, :, j] += (sa * b[i, j]) * a
                    else:
                        for p_, q_ in zip(*np.nonzero(a)):
                            v[p_, :, q_, :] += (sa * a[p_, q_]) * b
        if acc is not None:
            return acc
        ferm = (opA is not None and opA.fermionic) != (opB is not None and opB.fermionic)
        return BSOp(self.basis, self.basis, out, ferm)
