"""Wave-function transformation between consecutive finite-sweep steps (the initial guess of the
next Lanczos run), WaveFunctionTransfFactory / WftAccelBlocks / BlockDiagWf in the reference
(WftAccelBlocks.h:162-322, BlockDiagWf.h:131-420).

Moving the two-site center right, the old superblock S.s_l | s_{l+1}.E becomes
S'.s_{l+1} | s_{l+2}.E' where S' = truncation of S.s_l (transform V_L) and E = truncation of
s_{l+2}.E' (transform V_E, kept from when E was built). In matrix form, for every new patch and
every sub-block of the site that changes sides,

    Y = V_E . X_sub . V_L            (move right)
    Y = V_R^T . X_sub . V_S^T        (move left, the mirror image)

where X_sub is a contiguous slice of an old patch segment (rows when the site leaves the right
block, columns when it leaves the left block) and Y a slice of a new patch segment. Every site
state here carries its own quantum number (site groups of size 1), so each slice is a plain
column-major sub-matrix and the whole transformation is a list of GEMM chains: two
variable-size batched GEMM launches on the GPU (dmk_gemm_vbatch_d), each chain in its
cheaper association order (the reference's min-flop ordering, WftAccelBlocks.h:200-260).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Chain:
    """Y[out] = P . X[src] . Q with optional transposes of P and Q."""
    P: np.ndarray
    P_t: bool
    Q: np.ndarray
    Q_t: bool
    src: tuple   # (old patch, row offset, col offset, rows, cols)
    dst: tuple   # (new patch, row offset, col offset, rows, cols)


def _inverse(W: dict) -> dict:
    """new group -> (old group, V) of a truncation transform {old group: (V, (new group, k))}."""
    return {ng: (g, V) for g, (V, (ng, _)) in W.items()}


def chains_move_right(old, new, W_L: dict, W_E: dict, P_Rold, P_Lnew) -> list:
    """old/new: (patches [(gL, gR)], L sizes, R sizes). W_L truncates the old left block's
    groups into S'; W_E expanded E from the groups of s_{l+2}.E' (= the new right block)."""
    patches_o, _, _ = old
    patches_n, Ln, Rn = new
    pidx = {p: I for I, p in enumerate(patches_o)}
    invL = _inverse(W_L)
    out = []
    for In, (gLn, gRn) in enumerate(patches_n):
        if gRn not in W_E:
            continue
        V_E, (gE, kE) = W_E[gRn]
        for (ng, gs1), (G, offL) in P_Lnew.place.items():
            if G != gLn or ng not in invL:
                continue
            gL, V_L = invL[ng]
            pl = P_Rold.place.get((gs1, gE))
            if pl is None:
                continue
            gR, offR = pl
            Io = pidx.get((gL, gR))
            if Io is None:
                continue
            out.append(Chain(V_E, False, V_L, False,
                             (Io, offR, 0, kE, V_L.shape[0]),
                             (In, 0, offL, Rn[In], V_L.shape[1])))
    return out


def chains_move_left(old, new, W_R: dict, W_S: dict, P_Lold, P_Rnew) -> list:
    """W_R truncates the old right block's groups into E''; W_S expanded S from the groups of
    left[l-1].s_{l-1} (= the new left block)."""
    patches_o, _, _ = old
    patches_n, Ln, Rn = new
    pidx = {p: I for I, p in enumerate(patches_o)}
    invR = _inverse(W_R)
    out = []
    for In, (gLn, gRn) in enumerate(patches_n):
        if gLn not in W_S:
            continue
        V_S, (gS, kS) = W_S[gLn]
        for (gsl, ge), (G, offR) in P_Rnew.place.items():
            if G != gRn or ge not in invR:
                continue
            gR, V_R = invR[ge]
            pl = P_Lold.place.get((gS, gsl))
            if pl is None:
                continue
            gL, offL = pl
            Io = pidx.get((gL, gR))
            if Io is None:
                continue
            out.append(Chain(V_R, True, V_S, True,
                             (Io, 0, offL, V_R.shape[0], kS),
                             (In, offR, 0, V_R.shape[1], Ln[In])))
    return out


def apply_numpy(chains: list, x_old: np.ndarray, old_off, old_R, n_new: int, new_off,
                new_R) -> np.ndarray:
    """Host evaluation of the chains (test reference)."""
    y = np.zeros(n_new)
    for c in chains:
        Io, r0, c0, nr, nc = c.src
        X = x_old[old_off[Io]:old_off[Io + 1]].reshape((old_R[Io], -1), order="F")
        Xs = X[r0:r0 + nr, c0:c0 + nc]
        P = c.P.T if c.P_t else c.P
        Q = c.Q.T if c.Q_t else c.Q
        In, s0, t0, mr, mc = c.dst
        Y = y[new_off[In]:new_off[In + 1]].reshape((new_R[In], -1), order="F")
        Y[s0:s0 + mr, t0:t0 + mc] += P @ Xs @ Q
    return y


def apply_gpu(chains: list, x_old, old_off, old_R, n_new: int, new_off, new_R):
    """Device evaluation: x_old is a DeviceArray of the old superblock vector; returns a new
    DeviceArray. Two batched GEMM launches (first products into a workspace, then the second
    products straight into the new vector)."""
    from . import DeviceArray, gemm_vbatch

    y = DeviceArray.from_host(np.zeros(max(n_new, 1)))
    if not chains:
        return y
    # Upload every P and Q once, column-major, back to back.
    mats, moff = [], []
    pos = 0
    for c in chains:
        for M in (c.P, c.Q):
            moff.append(pos)
            mats.append(np.asfortranarray(M).ravel(order="F"))
            pos += M.size
    dmat = DeviceArray.from_host(np.concatenate(mats))
    # First products, in the cheaper association order per chain.
    g1, g2, wpos = [], [], 0
    plans = []
    for i, c in enumerate(chains):
        nr, nc = c.src[3], c.src[4]
        pm, pk = (c.P.shape[1], c.P.shape[0]) if c.P_t else c.P.shape       # P is pm x pk
        qk, qn = (c.Q.shape[1], c.Q.shape[0]) if c.Q_t else c.Q.shape       # Q is qk x qn
        assert pk == nr and qk == nc, (c.src, c.P.shape, c.Q.shape)
        left_first = (pm * nr * nc + pm * nc * qn) <= (nr * nc * qn + pm * nr * qn)
        plans.append((left_first, pm, pk, qk, qn, wpos))
        wpos += pm * nc if left_first else nr * qn
    dw = DeviceArray(max(wpos, 1))
    f8 = 8
    for i, c in enumerate(chains):
        left_first, pm, pk, qk, qn, w0 = plans[i]
        Io, r0, c0, nr, nc = c.src
        xptr = x_old.ptr + f8 * int(old_off[Io] + r0 + c0 * old_R[Io])
        ldx = int(old_R[Io])
        Pp = dmat.ptr + f8 * moff[2 * i]
        Qp = dmat.ptr + f8 * moff[2 * i + 1]
        ldP, ldQ = c.P.shape[0], c.Q.shape[0]
        In, s0, t0, mr, mc = c.dst
        yptr = y.ptr + f8 * int(new_off[In] + s0 + t0 * new_R[In])
        ldy = int(new_R[In])
        wptr = dw.ptr + f8 * w0
        if left_first:   # W = P X (pm x nc), Y = W Q
            g1.append(("T" if c.P_t else "N", "N", (pm, nc, nr, ldP, ldx, pm, [(Pp, xptr, wptr)])))
            g2.append(("N", "T" if c.Q_t else "N", (pm, qn, nc, pm, ldQ, ldy, [(wptr, Qp, yptr)])))
        else:            # W = X Q (nr x qn), Y = P W
            g1.append(("N", "T" if c.Q_t else "N", (nr, qn, nc, ldx, ldQ, nr, [(xptr, Qp, wptr)])))
            g2.append(("T" if c.P_t else "N", "N", (pm, qn, nr, ldP, nr, ldy, [(Pp, wptr, yptr)])))
    for stage in (g1, g2):
        by_trans = {}
        for ta, tb, grp in stage:
            by_trans.setdefault((ta, tb), []).append(grp)
        for (ta, tb), groups in by_trans.items():
            gemm_vbatch(ta, tb, groups, alpha=1.0, beta=0.0)
    dmat.free()
    dw.free()
    return y
