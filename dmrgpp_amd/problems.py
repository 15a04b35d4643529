"""Synthetic superblock Kron problems with the structure DMRG++ produces.

``qn_problem`` builds the decomposition InitKronHamiltonian produces for a superblock split into
a left block and a right block that each carry an abelian quantum number
(InitKronHamiltonian.h:114-283): patches are the (left group, right group) pairs whose numbers add
to the target (GenIjPatch.h:104-118); connection 0 is H_L (x) I, connection 1 is I (x) H_R
(InitKronHamiltonian.h:242-256); then one connection per link term crossing the cut, A_k on the
left block and B_k on the right block, A_k shifting the left number by +delta and B_k by -delta.
Block values are random (seeded) with the transpose pairing that makes H symmetric
(S+ (x) S- next to S- (x) S+, H_L and H_R symmetric), so Lanczos problems built from them have a
real spectrum. The group sizes are the performance-relevant part; they are given explicitly or
drawn from a Gaussian histogram.

``block_diagonal_problem`` is the driver-free shape of dmrg/Engine/tests/bench_BatchedGemm.cpp:
188-245 (npatches block-diagonal patches, one connection, U(-1,1) entries).
"""
from __future__ import annotations

import numpy as np

from .problem import KronProblem


def gaussian_histogram(m: int, sigma: float, qvals, min_count: int = 1) -> dict:
    """Distributes m states over quantum numbers qvals with weight exp(-q^2 / 2 sigma^2)."""
    qvals = list(qvals)
    w = np.exp(-np.asarray(qvals, dtype=float) ** 2 / (2.0 * sigma * sigma))
    w /= w.sum()
    counts = np.maximum(min_count, np.floor(w * m).astype(int))
    # Fix the total to exactly m by adjusting the largest bins.
    diff = m - int(counts.sum())
    order = np.argsort(-w)
    i = 0
    while diff != 0:
        j = order[i % len(order)]
        if diff > 0:
            counts[j] += 1
            diff -= 1
        elif counts[j] > min_count:
            counts[j] -= 1
            diff += 1
        i += 1
    return {int(q): int(c) for q, c in zip(qvals, counts) if c > 0}


def qn_problem(left: dict, right: dict, target: int, links: list, seed: int = 1234,
               scale_h: float = 1.0, name: str = "") -> KronProblem:
    """left/right: {qn: group size}; links: list of (delta, value) pairs, one per link term.

    Each link term contributes A_k (left, shifts qn by +delta) (x) B_k (right, shifts by -delta)
    with coefficient ``value`` folded into A as InitKronBase::calculateAhat does. For every
    term with delta != 0 the caller should also list its (-delta) partner; the partner's blocks
    are generated as transposes so that H is symmetric.
    """
    rng = np.random.default_rng(seed)
    lq = sorted(left)
    rq = sorted(right)
    patches = [(ql, target - ql) for ql in lq if (target - ql) in right]
    if not patches:
        raise ValueError("no patch matches the target quantum number")
    L = np.array([left[ql] for ql, _ in patches], dtype=np.int64)
    R = np.array([right[qr] for _, qr in patches], dtype=np.int64)
    nC = 2 + len(links)
    blocks = {}

    def sym(nn):
        a = rng.uniform(-1.0, 1.0, size=(nn, nn))
        return scale_h * 0.5 * (a + a.T)

    HL = {ql: sym(left[ql]) for ql in lq}
    HR = {qr: sym(right[qr]) for qr in rq}
    for i, (ql, qr) in enumerate(patches):
        blocks[(i, i, 0)] = (HL[ql], np.eye(R[i]))
        blocks[(i, i, 1)] = (np.eye(L[i]), HR[qr])

    # Operators per link term: left op O_L[q_out <- q_in] nonzero iff q_out = q_in + delta,
    # right op O_R[q_out <- q_in] nonzero iff q_out = q_in - delta. A term with delta < 0 takes
    # the transposes of the oldest unpaired term with +delta (S- = (S+)^T).
    pending = {}
    for t, (delta, value) in enumerate(links):
        k = 2 + t
        opsL, opsR = {}, {}
        if delta < 0:
            if not pending.get(-delta):
                raise ValueError("a link with delta < 0 needs an earlier partner with -delta")
            srcL, srcR = pending[-delta].pop(0)
            opsL = {(qi, qo): a.T for (qo, qi), a in srcL.items()}
            opsR = {(qi, qo): b.T for (qo, qi), b in srcR.items()}
        else:
            for ql in lq:
                if ql + delta in left:
                    a = rng.uniform(-1.0, 1.0, size=(left[ql + delta], left[ql]))
                    opsL[(ql + delta, ql)] = 0.5 * (a + a.T) if delta == 0 else a
            for qr in rq:
                if qr - delta in right:
                    b = rng.uniform(-1.0, 1.0, size=(right[qr - delta], right[qr]))
                    opsR[(qr - delta, qr)] = 0.5 * (b + b.T) if delta == 0 else b
            if delta > 0:
                pending.setdefault(delta, []).append((opsL, opsR))
        for I, (qlI, qrI) in enumerate(patches):
            for J, (qlJ, qrJ) in enumerate(patches):
                a = opsL.get((qlI, qlJ))
                b = opsR.get((qrI, qrJ))
                if a is None or b is None:
                    continue
                blocks[(I, J, k)] = (value * a, b)
    return KronProblem(L=L, R=R, nC=nC, blocks=blocks, name=name)


def heisenberg_links(nbonds: int, J: float = 1.0) -> list:
    """Link terms of nbonds Heisenberg bonds across the cut: J/2 (S+S- + S-S+) + J SzSz,
    with 2Sz as the quantum number (S+ shifts it by +2)."""
    out = []
    for _ in range(nbonds):
        out += [(2, 0.5 * J), (-2, 0.5 * J), (0, J)]
    return out


def ladder_m2048_problem(seed: int = 1234, m: int = 2048, sigma: float = 4.0) -> KronProblem:
    """Superblock of the Heisenberg 32x2 ladder at m = 2048 kept states, Sz_total = 0.

    Each of the left block S+s and the right block s+E holds 2m states (m kept block states
    times a spin-1/2 site) spread over 2Sz sectors; the cut of a two-leg ladder crosses two leg
    bonds or two legs and one rung, here three Heisenberg bonds (11 connections).
    """
    qS = list(range(-15, 16, 2))
    hist = gaussian_histogram(m, sigma, qS)
    block = {}
    for q, c in hist.items():
        for s in (-1, 1):
            block[q + s] = block.get(q + s, 0) + c
    return qn_problem(block, block, 0, heisenberg_links(3), seed=seed,
                      name=f"heisenberg-ladder-32x2-m{m}")


def block_diagonal_problem(npatches: int = 40, nsmall: int = 34, small: int = 60,
                           large: int = 900, seed: int = 42) -> KronProblem:
    """bench_BatchedGemm.cpp:188-245 shape: block-diagonal, one connection, U(-1,1) entries,
    small and large patches interleaved (shuffled)."""
    rng = np.random.default_rng(seed)
    sizes = np.array([small] * nsmall + [large] * (npatches - nsmall))
    rng.shuffle(sizes)
    blocks = {}
    for i, s in enumerate(sizes):
        blocks[(i, i, 0)] = (rng.uniform(-1, 1, size=(s, s)), rng.uniform(-1, 1, size=(s, s)))
    return KronProblem(L=sizes, R=sizes, nC=1, blocks=blocks, name="bench_BatchedGemm")


def random_problem(L, R, nC: int, density: float = 0.5, seed: int = 0,
                   dtype=np.float64) -> KronProblem:
    """Unstructured random problem: every (I, J, k) present with probability density."""
    rng = np.random.default_rng(seed)
    L = np.asarray(L)
    R = np.asarray(R)
    blocks = {}
    npch = len(L)
    for k in range(nC):
        for J in range(npch):
            for I in range(npch):
                if rng.random() < density:
                    a = rng.uniform(-1, 1, size=(L[I], L[J]))
                    b = rng.uniform(-1, 1, size=(R[I], R[J]))
                    if np.issubdtype(dtype, np.complexfloating):
                        a = a + 1j * rng.uniform(-1, 1, size=a.shape)
                        b = b + 1j * rng.uniform(-1, 1, size=b.shape)
                    blocks[(I, J, k)] = (a, b)
    return KronProblem(L=L, R=R, nC=nC, blocks=blocks, dtype=dtype, name="random")


def hermitian_problem(L, R, nC: int, density: float = 0.5, seed: int = 0,
                      dtype=np.complex128) -> KronProblem:
    """Random Hermitian (real symmetric for a real dtype) problem: every random term
    A_k(I,J) (x) B_k(I,J) comes with its adjoint A^H (x) B^H at (J, I, k + nC), since the (J, I)
    block of H^H is the adjoint of the (I, J) block."""
    base = random_problem(L, R, nC, density=density, seed=seed, dtype=dtype)
    blocks = dict(base.blocks)
    for (I, J, k), (a, b) in base.blocks.items():
        blocks[(J, I, k + nC)] = (np.conj(a.T), np.conj(b.T))
    return KronProblem(L=base.L, R=base.R, nC=2 * nC, blocks=blocks, dtype=dtype,
                       name="hermitian")


def comb_patch_sizes(left_size: int = 8, right_size: int = 8, target_up: int = 8,
                     target_down: int = 8, keep_left: int = 8000, keep_right: int = 2000):
    """Patch sizes estimated from state counting, as the plugin's own test driver builds them
    (GPUPlugin/gen_patches_comb.cpp): a patch per (n_up, n_down) split of the target between
    left and right; each side's share of kept states is proportional to
    C(size, n_up) C(size, n_down) among the valid splits, rounded; patches with an empty side
    are dropped."""
    from math import comb

    max_lu, max_ld = min(left_size, target_up), min(left_size, target_down)
    max_ru, max_rd = min(right_size, target_up), min(right_size, target_down)
    valid = [(nu, nd) for nd in range(max_ld + 1) for nu in range(max_lu + 1)
             if target_up - nu <= max_ru and target_down - nd <= max_rd
             and target_up - nu >= 0 and target_down - nd >= 0]
    lw = {(nu, nd): comb(left_size, nu) * comb(left_size, nd) for nu, nd in valid}
    rw = {(target_up - nu, target_down - nd): comb(right_size, target_up - nu)
          * comb(right_size, target_down - nd) for nu, nd in valid}
    tl, tr = float(sum(lw.values())), float(sum(rw.values()))
    L, R = [], []
    for nu, nd in valid:
        ls = int(np.floor(lw[(nu, nd)] / tl * keep_left + 0.5))
        rs = int(np.floor(rw[(target_up - nu, target_down - nd)] / tr * keep_right + 0.5))
        if ls >= 1 and rs >= 1:
            L.append(ls)
            R.append(rs)
    return L, R


def comb_problem(noperator: int = 2, left_size: int = 8, keep: int = 2000) -> KronProblem:
    """The plugin test driver's problem shape (GPUPlugin/test_vbatch.cpp:35-240): comb patch
    sizes with keep_left = 4 keep, keep_right = keep, every (I, J, k) block dense, with the
    values of GPUPlugin/setup_matrix.cpp: entry (ia, ja) of operator k of the assembled
    left (right) matrix is ia + (ja - 1) * total + (k - 1) * total^2 (1-based, total = the
    summed patch sizes of that side)."""
    L, R = comb_patch_sizes(left_size, left_size, left_size, left_size, 4 * keep, keep)
    npch = len(L)

    def side(sizes):
        start = np.concatenate([[0], np.cumsum(sizes)])
        total = int(start[-1])
        return start, total

    ls, lt = side(L)
    rs, rt = side(R)
    blocks = {}
    for k in range(noperator):
        for J in range(npch):
            for I in range(npch):
                ia = np.arange(ls[I], ls[I + 1])[:, None] + 1.0
                ja = np.arange(ls[J], ls[J + 1])[None, :] + 1.0
                a = ia + (ja - 1.0) * lt + float(k) * lt * lt
                ib = np.arange(rs[I], rs[I + 1])[:, None] + 1.0
                jb = np.arange(rs[J], rs[J + 1])[None, :] + 1.0
                b = ib + (jb - 1.0) * rt + float(k) * rt * rt
                blocks[(I, J, k)] = (np.asfortranarray(a), np.asfortranarray(b))
    return KronProblem(L=L, R=R, nC=noperator, blocks=blocks, name="test_vbatch-comb")
