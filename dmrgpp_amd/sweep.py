"""A minimal two-site DMRG sweep driver that produces the superblock Kron problems and the sweep
schedules of the benchmark configurations, with the eigensolve on the GPU through the C ABI.

This is the caller of the hot path, not part of it: DMRG++'s own driver (DmrgSolver.h:219-603,
Diagonalization.h, Truncation.h) is out of scope. The step structure follows it: the left block
S+s and the right block e+E (InitKronHamiltonian.h:114-145), connections H_L (x) I, I (x) H_R,
then one term per link crossing the cut with the fermion sign folded into A
(InitKronBase.h:392-413), the lowest state of the target sector by Lanczos, then the reduced
density matrix of the grown block diagonalized per quantum-number group in FP64 and truncated to
the m largest eigenvalues.

Infinite-size growth builds left blocks from site 0 and right blocks from site L-1
simultaneously; finite sweeps then move the two-site center right to the end, left to the
beginning and right back to the middle.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .blocks import Basis, BSOp, Product, sub_qn
from .models import Lattice, SiteModel
from .problem import KronProblem


@dataclass
class Block:
    sites: list
    basis: Basis
    H: BSOp
    ops: dict       # (site, name) -> BSOp
    parity: np.ndarray
    product: object = None  # grown blocks: the Product of the two blocks they were made of
    wft: dict = None        # truncated blocks: {grown group: (V, (new group, kept))}

    @property
    def dim(self):
        return self.basis.dim


def _parity_of(model: SiteModel, qn: tuple) -> float:
    if model.name == "hubbard":
        return -1.0 if sum(qn) % 2 else 1.0
    return 1.0


def site_block(model: SiteModel, site: int) -> Block:
    ops = {(site, name): op for name, op in model.ops.items()}
    return Block([site], model.basis, model.h_site, ops,
                 np.array([_parity_of(model, q) for q in model.basis.qns]))


def cross_bonds(lattice: Lattice, A: Block, B: Block, offset: int = 0) -> list:
    """Bonds (i, j, value) from a site of A to a site of B that both carry operators. B's sites
    sit at j - offset in the (possibly shorter) system the superblock represents."""
    sa = sorted({s for s, _ in A.ops})
    sb = sorted({s for s, _ in B.ops})
    out = []
    for i in sa:
        for j in sb:
            v = lattice.bond_value(i, j - offset)
            if v is not None and j - offset > i:
                out.append((i, j, v))
    return out


def grow(A: Block, B: Block, model: SiteModel, lattice: Lattice, keep) -> Block:
    """A (x) B with A's modes first: H, and the operators of the sites for which keep(s)."""
    P = Product(A.basis, B.basis)
    H = P.kron(A.H, None)
    P.kron(None, B.H, acc=H)
    for i, j, v in cross_bonds(lattice, A, B):
        for oi, oj, c, ferm in model.bond_terms(v):
            a = A.ops[(i, oi)]
            if ferm:
                a = a.times_parity_in(A.parity)
            P.kron(a, B.ops[(j, oj)], c, acc=H)
    ops = {}
    for (s, name), op in A.ops.items():
        if keep(s):
            ops[(s, name)] = P.kron(op, None)
            ops[(s, name)].fermionic = op.fermionic
    for (s, name), op in B.ops.items():
        if keep(s):
            # A fermion operator of B carries the parity of A's modes (Jordan-Wigner string).
            ops[(s, name)] = P.kron(None, op, signA=A.parity if op.fermionic else None)
            ops[(s, name)].fermionic = op.fermionic
    parity = np.array([_parity_of(model, q) for q in P.basis.qns])
    return Block(sorted(A.sites + B.sites), P.basis, H, ops, parity, product=P)


def kron_problem(Lb: Block, Rb: Block, model: SiteModel, lattice: Lattice, target: tuple,
                 name: str = "", offset: int = 0):
    """The superblock decomposition for target sector `target` (patches, A_k, B_k). The right
    block's sites sit at site - offset in the system the superblock represents."""
    patches = []
    for gL, qL in enumerate(Lb.basis.qns):
        gR = Rb.basis.index.get(sub_qn(target, qL))
        if gR is not None:
            patches.append((gL, gR))
    if not patches:
        raise ValueError(f"no patch for target {target}")
    pl = {gL: I for I, (gL, _) in enumerate(patches)}
    pr = {gR: I for I, (_, gR) in enumerate(patches)}
    L = [Lb.basis.sizes[gL] for gL, _ in patches]
    R = [Rb.basis.sizes[gR] for _, gR in patches]
    conns = [(Lb.H, None, 1.0), (None, Rb.H, 1.0)]
    for i, j, v in cross_bonds(lattice, Lb, Rb, offset):
        for oi, oj, c, ferm in model.bond_terms(v):
            a = Lb.ops[(i, oi)]
            if ferm:
                a = a.times_parity_in(Lb.parity)
            conns.append((a, Rb.ops[(j, oj)], c))
    eyes = {}

    def eye(nn):
        # Identity factors are shared between blocks (read-only, column-major).
        if nn not in eyes:
            eyes[nn] = np.asfortranarray(np.eye(nn))
        return eyes[nn]

    def fortran(a, c):
        return np.multiply(a, c, order="F") if c != 1.0 else np.asfortranarray(a)

    blocks = {}
    for k, (A, B, c) in enumerate(conns):
        if A is None:
            for (go, gi), b in B.blocks.items():
                I, J = pr.get(go), pr.get(gi)
                if I is None or J is None or I != J:
                    continue
                blocks[(I, J, k)] = (c * eye(L[I]) if c != 1.0 else eye(L[I]), fortran(b, 1.0))
            continue
        for (go, gi), a in A.blocks.items():
            I, J = pl.get(go), pl.get(gi)
            if I is None or J is None:
                continue
            if B is None:
                if I == J:
                    blocks[(I, J, k)] = (fortran(a, c), eye(R[I]))
                continue
            b = B.blocks.get((patches[I][1], patches[J][1]))
            if b is not None:
                blocks[(I, J, k)] = (fortran(a, c), fortran(b, 1.0))
    return KronProblem(L=L, R=R, nC=len(conns), blocks=blocks, name=name), patches


def host_eig(rhos: list) -> list:
    """Eigenpairs (ascending) of symmetric blocks with the host LAPACK."""
    return [np.linalg.eigh(r) for r in rhos]


class GpuDensityEig:
    """Eigenpairs of the reduced density-matrix blocks: blocks of at least min_size rows on the
    GPU (dmk_density_diag_d, rocSOLVER dsyevd, several blocks at once), the smaller ones with
    the host LAPACK, which is faster for them (n = 136: 0.9 ms on the host against 2.1 ms on the
    device; n = 1093: 35.6 against 23.1 ms). The host blocks are solved while the device call
    runs on a second thread (the C call releases the interpreter lock)."""

    def __init__(self, min_size: int = 512):
        self.min_size = int(min_size)

    def __call__(self, rhos: list) -> list:
        import threading

        from . import density_diag

        big = [i for i, r in enumerate(rhos) if r.shape[0] >= self.min_size]
        out = [None] * len(rhos)
        res, err = [], []

        def device():
            try:
                res.extend(density_diag([rhos[i] for i in big]))
            except BaseException as e:  # re-raised in the caller's thread
                err.append(e)

        th = threading.Thread(target=device) if big else None
        if th is not None:
            th.start()
        try:
            for i, r in enumerate(rhos):
                if r.shape[0] < self.min_size:
                    out[i] = np.linalg.eigh(r)
        finally:
            if th is not None:
                th.join()
        if err:
            raise err[0]
        for i, wv in zip(big, res):
            out[i] = wv
        return out

    # The same from a device-resident superblock vector: the density-matrix blocks are formed
    # on the device (rho_device, one batched GEMM), the large ones stay there for the eigensolver
    # and only the small ones are downloaded for the host LAPACK (solve_device).
    def rho_device(self, dev_psi, side: str):
        from . import density_matrices_device

        z_ptr, offsets, R, L, scale = dev_psi
        return density_matrices_device(z_ptr, offsets, R, L, side, scale)

    def solve_device(self, rho) -> list:
        import threading

        from . import _check, density_diag_device, lib

        dev, blocks = rho
        try:
            big = [i for i, (_, n) in enumerate(blocks) if n >= self.min_size]
            out = [None] * len(blocks)
            res, err = [], []

            def device():
                try:
                    res.extend(density_diag_device([blocks[i] for i in big]))
                except BaseException as e:  # re-raised in the caller's thread
                    err.append(e)

            th = threading.Thread(target=device) if big else None
            if th is not None:
                th.start()
            try:
                for i, (p, n) in enumerate(blocks):
                    if n < self.min_size:
                        r = np.empty((n, n), order="F")
                        _check(lib().dmk_memcpy_d2h(r.ctypes.data, p, r.nbytes),
                               "dmk_memcpy_d2h")
                        out[i] = np.linalg.eigh(0.5 * (r + r.T))
            finally:
                if th is not None:
                    th.join()
            if err:
                raise err[0]
            for i, wv in zip(big, res):
                out[i] = wv
            return out
        finally:
            dev.free()


class SvdDensityEig:
    """The SVD form of the truncation's eigenpairs (DensityMatrixSvd::diag,
    DensityMatrixSvd.h:517-554; ParallelSvd::doTask, :382-410). It takes each group's
    wave-function matrix M (the group's states x its partner states) instead of the density
    matrix M M^T: the left singular vectors of M are the eigenvectors and the squared singular
    values the eigenvalues (zero past min(rows, cols)), so the small eigenvalues keep their
    relative accuracy instead of sinking into the rounding of M M^T. With device set, blocks of
    at least min_size rows go to the GPU (dmk_density_svd_d, rocSOLVER dgesvd) and the rest to
    the host LAPACK meanwhile; by default all stay on the host, which is faster here (on the
    m = 2048 ladder the SVDs of a step take 0.28 s on the host and 1.37 s with the large blocks
    on the device; the eigenpairs of the density matrices take 0.05 s with GpuDensityEig).
    Returns the pairs in the ascending order of host_eig."""

    from_wavefunction = True

    def __init__(self, min_size: int = 512, device: bool = False):
        self.min_size, self.device = int(min_size), bool(device)

    @staticmethod
    def _pairs(s, u):
        w = np.zeros(u.shape[0])
        w[:s.size] = s * s
        return w[::-1].copy(), np.asfortranarray(u[:, ::-1])

    def __call__(self, mats: list) -> list:
        import threading

        big = [i for i, a in enumerate(mats) if self.device and a.shape[0] >= self.min_size]
        out = [None] * len(mats)
        res, err = [], []

        def device():
            from . import density_svd

            try:
                res.extend(density_svd([mats[i] for i in big]))
            except BaseException as e:  # re-raised in the caller's thread
                err.append(e)

        th = threading.Thread(target=device) if big else None
        if th is not None:
            th.start()
        try:
            bigset = set(big)
            for i, a in enumerate(mats):
                if i not in bigset:
                    u, sv, _ = np.linalg.svd(a, full_matrices=True)
                    out[i] = self._pairs(sv, u)
        finally:
            if th is not None:
                th.join()
        if err:
            raise err[0]
        for i, (sv, u) in zip(big, res):
            out[i] = self._pairs(sv, u)
        return out


class GpuBasisChange:
    """The truncation's change of basis O' = W_out^T O W_in of H and of every kept operator on
    the GPU: all blocks in two variable-size batched GEMM calls (dmk_gemm_vbatch_d: T = O W_in,
    then O' = W_out^T T), with one upload of the blocks and the W matrices and one download of
    the results."""

    def __call__(self, ops: list, W: dict, nb) -> list:
        from . import DeviceArray, gemm_vbatch, lib, _check

        outs = [{} for _ in ops]
        jobs = []  # (op index, new group pair, block, W_out, W_in)
        for t, op in enumerate(ops):
            for (go, gi), b in op.blocks.items():
                if go not in W or gi not in W:
                    continue
                wo, (ngo, ko) = W[go]
                wi, (ngi, ki) = W[gi]
                if b.size == 0 or ko == 0 or ki == 0:
                    outs[t][(ngo, ngi)] = np.zeros((ko, ki), order="F")
                else:
                    jobs.append((t, (ngo, ngi), b, wo, wi))
        if jobs:
            # Device layout, column major throughout: the distinct W matrices, the blocks, the
            # products T (no x ki) and the results (ko x ki), one after another.
            host, at = [], 0

            def put(a):
                nonlocal at
                host.append(np.asarray(a, dtype=np.float64).ravel(order="F"))
                at += a.size
                return at - a.size

            woff = {}
            for *_, wo, wi in jobs:
                for w in (wo, wi):
                    if id(w) not in woff:
                        woff[id(w)] = put(w)
            boff = [put(b) for _, _, b, _, _ in jobs]
            toff, ooff = [], []
            for *_, b, wo, wi in jobs:
                toff.append(at)
                at += b.shape[0] * wi.shape[1]
            r0 = at
            for *_, b, wo, wi in jobs:
                ooff.append(at)
                at += wo.shape[1] * wi.shape[1]
            dev = DeviceArray(at)
            try:
                dev.upload(np.concatenate(host))
                adr = lambda o: dev.ptr + 8 * o  # noqa: E731
                g1, g2 = [], []
                for j, (_, _, b, wo, wi) in enumerate(jobs):
                    no, ni = b.shape
                    ko, ki = wo.shape[1], wi.shape[1]
                    g1.append((no, ki, ni, no, ni, no,
                               [(adr(boff[j]), adr(woff[id(wi)]), adr(toff[j]))]))
                    g2.append((ko, ki, no, no, no, ko,
                               [(adr(woff[id(wo)]), adr(toff[j]), adr(ooff[j]))]))
                gemm_vbatch("N", "N", g1)
                gemm_vbatch("T", "N", g2)
                res = np.empty(at - r0)
                _check(lib().dmk_memcpy_d2h(res.ctypes.data, adr(r0), res.nbytes),
                       "dmk_memcpy_d2h")
            finally:
                dev.free()
            for j, (t, key, b, wo, wi) in enumerate(jobs):
                o, ko, ki = ooff[j] - r0, wo.shape[1], wi.shape[1]
                outs[t][key] = res[o:o + ko * ki].reshape((ko, ki), order="F")
        return [BSOp(nb, nb, outs[t], op.fermionic) for t, op in enumerate(ops)]


def truncate(block: Block, psi_blocks: list, patches: list, side: str, m: int,
             model: SiteModel, eig=None, times: dict | None = None, rotate=None,
             dev_psi=None):
    """Keeps the m largest eigenvalues of the block's reduced density matrix.

    psi_blocks[I] is the R_I x L_I ground-state segment of patch I. For the left block
    rho_g = X^T X, for the right block rho_g = X X^T (each group meets one patch). eig maps the
    list of symmetric blocks to their eigenpairs (host_eig by default; GpuDensityEig runs the
    large ones on the GPU; an eig with from_wavefunction set, SvdDensityEig, gets the
    wave-function blocks instead); rotate, when given, does the change of basis of H and the kept operators
    (GpuBasisChange runs it on the GPU; by default the host does it). times, when given, accumulates the milliseconds of the density
    matrices (rho_ms), their eigenpairs (eig_ms) and the change of basis (basis_ms). dev_psi,
    when given with an eig that has rho_device / solve_device (GpuDensityEig), is the device
    copy of the ground state, (address, patch offsets, R, L, 1 / |z|^2): the density matrices
    are then formed on the device and the large ones never leave it. Returns (new block,
    discarded weight)."""
    t0 = time.perf_counter()
    cand = []
    vecs = {}
    groups = [gL if side == "left" else gR for gL, gR in patches]
    if dev_psi is not None and hasattr(eig, "rho_device"):
        rho = eig.rho_device(dev_psi, side)
        t1 = time.perf_counter()
        pairs = eig.solve_device(rho)
    else:
        rhos = []
        for I in range(len(patches)):
            X = psi_blocks[I]
            if getattr(eig, "from_wavefunction", False):
                rho = X.T if side == "left" else X
            else:
                rho = X.T @ X if side == "left" else X @ X.T
                rho = 0.5 * (rho + rho.T)
            rhos.append(rho)
        t1 = time.perf_counter()
        pairs = (eig or host_eig)(rhos)
    t2 = time.perf_counter()
    for g, (w, v) in zip(groups, pairs):
        vecs[g] = v
        for t, val in enumerate(w):
            cand.append((val, g, t))
    cand.sort(key=lambda t: -t[0])
    kept = cand[:m]
    # Weight of the dropped states, summed directly (1 - kept cancels at m = 2048).
    discarded = float(sum(max(c[0], 0.0) for c in cand[m:]))
    per = {}
    for val, g, t in kept:
        per.setdefault(g, []).append(t)
    groups = sorted(per, key=lambda g: block.basis.qns[g])
    nb = Basis([block.basis.qns[g] for g in groups], [len(per[g]) for g in groups])
    W = {}
    for ng, g in enumerate(groups):
        cols = sorted(per[g])
        W[g] = (vecs[g][:, cols], (ng, len(cols)))
    if rotate is None:
        H = block.H.transform(W, nb, W, nb)
        ops = {}
        for key, op in block.ops.items():
            ops[key] = op.transform(W, nb, W, nb)
    else:
        keys = list(block.ops)
        done = rotate([block.H] + [block.ops[k] for k in keys], W, nb)
        H, ops = done[0], dict(zip(keys, done[1:]))
    parity = np.array([_parity_of(model, q) for q in nb.qns])
    if times is not None:
        t3 = time.perf_counter()
        for k, dt in (("rho_ms", t1 - t0), ("eig_ms", t2 - t1), ("basis_ms", t3 - t2)):
            times[k] = times.get(k, 0.0) + 1e3 * dt
    return Block(block.sites, nb, H, ops, parity, wft=W), discarded


@dataclass
class StepRecord:
    phase: str
    center: int
    m: int
    n: int
    npatches: int
    nterms: int
    energy: float
    lanczos_steps: int
    matvecs: int
    ms_lanczos: float
    ms_matvec: float
    discarded: float
    info: dict = field(default_factory=dict)


class GpuSolver:
    """Ground state of a Kron problem on the GPU (C-ABI plugin + device Lanczos).

    A mixed precision (fp32, bf16x3) runs the cold-start solves (no transformed wave function
    to start from) in that precision with an FP64 refinement; with mixed_warm=False (default)
    the solves started from the previous step's transformed wave function run in FP64: they
    take 5-16 matvecs, fewer than a mixed phase plus an FP64 refinement."""

    def __init__(self, precision="fp64", refine=2, max_steps=200, min_steps=4, tol=1e-12,
                 seed=1234, graph=1, mixed_warm=False):
        self.precision, self.refine, self.graph = precision, refine, graph
        self.max_steps, self.min_steps, self.tol, self.seed = max_steps, min_steps, tol, seed
        self.mixed_warm = mixed_warm

    def __call__(self, prob: KronProblem, init=None):
        """init: None (uniform random from the seed), a host vector, or a DeviceArray (the
        transformed wave function of the previous step). The returned report carries the
        device ground vector as r["z_dev"] (the caller frees it)."""
        from . import BatchedGemm, DeviceArray

        own_init = not isinstance(init, DeviceArray)
        warm = init is not None
        if init is None:
            init = np.random.default_rng(self.seed).uniform(-1.0, 1.0, prob.n)
        precision = self.precision if (not warm or self.mixed_warm) else "fp64"
        t0 = time.perf_counter()
        op = BatchedGemm(prob, "fp64")
        info = op.info()
        t1 = time.perf_counter()
        dinit = DeviceArray.from_host(init) if own_init else init
        dz = DeviceArray(prob.n)
        r = op.lanczos_ground(dinit, dz, max_steps=self.max_steps, min_steps=self.min_steps,
                              tolerance=self.tol, precision=precision, refine=self.refine,
                              refine_tolerance=self.tol, graph=self.graph)
        z = dz.download()
        t2 = time.perf_counter()
        op.close()
        if own_init:
            dinit.free()
        r["z_dev"] = dz
        r["plan"] = {k: info[k] for k in ("nterms", "flops_algorithmic", "stage1_tiles",
                                          "stage2_tiles")}
        r["timing"] = {"plan_ms": 1e3 * (t1 - t0), "lanczos_ms": 1e3 * (t2 - t1),
                       "plan_scan_ms": info["ms_scan"], "plan_stage_ms": info["ms_stage"],
                       "plan_upload_ms": info["ms_upload"],
                       "plan_raw_mb": info["raw_bytes"] / 2**20}
        return r["energy"], z, r


def clone_state(state: dict) -> dict:
    """A copy of the block stores that a later run can advance without touching the original
    (blocks themselves are never modified in place)."""
    return {"left": dict(state["left"]), "right": dict(state["right"])}


def run(model: SiteModel, lattice: Lattice, target_of, m_infinite: int, sweeps: list,
        solver, log=print, stop_after=None, state: dict | None = None, resume: bool = False,
        wft=None, eig=None, rotate=None):
    """Infinite growth to the full lattice, then finite sweeps.

    target_of(nsites) gives the target quantum number of a superblock of nsites sites;
    sweeps is a list of m values, one per finite sweep (right to the end, left to the
    beginning, right back to the middle). state, when given, receives the left/right block
    stores; with resume=True the infinite growth is skipped and the sweeps start from them.
    eig and rotate are the truncation's density-matrix eigensolver and change of basis (see
    truncate). Returns (final energy,
    records)."""
    Ls = lattice.nsites
    assert Ls % 2 == 0 and Ls >= 4
    sites = [site_block(model, s) for s in range(Ls)]
    if state is None:
        state = {}
    if resume:
        left, right = state["left"], state["right"]
    else:
        left = {1: sites[0]}
        right = {1: sites[Ls - 1]}
        state["left"], state["right"] = left, right
    records = []
    energy = None

    W = lattice.max_range
    prev = {}
    if wft is True:
        wft = GpuWft()

    def step(phase, lsize, rsize, m, trunc):
        nonlocal energy
        S = left[lsize]
        E = right[rsize]
        l = lsize  # the left block's added site
        r = Ls - rsize - 1  # the right block's added site
        # Operators are kept for sites within the bond range of the block's open end.
        tg = time.perf_counter()
        Lb = grow(S, sites[l], model, lattice, lambda s: l - s < W)
        Rb = grow(sites[r], E, model, lattice, lambda s: s - r < W)
        nsuper = lsize + rsize + 2
        tk0 = time.perf_counter()
        target = target_of(lsize + rsize + 2)
        prob, patches = kron_problem(Lb, Rb, model, lattice, target,
                                     name=f"{model.name}-{phase}-{l}", offset=Ls - nsuper)
        tw = time.perf_counter()
        init = _wft_init(wft, prev, phase, l, Lb, Rb, prob, patches, left, right, lsize,
                         rsize) if wft else None
        t0 = time.perf_counter()
        e, z, rep = solver(prob, init=init) if wft else solver(prob)
        t1 = time.perf_counter()
        rep = dict(rep)
        if init is not None:
            wft.free(init)
        zd = rep.pop("z_dev", None)
        timing = {"grow_ms": 1e3 * (tk0 - tg), "kron_ms": 1e3 * (tw - tk0),
                  "wft_ms": 1e3 * (t0 - tw)}
        off = prob.offsets
        psi = [z[off[I]:off[I + 1]].reshape((prob.R[I], prob.L[I]), order="F")
               for I in range(prob.npatches)]
        nrm = np.linalg.norm(z)
        psi = [p / nrm for p in psi]
        # The device copy of the ground state, for density matrices formed on the device.
        dev_psi = (zd.ptr, off, prob.R, prob.L, 1.0 / nrm ** 2) \
            if zd is not None and getattr(zd, "ptr", None) else None
        disc = 0.0
        tparts = {}
        if "left" in trunc:
            left[lsize + 1], d = truncate(Lb, psi, patches, "left", m, model, eig, tparts,
                                          rotate, dev_psi)
            disc = max(disc, d)
        if "right" in trunc:
            right[rsize + 1], d = truncate(Rb, psi, patches, "right", m, model, eig, tparts,
                                           rotate, dev_psi)
            disc = max(disc, d)
        if wft:
            old = prev.pop("z_dev", None)
            if old is not None:
                wft.free(old)
            prev.update(phase=phase, l=l, Lb=Lb, Rb=Rb, prob=prob, patches=patches, z_dev=zd)
            rep["wft_init"] = init is not None
        elif zd is not None and hasattr(zd, "free"):
            zd.free()
        energy = e
        timing["truncate_ms"] = 1e3 * (time.perf_counter() - t1)
        timing.update({"truncate_" + k: v for k, v in tparts.items()})
        timing.update(rep.get("timing", {}))
        rep["timing"] = timing
        rec = StepRecord(phase, l, m, prob.n, prob.npatches, len(prob.blocks), e,
                         rep.get("steps", 0), rep.get("matvecs", 0), 1e3 * (t1 - t0),
                         rep.get("ms_matvec", 0.0), disc, rep)
        records.append(rec)
        tm = rep.get("timing", {})
        log(f"{phase:8s} center={l:3d} m={m:5d} n={prob.n:9d} patches={prob.npatches:4d} "
            f"terms={len(prob.blocks):5d} E={e:.12f} steps={rec.lanczos_steps:3d} "
            f"solve={rec.ms_lanczos:7.1f}ms (plan {tm.get('plan_ms', 0):6.1f} "
            f"[scan {tm.get('plan_scan_ms', 0):.1f} stage {tm.get('plan_stage_ms', 0):.1f} "
            f"upload {tm.get('plan_upload_ms', 0):.1f} raw {tm.get('plan_raw_mb', 0):.0f} MiB] "
            f"lanczos "
            f"{tm.get('lanczos_ms', 0):6.1f}) grow={tm['grow_ms']:6.1f} kron={tm['kron_ms']:6.1f} "
            f"wft={tm['wft_ms']:6.1f} trunc={tm['truncate_ms']:6.1f} disc={disc:.2e}")
        return rec

    # Infinite-size growth: superblock of 2*lsize + 2 sites.
    lsize = 1
    while not resume and 2 * lsize + 2 <= Ls:
        step("infinite", lsize, lsize, m_infinite, ("left", "right"))
        lsize += 1
        if stop_after and len(records) >= stop_after:
            return energy, records
    # Finite sweeps. The center's left block has lsize sites, right block Ls - lsize - 2.
    half = Ls // 2 - 1
    schedule = []
    for m in sweeps:
        schedule += [("sweep>", l, m, ("left",)) for l in range(half, Ls - 3)]   # to the right end
        schedule += [("sweep<", l, m, ("right",)) for l in range(Ls - 3, 1, -1)]  # to the left end
        schedule += [("sweep>", l, m, ("left",)) for l in range(1, half)]       # back to the middle
    for phase, l, m, trunc in schedule:
        step(phase, l, Ls - l - 2, m, trunc)
        if stop_after and len(records) >= stop_after:
            break
    if wft and prev.get("z_dev") is not None:
        wft.free(prev["z_dev"])
    return energy, records


class GpuWft:
    """Executes the wave-function transformation on the device (DeviceArray vectors)."""

    def copy(self, z, n):
        from . import DeviceArray, _check, lib
        y = DeviceArray(n)
        _check(lib().dmk_device_copy_d(y.ptr, z.ptr, n), "dmk_device_copy_d")
        return y

    def apply(self, chains, z, old_off, old_R, n, new_off, new_R):
        from . import wft as W
        return W.apply_gpu(chains, z, old_off, old_R, n, new_off, new_R)

    def free(self, z):
        z.free()


def _wft_init(ex, prev, phase, l, Lb, Rb, prob, patches, left, right, lsize, rsize):
    """The previous ground state carried to this superblock, or None (infinite growth, first
    step). A step at the same center (end of the infinite growth) reuses it unchanged."""
    from . import wft as W

    if not prev or phase == "infinite" or prev.get("z_dev") is None:
        return None
    pprob, pl = prev["prob"], prev["l"]
    old = (prev["patches"], pprob.L, pprob.R)
    new = (patches, prob.L, prob.R)
    if l == pl:
        if prob.n != pprob.n or list(patches) != list(prev["patches"]):
            return None
        return ex.copy(prev["z_dev"], prob.n)
    if l == pl + 1:
        # S' = left[lsize] was truncated from the old left block; E = right[rsize + 1] expands
        # into s_{l+1}.right[rsize] = this step's right block.
        Sp, E = left[lsize], right[rsize + 1]
        if Sp.wft is None or E.wft is None:
            return None
        chains = W.chains_move_right(old, new, Sp.wft, E.wft, prev["Rb"].product, Lb.product)
    elif l == pl - 1:
        Epp, S = right[rsize], left[lsize + 1]
        if Epp.wft is None or S.wft is None:
            return None
        chains = W.chains_move_left(old, new, Epp.wft, S.wft, prev["Lb"].product, Rb.product)
    else:
        return None
    return ex.apply(chains, prev["z_dev"], pprob.offsets, pprob.R, prob.n, prob.offsets, prob.R)
