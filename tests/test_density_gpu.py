"""The truncation's density-matrix eigensolver on the GPU (dmk_density_diag_d, the replacement of
DiagBlockDiagMatrix::diagonalise): eigenpairs against the host LAPACK, degenerate spectra, and a
sweep whose truncation runs on the GPU against the same sweep on the host. Also the truncation's
change of basis on the GPU (sweep.GpuBasisChange) against the host's BSOp.transform, and the SVD
form of the truncation (dmk_density_svd_d, the replacement of DensityMatrixSvd::diag) against the
host LAPACK."""
import ctypes

import numpy as np
import pytest

import dmrgpp_amd as dm
from dmrgpp_amd import configs, models, sweep
from dmrgpp_amd.blocks import Basis, BSOp


def _density_like(n, rank, seed):
    # rho = X^T X of a random R x n segment: symmetric, positive semidefinite, rank min(R, n)
    # (zero eigenvalues repeated when R < n), scaled to unit trace as a density matrix is.
    x = np.random.default_rng(seed).standard_normal((rank, n))
    rho = x.T @ x
    return rho / np.trace(rho)


def test_density_diag_rejects_bad_arguments():
    # Argument checks run before any device call.
    lib = dm.lib()
    sizes = (ctypes.c_int64 * 1)(-3)
    ptrs = (ctypes.c_void_p * 1)(None)
    eigs = (ctypes.c_double * 1)()
    rc = lib.dmk_density_diag_d(1, ctypes.addressof(sizes), ctypes.addressof(ptrs),
                                ctypes.addressof(eigs))
    assert rc != 0 and b"out of range" in lib.dmk_last_error()
    assert lib.dmk_density_diag_d(-1, None, None, None) != 0
    # No blocks, or only empty ones: nothing to do.
    assert lib.dmk_density_diag_d(0, None, None, None) == 0
    sizes[0] = 0
    assert lib.dmk_density_diag_d(1, ctypes.addressof(sizes), ctypes.addressof(ptrs),
                                  ctypes.addressof(eigs)) == 0


@pytest.mark.gpu
def test_gpu_density_diag_matches_lapack():
    # Sizes from 1 to 700, full-rank and rank-deficient blocks (a zero eigenvalue repeated
    # hundreds of times), in one call.
    specs = [(1, 1), (5, 5), (64, 64), (300, 40), (700, 700), (513, 100)]
    blocks = [_density_like(n, r, seed=i) for i, (n, r) in enumerate(specs)]
    out = dm.density_diag(blocks)
    for a, (w, v) in zip(blocks, out):
        n = a.shape[0]
        want = np.linalg.eigvalsh(a)
        assert np.all(np.diff(w) >= -1e-15)                      # ascending
        assert np.max(np.abs(w - want)) < 1e-13                  # unit trace: absolute = relative
        assert np.linalg.norm(a @ v - v * w) < 1e-12
        assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-11 * np.sqrt(n)


@pytest.mark.gpu
def test_gpu_density_diag_padded_batch_classes():
    # Blocks of similar size are solved as one padded strided batch (sizes down to half the
    # class's largest): density blocks, a rank-deficient one, an exactly degenerate one and a
    # general symmetric block with negative eigenvalues (the padding must stay below every
    # spectrum), next to a second class and a single block. Every block against LAPACK.
    rng = np.random.default_rng(11)
    q, _ = np.linalg.qr(rng.standard_normal((256, 256)))
    degenerate = q @ np.diag(np.repeat([0.5, 0.25, 0.125, 0.0625], 64)) @ q.T
    general = rng.uniform(-1, 1, (300, 300))
    general = 3.0 * (general + general.T)
    blocks = [_density_like(400, 400, 1), _density_like(390, 60, 2), degenerate, general,
              _density_like(201, 201, 3), _density_like(200, 200, 4),
              _density_like(120, 30, 5), _density_like(90, 90, 6), _density_like(17, 17, 7)]
    out = dm.density_diag(blocks)
    for a, (w, v) in zip(blocks, out):
        n = a.shape[0]
        w0, v0 = np.linalg.eigh(a)
        scale = max(1.0, np.max(np.abs(w0)))
        assert np.max(np.abs(w - w0)) < 1e-13 * scale, n
        assert np.linalg.norm(a @ v - v * w) < 1e-12 * scale * np.sqrt(n), n
        assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-11 * np.sqrt(n), n
    # The degenerate block's eigenspaces.
    (w, v), a = out[2], blocks[2]
    w0, v0 = np.linalg.eigh(a)
    for lam in (0.5, 0.25, 0.125, 0.0625):
        s = np.abs(w0 - lam) < 1e-10
        assert s.sum() == 64
        assert np.linalg.norm(v[:, s] @ v[:, s].T - v0[:, s] @ v0[:, s].T) < 1e-12


@pytest.mark.gpu
def test_gpu_density_matrices_formed_on_the_device():
    # rho_I = s X_I^T X_I (left) and s X_I X_I^T (right) from a device vector holding the
    # patches' R_I x L_I segments, against numpy; then their eigenpairs from device memory
    # (dmk_density_diag_dev_d) against the host-memory path and LAPACK.
    rng = np.random.default_rng(8)
    R, L = [3, 70, 1, 300, 129], [40, 1, 90, 260, 128]
    off = np.concatenate([[0], np.cumsum(np.array(R) * np.array(L))])
    z = rng.standard_normal(off[-1])
    dz = dm.DeviceArray.from_host(z)
    scale = 1.0 / np.dot(z, z)
    for side in ("left", "right"):
        dev, blocks = dm.density_matrices_device(dz.ptr, off, R, L, side, scale)
        want = []
        for I in range(len(R)):
            X = z[off[I]:off[I + 1]].reshape((R[I], L[I]), order="F")
            want.append(scale * (X.T @ X if side == "left" else X @ X.T))
        for (p, n), w in zip(blocks, want):
            got = np.empty((n, n), order="F")
            dm._check(dm.lib().dmk_memcpy_d2h(got.ctypes.data, p, got.nbytes), "d2h")
            assert np.max(np.abs(got - w)) < 1e-14 * max(1.0, np.max(np.abs(w)))
        pairs = dm.density_diag_device(blocks)
        host = dm.density_diag([0.5 * (w + w.T) for w in want])
        for w, (ev, v), (eh, _) in zip(want, pairs, host):
            a = 0.5 * (w + w.T)
            n = a.shape[0]
            assert np.max(np.abs(ev - eh)) < 1e-14
            assert np.max(np.abs(ev - np.linalg.eigvalsh(a))) < 1e-14
            assert np.linalg.norm(a @ v - v * ev) < 1e-13 * np.sqrt(n)
            assert np.linalg.norm(v.T @ v - np.eye(n)) < 1e-11 * np.sqrt(n)
        dev.free()
    dz.free()


@pytest.mark.gpu
def test_gpu_density_diag_degenerate_spectrum():
    # Exactly degenerate eigenvalues (a Kronecker structure gives each one several times): the
    # eigenspace projectors agree with LAPACK's even though the vectors inside them may not.
    d = np.diag([0.4, 0.3, 0.2, 0.1])
    q, _ = np.linalg.qr(np.random.default_rng(3).standard_normal((16, 16)))
    a = q @ np.kron(d, d) @ q.T
    (w, v), = dm.density_diag([a])
    w0, v0 = np.linalg.eigh(a)
    assert np.max(np.abs(w - w0)) < 1e-14
    for lam in np.unique(np.round(w0, 12)):
        s = np.abs(w0 - lam) < 1e-10
        assert np.linalg.norm(v[:, s] @ v[:, s].T - v0[:, s] @ v0[:, s].T) < 1e-12


@pytest.mark.gpu
def test_gpu_truncation_on_gpu_matches_host():
    # A Hubbard chain sweep with every density-matrix block on the GPU gives the host-truncation
    # energies and discarded weights.
    runs = {}
    for key, min_rows in (("host", 0), ("gpu", 1)):
        e, recs, _ = configs.run_config("hubbard-chain-8-U0", log=lambda s: None,
                                        gpu_trunc_min=min_rows)
        runs[key] = (e, [r.energy for r in recs], [r.discarded for r in recs])
    (eh, sh, dh), (eg, sg, dg) = runs["host"], runs["gpu"]
    assert abs(eh - eg) < 1e-10 * abs(eh)
    assert np.max(np.abs(np.array(sh) - np.array(sg))) < 1e-9
    assert np.max(np.abs(np.array(dh) - np.array(dg))) < 1e-10


@pytest.mark.gpu
def test_gpu_basis_change_matches_host_transform():
    # Three groups of 40, 7 and 5 states truncated to 12, 0 and 5: blocks between kept groups,
    # blocks touching the dropped group (left out), and one W matrix shared by several blocks
    # and operators.
    rng = np.random.default_rng(11)
    old = Basis([(0,), (1,), (2,)], [40, 7, 5])
    nb = Basis([(0,), (2,)], [12, 5])
    q0, _ = np.linalg.qr(rng.standard_normal((40, 40)))
    q2, _ = np.linalg.qr(rng.standard_normal((5, 5)))
    W = {0: (np.asfortranarray(q0[:, :12]), (0, 12)), 2: (np.asfortranarray(q2), (1, 5))}
    sizes = old.sizes
    blk = lambda a, b: np.asfortranarray(rng.standard_normal((sizes[a], sizes[b])))  # noqa: E731
    ops = [BSOp(old, old, {(0, 0): blk(0, 0), (2, 2): blk(2, 2), (1, 1): blk(1, 1)}),
           BSOp(old, old, {(0, 2): blk(0, 2), (2, 0): blk(2, 0), (1, 0): blk(1, 0)}, True),
           BSOp(old, old, {(0, 0): blk(0, 0)})]
    got = sweep.GpuBasisChange()(ops, W, nb)
    for op, g in zip(ops, got):
        want = op.transform(W, nb, W, nb)
        assert g.fermionic == op.fermionic
        assert set(g.blocks) == set(want.blocks)
        for key, b in want.blocks.items():
            assert g.blocks[key].shape == b.shape
            assert np.max(np.abs(g.blocks[key] - b)) < 1e-13 * max(1.0, np.max(np.abs(b)))


@pytest.mark.gpu
def test_gpu_basis_change_sweep_matches_host():
    # The same Hubbard sweep with the change of basis on the GPU and on the host.
    runs = {}
    for key, on_gpu in (("host", False), ("gpu", True)):
        e, recs, _ = configs.run_config("hubbard-chain-8-U0", log=lambda s: None,
                                        gpu_basis=on_gpu)
        runs[key] = (e, [r.energy for r in recs], [r.discarded for r in recs])
    (eh, sh, dh), (eg, sg, dg) = runs["host"], runs["gpu"]
    assert abs(eh - eg) < 1e-10 * abs(eh)
    assert np.max(np.abs(np.array(sh) - np.array(sg))) < 1e-9
    assert np.max(np.abs(np.array(dh) - np.array(dg))) < 1e-10


@pytest.mark.gpu
def test_gpu_density_eig_splits_host_and_device():
    # Blocks above the threshold go to the device (several streams at once), the rest to the
    # host LAPACK on the calling thread meanwhile; every block comes back in its own slot.
    sizes = [5, 700, 70, 300, 1, 530, 30, 64]
    rhos = [_density_like(n, max(1, n // 2), seed=20 + i) for i, n in enumerate(sizes)]
    got = sweep.GpuDensityEig(64)(rhos)
    for a, (w, v) in zip(rhos, got):
        n = a.shape[0]
        assert w.shape == (n,) and v.shape == (n, n)
        assert np.max(np.abs(w - np.linalg.eigvalsh(a))) < 1e-13
        assert np.linalg.norm(a @ v - v * w) < 1e-12


def test_density_svd_rejects_bad_arguments():
    # Argument checks of dmk_density_svd_d run before any device call.
    lib = dm.lib()
    rows = (ctypes.c_int64 * 1)(4)
    cols = (ctypes.c_int64 * 1)(-1)
    ptrs = (ctypes.c_void_p * 1)(None)
    s = (ctypes.c_double * 4)()
    rc = lib.dmk_density_svd_d(1, ctypes.addressof(rows), ctypes.addressof(cols),
                               ctypes.addressof(ptrs), ctypes.addressof(ptrs), ctypes.addressof(s))
    assert rc != 0 and b"out of range" in lib.dmk_last_error()
    cols[0] = 3
    rc = lib.dmk_density_svd_d(1, ctypes.addressof(rows), ctypes.addressof(cols),
                               ctypes.addressof(ptrs), ctypes.addressof(ptrs), ctypes.addressof(s))
    assert rc != 0 and b"null" in lib.dmk_last_error()
    assert lib.dmk_density_svd_d(-1, None, None, None, None, None) != 0
    assert lib.dmk_density_svd_d(0, None, None, None, None, None) == 0


@pytest.mark.gpu
def test_gpu_density_svd_matches_lapack():
    # Tall, wide, square, rank-deficient, single-row and single-column blocks, and one with no
    # columns (identity), in one call: singular values against LAPACK, U orthogonal, and
    # M = U S V^T reproduced through the leading columns of U.
    rng = np.random.default_rng(7)
    shapes = [(700, 300), (200, 640), (513, 513), (64, 1), (1, 90), (40, 0), (333, 333)]
    mats = [rng.standard_normal(s) for s in shapes]
    mats[-1] = mats[-1][:, :20] @ rng.standard_normal((20, 333))  # rank 20
    out = dm.density_svd(mats)
    for m, (sv, u) in zip(mats, out):
        r, c = m.shape
        k = min(r, c)
        assert sv.shape == (k,) and u.shape == (r, r)
        assert np.linalg.norm(u.T @ u - np.eye(r)) < 1e-11 * np.sqrt(r)
        if k == 0:
            assert np.array_equal(u, np.eye(r))
            continue
        s0 = np.linalg.svd(m, compute_uv=False)
        assert np.all(np.diff(sv) <= 1e-12 * s0[0])
        assert np.max(np.abs(sv - s0)) < 1e-12 * s0[0]
        # The leading k columns of U carry M: U_k^T M has the singular values as row norms.
        proj = u[:, :k].T @ m
        assert np.max(np.abs(np.linalg.norm(proj, axis=1) - sv)) < 1e-11 * s0[0]
        assert np.linalg.norm(m - u[:, :k] @ proj) < 1e-11 * np.linalg.norm(m)


@pytest.mark.gpu
def test_gpu_svd_truncation_matches_host():
    # The SVD form of the truncation with every block on the GPU: the host-eigensolver sweep's
    # energies and discarded weights.
    model, lat = models.hubbard(0.0), models.chain(8)
    runs = {}
    for key, eig in (("host", None), ("svd", sweep.SvdDensityEig(min_size=1, device=True))):
        e, recs = sweep.run(model, lat, lambda N: (N // 2, N // 2), 256, [256],
                            sweep.GpuSolver(precision="fp64"), log=lambda s: None, eig=eig)
        runs[key] = (e, [r.energy for r in recs], [r.discarded for r in recs])
    (eh, sh, dh), (es, ss, ds) = runs["host"], runs["svd"]
    assert abs(eh - es) < 1e-10 * abs(eh)
    assert np.max(np.abs(np.array(sh) - np.array(ss))) < 1e-9
    assert np.max(np.abs(np.array(dh) - np.array(ds))) < 1e-10
