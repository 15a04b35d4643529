"""The sweep driver that produces real superblock Kron problems: exact energies on small
systems (ED, free-fermion closed form), structure of the decomposition, and on the GPU the same
energies through the C-ABI plugin and device Lanczos."""
import numpy as np
import pytest

from dmrgpp_amd import configs, models, sweep
from oracle import oracle


class OracleSolver:
    """Lanczos of the oracle (LanczosCore restatement) on the plugin-layout problem."""

    def __call__(self, prob):
        init = np.random.default_rng(1234).uniform(-1, 1, prob.n)
        e, z, s, _, _ = oracle.lanczos(prob, init)
        return e, z, {"steps": s, "matvecs": s}


E_HEIS_L8 = -3.374932598687897        # ED (dmrg/CMakeLists.txt:199-200 pins -3.37493)
E_LADDER_4x2 = -4.293066457           # ED (dmrg/CMakeLists.txt:212-213 pins -4.29307)
E_HUB_U0_L8 = 2 * sum(-2 * np.cos(k * np.pi / 9) for k in range(1, 5))  # -9.517540966


def test_sweep_heisenberg_chain8_exact():
    e, recs, _ = configs.run_config("heisenberg-chain-8", solver=OracleSolver(), log=lambda s: None)
    assert abs(e - E_HEIS_L8) < 1e-10


def test_sweep_heisenberg_ladder_exact():
    e, recs, _ = configs.run_config("heisenberg-ladder-4x2", solver=OracleSolver(),
                                    log=lambda s: None)
    assert abs(e - E_LADDER_4x2) < 1e-8


def test_sweep_hubbard_free_fermions_exact():
    # U = 0, 8 sites, 4 up 4 down: 2 * sum_{k=1..4} -2 cos(k pi / 9) (input2 / input2_batchedgemm).
    e, recs, _ = configs.run_config("hubbard-chain-8-U0", solver=OracleSolver(), log=lambda s: None)
    assert abs(e - E_HUB_U0_L8) < 1e-10


def _ed_hubbard_chain(L, U, t=1.0):
    nm = 2 * L
    dim = 2 ** nm

    def c(mode):
        M = np.zeros((dim, dim))
        for s in range(dim):
            if s >> mode & 1:
                M[s ^ (1 << mode), s] = (-1) ** bin(s & ((1 << mode) - 1)).count("1")
        return M

    C = [c(m) for m in range(nm)]
    H = np.zeros((dim, dim))
    for i in range(L - 1):
        for sp in range(2):
            a, b = C[2 * i + sp], C[2 * (i + 1) + sp]
            H += -t * (a.T @ b + b.T @ a)
    for i in range(L):
        H += U * (C[2 * i].T @ C[2 * i]) @ (C[2 * i + 1].T @ C[2 * i + 1])
    nup = sum(C[2 * i].T @ C[2 * i] for i in range(L)).diagonal()
    ndn = sum(C[2 * i + 1].T @ C[2 * i + 1] for i in range(L)).diagonal()
    idx = np.where((nup == L // 2) & (ndn == L // 2))[0]
    return np.linalg.eigvalsh(H[np.ix_(idx, idx)])[0]


def test_sweep_hubbard_interacting_matches_ed():
    e_ed = _ed_hubbard_chain(6, 4.0)
    e, _ = sweep.run(models.hubbard(4.0), models.chain(6), lambda N: (N // 2, N // 2), 64, [64],
                     OracleSolver(), log=lambda s: None)
    assert abs(e - e_ed) < 1e-10


def test_kron_problem_structure_and_hermiticity():
    # A mid-sweep superblock of the Heisenberg ladder: connections H_L (x) I, I (x) H_R and three
    # bonds x three terms; the assembled H is symmetric and its blocks respect the patches.
    model, lat = models.heisenberg(), models.ladder(4, 2)
    captured = {}

    class Capture(OracleSolver):
        def __call__(self, prob):
            captured.setdefault("prob", prob)
            return super().__call__(prob)

    sweep.run(model, lat, lambda N: (0,), 16, [], Capture(), log=lambda s: None, stop_after=2)
    prob = captured["prob"]
    H = oracle.kron_form(prob)
    assert np.abs(H - H.T).max() < 1e-12
    for (I, J, k), (a, b) in prob.blocks.items():
        if k == 0:
            assert I == J and np.array_equal(b, np.eye(prob.R[I]))
        if k == 1:
            assert I == J and np.array_equal(a, np.eye(prob.L[I]))


def test_lattice_bonds():
    lat = models.ladder(3, 2)
    assert sorted(lat.bonds) == [(0, 1, 1.0), (0, 2, 1.0), (1, 3, 1.0), (2, 3, 1.0), (2, 4, 1.0),
                                 (3, 5, 1.0), (4, 5, 1.0)]
    cyl = models.ladder(2, 4, periodic_y=True)
    assert (0, 3, 1.0) in cyl.bonds and (4, 7, 1.0) in cyl.bonds and len(cyl.bonds) == 12


@pytest.mark.gpu
def test_gpu_sweep_small_systems_exact():
    e, _, _ = configs.run_config("heisenberg-chain-8", log=lambda s: None)
    assert abs(e - E_HEIS_L8) < 1e-10
    e, _, _ = configs.run_config("hubbard-chain-8-U0", log=lambda s: None)
    assert abs(e - E_HUB_U0_L8) < 1e-10


@pytest.mark.gpu
def test_gpu_sweep_heisenberg_chain32_reference_pin():
    # The reference's own pin of this model: input3100, -13.9973 at 2e-5 (dmrg/CMakeLists.txt:343).
    e, recs, _ = configs.run_config("heisenberg-chain-32", sweeps=1, log=lambda s: None)
    assert abs(e - (-13.9973)) < 2e-5
    e32, _, _ = configs.run_config("heisenberg-chain-32", precision="fp32", sweeps=1,
                                   log=lambda s: None)
    assert abs(e32 - e) < 1e-10 * abs(e)


class HostWft:
    """Host executor of the wave-function transformation chains (test reference)."""

    def copy(self, z, n):
        return np.array(z)

    def apply(self, chains, z, old_off, old_R, n, new_off, new_R):
        from dmrgpp_amd import wft
        return wft.apply_numpy(chains, z, old_off, old_R, n, new_off, new_R)

    def free(self, z):
        pass


class OracleSolverInit:
    """Oracle Lanczos that starts from a given vector and reports the overlap of that start with
    the ground state it finds."""

    def __init__(self):
        self.overlaps = []

    def __call__(self, prob, init=None):
        x0 = np.random.default_rng(1234).uniform(-1, 1, prob.n) if init is None else init
        e, z, s, _, _ = oracle.lanczos(prob, x0)
        zn = z / np.linalg.norm(z)
        if init is not None:
            self.overlaps.append((s, abs(zn @ init) / np.linalg.norm(init)))
        return e, z, {"steps": s, "matvecs": s, "z_dev": zn}


@pytest.mark.parametrize("name,model,lat,target,m,tol", [
    # No truncation (m covers the blocks): the transformation is exact.
    ("heisenberg-chain-10", models.heisenberg(), models.chain(10), lambda N: (0,), 64, 1e-12),
    ("heisenberg-ladder-4x2", models.heisenberg(), models.ladder(4, 2), lambda N: (0,), 16, 1e-12),
    # Truncated fermions (Jordan-Wigner signs in the transformed state): overlap near 1.
    ("hubbard-chain-8", models.hubbard(4.0), models.chain(8), lambda N: (N // 2, N // 2), 64, 1e-7),
])
def test_wft_carries_ground_state_between_steps(name, model, lat, target, m, tol):
    e_ref, recs_ref = sweep.run(model, lat, target, m, [m, m], OracleSolver(), log=lambda s: None)
    sol = OracleSolverInit()
    e, recs = sweep.run(model, lat, target, m, [m, m], sol, log=lambda s: None, wft=HostWft())
    assert abs(e - e_ref) < 1e-10
    # Every finite step after the first starts from the transformed previous ground state.
    nfinite = sum(r.phase != "infinite" for r in recs)
    assert len(sol.overlaps) == nfinite
    assert min(ov for _, ov in sol.overlaps) > 1 - tol
    steps_wft = sum(r.lanczos_steps for r in recs if r.phase != "infinite")
    steps_ref = sum(r.lanczos_steps for r in recs_ref if r.phase != "infinite")
    assert steps_wft < 0.6 * steps_ref, (steps_wft, steps_ref)


@pytest.mark.gpu
def test_gpu_wft_sweep_matches_host_chains():
    # The same sweep with the device executor (batched GEMMs) and the host one: equal energies,
    # and the GPU-transformed start vectors agree with the host evaluation of the same chains.
    from dmrgpp_amd import DeviceArray, wft

    model, lat, target, m = models.heisenberg(), models.ladder(6, 2), (lambda N: (0,)), 64
    e_host, _ = sweep.run(model, lat, target, m, [m], OracleSolverInit(), log=lambda s: None,
                          wft=HostWft())
    checked = []

    class CheckingGpuWft(sweep.GpuWft):
        def apply(self, chains, z, old_off, old_R, n, new_off, new_R):
            y = super().apply(chains, z, old_off, old_R, n, new_off, new_R)
            want = wft.apply_numpy(chains, z.download(), old_off, old_R, n, new_off, new_R)
            got = y.download()
            checked.append(np.abs(got - want).max() / np.abs(want).max())
            return y

    e_gpu, recs = sweep.run(model, lat, target, m, [m], sweep.GpuSolver(), log=lambda s: None,
                            wft=CheckingGpuWft())
    assert abs(e_gpu - e_host) < 1e-9
    assert checked and max(checked) < 1e-13
    assert all(r.info.get("wft_init") for r in recs if r.phase != "infinite")


def test_svd_truncation_pairs_match_density_eigenpairs():
    # The SVD form of the truncation (DensityMatrixSvd) on the host: the eigenpairs built from
    # the wave-function blocks equal those of M M^T, for tall, wide, square, rank-deficient and
    # single-column blocks; the zero eigenvalues past min(rows, cols) are exact zeros.
    rng = np.random.default_rng(5)
    shapes = [(7, 3), (3, 7), (6, 6), (9, 1), (1, 4), (5, 5)]
    mats = [rng.standard_normal(s) for s in shapes]
    mats[-1] = mats[-1][:, :2] @ rng.standard_normal((2, 5))   # rank 2
    got = sweep.SvdDensityEig(device=False)(mats)
    want = sweep.host_eig([m @ m.T for m in mats])
    for m, (w, v), (w0, v0) in zip(mats, got, want):
        r = m.shape[0]
        assert w.shape == (r,) and v.shape == (r, r)
        assert np.all(np.diff(w) >= 0)
        assert np.max(np.abs(w - w0)) < 1e-12 * max(1.0, w0[-1])
        assert np.linalg.norm(v.T @ v - np.eye(r)) < 1e-12
        assert np.linalg.norm((m @ m.T) @ v - v * w) < 1e-11 * max(1.0, w0[-1])
        k = min(m.shape)
        if k < r:
            assert np.all(w[:r - k] == 0.0)


def test_sweep_with_svd_truncation_exact():
    # The whole sweep with the SVD form of the truncation: the exact Heisenberg L = 8 energy.
    e, _ = sweep.run(models.heisenberg(1.0), models.chain(8), lambda N: (0,), 64, [64],
                     OracleSolver(), log=lambda s: None, eig=sweep.SvdDensityEig(device=False))
    assert abs(e - E_HEIS_L8) < 1e-10


def _product_permutation(P):
    # Position of each product-basis state in the full Kronecker product kron(B, A) (B's index
    # major, the order numpy.kron uses for kron(B_dense, A_dense)).
    A, B = P.A, P.B
    perm = np.empty(A.dim * B.dim, dtype=np.int64)
    for (ga, gb), (G, o) in P.place.items():
        na, nb = A.sizes[ga], B.sizes[gb]
        for i in range(nb):
            for p in range(na):
                perm[P.basis.offsets[G] + o + i * na + p] = \
                    (B.offsets[gb] + i) * A.dim + A.offsets[ga] + p
    return perm


@pytest.mark.parametrize("big_a", [True, False])
def test_product_kron_matches_numpy_kron(big_a):
    # Product.kron against numpy.kron of the dense operators, permuted into the product basis:
    # identity factors on either side, the fermion-parity form of an identity (signA), dense
    # pairs with sparse entries, and accumulation into an existing operator.
    from dmrgpp_amd.blocks import Basis, BSOp, Product

    rng = np.random.default_rng(8)
    big = Basis([(0,), (1,), (2,)], [7, 11, 4])
    small = Basis([(0,), (1,)], [1, 2])
    A, B = (big, small) if big_a else (small, big)

    def op(basis):
        blocks = {}
        for go, so in enumerate(basis.sizes):
            for gi, si in enumerate(basis.sizes):
                if (go + gi) % 2 == 0 or go == gi + 1:
                    m = rng.standard_normal((so, si))
                    m[rng.random((so, si)) < 0.4] = 0.0
                    blocks[(go, gi)] = np.asfortranarray(m)
        return BSOp(basis, basis, blocks)

    a, b = op(A), op(B)
    sign = np.array([1.0, -1.0, -1.0][:A.ngroups])
    P = Product(A, B)
    perm = _product_permutation(P)
    Ia, Ib = np.eye(A.dim), np.eye(B.dim)
    Sa = np.diag(np.repeat(sign, A.sizes))

    def want(Ad, Bd, c=1.0):
        return c * np.kron(Bd, Ad)[np.ix_(perm, perm)]

    for got, ref in ((P.kron(a, None), want(a.dense(), Ib)),
                     (P.kron(None, b, 0.5), want(Ia, b.dense(), 0.5)),
                     (P.kron(a, b, -1.5), want(a.dense(), b.dense(), -1.5)),
                     (P.kron(None, None, 2.0), want(Ia, Ib, 2.0)),
                     (P.kron(None, b, signA=sign), want(Sa, b.dense()))):
        assert np.max(np.abs(got.dense() - ref)) < 1e-13
    H = P.kron(a, None)
    P.kron(None, b, acc=H)
    P.kron(a, b, 0.25, acc=H)
    ref = want(a.dense(), Ib) + want(Ia, b.dense()) + want(a.dense(), b.dense(), 0.25)
    assert np.max(np.abs(H.dense() - ref)) < 1e-13
