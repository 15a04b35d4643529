"""Site Hilbert spaces, operators and lattices of the models the benchmark configurations use:
the spin-1/2 Heisenberg model (2Sz conserved) and the one-band Fermi-Hubbard model (N_up, N_down
conserved) on chains, two-leg ladders and four-leg cylinders.

DMRG++'s own models (dmrg/Models/Heisenberg, dmrg/Models/HubbardOneBand) are out of scope; this
module only produces the site operators and the bonds whose products cross the superblock cut.
Operators are real; S- = (S+)^T and c^dagger = c^T.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .blocks import Basis, BSOp


@dataclass
class SiteModel:
    name: str
    basis: Basis
    ops: dict              # name -> BSOp on one site
    h_site: BSOp           # on-site Hamiltonian
    bond_terms: callable   # value -> list of (op_i, op_j, coef, fermionic)
    parity: np.ndarray     # fermion parity (+1/-1) per site basis group

    def qn_of_filling(self, nsites: int, target: dict) -> tuple:
        raise NotImplementedError


def _site_op(basis: Basis, dense: np.ndarray, fermionic=False) -> BSOp:
    blocks = {}
    for go in range(basis.ngroups):
        for gi in range(basis.ngroups):
            b = dense[basis.offsets[go]:basis.offsets[go + 1], basis.offsets[gi]:basis.offsets[gi + 1]]
            if np.any(b != 0):
                blocks[(go, gi)] = np.asfortranarray(b)
    return BSOp(basis, basis, blocks, fermionic)


def heisenberg(J: float = 1.0) -> SiteModel:
    # States ordered by 2Sz: |down> (-1), |up> (+1).
    basis = Basis([(-1,), (1,)], [1, 1])
    sz = np.diag([-0.5, 0.5])
    sp = np.array([[0.0, 0.0], [1.0, 0.0]])  # |up><down|
    ops = {"Sz": _site_op(basis, sz), "Sp": _site_op(basis, sp), "Sm": _site_op(basis, sp.T)}
    h = _site_op(basis, np.zeros((2, 2)))

    def bond(value):
        return [("Sz", "Sz", value, False), ("Sp", "Sm", 0.5 * value, False),
                ("Sm", "Sp", 0.5 * value, False)]

    return SiteModel("heisenberg", basis, ops, h, bond, np.ones(2))


def hubbard(U: float) -> SiteModel:
    # Fock states |n_up n_dn>, modes ordered (up, down) inside a site:
    # 0:|0>, 1:|up>=c_up^+|0>, 2:|dn>=c_dn^+|0>, 3:|up dn>=c_up^+ c_dn^+|0>.
    # Quantum numbers (N_up, N_dn) sort them as |0>(0,0), |dn>(0,1), |up>(1,0), |updn>(1,1).
    qn = {0: (0, 0), 1: (1, 0), 2: (0, 1), 3: (1, 1)}
    order = sorted(range(4), key=lambda s: qn[s])
    pos = {s: i for i, s in enumerate(order)}
    basis = Basis([qn[s] for s in order], [1, 1, 1, 1])
    cu = np.zeros((4, 4))
    cd = np.zeros((4, 4))
    cu[pos[0], pos[1]] = 1.0    # c_up |up> = |0>
    cu[pos[2], pos[3]] = 1.0    # c_up c_up^+ c_dn^+ |0> = c_dn^+|0>
    cd[pos[0], pos[2]] = 1.0    # c_dn |dn> = |0>
    cd[pos[1], pos[3]] = -1.0   # c_dn c_up^+ c_dn^+ |0> = -c_up^+|0>
    nup = cu.T @ cu
    ndn = cd.T @ cd
    ops = {"cu": _site_op(basis, cu, True), "cd": _site_op(basis, cd, True),
           "cdu": _site_op(basis, cu.T.copy(), True), "cdd": _site_op(basis, cd.T.copy(), True)}
    h = _site_op(basis, U * (nup @ ndn))
    parity = np.array([1.0 if sum(q) % 2 == 0 else -1.0 for q in basis.qns])

    def bond(t):
        # -t (c+_i c_j + c+_j c_i) per spin; with i in the left part the second product is
        # -c_i c+_j, which becomes +t (c_i P) (x) c+_j.
        return [("cdu", "cu", -t, True), ("cu", "cdu", t, True),
                ("cdd", "cd", -t, True), ("cd", "cdd", t, True)]

    return SiteModel("hubbard", basis, ops, h, bond, parity)


@dataclass
class Lattice:
    """legs x Lx lattice numbered column by column (site = legs * x + y); legs = 1 is a chain.

    Bonds follow a translation-invariant rule, so the same rule also couples the left block to
    the right block of a shorter superblock during infinite-size growth."""
    nsites: int
    legs: int = 1
    leg_value: float = 1.0
    rung_value: float = 1.0
    periodic_y: bool = False

    def bond_value(self, i: int, j: int):
        """Coupling of sites i < j, or None."""
        if j < i:
            i, j = j, i
        d = j - i
        y = i % self.legs
        if d == self.legs:
            return self.leg_value
        if self.legs > 1 and d == 1 and y < self.legs - 1:
            return self.rung_value
        if self.periodic_y and self.legs > 2 and d == self.legs - 1 and y == 0:
            return self.rung_value
        return None

    @property
    def bonds(self) -> list:
        out = []
        for i in range(self.nsites):
            for j in range(i + 1, min(self.nsites, i + self.legs + 1)):
                v = self.bond_value(i, j)
                if v is not None:
                    out.append((i, j, v))
        return out

    @property
    def max_range(self) -> int:
        return self.legs


def chain(L: int, value: float = 1.0) -> Lattice:
    return Lattice(L, 1, value, value)


def ladder(Lx: int, legs: int = 2, leg_value: float = 1.0, rung_value: float = 1.0,
           periodic_y: bool = False) -> Lattice:
    """legs x Lx; for legs > 2 with periodic_y the rungs close around a cylinder."""
    return Lattice(legs * Lx, legs, leg_value, rung_value, periodic_y)
