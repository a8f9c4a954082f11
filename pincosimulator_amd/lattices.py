"""Lattice couplings for matrix hopping.

``lieb_coupling`` builds the Lieb-chain coupling of TWMCLiebSolver::Setup
(src/TWMC/Solvers/TWMCLiebSolver.cpp:52-85): per cell A(3i)-B(3i+1) J_AB,
B(3i+1)-C(3i+2) J_BC, inter-cell C(3i+2)-B(3i+4) J_BC, and with PBC the
closing C(last)-B(0) bond.  Optional Gaussian bond disorder J_b + sigma N(0,1)
(an extension: the reference has no hopping disorder) uses the engine's
Philox draws with tag 7 on trajectory 0, one draw per bond in that order --
the same numbers the C++ host (LiebCoupling) produces.
"""
from __future__ import annotations

import numpy as np

from .engine import noisy_matrix

TAG_BOND = 7
TAG_OMEGA = 3


def lieb_bonds(cells, pbc=False):
    bonds = []
    for i in range(cells):
        bonds += [(3 * i, 3 * i + 1, "AB"), (3 * i + 1, 3 * i + 2, "BC")]
        if i == cells - 1:
            if pbc:
                bonds.append((3 * i + 2, 1, "BC"))
            break
        bonds.append((3 * i + 2, 3 * (i + 1) + 1, "BC"))
    return bonds


def lieb_coupling(cells, J_AB=1.0, J_BC=1.0, pbc=False, sigma=0.0, seed=0):
    bonds = lieb_bonds(cells, pbc)
    z = np.zeros(len(bonds), dtype=np.complex128)
    if sigma != 0.0:
        z = noisy_matrix({"base": 0.0, "noise": ("gaussian", sigma)}, len(bonds), seed, 0, TAG_BOND)
    n = 3 * cells
    C = np.zeros((n, n))
    for b, (i, j, kind) in enumerate(bonds):
        J = (J_AB if kind == "AB" else J_BC) + z[b].real
        C[i, j] += J
        C[j, i] += J
    return C


def onsite_disorder(n, sigma, seed, base=0.0):
    """One shared Gaussian realisation of an on-site energy (E), real."""
    return noisy_matrix({"base": base, "noise": ("gaussian", sigma)}, n, seed, 0, TAG_OMEGA).real
