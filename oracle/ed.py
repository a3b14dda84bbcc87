"""Exact diagonalisation of the Fermi-Hubbard model for small systems (a physics check of fTWA).

TEST INFRASTRUCTURE ONLY, like the rest of oracle/: imported by tests/ as a reference for the
short-time dynamics, never by the product package.

H = sum_{ij,s} tM_ij c^dagger_is c_js + sum_i U_ii n_iu n_id on the fixed-(N_up, N_dn) sector, with the
Jordan-Wigner order "all up orbitals, then all down orbitals" (so the two spin species' hoppings
commute and carry signs only from same-spin orbitals between i and j).  The time evolution of a
product state uses scipy's expm_multiply.  This is the textbook many-body problem that fTWA
approximates (Davidson, Sels & Polkovnikov, Ann. Phys. 384, 128 (2017)); the reference's fTWA
is exact for quadratic Hamiltonians and agrees with ED at short times when interacting.
"""
from __future__ import annotations

from itertools import combinations

import numpy as np
import scipy.sparse as sp
from scipy.sparse.linalg import expm_multiply


def sector(L: int, n: int) -> np.ndarray:
    """Bit patterns of L orbitals with n particles, ascending."""
    return np.array(sorted(sum(1 << i for i in c) for c in combinations(range(L), n)), dtype=np.int64)


def hopping_matrix(L: int, states: np.ndarray, tM) -> sp.csr_matrix:
    """sum_ij tM_ij c^dagger_i c_j on one spin species (Jordan-Wigner signs)."""
    index = {int(s): k for k, s in enumerate(states)}
    rows, cols, vals = [], [], []
    tM = np.asarray(tM, dtype=np.float64)
    for k, s in enumerate(states):
        s = int(s)
        for i in range(L):
            for j in range(L):
                t = tM[i, j]
                if t == 0.0:
                    continue
                if i == j:
                    if s >> i & 1:
                        rows.append(k), cols.append(k), vals.append(t)
                    continue
                if not (s >> j & 1) or (s >> i & 1):
                    continue
                lo, hi = min(i, j), max(i, j)
                between = bin(s & (((1 << hi) - 1) & ~((1 << (lo + 1)) - 1))).count("1")
                s2 = (s & ~(1 << j)) | (1 << i)
                rows.append(index[s2]), cols.append(k), vals.append(t * (-1) ** between)
    n = len(states)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def hubbard(L: int, tM, U_diag, n_up: int, n_dn: int):
    """Sparse H on the (n_up, n_dn) sector and the sector bases (index = iu * len(dn) + id)."""
    up, dn = sector(L, n_up), sector(L, n_dn)
    Hu, Hd = hopping_matrix(L, up, tM), hopping_matrix(L, dn, tM)
    Iu, Id = sp.identity(len(up), format="csr"), sp.identity(len(dn), format="csr")
    occ_u = ((up[:, None] >> np.arange(L)) & 1).astype(np.float64)  # [n_up states][L]
    occ_d = ((dn[:, None] >> np.arange(L)) & 1).astype(np.float64)
    inter = (occ_u * np.asarray(U_diag, dtype=np.float64)) @ occ_d.T  # [iu][id] sum_i U_i n_iu n_id
    H = sp.kron(Hu, Id, format="csr") + sp.kron(Iu, Hd, format="csr") + sp.diags(inter.ravel())
    return H, up, dn, occ_u, occ_d


def evolve_product_state(L: int, tM, U_diag, occ_up, occ_dn, t_max: float, n_times: int) -> dict:
    """<n_iu>(t), <n_id>(t), <n_iu n_id>(t) on t = linspace(0, t_max, n_times) from the product state
    with the given 0/1 occupations."""
    occ_up = np.asarray(occ_up, dtype=int)
    occ_dn = np.asarray(occ_dn, dtype=int)
    H, up, dn, ou, od = hubbard(L, tM, U_diag, int(occ_up.sum()), int(occ_dn.sum()))
    su = sum(1 << i for i in range(L) if occ_up[i])
    sd = sum(1 << i for i in range(L) if occ_dn[i])
    psi0 = np.zeros(H.shape[0], dtype=np.complex128)
    psi0[np.searchsorted(up, su) * len(dn) + np.searchsorted(dn, sd)] = 1.0
    psis = expm_multiply(-1j * H, psi0, start=0.0, stop=t_max, num=n_times, endpoint=True)
    out = {"t": np.linspace(0.0, t_max, n_times), "n_up": [], "n_dn": [], "nund": []}
    for psi in psis:
        p = (np.abs(psi) ** 2).reshape(len(up), len(dn))
        pu, pd = p.sum(axis=1), p.sum(axis=0)
        out["n_up"].append(pu @ ou)
        out["n_dn"].append(pd @ od)
        out["nund"].append(np.einsum("ab,ai,bi->i", p, ou, od))
    for k in ("n_up", "n_dn", "nund"):
        out[k] = np.array(out[k]).T  # [L][n_times]
    return out
