"""Exact diagonalisation of the same spin-1/2 Hamiltonian (oracle truth anchor).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

H = sum_{i<j} J_ij S_i . S_j with J_ij = coup[hmap[i][j]] from the oracle's
Lattice (dmrg_su2.cpp:97-260), built in the S_z = M sector of the 2^ns
product basis; the lowest level of total spin S is found by diagonalising
H + lambda * (S^2 - S(S+1))^2-free shift: we take the lowest eigenvalues in
S_z = S and keep the first whose <S^2> equals S(S+1).
"""
from __future__ import annotations

import itertools

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def _basis(ns, twoM):
    nup = (ns + twoM) // 2
    states = []
    for ups in itertools.combinations(range(ns), nup):
        s = 0
        for u in ups:
            s |= 1 << u
        states.append(s)
    states.sort()
    return np.array(states, dtype=np.int64)


def _heisenberg(ns, bonds, states):
    index = {int(s): i for i, s in enumerate(states)}
    rows, cols, vals = [], [], []
    diag = np.zeros(len(states))
    for a, s in enumerate(states):
        s = int(s)
        for (i, j, J) in bonds:
            si = (s >> i) & 1
            sj = (s >> j) & 1
            if si == sj:
                diag[a] += 0.25 * J
            else:
                diag[a] -= 0.25 * J
                t = s ^ ((1 << i) | (1 << j))
                rows.append(a)
                cols.append(index[t])
                vals.append(0.5 * J)
    n = len(states)
    H = sp.csr_matrix((vals, (rows, cols)), shape=(n, n)) + sp.diags(diag)
    return H


def _s2(ns, states, vec):
    # <S^2> = sum_ij <S_i . S_j>
    bonds = [(i, j, 2.0) for i in range(ns) for j in range(i + 1, ns)]
    H = _heisenberg(ns, bonds, states)
    return float(vec @ (H @ vec)) + 0.75 * ns


def ground_energy(lat, twoS=0, k=6):
    """Lowest energy with total spin S=twoS/2 on the oracle lattice."""
    ns = lat.ns
    states = _basis(ns, twoS)
    H = _heisenberg(ns, lat.bonds(), states)
    n = H.shape[0]
    if n <= 400:
        w, v = np.linalg.eigh(H.toarray())
    else:
        w, v = spla.eigsh(H, k=min(k, n - 2), which="SA", tol=1e-13)
        order = np.argsort(w)
        w, v = w[order], v[:, order]
    target = 0.25 * twoS * (twoS + 2)
    for e, vec in zip(w, v.T):
        if abs(_s2(ns, states, vec) - target) < 1e-6:
            return float(e)
    raise RuntimeError("no level with the requested total spin among the lowest")


def sector_energies(lat, twoS=0, count=2, k=16):
    """The `count` lowest energies with total spin S = twoS/2 (multiplets of
    that spin only), for excited-state checks.  Degenerate levels of different
    spin are separated by diagonalising S^2 inside each degenerate group."""
    ns = lat.ns
    states = _basis(ns, twoS)
    H = _heisenberg(ns, lat.bonds(), states)
    n = H.shape[0]
    if n <= 600:
        w, v = np.linalg.eigh(H.toarray())
    else:
        w, v = spla.eigsh(H, k=min(k, n - 2), which="SA", tol=1e-13)
        order = np.argsort(w)
        w, v = w[order], v[:, order]
    bonds = [(i, j, 2.0) for i in range(ns) for j in range(i + 1, ns)]
    S2op = _heisenberg(ns, bonds, states)
    target = 0.25 * twoS * (twoS + 2)
    out = []
    i = 0
    while i < len(w) and len(out) < count:
        grp = [i]
        while grp[-1] + 1 < len(w) and abs(w[grp[-1] + 1] - w[i]) < 1e-9 * max(1.0, abs(w[i])):
            grp.append(grp[-1] + 1)
        V = v[:, grp]
        s2 = V.T @ (S2op @ V) + 0.75 * ns * np.eye(len(grp))
        for val in np.linalg.eigvalsh(0.5 * (s2 + s2.T)):
            if abs(val - target) < 1e-6 and len(out) < count:
                out.append(float(w[i]))
        i = grp[-1] + 1
    if len(out) < count:
        raise RuntimeError("not enough levels with the requested total spin among the lowest")
    return out


def jq_model(lx, ly, j1, j2, q):
    """Bonds and plaquettes of the J1-J2-Q model on an lx (open) x ly
    (periodic) cylinder, from coordinates alone (site x*ly + y):
    H = j1 sum_nn S.S + j2 sum_diag S.S
        - q sum_plaquettes (P_ij P_kl + P_il P_jk),  P = 1/4 - S.S,
    with plaquette (i, j, k, l) = (x,y), (x,y+1), (x+1,y+1), (x+1,y) for
    x < lx-1 and every y (ly >= 3 so that no bond or plaquette repeats)."""
    if ly < 3:
        raise ValueError("jq_model needs ly >= 3")

    def s(x, y):
        return x * ly + (y % ly)

    nn, diag, plq = set(), set(), []
    for x in range(lx):
        for y in range(ly):
            nn.add(tuple(sorted((s(x, y), s(x, y + 1)))))
            if x + 1 < lx:
                nn.add(tuple(sorted((s(x, y), s(x + 1, y)))))
                diag.add(tuple(sorted((s(x, y), s(x + 1, y + 1)))))
                diag.add(tuple(sorted((s(x + 1, y), s(x, y + 1)))))
                plq.append((s(x, y), s(x, y + 1), s(x + 1, y + 1), s(x + 1, y)))
    bonds = [(i, j, j1) for i, j in sorted(nn)] + [(i, j, j2) for i, j in sorted(diag)]
    return bonds, plq


def jq_energies(lx, ly, j1, j2, q, twoS=0, count=1, k=8):
    """Lowest `count` energies of total spin S = twoS/2 of jq_model."""
    ns = lx * ly
    states = _basis(ns, twoS)
    bonds, plq = jq_model(lx, ly, j1, j2, q)
    H = _heisenberg(ns, bonds, states)
    n = len(states)
    one = sp.identity(n, format="csr")

    def proj(i, j):
        return 0.25 * one - _heisenberg(ns, [(i, j, 1.0)], states)

    for (i, j, kk, l) in plq:
        H = H - q * (proj(i, j) @ proj(kk, l) + proj(i, l) @ proj(j, kk))
    if n <= 600:
        w, v = np.linalg.eigh(H.toarray())
    else:
        w, v = spla.eigsh(H, k=min(k, n - 2), which="SA", tol=1e-13)
        order = np.argsort(w)
        w, v = w[order], v[:, order]
    target = 0.25 * twoS * (twoS + 2)
    out = []
    for e, vec in zip(w, v.T):
        if abs(_s2(ns, states, vec) - target) < 1e-6:
            out.append(float(e))
            if len(out) == count:
                return out
    raise RuntimeError("no level with the requested total spin among the lowest")
