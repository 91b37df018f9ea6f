"""Plaquette (Q) terms of the J-Q-J2 model in the DMRG environments (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

The reference adds, for every plaquette m of sites p0 < p1 < p2 < p3 (sorted
snake positions, dmrg_su2.cpp:200-232), the term
    -Q [ (S_p0.S_p1)(S_p2.S_p3) + (S_p0.S_p3)(S_p1.S_p2) ]
(the singlet projectors' bilinear parts are folded into coup[1] and coup[3],
dmrg_su2.cpp:258-260, the constant dropped).  The two channels are grown site
by site as partial operators plqopr[i][m + ns*ch] (dmrg_su2_omp.cpp:411-447
from the left, 553-590 from the right) and enter H_eff on adjacent sites
(dmrg_su2.cpp:635-655).

Operators here are dicts ``K -> {(j_in, j_out): block}`` with the operator leg
2K in {0, 2, 4}: rank 0 is the reference's scalar (two-leg) operator, since
its recoupling factors equal those of a two-leg one (fac_hamilt_vec(j,0,j) =
1, the rank-0 transformation factor is 1).  Channel 0 is grown with the
reference's own kernels (pair-up, scalar times site spin, transformations);
channel 1 couples S_p0 (x) s_p1 to the ranks K = 0, 1, 2 at p1 and reduces
them with s_p2 to a vector at p2 (``recoupling.couple_*``/``reduce_*``).

Pinned by exact diagonalisation of J-Q-J2 cylinders (tests/test_oracle_plaquette.py).
"""
from __future__ import annotations

import numpy as np

from . import recoupling as rc

PHYS = 1
RANKS = (0, 2, 4)


def _tri(a, b, c):
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


def _zeros(fd, K):
    return {(a, b): np.zeros((fd[a], fd[b])) for a in fd for b in fd if _tri(a, K, b)}


def _prune(op):
    return {k: v for k, v in op.items() if np.any(v != 0)} or op


def transform(E, op, K, left):
    """O (x) 1 for a rank-K block operator (operator_transformation /
    overlap_transformation, tensor_su2_dmrg_src.cpp:221-402, 531-688)."""
    fd = E.dims
    out = _zeros(fd, K)
    for (jf, jfp), blk in out.items():
        for jb, o, d in E.parts(jf):
            for jbp, op_, dp in E.parts(jfp):
                src = op.get((jb, jbp))
                if src is None:
                    continue
                f = (rc.op_transform_left(jb, jf, jbp, jfp, K, PHYS) if left
                     else rc.op_transform_right(jb, jf, jbp, jfp, K, PHYS))
                if abs(f) < 1e-8:
                    continue
                blk[o:o + d, op_:op_ + dp] += f * src
    return out


def pair(E, vec, sig, left):
    """Block vector operator . site spin -> rank 0 (operator_pairup,
    tensor_su2_dmrg_src.cpp:691-890)."""
    fd = E.dims
    out = {(j, j): np.zeros((d, d)) for j, d in fd.items()}
    for jf in fd:
        blk = out[(jf, jf)]
        for jb, o, d in E.parts(jf):
            for jbp, op_, dp in E.parts(jf):
                src = vec.get((jb, jbp))
                if src is None:
                    continue
                f = (rc.pairup_left(jf, jb, jbp, 2, PHYS, PHYS) if left
                     else rc.pairup_right(jf, jb, jbp, 2, PHYS, PHYS))
                if abs(f) < 1e-8:
                    continue
                blk[o:o + d, op_:op_ + dp] += f * sig * src
    return out


def scalar_times_site(E, sca, sig, left):
    """Block scalar operator (x) site spin -> rank 1 (operator_initial with a
    block operator, tensor_su2_dmrg_src.cpp:27-218)."""
    fd = E.dims
    out = _zeros(fd, 2)
    for (jf, jfp), blk in out.items():
        for jb, o, d in E.parts(jf):
            for jbp, op_, dp in E.parts(jfp):
                if jbp != jb:
                    continue
                src = sca.get((jb, jb))
                if src is None:
                    continue
                f = (rc.op_initial_left(jb, jf, jfp, 2, PHYS, PHYS) if left
                     else rc.op_initial_right(jb, jf, jfp, 2, PHYS, PHYS))
                if abs(f) < 1e-8:
                    continue
                blk[o:o + d, op_:op_ + dp] += f * sig * src
    return out


def couple(E, vec, sig, K, left):
    """Block vector operator (x) site spin coupled to rank K."""
    fd = E.dims
    out = _zeros(fd, K)
    for (jf, jfp), blk in out.items():
        for jb, o, d in E.parts(jf):
            for jbp, op_, dp in E.parts(jfp):
                src = vec.get((jb, jbp))
                if src is None:
                    continue
                f = (rc.couple_left(jb, jf, jbp, jfp, K) if left
                     else rc.couple_right(jb, jf, jbp, jfp, K))
                if abs(f) < 1e-12:
                    continue
                blk[o:o + d, op_:op_ + dp] += f * sig * src
    return out


def reduce(E, ranks, sig, left):
    """sum_K z_K [X^K (x) site spin]^1 -> rank 1."""
    fd = E.dims
    out = _zeros(fd, 2)
    for K, op in ranks.items():
        z = rc.REDUCE_WEIGHT[K]
        for (jf, jfp), blk in out.items():
            for jb, o, d in E.parts(jf):
                for jbp, op_, dp in E.parts(jfp):
                    src = op.get((jb, jbp))
                    if src is None:
                        continue
                    f = (rc.reduce_left(jb, jf, jbp, jfp, K) if left
                         else rc.reduce_right(jb, jf, jbp, jfp, K))
                    if abs(f) < 1e-12:
                        continue
                    blk[o:o + d, op_:op_ + dp] += z * f * sig * src
    return out


def grow(lat, E, i, Pprev, Sprev, left, sig_site, sig_pair):
    """Partial plaquette operators after adding site i to a block.

    Pprev: {(m, ch): {K: op}} of the block before site i; Sprev: the block's
    spin operators.  Returns (P, H_add) with H_add the completed plaquettes'
    -Q (...) as a rank-0 operator dict (or None).
    """
    P = {}
    hadd = None
    Q = lat.qcoup
    for m, pos in enumerate(lat.plq):
        p0, p1, p2, p3 = pos
        if not (p0 <= i <= p3):
            continue
        # from the left the chain runs p0 -> p3, from the right p3 -> p0
        a, b, c, d = (p0, p1, p2, p3) if left else (p3, p2, p1, p0)
        beyond_b = (i > b) if left else (i < b)
        if i != b and not beyond_b:
            continue
        for ch in (0, 1):
            key = (m, ch)
            if i == b:
                if ch == 0:
                    P[key] = {0: pair(E, Sprev[a], sig_pair, left)}
                else:
                    P[key] = {K: couple(E, Sprev[a], sig_site, K, left) for K in RANKS}
            elif i == c:
                prev = Pprev[key]
                if ch == 0:
                    P[key] = {2: scalar_times_site(E, prev[0], sig_site, left)}
                else:
                    P[key] = {2: reduce(E, prev, sig_site, left)}
            elif i == d:
                done = pair(E, Pprev[key][2], sig_pair, left)
                if hadd is None:
                    hadd = {k: -Q * v for k, v in done.items()}
                else:
                    for k, v in done.items():
                        hadd[k] = hadd[k] - Q * v
            else:
                P[key] = {K: transform(E, op, K, left) for K, op in Pprev[key].items()}
    return P, hadd


def heff_terms(lat, il, ir, Sl, Sr, Pl, Pr):
    """Plaquette terms of H_eff at adjacent (il, ir) as (left, right, K)
    (dmrg_su2.cpp:635-655): the right side carries -Q."""
    out = []
    if il + 1 != ir:
        return out
    Q = lat.qcoup
    for m, (p0, p1, p2, p3) in enumerate(lat.plq):
        if not (p0 <= il and ir <= p3):
            continue
        for ch in (0, 1):
            key = (m, ch)
            if ir <= p1:
                out.append((Sl[p0], _scale(Pr[key][2], -Q), 2))
            elif il >= p2:
                out.append((Pl[key][2], _scale(Sr[p3], -Q), 2))
            else:
                for K, op in Pl[key].items():
                    out.append((op, _scale(Pr[key][K], -Q), K))
    return out


def _scale(op, c):
    return {k: c * v for k, v in op.items()}


def project(P, ops):
    """U^T O U for every rank of every partial operator."""
    return {key: {K: {(a, b): P[a].T @ blk @ P[b] for (a, b), blk in op.items()
                      if a in P and b in P}
                  for K, op in ranks.items()}
            for key, ranks in ops.items()}
