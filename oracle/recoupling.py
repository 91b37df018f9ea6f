"""Recoupling scalars of the environment kernels (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

The reference computes each table numerically: it contracts dense CG tensors
in the leg pattern of the matching ``tensor::contract_dmrg_*`` routine and
takes the ratio to the output CG tensor with ``is_proportional_to``
(dmrg_su2.cpp:777-943).  These functions restate each contraction as an
einsum whose index letters follow the leg bookkeeping of those routines
(derived in the docstrings), then take the same ratio.

Labels: l/l' = bond of the block, s/s' = site spin, a/a' = fused (enlarged)
leg, r/r' = right bond, k = operator leg (2j).  Pinned by the exact
diagonalisation of small lattices in tests/test_oracle_dmrg.py: every table
enters the environments, so a wrong factor moves the DMRG energy.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

from .su2 import cgc, identity, is_proportional_to, singlet


def _ratio(t, ref):
    r = is_proportional_to(t, ref)
    if r is None:
        raise ValueError("recoupling: contraction is not proportional to the output CG")
    return r


def _tri(a, b, c):
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


@lru_cache(maxsize=None)
def op_transform_left(jl, ja, jlp, jap, k, s):
    """fac_operator_transformation_left[ja][k][jap][jl][jlp][s] (dmrg_su2.cpp:862-869).

    contract_dmrg_operator_transformation(flag=0) (tensor_dmrg_src.cpp:318-325):
    out[a,q,a'] = sum C(jl,s,ja)[l,x,a] C(jl,k,jlp)[l,q,m] C(jlp,s,jap)[m,x,b].
    """
    if not (_tri(jl, s, ja) and _tri(jlp, s, jap) and _tri(jl, k, jlp) and _tri(ja, k, jap)):
        return 0.0
    t = np.einsum("lxa,lqm,mxb->aqb", cgc(jl, s, ja), cgc(jl, k, jlp), cgc(jlp, s, jap))
    return _ratio(t, cgc(ja, k, jap))


@lru_cache(maxsize=None)
def op_transform_right(jr, ja, jrp, jap, k, s):
    """fac_operator_transformation_rght[ja][k][jap][jr][jrp][s] (dmrg_su2.cpp:871-876).

    flag=1 branch (tensor_dmrg_src.cpp:326-333): the uu CG is C(s,jr,ja),
    out[a,q,a'] = sum C(s,jr,ja)[x,r,a] C(jr,k,jrp)[r,q,m] C(s,jrp,jap)[x,m,b].
    """
    if not (_tri(s, jr, ja) and _tri(s, jrp, jap) and _tri(jr, k, jrp) and _tri(ja, k, jap)):
        return 0.0
    t = np.einsum("xra,rqm,xmb->aqb", cgc(s, jr, ja), cgc(jr, k, jrp), cgc(s, jrp, jap))
    return _ratio(t, cgc(ja, k, jap))


@lru_cache(maxsize=None)
def op_initial_left(jl, ja, jap, k, s0, s1):
    """fac_operator_onsite_left[ja][k][jap][jl][s0+s1*3] (dmrg_su2.cpp:817-822).

    contract_dmrg_operator_initial(flag=0) (tensor_dmrg_src.cpp:85-92):
    out[a,q,a'] = sum C(jl,s0,ja)[l,x,a] C(s0,k,s1)[x,q,y] C(jl,s1,jap)[l,y,b].
    """
    if not (_tri(jl, s0, ja) and _tri(jl, s1, jap) and _tri(s0, k, s1) and _tri(ja, k, jap)):
        return 0.0
    t = np.einsum("lxa,xqy,lyb->aqb", cgc(jl, s0, ja), cgc(s0, k, s1), cgc(jl, s1, jap))
    return _ratio(t, cgc(ja, k, jap))


@lru_cache(maxsize=None)
def op_initial_right(jr, ja, jap, k, s0, s1):
    """fac_operator_onsite_rght[ja][k][jap][jr][s0+s1*3] (dmrg_su2.cpp:825-830).

    flag=1 branch (tensor_dmrg_src.cpp:93-100): contract over the site leg of
    the first CG with the operator, cyclic shift, then over (s1, r):
    out[a,q,a'] = sum C(s0,jr,ja)[x,r,a] C(s0,k,s1)[x,q,y] C(s1,jr,jap)[y,r,b].
    """
    if not (_tri(s0, jr, ja) and _tri(s1, jr, jap) and _tri(s0, k, s1) and _tri(ja, k, jap)):
        return 0.0
    t = np.einsum("xra,xqy,yrb->aqb", cgc(s0, jr, ja), cgc(s0, k, s1), cgc(s1, jr, jap))
    return _ratio(t, cgc(ja, k, jap))


@lru_cache(maxsize=None)
def _turned_site_op(s0, k, s1):
    """The site operator leg turned by a singlet, times sqrt(k+1).

    dmrg_su2.cpp:924-926: contract(C(s0,k,s1), 1, singlet(k), 1) gives legs
    [s1, s0, q'], shift(0,2) reorders them to [s0, q', s1].
    """
    t = np.einsum("xqy,pq->xpy", cgc(s0, k, s1), singlet(k))
    return t * math.sqrt(k + 1)


@lru_cache(maxsize=None)
def pairup_left(ja, jl, jlp, k, s0, s1):
    """fac_operator_pairup_left[ja][jl][jlp][k][s0+s1*3] (dmrg_su2.cpp:927-931).

    contract_dmrg_operator_pairup(flag=0) (tensor_dmrg_src.cpp:180-192):
    out[a,a'] = sum C(jl,s0,ja)[l,x,a] C(jl,k,jlp)[l,q,m] T[x,q,y] C(jlp,s1,ja)[m,y,b]
    with T the turned site operator; ratio against identity(ja).
    """
    if not (_tri(jl, s0, ja) and _tri(jlp, s1, ja) and _tri(jl, k, jlp) and _tri(s0, k, s1)):
        return 0.0
    t = np.einsum("lxa,lqm,xqy,myb->ab", cgc(jl, s0, ja), cgc(jl, k, jlp),
                  _turned_site_op(s0, k, s1), cgc(jlp, s1, ja))
    return _ratio(t, identity(ja))


@lru_cache(maxsize=None)
def pairup_right(ja, jr, jrp, k, s0, s1):
    """fac_operator_pairup_rght[ja][jr][jrp][k][s0+s1*3] (dmrg_su2.cpp:932-936).

    flag=1 branch (tensor_dmrg_src.cpp:193-206):
    out[a,a'] = sum C(s0,jr,ja)[x,r,a] C(jr,k,jrp)[r,q,m] T[x,q,y] C(s1,jrp,ja)[y,m,b].
    """
    if not (_tri(s0, jr, ja) and _tri(s1, jrp, ja) and _tri(jr, k, jrp) and _tri(s0, k, s1)):
        return 0.0
    t = np.einsum("xra,rqm,xqy,ymb->ab", cgc(s0, jr, ja), cgc(jr, k, jrp),
                  _turned_site_op(s0, k, s1), cgc(s1, jrp, ja))
    return _ratio(t, identity(ja))


@lru_cache(maxsize=None)
def left_to_right_psi(jl, s, jr):
    """Factor that re-couples a singlet wavefunction across one site.

    psi[(l),(s r)F=l] with the right leg built as C(s,jr,F) becomes
    psi'[(l s)E=r,(r)] with the left leg built as C(jl,s,E): the dense
    tensors are equal when psi'_red = f * psi_red (both sides carry the unit
    singlet between the two legs).  Not a reference table: the reference
    re-labels the MPS tensor with a fresh make_cgc and no factor when it forms
    the Lanczos start vector (tensor_su2.cpp:749-836, then overlap_initial and
    contract, dmrg_su2_omp.cpp:229-254).  For a spin-1/2 site (s = 1) this
    factor is exactly 1 for every (jl, jr), so that re-labelling is exact and
    the start vectors agree.
    """
    src = np.einsum("lf,xrf->lxr", singlet(jl), cgc(s, jr, jl))
    dst = np.einsum("lxe,er->lxr", cgc(jl, s, jr), singlet(jr))
    return float(np.sum(src * dst) / np.sum(dst * dst))


@lru_cache(maxsize=None)
def right_to_left_psi(jl, s, jr):
    """Inverse direction of ``left_to_right_psi``: psi[(l s)E=r,(r)] ->
    psi'[(l),(s r)F=l]."""
    src = np.einsum("lxe,er->lxr", cgc(jl, s, jr), singlet(jr))
    dst = np.einsum("lf,xrf->lxr", singlet(jl), cgc(s, jr, jl))
    return float(np.sum(src * dst) / np.sum(dst * dst))


# --------------------------------------------------------------------------
# plaquette (Q) terms: two spin-1 operators coupled to rank K, and back
# --------------------------------------------------------------------------
# The reference carries the nested plaquette product (S_p0.S_p3)(S_p1.S_p2)
# between p1 and p2 as one operator whose operator leg holds the ranks 0, 1
# and 2 of S_p0 (x) s_p1 (permutation with the qterm tensors,
# tensor_su2_hamiltonian.cpp:19-76, 871-1108; dmrg_su2_omp.cpp:411-447).
# Here each rank K (2K = 0, 2, 4) is an operator of its own:
#   X^K_Q = sum_qp C(2,2,K)[q,p,Q] A_q s_p          (couple)
#   Z_q   = sum_K z_K sum_Qp C(K,2,2)[Q,p,q] X^K_Q s'_p   (reduce to rank 1)
# with A_q[out,in] = A[in,q,out] the block operator and s the site spin.
# REDUCE_WEIGHT holds z_K such that Z = A (s.s') in the normalisation of the
# environment operators, and the H_eff contraction of X^K with its right
# counterpart Y^K (fac_hamilt_vec) sums to (A.D)(s.s') with weight 1 per K;
# both are fixed by dense four-spin operator identities in
# tests/test_oracle_plaquette.py.
REDUCE_WEIGHT = {0: -1.0 / math.sqrt(3.0), 2: 1.0, 4: -math.sqrt(5.0 / 3.0)}


@lru_cache(maxsize=None)
def couple_left(jl, ja, jlp, jap, K, s=1):
    """Block vector operator (x) site spin coupled to rank K, left enlarged leg:
    out[a,Q,a'] = sum C(jl,s,ja)[l,x,a] C(jl,2,jlp)[l,q,m] C(s,2,s)[x,p,y]
                      C(2,2,K)[q,p,Q] C(jlp,s,jap)[m,y,b]."""
    if not (_tri(jl, s, ja) and _tri(jlp, s, jap) and _tri(jl, 2, jlp) and _tri(ja, K, jap)):
        return 0.0
    t = np.einsum("lxa,lqm,xpy,qpQ,myb->aQb", cgc(jl, s, ja), cgc(jl, 2, jlp), cgc(s, 2, s),
                  cgc(2, 2, K), cgc(jlp, s, jap))
    return _ratio(t, cgc(ja, K, jap))


@lru_cache(maxsize=None)
def couple_right(jr, ja, jrp, jap, K, s=1):
    """Right enlarged leg (fusion C(s,jr,ja)) version of couple_left."""
    if not (_tri(s, jr, ja) and _tri(s, jrp, jap) and _tri(jr, 2, jrp) and _tri(ja, K, jap)):
        return 0.0
    t = np.einsum("xra,rqm,xpy,qpQ,ymb->aQb", cgc(s, jr, ja), cgc(jr, 2, jrp), cgc(s, 2, s),
                  cgc(2, 2, K), cgc(s, jrp, jap))
    return _ratio(t, cgc(ja, K, jap))


@lru_cache(maxsize=None)
def reduce_left(jl, ja, jlp, jap, K, s=1):
    """Block rank-K operator (x) site spin coupled to rank 1, left enlarged leg:
    out[a,q,a'] = sum C(jl,s,ja)[l,x,a] C(jl,K,jlp)[l,Q,m] C(s,2,s)[x,p,y]
                      C(K,2,2)[Q,p,q] C(jlp,s,jap)[m,y,b]."""
    if not (_tri(jl, s, ja) and _tri(jlp, s, jap) and _tri(jl, K, jlp) and _tri(ja, 2, jap)):
        return 0.0
    t = np.einsum("lxa,lQm,xpy,Qpq,myb->aqb", cgc(jl, s, ja), cgc(jl, K, jlp), cgc(s, 2, s),
                  cgc(K, 2, 2), cgc(jlp, s, jap))
    return _ratio(t, cgc(ja, 2, jap))


@lru_cache(maxsize=None)
def reduce_right(jr, ja, jrp, jap, K, s=1):
    """Right enlarged leg version of reduce_left."""
    if not (_tri(s, jr, ja) and _tri(s, jrp, jap) and _tri(jr, K, jrp) and _tri(ja, 2, jap)):
        return 0.0
    t = np.einsum("xra,rQm,xpy,Qpq,ymb->aqb", cgc(s, jr, ja), cgc(jr, K, jrp), cgc(s, 2, s),
                  cgc(K, 2, 2), cgc(s, jrp, jap))
    return _ratio(t, cgc(ja, 2, jap))
