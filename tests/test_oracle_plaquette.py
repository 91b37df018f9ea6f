"""Plaquette (Q) terms of the oracle: the rank-K recoupling weights against
dense four-spin operator identities, the rank-K H_eff term against a dense
contraction, and the whole J-Q-J2 DMRG against exact diagonalisation of the
physical model built from coordinates (CPU only)."""
import itertools
import math

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import ed, heff
from oracle import recoupling as rc
from oracle.su2 import cgc, fac_hamilt_vec, singlet

# --------------------------------------------------------------------------
# dense four-spin operators in the library's spherical convention:
# M_q[out, in] = sqrt(3)/2 C(1,2,1)[in, q, out]
# --------------------------------------------------------------------------
_M = [math.sqrt(3) / 2 * cgc(1, 2, 1)[:, q, :].T for q in range(3)]


def _site(op, k, n=4):
    out = np.eye(1)
    for i in range(n):
        out = np.kron(out, op if i == k else np.eye(2))
    return out


_S = {k: [_site(_M[q], k) for q in range(3)] for k in range(4)}


def _dot(A, B):
    # the physical S.S in this convention (pinned by its spectrum below)
    s2 = singlet(2)
    return -math.sqrt(3) * sum(s2[q, p] * A[q] @ B[p] for q in range(3) for p in range(3))


def _couple(A, B, K):
    C = cgc(2, 2, K)
    return [sum(C[q, p, Q] * A[q] @ B[p] for q in range(3) for p in range(3))
            for Q in range(K + 1)]


def _reduce(X, K, B):
    C = cgc(K, 2, 2)
    return [sum(C[Q, p, q] * X[Q] @ B[p] for Q in range(K + 1) for p in range(3))
            for q in range(3)]


def _heff_dense(L, R, K):
    # operator legs joined by sqrt(2K+1) singlet(K), as fac_hamilt_vec does
    sK = singlet(K)
    return math.sqrt(K + 1) * sum(sK[Q, P] * L[Q] @ R[P] for Q in range(K + 1)
                                  for P in range(K + 1))


def test_dot_is_the_heisenberg_coupling():
    ev = np.linalg.eigvalsh(_dot(_S[0], _S[1]))
    np.testing.assert_allclose(ev, [-0.75] * 4 + [0.25] * 12, atol=1e-14)


def test_reduce_weights_build_a_times_dot():
    # left: block operator +S_a, site spins with sigma[0] = +sqrt(3)/2
    a, b, c = 0, 1, 2
    dbc = _dot(_S[b], _S[c])
    Z = [np.zeros((16, 16)) for _ in range(3)]
    for K in (0, 2, 4):
        X = _couple(_S[a], _S[b], K)
        R = _reduce(X, K, _S[c])
        for q in range(3):
            Z[q] += rc.REDUCE_WEIGHT[K] * R[q]
    for q in range(3):
        np.testing.assert_allclose(Z[q], _S[a][q] @ dbc, atol=1e-14)
    # right: environment operators and site spins carry sigma[1] = -sqrt(3)/2
    d = 3
    neg = lambda ops: [-m for m in ops]
    Z = [np.zeros((16, 16)) for _ in range(3)]
    for K in (0, 2, 4):
        Y = _couple(neg(_S[d]), neg(_S[c]), K)
        R = _reduce(Y, K, neg(_S[b]))
        for q in range(3):
            Z[q] += rc.REDUCE_WEIGHT[K] * R[q]
    for q in range(3):
        np.testing.assert_allclose(Z[q], -(_S[d][q] @ dbc), atol=1e-14)


def test_rank_channels_sum_to_the_nested_product():
    a, b, c, d = 0, 1, 2, 3
    target = _dot(_S[a], _S[d]) @ _dot(_S[b], _S[c])
    neg = lambda ops: [-m for m in ops]
    tot = np.zeros((16, 16))
    for K in (0, 2, 4):
        X = _couple(_S[a], _S[b], K)
        Y = _couple(neg(_S[d]), neg(_S[c]), K)
        tot += _heff_dense(X, Y, K)
    np.testing.assert_allclose(tot, target, atol=1e-14)
    # no single channel does it
    for K in (0, 2, 4):
        X = _couple(_S[a], _S[b], K)
        Y = _couple(neg(_S[d]), neg(_S[c]), K)
        assert np.abs(_heff_dense(X, Y, K) - target).max() > 1e-3


@pytest.mark.parametrize("K", [0, 2, 4])
def test_rank_k_term_matches_dense_contraction(K):
    # reduced term formula with fac_hamilt_vec(j, 2K, j') against the dense
    # contraction of the reduced blocks (x) CG tensors, joined by sqrt(2K+1)
    # singlet(2K) on the operator legs and projected back on the singlet
    rng = np.random.default_rng(7 + K)
    twoj = [0, 2, 4]
    da = {0: 2, 2: 3, 4: 1}
    db = {0: 1, 2: 2, 4: 2}
    psi = {j: rng.standard_normal((da[j], db[j])) for j in twoj}
    L, R = {}, {}
    for j, jp in itertools.product(twoj, twoj):
        if abs(j - jp) <= K <= j + jp:
            L[(j, jp)] = rng.standard_normal((da[j], da[jp]))
            R[(j, jp)] = rng.standard_normal((db[j], db[jp]))
    red = heff.term_apply(psi, L, R, K)
    sK = singlet(K) * math.sqrt(K + 1)
    for jp in twoj:
        dense = np.zeros((da[jp], jp + 1, db[jp], jp + 1))
        for j in twoj:
            if (j, jp) not in L:
                continue
            # psi dense: [a, m, b, n] = psi[a,b] singlet(j)[m,n]
            P = np.einsum("ab,mn->ambn", psi[j], singlet(j))
            Ld = np.einsum("xy,mqu->xmqyu", L[(j, jp)], cgc(j, K, jp))
            Rd = np.einsum("xy,nru->xnryu", R[(j, jp)], cgc(j, K, jp))
            dense += np.einsum("ambn,amqxu,bnryv,qr->xuyv", P, Ld, Rd, sK)
        got = np.einsum("xuyv,uv->xy", dense, singlet(jp)) / np.sum(singlet(jp) ** 2)
        want = red.get(jp, np.zeros((da[jp], db[jp])))
        np.testing.assert_allclose(got, want, atol=1e-12)


def test_plaquette_table():
    lat = od.Lattice(4, 4, jcoup=1.0, qcoup=0.5, bdry=2)
    assert len(lat.plq) == 3 * 4
    for p0, p1, p2, p3 in lat.plq:
        assert p0 < p1 < p2 < p3
        # both channels pair nearest neighbours of the cylinder
        for i, j in ((p0, p1), (p2, p3), (p0, p3), (p1, p2)):
            assert lat.hmap[i, j] in (1, 3)
    assert od.Lattice(4, 4, jcoup=1.0, qcoup=0.0).plq == []


# (lx, ly, jcoup, qcoup, 2S): J-Q (J2 = 1 - |Q| = 0), J-Q-J2, negative Q, S > 0
JQ_CASES = [(3, 4, 1.0, 0.5, 0), (4, 4, 1.0, 1.0, 0), (4, 3, 0.7, 0.3, 2), (3, 4, 1.2, -0.4, 4)]


@pytest.mark.parametrize("lx,ly,jc,q,twoS", JQ_CASES)
def test_jq_dmrg_matches_ed(lx, ly, jc, q, twoS):
    # untruncated DMRG of the reference's (jcoup, qcoup) Hamiltonian equals the
    # physical J1 = jcoup, J2 = 1 - |Q|, Q model plus Q/8 per plaquette
    lat = od.Lattice(lx, ly, jcoup=jc, qcoup=q, bdry=2)
    e_ed = ed.jq_energies(lx, ly, jc, 1 - abs(q), q, twoS)[0] + q * len(lat.plq) / 8
    o = od.DMRG(lat, twoS, 200)
    o.run(nsweep=6, stop=None)
    assert abs(o.energies[-1] - e_ed) < 1e-10 * abs(e_ed)
    # every sweep step is variational
    nstep = 2 * (lat.ns - 6)
    assert min(o.energies[-nstep:]) > e_ed - 1e-10 * abs(e_ed)


def test_jq_truncated_dmrg_is_variational():
    lat = od.Lattice(4, 4, jcoup=1.0, qcoup=0.8, bdry=2)
    e_ed = ed.jq_energies(4, 4, 1.0, 0.2, 0.8, 0)[0] + 0.8 * len(lat.plq) / 8
    o = od.DMRG(lat, 0, 16)
    o.run(nsweep=3, stop=None)
    e = o.energies[-1]
    assert e > e_ed and e - e_ed < 0.05 * abs(e_ed)
