"""Environment kernels and rank-K H_eff terms through the C ABI (GPU).

su2_env_transform runs one of the reference's environment kernels
(operator_transformation, overlap_transformation, operator_pairup,
operator_initial; tensor_su2_dmrg_src.cpp:27-890) on one block operator,
then projects it on a truncated basis; compared with the oracle's
enlarge/project at 1e-10.  su2_heff_create with rank-0/2/4 terms (the J-Q
plaquette parts, dmrg_su2.cpp:635-655) is compared with the oracle's H_eff.
"""
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import heff as oheff
from oracle.su2 import Bond

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from su2dmrg_amd import lib as L
    c = L.Context(0)
    yield c
    c.close()


def _rel(got, ref):
    num = sum(float(np.sum((got[k] - ref[k]) ** 2)) for k in ref)
    den = sum(float(np.sum(ref[k] ** 2)) for k in ref)
    extra = sum(float(np.sum(got[k] ** 2)) for k in got if k not in ref)
    return np.sqrt((num + extra) / max(den, 1e-300))


def _blocks(t, nleg):
    """device tensor -> {labels: ndarray} (2 legs: (a, b); 3 legs: (a, c))"""
    flat = t.download()
    out = {}
    for _, tj, dm, off in t.blocks:
        if nleg == 2:
            out[(tj[0], tj[1])] = flat[off:off + dm[0] * dm[1]].reshape((dm[0], dm[1]),
                                                                        order="F")
        else:
            out[(tj[0], tj[2])] = flat[off:off + dm[0] * dm[2]].reshape((dm[0], dm[2]),
                                                                        order="F")
    return out


def _scalar(ctx, L, twoj, dims, h):
    return L.scalar_operator(ctx, twoj, dims, h)


def _projector(ctx, L, E, rng, keep_frac=0.6):
    """random orthonormal-column projector per enlarged sector, as a device
    tensor [enlarged (dir +1), bond (dir -1)] and the oracle's {jf: matrix}"""
    bt, bd, P = [], [], {}
    for jf, d in zip(E.bond.twoj, E.bond.dims):
        k = max(1, int(keep_frac * d))
        q, _ = np.linalg.qr(rng.standard_normal((d, k)))
        P[jf] = q
        bt.append(jf)
        bd.append(k)
    t = L.Tensor(ctx, [(E.bond.twoj, E.bond.dims, 1), (bt, bd, -1)], locspin=0)
    flat = np.zeros(t.size)
    for _, tj, dm, off in t.blocks:
        flat[off:off + dm[0] * dm[1]] = P[tj[0]].ravel(order="F")
    t.upload(flat)
    return t, P


@pytest.mark.parametrize("left", [True, False])
@pytest.mark.parametrize("twoj,dims", [((0, 2, 4), (5, 7, 3)), ((1, 3, 5, 7), (6, 9, 4, 2))])
def test_env_transform_matches_oracle(ctx, left, twoj, dims):
    from su2dmrg_amd import lib as L
    from su2dmrg_amd import synthetic
    _say(f"env transform left={left} twoj={twoj} dims={dims}")
    rng = np.random.default_rng(3 + len(twoj) + int(left))
    block = Bond(list(twoj), list(dims))
    E = od.Enlarged(block, left)
    H = {j: (lambda m: m + m.T)(rng.standard_normal((d, d))) for j, d in zip(twoj, dims)}
    S = {(a, b): rng.standard_normal((dict(zip(twoj, dims))[a], dict(zip(twoj, dims))[b]))
         for (a, b) in synthetic.transitions(list(twoj))}
    dH = _scalar(ctx, L, list(twoj), list(dims), H)
    dS = L.vector_operator(ctx, list(twoj), list(dims), S)
    dP, P = _projector(ctx, L, E, rng)
    J = 0.7
    cases = {
        "overlap": (dH, od.enlarge(E, H, {}, [], left)[0], 2),
        "pairup": (dS, od.enlarge(E, None, {}, [(J, S)], left)[0], 2),
        "op_transform": (dS, od.enlarge(E, None, {0: S}, [], left)[1][0], 3),
        "initial": (None, od.enlarge(E, None, {}, [], left)[2], 3),
    }
    for kind, (op, ref_big, nleg) in cases.items():
        if nleg == 2:
            ref_big = {(j, j): b for j, b in ref_big.items()}
        # on the enlarged basis
        t = L.env_transform(ctx, kind, left, list(twoj), list(dims), None, op, J)
        got = _blocks(t, nleg)
        err = _rel(got, ref_big)
        assert err < RTOL, (kind, "enlarged", err)
        # projected
        if nleg == 2:
            ref = {(j, j): P[j].T @ b @ P[j] for (j, _), b in ref_big.items() if j in P}
        else:
            ref = {(a, b): P[a].T @ m @ P[b] for (a, b), m in ref_big.items()
                   if a in P and b in P}
        t = L.env_transform(ctx, kind, left, list(twoj), list(dims), dP, op, J)
        got = _blocks(t, nleg)
        err = _rel(got, ref)
        _say(f"  {kind}: projected rel err {err:.2e}")
        assert err < RTOL, (kind, "projected", err)


def test_env_transform_rejects_mismatched_legs(ctx):
    from su2dmrg_amd import lib as L
    H = {0: np.eye(3), 2: np.eye(4)}
    dH = L.scalar_operator(ctx, [0, 2], [3, 4], H)
    with pytest.raises(L.Su2Error):
        L.env_transform(ctx, "overlap", True, [0, 2], [3, 5], None, dH)
    with pytest.raises(L.Su2Error):
        L.env_transform(ctx, "op_transform", True, [0, 2], [3, 4], None, dH)


def _rank_transitions(twoj, K):
    s = set(twoj)
    return [(a, b) for a in twoj for b in twoj
            if b in s and abs(a - b) <= K <= a + b and (a + b + K) % 2 == 0 and (a % 2) == (b % 2)]


def test_heff_rank_k_terms_match_oracle(ctx):
    from su2dmrg_amd import lib as L
    from su2dmrg_amd import synthetic
    _say("heff with rank 0, 2 and 4 terms")
    twoj = [1, 3, 5, 7]
    dims = [40, 55, 30, 6]
    work = synthetic.make_workload(twoj, dims, dims, nterm=0, seed=12)
    rng = np.random.default_rng(12)
    dd = dict(zip(twoj, dims))
    terms, dev_terms, ranks, keep = [], [], [], []
    for K in (0, 2, 4, 2):
        tr = _rank_transitions(twoj, K)
        Lo = {k: rng.standard_normal((dd[k[0]], dd[k[1]])) for k in tr}
        Ro = {k: rng.standard_normal((dd[k[0]], dd[k[1]])) for k in tr}
        terms.append((Lo, Ro, K))
        tl = L.vector_operator(ctx, twoj, dims, Lo, rank=K)
        tr_ = L.vector_operator(ctx, twoj, dims, Ro, rank=K)
        keep += [tl, tr_]
        dev_terms.append(([(1.0, tl)], [(1.0, tr_)]))
        ranks.append(K)
    psi = L.wavefunction(ctx, twoj, dims, dims, work["psi"])
    out = L.wavefunction(ctx, twoj, dims, dims)
    hl = L.scalar_operator(ctx, twoj, dims, work["hl"])
    hr = L.scalar_operator(ctx, twoj, dims, work["hr"])
    plan = L.HeffPlan(ctx, psi, hl, hr, dev_terms, ranks=ranks)
    plan.apply(psi, out)
    got = L.unpack_wavefunction(out, out.download())
    ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], terms)
    num = sum(float(np.sum((got[j] - ref[j]) ** 2)) for j in ref)
    den = sum(float(np.sum(ref[j] ** 2)) for j in ref)
    assert np.sqrt(num / den) < RTOL
