"""Truncated SVD and the full two-site DMRG on the MI355X versus the oracle
and exact diagonalisation (GPU).

Step-by-step energy parity is asserted on lattices whose DMRG trajectory is
well conditioned: perturbing every Gram matrix of the oracle, and every
wavefunction before its SVD, by 1e-15 relative moves no energy by more than
about 1e-11 (pinned on the CPU in test_oracle_dmrg.py).  On the
periodic cylinders the lattice symmetries make Gram eigenvalues degenerate;
the warm-up's all-ones start vector (tensor_su2.cpp:2367) in that arbitrary
basis, fed to a Lanczos run of at most five vectors, then gives
eigensolver-dependent intermediate energies in any implementation, the
reference's included.  There only the converged energies are compared, with
exact diagonalisation.
"""
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import ed

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from su2dmrg_amd import lib as L
    c = L.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("p1,p2,dcut", [(1, 0, 60), (0, 1, 60), (1, 0, 10000), (0, 1, 25)])
def test_truncated_svd_matches_oracle(ctx, p1, p2, dcut):
    from su2dmrg_amd import lib as L
    _say(f"svd p1={p1} p2={p2} dcut={dcut}")
    rng = np.random.default_rng(4)
    twoj = [1, 3, 5, 7]
    da = [150, 40, 7, 1]
    db = [90, 70, 3, 2]
    # decaying spectra like a DMRG wavefunction
    psi = {}
    for j, a, b in zip(twoj, da, db):
        u, _ = np.linalg.qr(rng.standard_normal((a, a)))
        v, _ = np.linalg.qr(rng.standard_normal((b, b)))
        k = min(a, b)
        s = np.exp(-0.15 * np.arange(k)) * (1 + 0.01 * j)
        psi[j] = (u[:, :k] * s) @ v[:, :k].T
    t = L.wavefunction(ctx, twoj, da, db, psi)
    U, V, kept, disc = L.svd_truncate(ctx, t, p1, p2, dcut)
    Uo, Vo, bond, keptd = od.truncated_svd(psi, p1, p2, dcut)
    ko = np.concatenate([keptd[j] for j in sorted(keptd)])
    # the Gram-matrix route (the reference's) fixes a singular value s to
    # about eps*s_max^2/s: compare absolutely against 1e-11 of the largest
    np.testing.assert_allclose(np.array(kept), ko, rtol=0, atol=1e-11 * ko.max())
    assert sorted(U) == sorted(Uo)
    for j in Uo:
        np.testing.assert_allclose(U[j] @ V[j].T, Uo[j] @ Vo[j].T, atol=1e-10)
        # orthonormal side
        W = V[j] if p1 == 1 else U[j]
        np.testing.assert_allclose(W.T @ W, np.eye(W.shape[1]), atol=1e-11)


@pytest.mark.parametrize("block,inner", [(32, 1), (32, 3), (16, 1)])
@pytest.mark.parametrize("m,n,decades", [(300, 200, 20), (180, 420, 12), (520, 500, 6),
                                          (700, 560, 14), (600, 1100, 8)])
def test_svd_wide_spectrum_against_lapack(ctx, m, n, decades, block, inner):
    # Singular values spread over many decades, as in a converged DMRG
    # wavefunction: every value to 1e-12 of the largest (the Gram route of the
    # reference reaches only eps s_max^2 / s), the reconstruction to the
    # weight of the columns zeroed at s <= 1e-11 s_max, and the left vectors
    # orthogonal to 1e-13 s_max / min(s_i, s_j) where s > 1e-9 s_max.
    from su2dmrg_amd import lib as L
    _say(f"svd wide spectrum {m}x{n}, {decades} decades, block {block}, inner {inner}")
    ctx.set_svd_tuning(block, inner)
    rng = np.random.default_rng(m + n)
    k = min(m, n)
    u, _ = np.linalg.qr(rng.standard_normal((m, k)))
    v, _ = np.linalg.qr(rng.standard_normal((n, k)))
    A = (u * np.logspace(0, -decades, k)) @ v.T
    t = L.wavefunction(ctx, [1], [m], [n], {1: A})
    U, V, kept, disc = L.svd_truncate(ctx, t, 0, 1, 100000)
    ctx.set_svd_tuning(32, 1)
    ref = np.sort(np.linalg.svd(A, compute_uv=False))
    s = np.array(kept)
    assert len(s) == k
    np.testing.assert_allclose(s, ref, rtol=0, atol=1e-12 * ref[-1])
    # p1 = 0, p2 = 1: U orthonormal, V carries s
    cut = s[s <= 1e-11 * s[-1]].sum()
    np.testing.assert_allclose(U[1] @ V[1].T, A, rtol=0, atol=1e-13 * ref[-1] + cut)
    big = s > 1e-9 * s[-1]
    Ub = U[1][:, big]
    sb = s[big]
    bound = 1e-13 * s[-1] / np.minimum.outer(sb, sb) + 1e-13
    assert np.all(np.abs(Ub.T @ Ub - np.eye(big.sum())) <= bound)


@pytest.mark.parametrize("m,n,decades", [(700, 560, 14), (600, 1100, 8), (1300, 1290, 16)])
def test_svd_preconditioned_equals_plain(ctx, m, n, decades):
    # QR preconditioning (here of every block from min(m, n) = 48) changes
    # where the Jacobi iteration starts, not its accuracy: both runs match
    # LAPACK, and each other, to 1e-12 of the largest value
    from su2dmrg_amd import lib as L
    _say(f"svd {m}x{n}, {decades} decades, preconditioned and plain")
    rng = np.random.default_rng(m * n)
    k = min(m, n)
    u, _ = np.linalg.qr(rng.standard_normal((m, k)))
    v, _ = np.linalg.qr(rng.standard_normal((n, k)))
    A = (u * np.logspace(0, -decades, k)) @ v.T
    ref = np.sort(np.linalg.svd(A, compute_uv=False))
    out = {}
    for pre in (True, False):
        ctx.set_svd_precondition(48 if pre else 0)
        t = L.wavefunction(ctx, [1], [m], [n], {1: A})
        out[pre] = L.svd_truncate(ctx, t, 0, 1, 100000)
    ctx.set_svd_precondition(256)   # the context's default
    for pre in (True, False):
        U, V, kept, _ = out[pre]
        s = np.array(kept)
        np.testing.assert_allclose(s, ref, rtol=0, atol=1e-12 * ref[-1])
        cut = s[s <= 1e-11 * s[-1]].sum()
        np.testing.assert_allclose(U[1] @ V[1].T, A, rtol=0, atol=1e-13 * ref[-1] + cut)
    np.testing.assert_allclose(np.array(out[True][2]), np.array(out[False][2]), rtol=0,
                               atol=1e-12 * ref[-1])


@pytest.mark.parametrize("m,n,rank,gap", [(700, 720, 400, True), (530, 512, 300, False)])
def test_svd_rank_deficient_blocks(ctx, m, n, rank, gap):
    # A two-site wavefunction sector is rank-deficient: its null directions
    # reach the rounding level after preconditioning, leave the sweeps
    # (deflation), and Rc's zero rows end the Householder QR early.  Values
    # against LAPACK, the reconstruction, orthonormality of U over the
    # nonzero values, and the preconditioned run against the plain one.
    # With a gap the nonzero values stop at 1e-10 (the null space is at
    # rounding level, as in a wavefunction sector); without one they run
    # down to 1e-14, below the Cholesky pivot floor.
    from su2dmrg_amd import lib as L
    _say(f"svd {m}x{n} of rank {rank}, gap {gap}")
    rng = np.random.default_rng(m + 7 * n)
    u, _ = np.linalg.qr(rng.standard_normal((m, rank)))
    v, _ = np.linalg.qr(rng.standard_normal((n, rank)))
    A = (u * np.logspace(0, -10 if gap else -14, rank)) @ v.T
    ref = np.sort(np.linalg.svd(A, compute_uv=False))
    out = {}
    for pre in (True, False):
        ctx.set_svd_precondition(64 if pre else 0)
        t = L.wavefunction(ctx, [1], [m], [n], {1: A})
        out[pre] = L.svd_truncate(ctx, t, 0, 1, 100000)
    ctx.set_svd_precondition(256)   # the context's default
    for pre in (True, False):
        U, V, kept, _ = out[pre]
        s = np.array(kept)
        assert len(s) == min(m, n)
        np.testing.assert_allclose(s, ref, rtol=0, atol=1e-12 * ref[-1])
        cut = s[s <= 1e-11 * s[-1]].sum()
        np.testing.assert_allclose(U[1] @ V[1].T, A, rtol=0, atol=1e-13 * ref[-1] + cut)
        big = s > 1e-11 * s[-1]   # U's columns below are zero by convention
        if gap:
            assert big.sum() == rank
        Ub = U[1][:, big]
        sb = s[big]
        bound = 1e-13 * s[-1] / np.minimum.outer(sb, sb) + 1e-13
        assert np.all(np.abs(Ub.T @ Ub - np.eye(big.sum())) <= bound)
    np.testing.assert_allclose(np.array(out[True][2]), np.array(out[False][2]), rtol=0,
                               atol=1e-12 * ref[-1])


def _run_both(ctx, lx, ly, jcoup, bdry, D, nsweep, totspin=0):
    from su2dmrg_amd import lib as L
    lat = od.Lattice(lx, ly, jcoup=jcoup, bdry=bdry)
    o = od.DMRG(lat, totspin=totspin, max_dcut=D)
    o.run(nsweep=nsweep, stop=None)
    g = L.Dmrg(ctx, lx, ly, jcoup, 0.0, bdry, totspin, D)
    g.run(nsweep=nsweep, stop=None)
    return lat, o, g


# (lx, ly, J2/J1 as jcoup, bdry, D, 2S): open lattices, the last three truncate
STEP_CASES = [(4, 2, 1.0, 0, 100, 0), (3, 4, 2.0, 0, 20, 2), (6, 2, 2.0, 0, 10, 0),
              (4, 4, 2.0, 0, 12, 0)]


@pytest.mark.parametrize("lx,ly,jcoup,bdry,D,twoS", STEP_CASES)
def test_dmrg_matches_oracle_step_by_step(ctx, lx, ly, jcoup, bdry, D, twoS):
    _say(f"dmrg step parity {lx}x{ly} jcoup={jcoup} bdry={bdry} D={D} 2S={twoS}")
    lat, o, g = _run_both(ctx, lx, ly, jcoup, bdry, D, 2, twoS)
    eg = np.array(g.energies())
    eo = np.array(o.energies)
    assert len(eg) == len(eo)
    np.testing.assert_allclose(eg, eo, rtol=1e-10, atol=1e-10)
    if lx * ly <= 12 and D >= 100:
        # untruncated: the last energy is exact
        e_ed = ed.ground_energy(lat, twoS)
        assert abs(eg[-1] - e_ed) < 1e-10 * abs(e_ed)


@pytest.mark.parametrize("lx,ly,jcoup,twoS", [(4, 2, 1.0, 0), (3, 4, 2.0, 0), (3, 2, 2.0, 4)])
def test_dmrg_cylinder_converges_to_ed(ctx, lx, ly, jcoup, twoS):
    from su2dmrg_amd import lib as L
    _say(f"dmrg cylinder {lx}x{ly} jcoup={jcoup} 2S={twoS} vs ED")
    lat = od.Lattice(lx, ly, jcoup=jcoup, bdry=2)
    e_ed = ed.ground_energy(lat, twoS)
    g = L.Dmrg(ctx, lx, ly, jcoup, 0.0, 2, twoS, 100)
    g.run(nsweep=6, stop=None)
    e = g.energies()
    assert abs(e[-1] - e_ed) < 1e-10 * abs(e_ed)
    # variational: no sweep step (each treats the whole lattice; 2 (ns - 6)
    # steps per sweep with nfree = 2) lies below the exact ground state
    nstep = 6 * 2 * (lx * ly - 6)
    if nstep:
        assert min(e[-nstep:]) > e_ed - 1e-10 * abs(e_ed)


@pytest.mark.parametrize("lx,ly,jcoup,twoS,nsweep", [(4, 4, 0.8, 0, 4), (4, 2, 1.5, 0, 4),
                                                    (3, 4, 1.5, 2, 10)])
def test_shastry_sutherland_converges_to_ed(ctx, lx, ly, jcoup, twoS, nsweep):
    # boundary code 3: square bonds 1, orthogonal dimers jcoup; periodic
    # cylinders, so only the converged energy is compared
    from su2dmrg_amd import lib as L
    _say(f"dmrg Shastry-Sutherland {lx}x{ly} jcoup={jcoup} 2S={twoS} vs ED")
    lat = od.Lattice(lx, ly, jcoup=jcoup, bdry=3)
    e_ed = ed.ground_energy(lat, twoS)
    g = L.Dmrg(ctx, lx, ly, jcoup, 0.0, 3, twoS, 200)
    g.run(nsweep=nsweep, stop=None)
    assert abs(g.energies()[-1] - e_ed) < 1e-10 * abs(e_ed)


def test_dmrg_config1_4x4_d200(ctx):
    # J1-J2 at J2/J1 = 0.5 on 4 (open) x 4 (periodic), S = 0, D = 200, 5 sweeps
    # with the reference's |dE| < 1e-6 stop rule: exact up to Lanczos convergence
    from su2dmrg_amd import lib as L
    _say("dmrg 4x4 D=200 (config 1)")
    lat = od.Lattice(4, 4, jcoup=2.0, bdry=2)
    e_ed = ed.ground_energy(lat, 0)
    g = L.Dmrg(ctx, 4, 4, 2.0, 0.0, 2, 0, 200)
    per = g.run(nsweep=5, stop=1e-6)
    assert abs(per[-1] - e_ed) < 1e-10 * abs(e_ed)
    assert abs(0.5 * per[-1] / 16 - (-0.5163270351893605)) < 1e-9
