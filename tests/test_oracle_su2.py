"""Pins of the oracle's structure and Clebsch-Gordan layer (CPU only)."""
import itertools
import math
import random

import numpy as np
import pytest
from sympy import Rational
from sympy.physics.quantum.cg import CG

from oracle import su2


def test_isort2_sorts_and_carries():
    rng = random.Random(1)
    for n in range(0, 40):
        a = [rng.randint(0, 30) for _ in range(n)]
        b = list(range(n))
        sa, sb = su2.isort2(a, b)
        assert sa == sorted(a)
        assert sorted(zip(sa, sb)) == sorted(zip(a, b))
        assert all(a[i] == k for k, i in zip(sa, sb))


def _shell_trace(keys):
    """Literal re-trace of the gap sequence N/2, N/4.. with strict '>'."""
    keys = list(keys)
    pos = list(range(len(keys)))
    n = len(keys)
    g = n // 2
    while g:
        for i in range(g, n):
            j = i - g
            while j >= 0 and keys[j] > keys[j + g]:
                keys[j], keys[j + g] = keys[j + g], keys[j]
                pos[j], pos[j + g] = pos[j + g], pos[j]
                j -= g
        g //= 2
    return keys, pos


def test_dsort2_matches_shell_trace_with_ties():
    rng = random.Random(7)
    for n in [1, 2, 5, 17, 64, 200]:
        w = [rng.choice([0.5, 1.0, 1.5, 2.0, rng.random()]) for _ in range(n)]
        sw, si = su2.dsort2(w, list(range(n)))
        tw, tp = _shell_trace(w)
        assert sw == sorted(w)
        assert sw == tw and si == tp


def test_fuse_counts_multiplets_brute_force():
    # multiplicities from m-state counting for tiny legs
    b1 = su2.Bond([0, 1, 2], [2, 1, 3], 1)
    b2 = su2.Bond([1, 3], [1, 2], 1)
    f = su2.fuse_bonds(b1, b2)
    assert f.direction == -1
    # count states with each total m, then peel multiplets off from the top
    mcount = {}
    for j1, d1 in zip(b1.twoj, b1.dims):
        for j2, d2 in zip(b2.twoj, b2.dims):
            for m1 in range(-j1, j1 + 1, 2):
                for m2 in range(-j2, j2 + 1, 2):
                    mcount[m1 + m2] = mcount.get(m1 + m2, 0) + d1 * d2
    mult = {}
    for twoj in sorted((m for m in mcount if m >= 0), reverse=True):
        mult[twoj] = mcount[twoj] - mcount.get(twoj + 2, 0)
    expect = {k: v for k, v in mult.items() if v}
    assert dict(zip(f.twoj, f.dims)) == expect
    assert f.twoj == sorted(f.twoj)


def test_fuse_caps_at_38():
    f = su2.fuse_bonds(su2.Bond([24], [1], 1), su2.Bond([24], [1], 1))
    assert max(f.twoj) == 38 and min(f.twoj) == 0 and len(f.twoj) == 20


def test_block_index_leg0_fastest_and_validity():
    bonds = [su2.Bond([0, 2], [1, 1]), su2.Bond([1], [1]), su2.Bond([1, 3], [1, 1])]
    labs = su2.valid_blocks(bonds, 0)
    # 3-leg locspin 0: third leg must be in j1 x j2
    assert labs == [(0, 1, 1), (2, 1, 1), (2, 1, 3)]
    for lab in labs:
        assert su2.block_labels(bonds, su2.block_index(bonds, lab)) == lab
    assert su2.block_index(bonds, (2, 1, 3)) == 1 + 2 * (0 + 1 * 1)


@pytest.mark.parametrize("j1,j2", [(0, 1), (1, 1), (2, 1), (2, 2), (3, 2), (4, 3), (6, 2), (5, 5)])
def test_cgc_matches_sympy(j1, j2):
    for j3 in range(abs(j1 - j2), j1 + j2 + 1, 2):
        t = su2.cgc(j1, j2, j3)
        for i, j, k in itertools.product(range(j1 + 1), range(j2 + 1), range(j3 + 1)):
            m1, m2, m3 = 2 * i - j1, 2 * j - j2, 2 * k - j3
            ref = float(CG(Rational(j1, 2), Rational(m1, 2), Rational(j2, 2), Rational(m2, 2),
                           Rational(j3, 2), Rational(m3, 2)).doit())
            assert abs(t[i, j, k] - ref) < 1e-14


def test_cgc_orthonormal_large_j():
    j1, j2 = 24, 2
    for j3 in range(22, 27, 2):
        for j3b in range(22, 27, 2):
            a = su2.cgc(j1, j2, j3)
            b = su2.cgc(j1, j2, j3b)
            ov = np.einsum("ijk,ijl->kl", a, b)
            ref = np.eye(j3 + 1) if j3 == j3b else np.zeros((j3 + 1, j3b + 1))
            assert np.abs(ov - ref).max() < 1e-12


def test_singlet_is_cg_to_zero():
    for j in range(0, 9):
        s = su2.singlet(j)
        c = su2.cgc(j, j, 0)[:, :, 0]
        np.testing.assert_allclose(s, c, atol=1e-15)
        assert abs(np.sum(s * s) - 1) < 1e-14


def test_contract_leg_order():
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2, 3, 4))
    b = rng.standard_normal((3, 5))
    c = su2.contract(a, 1, b, 0)          # legs [a2, a0, b1]
    np.testing.assert_allclose(c, np.einsum("ijk,jl->kil", a, b))
    e = rng.standard_normal((6, 4))
    d = su2.contract(a, 2, e, 1)          # legs [a0, a1, e0]
    np.testing.assert_allclose(d, np.einsum("ijk,lk->ijl", a, e))


def test_fac_hamilt_vec_closed_form():
    # pinned further by the dense brute force in test_oracle_heff; here the
    # closed form sqrt((2j'+1)/(2j+1)) it reduces to for a spin-1 operator
    for j in range(0, 24):
        for jp in (j - 2, j, j + 2):
            if jp < 0:
                continue
            f = su2.fac_hamilt_vec(j, 2, jp)
            if j == 0 and jp == 0:
                assert f == 0.0
                continue
            assert abs(f - math.sqrt((jp + 1) / (j + 1))) < 1e-12
    assert su2.fac_hamilt_vec(1, 2, 2) == 0.0   # parity mismatch


def test_is_proportional_to():
    c = su2.cgc(2, 2, 2)
    assert abs(su2.is_proportional_to(3.5 * c, c) - 3.5) < 1e-14
    assert su2.is_proportional_to(c + 1e-3, c) is None
