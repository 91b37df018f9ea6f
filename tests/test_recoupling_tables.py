"""The library's recoupling scalars equal the oracle's restatement of the
reference tables (host code of the C ABI; CPU only)."""
import itertools

import pytest

from oracle import recoupling as orc
from oracle import su2 as osu2
from su2dmrg_amd import lib as L

SPINS = range(0, 13)


def _close(a, b):
    return abs(a - b) <= 1e-12 * max(1.0, abs(b))


def test_transform_and_initial_tables():
    n = 0
    for jl, ja, jlp, jap in itertools.product(SPINS, repeat=4):
        if abs(ja - jl) != 1 or abs(jap - jlp) != 1:
            continue
        for name in ("op_transform_left", "op_transform_right"):
            a = L.recoupling(name, jl, ja, jlp, jap, 2, 1)
            b = getattr(orc, name)(jl, ja, jlp, jap, 2, 1)
            assert _close(a, b), (name, jl, ja, jlp, jap, a, b)
            n += b != 0
    for jl, ja, jap in itertools.product(SPINS, repeat=3):
        for name in ("op_initial_left", "op_initial_right"):
            a = L.recoupling(name, jl, ja, jap, 2, 1, 1)
            b = getattr(orc, name)(jl, ja, jap, 2, 1, 1)
            assert _close(a, b), (name, jl, ja, jap)
    assert n > 100


def test_pairup_and_psi_tables():
    for ja, jl, jlp in itertools.product(SPINS, repeat=3):
        for name in ("pairup_left", "pairup_right"):
            a = L.recoupling(name, ja, jl, jlp, 2, 1, 1)
            b = getattr(orc, name)(ja, jl, jlp, 2, 1, 1)
            assert _close(a, b), (name, ja, jl, jlp)
    for jl, jr in itertools.product(SPINS, repeat=2):
        if abs(jl - jr) != 1:
            continue
        assert _close(L.recoupling("psi_left_to_right", jl, 1, jr), orc.left_to_right_psi(jl, 1, jr))
        assert _close(L.recoupling("psi_right_to_left", jl, 1, jr), orc.right_to_left_psi(jl, 1, jr))


@pytest.mark.parametrize("j", range(0, 24))
def test_heff_table(j):
    for jp in (j - 2, j, j + 2):
        if jp < 0:
            continue
        assert _close(L.recoupling("hamilt_vec", j, 2, jp), osu2.fac_hamilt_vec(j, 2, jp))


@pytest.mark.parametrize("K", [0, 2, 4])
def test_plaquette_coupling_tables(K):
    n = 0
    for jl, ja, jlp, jap in itertools.product(range(0, 11), repeat=4):
        if abs(ja - jl) != 1 or abs(jap - jlp) != 1:
            continue
        for name in ("couple_left", "couple_right", "reduce_left", "reduce_right"):
            a = L.recoupling(name, jl, ja, jlp, jap, K, 1)
            b = getattr(orc, name)(jl, ja, jlp, jap, K, 1)
            assert _close(a, b), (name, jl, ja, jlp, jap, K, a, b)
            n += b != 0
    assert n > 50
