"""Closed forms (sympy's exact Wigner 6j and 9j symbols) for every recoupling
table of the environment kernels, over the reference's label range 2j < 25
and beyond (CPU only).

The library (and the oracle, equal to it in test_recoupling_tables.py)
computes each table as the reference does: a contraction of dense CG tensors
and its ratio to the output CG (dmrg_su2.cpp:714-1020).  Angular-momentum
algebra fixes the same numbers independently:

* an operator of rank k acting on the block (left: (jl s) ja, right:
  (s jr) ja) gives sqrt((ja+1)(jl'+1)) times the 6j {jl s ja; ja' k jl'};
* on the site: sqrt(2(ja+1)) times {s ja jl; ja' s k};
* the scalar pair-up S_block . s_site: sqrt(2(jl'+1)) times
  {jl s ja; s jl' k};
* two rank-1 operators coupled to rank K (the plaquette tables) and a rank-K
  operator reduced with the site spin to rank 1: sqrt(2(K+1)(ja+1)(jl'+1))
  and sqrt(6(ja+1)(jl'+1)) times 9j symbols,
each with a phase (-1)^(sum of labels / 2).  The phases and constants were
read off small labels; that they then hold for every label up to 2j = 24
(and at the large spins where the tables once failed their proportionality
check) is what pins the tables: a dropped factor, a wrong index or a
transposed operand breaks the match.
"""
import itertools
import math

import pytest
from sympy import Rational
from sympy.physics.wigner import wigner_6j, wigner_9j

from su2dmrg_amd import lib as L

H = lambda x: Rational(x, 2)  # noqa: E731
MAX2J = 24


def w6(*a):
    try:
        return float(wigner_6j(*[H(x) for x in a]))
    except ValueError:
        return 0.0


def w9(*a):
    try:
        return float(wigner_9j(*[H(x) for x in a]))
    except ValueError:
        return 0.0


def sg(e):
    return -1.0 if (e // 2) % 2 else 1.0


def phased(e, mag, w):
    """(-1)^(e/2) mag w; e is even wherever the symbol w is nonzero"""
    if w == 0.0:
        return 0.0
    assert e % 2 == 0, e
    return sg(e) * math.sqrt(mag) * w


def _close(a, b):
    return abs(a - b) <= 1e-11 * max(1.0, abs(b))


# s = 1 (spin-1/2 site), k = 2 (spin operators)
def op_transform_left(jl, ja, jlp, jap):
    return phased(ja + jlp + 3, (ja + 1) * (jlp + 1), w6(jl, 1, ja, jap, 2, jlp))


def op_transform_right(jr, ja, jrp, jap):
    return phased(jr + jap + 3, (ja + 1) * (jrp + 1), w6(jr, 1, ja, jap, 2, jrp))


def op_initial_left(jl, ja, jap):
    return phased(jl + jap + 3, 2 * (ja + 1), w6(1, ja, jl, jap, 1, 2))


def op_initial_right(jr, ja, jap):
    return phased(jr + ja + 3, 2 * (ja + 1), w6(1, ja, jr, jap, 1, 2))


def pairup_left(ja, jl, jlp):
    return phased(ja + jl + 3, 2 * (jlp + 1), w6(jl, 1, ja, 1, jlp, 2))


def pairup_right(ja, jr, jrp):
    return phased(ja + jrp + 3, 2 * (jrp + 1), w6(jr, 1, ja, 1, jrp, 2))


def couple_left(jl, ja, jlp, jap, K):
    return phased(0, 2 * (K + 1) * (ja + 1) * (jlp + 1), w9(jlp, jl, 2, 1, 1, 2, jap, ja, K))


def couple_right(jr, ja, jrp, jap, K):
    return phased(K, 2 * (K + 1) * (ja + 1) * (jrp + 1), w9(1, 1, 2, jrp, jr, 2, jap, ja, K))


def reduce_left(jl, ja, jlp, jap, K):
    return phased(0, 6 * (ja + 1) * (jlp + 1), w9(jlp, jl, K, 1, 1, 2, jap, ja, 2))


def reduce_right(jr, ja, jrp, jap, K):
    return phased(K, 6 * (ja + 1) * (jrp + 1), w9(1, 1, 2, jrp, jr, K, jap, ja, 2))


def _quads(blocks, dk):
    for jl in blocks:
        for ja in (jl - 1, jl + 1):
            for jlp in range(max(0, jl - dk), jl + dk + 1):
                for jap in (jlp - 1, jlp + 1):
                    if ja >= 0 and jap >= 0:
                        yield jl, ja, jlp, jap


def test_block_operator_tables_every_label():
    n = 0
    for jl, ja, jlp, jap in _quads(range(0, MAX2J + 1), 2):
        for name, f in (("op_transform_left", op_transform_left),
                        ("op_transform_right", op_transform_right)):
            a = L.recoupling(name, jl, ja, jlp, jap, 2, 1)
            b = f(jl, ja, jlp, jap)
            assert _close(a, b), (name, jl, ja, jlp, jap, a, b)
            n += b != 0
    assert n > 400


def test_site_and_pairup_tables_every_label():
    n = 0
    for jl in range(0, MAX2J + 1):
        for ja, jap in itertools.product((jl - 1, jl + 1), repeat=2):
            if ja < 0 or jap < 0:
                continue
            for name, f in (("op_initial_left", op_initial_left),
                            ("op_initial_right", op_initial_right)):
                a = L.recoupling(name, jl, ja, jap, 2, 1, 1)
                b = f(jl, ja, jap)
                assert _close(a, b), (name, jl, ja, jap, a, b)
                n += b != 0
        for ja in (jl - 1, jl + 1):
            for jlp in (ja - 1, ja + 1):
                if ja < 0 or jlp < 0:
                    continue
                for name, f in (("pairup_left", pairup_left), ("pairup_right", pairup_right)):
                    a = L.recoupling(name, ja, jl, jlp, 2, 1, 1)
                    b = f(ja, jl, jlp)
                    assert _close(a, b), (name, ja, jl, jlp, a, b)
                    n += b != 0
    assert n > 150


# a spread of block spins up to 2j = 24 and the large spins (2j >= 30) of the
# 12x6 scan (9j symbols are slow in exact arithmetic)
PLAQ_BLOCKS = [0, 1, 2, 5, 6, 11, 12, 17, 23, 24, 31, 38, 45]


@pytest.mark.parametrize("K", [0, 2, 4])
def test_plaquette_tables(K):
    n = 0
    for jl, ja, jlp, jap in _quads(PLAQ_BLOCKS, 2):
        if abs(jlp - jl) > 2 and abs(jlp - jl) > K:
            continue
        for name, f in (("couple_left", couple_left), ("couple_right", couple_right)):
            if abs(jl - jlp) <= 2:
                a = L.recoupling(name, jl, ja, jlp, jap, K, 1)
                b = f(jl, ja, jlp, jap, K)
                assert _close(a, b), (name, jl, ja, jlp, jap, K, a, b)
                n += b != 0
        for name, f in (("reduce_left", reduce_left), ("reduce_right", reduce_right)):
            if abs(jl - jlp) <= K:
                a = L.recoupling(name, jl, ja, jlp, jap, K, 1)
                b = f(jl, ja, jlp, jap, K)
                assert _close(a, b), (name, jl, ja, jlp, jap, K, a, b)
                n += b != 0
    assert n > 50
