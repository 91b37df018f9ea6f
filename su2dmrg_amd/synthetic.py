"""Synthetic SU(2) block-sparse H_eff workloads (matvec microbench inputs).

These generate the reduced blocks of a two-leg wavefunction, two block
Hamiltonians and a list of spin-1 operator term pairs -- the operands of
``dmrg_su2::hamiltonian_vector_multiplication_idmrg`` (dmrg_su2.cpp:543-664)
-- with i.i.d. N(0,1) entries from a seeded generator.  Sector layouts:

* ``cyl8x4_d1000``: bond spins 2j = 0..12 (integer and half-integer),
  multiplet counts from a Gaussian in j centred at j=1 with sigma=1.5,
  at least one multiplet per sector, summing to D=1000 on each leg; 12 term
  pairs (MPO bond dimension 14 = 12 spin-1 terms + the two block
  Hamiltonians).
* ``cyl16x8_d2000``: the centre bond of a 16x8 cylinder at D=2000: enlarged
  half-integer sectors 2j = 1,3,5,7,9 with 1362,1362,629,124,9 multiplets on
  both legs (bond D=2000 spread 2j=0..8 as 514,848,514,115,9, fused with a
  spin-1/2 site), and 8 term pairs (the 8 sites of the last left column
  couple across the cut through x and diagonal bonds).

The layout is plain Python: ``psi``/``hl``/``hr`` map 2j -> ndarray and an
operator maps (2j_in, 2j_out) -> ndarray, rows/cols = multiplets.
"""
from __future__ import annotations

import math

import numpy as np

OP_TWOJ = 2


def gaussian_sector_dims(twoj, total, center_j, sigma_j, min_dim=1):
    """Discretised Gaussian multiplet counts over sectors ``twoj`` summing to ``total``."""
    w = np.array([math.exp(-((t / 2.0 - center_j) ** 2) / (2.0 * sigma_j ** 2)) for t in twoj])
    raw = w / w.sum() * total
    dims = np.maximum(min_dim, np.floor(raw)).astype(int)
    # hand out the remainder (or take back the excess) by largest fraction
    diff = total - int(dims.sum())
    frac = raw - np.floor(raw)
    order = np.argsort(-frac, kind="stable")
    i = 0
    while diff != 0:
        k = order[i % len(order)]
        if diff > 0:
            dims[k] += 1
            diff -= 1
        elif dims[k] > min_dim:
            dims[k] -= 1
            diff += 1
        i += 1
    return [int(d) for d in dims]


def transitions(twoj):
    """(j, j') pairs a spin-1 operator connects with nonzero recoupling.

    Triangle |j-j'| <= 1 <= j+j' and equal parity of 2j (dmrg_su2.cpp:759-762).
    """
    s = set(twoj)
    out = []
    for a in twoj:
        for b in twoj:
            if b in s and abs(a - b) <= OP_TWOJ <= a + b and (a % 2) == (b % 2):
                out.append((a, b))
    return out


def _sym(rng, d):
    m = rng.standard_normal((d, d))
    return 0.5 * (m + m.T)


def make_workload(twoj, dims_a, dims_b, nterm, seed=0, scale_ops=None):
    """Random psi (normalised), symmetric hl/hr and ``nterm`` operator pairs."""
    rng = np.random.default_rng(seed)
    da = dict(zip(twoj, dims_a))
    db = dict(zip(twoj, dims_b))
    psi = {j: rng.standard_normal((da[j], db[j])) for j in twoj}
    nrm = math.sqrt(sum(float(np.sum(b * b)) for b in psi.values()))
    psi = {j: b / nrm for j, b in psi.items()}
    hl = {j: _sym(rng, da[j]) for j in twoj}
    hr = {j: _sym(rng, db[j]) for j in twoj}
    tr = transitions(twoj)
    terms = []
    for _ in range(nterm):
        L = {(a, b): rng.standard_normal((da[a], da[b])) for (a, b) in tr}
        R = {(a, b): rng.standard_normal((db[a], db[b])) for (a, b) in tr}
        if scale_ops is not None:
            L = {k: v * scale_ops for k, v in L.items()}
            R = {k: v * scale_ops for k, v in R.items()}
        terms.append((L, R))
    return {"twoj": list(twoj), "dims_a": list(dims_a), "dims_b": list(dims_b),
            "psi": psi, "hl": hl, "hr": hr, "terms": terms}


def cyl8x4_d1000(seed=0, D=1000):
    twoj = list(range(0, 13))
    dims = gaussian_sector_dims(twoj, D, center_j=1.0, sigma_j=1.5)
    return make_workload(twoj, dims, dims, nterm=12, seed=seed)


CYL16X8_TWOJ = [1, 3, 5, 7, 9]
CYL16X8_DIMS = [1362, 1362, 629, 124, 9]


def cyl16x8_d2000(seed=0, nterm=8):
    return make_workload(CYL16X8_TWOJ, CYL16X8_DIMS, CYL16X8_DIMS, nterm=nterm, seed=seed)


def small(seed=0, twoj=(0, 1, 2, 3, 4), dims=(3, 4, 5, 2, 1), nterm=3):
    return make_workload(list(twoj), list(dims), list(dims), nterm=nterm, seed=seed)


def flops(work):
    """Algorithmic GEMM flops of one matvec (reference order, both H blocks)."""
    da = dict(zip(work["twoj"], work["dims_a"]))
    db = dict(zip(work["twoj"], work["dims_b"]))
    f = 0
    for L, R in work["terms"]:
        for (a, b) in L:
            if (a, b) not in R:
                continue
            f += 2 * db[a] * da[a] * da[b] + 2 * da[b] * db[a] * db[b]
    for j in work["twoj"]:
        f += 2 * da[j] * da[j] * db[j] + 2 * da[j] * db[j] * db[j]
    return f
