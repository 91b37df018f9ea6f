"""Block-sparse SU(2) H_eff matvec and Lanczos vector algebra (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Data model (plain structures shared with the tests):
  * a two-leg wavefunction ``psi`` is a dict ``twoj -> ndarray[D_a(j), D_b(j)]``
    holding the reduced blocks of the (left enlarged, right enlarged) legs; the
    CG tensor of every block is the unit-norm singlet (tensor_su2.cpp:2334-2351);
  * a scalar block operator ``h`` (``hh[i]``) is ``twoj -> ndarray[D(j), D(j)]``
    with identity CG (leg 0 is contracted with the wavefunction);
  * a vector operator ``op`` (``opr[i][k]``, operator leg 2j=2 of multiplicity 1)
    is ``(twoj_in, twoj_out) -> ndarray[D(j_in), D(j_out)]``; its CG is
    make_cgc(j_in, 2, j_out) (tensor_su2_dmrg_src.cpp:163-165).

Pins (tests/test_oracle_heff.py):
  * the reduced term formula against a dense brute-force contraction of the
    full tensors (reduced blocks (x) CG tensors, operator legs joined by the
    spin-1 scalar product sqrt(3)*singlet(2)), projected on the singlet;
  * H_eff symmetry <u|H v> = <H u|v>, linearity;
  * the whole matvec against a dense spin Hamiltonian on tiny lattices in
    tests/test_oracle_dmrg.py (through the DMRG environments).
"""
from __future__ import annotations

import numpy as np

from .su2 import fac_hamilt_vec

OP_TWOJ = 2            # spin-1 operator leg (sigma operators, dmrg_su2.cpp:503-506)
FAC_SKIP = 1e-12       # |fac| below this skips the block (tensor_su2_dmrg_src.cpp:996)


def presum(ops):
    """op = sum_k c_k * opr_k in list order (dmrg_su2.cpp:585-599).

    ``ops`` is a list of (coef, op).  Blocks missing from an operator are
    treated as absent; the sum keeps the union of blocks.
    """
    out = {}
    for c, op in ops:
        for key, blk in op.items():
            t = blk * c
            if key in out:
                out[key] = out[key] + t
            else:
                out[key] = t
    return out


def term_apply(psi, opl, opr, K=OP_TWOJ):
    """tensor_su2::hamiltonian_vector_multiplication (tensor_su2_dmrg_src.cpp:893-1041).

    For every left block (j -> j') and wavefunction block j, step1 = psi^T L
    (tensor.cpp:521-583), then for the right block (j -> j') the product
    step1^T R (tensor_dmrg_src.cpp:536-575) is scaled by fac_hamilt_vec[j][2][j']
    and accumulated into output block j'.
    """
    out = {}
    for (ja, jap), L in sorted(opl.items(), key=lambda kv: (kv[0][1], kv[0][0])):
        if ja not in psi:
            continue
        step1 = psi[ja].T @ L                       # [D_b(j), D_a(j')]
        R = opr.get((ja, jap))
        if R is None:
            continue
        fac = fac_hamilt_vec(ja, K, jap)
        if abs(fac) < FAC_SKIP:
            continue
        tmp = (step1.T @ R) * fac                   # [D_a(j'), D_b(j')]
        if jap in out:
            out[jap] = out[jap] + tmp
        else:
            out[jap] = tmp
    return out


def heff_apply(psi, hl, hr, terms):
    """dmrg_su2::hamiltonian_vector_multiplication_idmrg (dmrg_su2.cpp:543-664).

    ``hl``/``hr`` are the block Hamiltonians (or None), ``terms`` a list of
    (left op, right op) pairs whose right (or left) side has already been
    summed with its couplings (``presum``), or (left, right, K) triples whose
    operator leg is 2K instead of 2 (plaquette terms; K = 0 is the
    reference's two-leg product L^T psi R, dmrg_su2.cpp:902-906).  Returns psi' with the same blocks
    as ``psi``, which must list every sector of the two-site space (zeros where
    the reference holds a null block): the reference's output lives on the
    operators' full legs (tensor_su2_dmrg_src.cpp:921-928) and its operator+=
    (tensor_su2.cpp:157-190) creates blocks the left operand lacks.
    """
    out = {j: np.zeros_like(b) for j, b in psi.items()}
    # dmrg_su2.cpp:558-570: hh[il] contracted on leg 0, hh[ir] on leg 1.
    if hl is not None:
        for j, b in psi.items():
            if j in hl:
                out[j] = out[j] + hl[j].T @ b
    if hr is not None:
        for j, b in psi.items():
            if j in hr:
                out[j] = out[j] + b @ hr[j]
    acc = {}
    for term in terms:                              # dmrg_su2.cpp:582-610, 635-655
        t = term_apply(psi, *term)
        for j, b in t.items():
            acc[j] = acc[j] + b if j in acc else b
    for j, b in acc.items():                        # dmrg_su2.cpp:657-660
        if j in out:
            out[j] = out[j] + b
    return out


# --------------------------------------------------------------------------
# vector algebra (tensor_su2.cpp:96-259, 324-356, 663-684)
# --------------------------------------------------------------------------
def inner(u, v):
    """tensor_su2::inner_prod: sum over blocks of reduced dot products."""
    s = 0.0
    for j in sorted(u):
        s += float(np.sum(u[j] * v[j]))
    return s


def axpy(a, x, y):
    return {j: y[j] + a * x[j] for j in y}


def scal(a, x):
    return {j: a * b for j, b in x.items()}


def norm(u):
    """tensor_su2::normalize_vector norm (reduced elements only, 340-341)."""
    return float(np.sqrt(inner(u, u)))


def heff_flops(psi_dims_a, psi_dims_b, terms_keys, with_hl=True, with_hr=True):
    """Algorithmic GEMM flops of one matvec in the reference's order.

    ``terms_keys`` is a list (one per term) of the (j, j') transition keys
    present in both operators.  Counts 2*m*n*k per dgemm: step1 (D_b x D_a
    x D_a') and step2 (D_a' x D_b x D_b') per transition, plus the two
    block-Hamiltonian products (tensor.cpp:567, tensor_dmrg_src.cpp:563).
    """
    f = 0
    for keys in terms_keys:
        for ja, jap in keys:
            if ja not in psi_dims_a or jap not in psi_dims_a:
                continue
            da, db = psi_dims_a[ja], psi_dims_b[ja]
            dap, dbp = psi_dims_a[jap], psi_dims_b[jap]
            f += 2 * db * da * dap + 2 * dap * db * dbp
    for j in psi_dims_a:
        da, db = psi_dims_a[j], psi_dims_b[j]
        if with_hl:
            f += 2 * da * da * db
        if with_hr:
            f += 2 * da * db * db
    return f
