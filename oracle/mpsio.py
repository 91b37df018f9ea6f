"""MPS, energy and singular-value files of the reference (oracle: test
infrastructure only).

Restates the reference's checkpoint format so the device library's files can
be checked field by field:

* ``uu-{ns}-{D}-{2S}-{ky}-{exci}-{i}.dat`` (dmrg_su2::save_mps2,
  dmrg_su2_omp.cpp:864-941): ``nbond<TAB>locspin``, then per leg
  ``dir<TAB>nmoment``, the 2j labels and the multiplet counts, each
  tab-terminated on one line;
* ``...-{i}.bin``: int32 flag per block index (1: block present,
  tensor_su2.cpp:566-583), the reduced doubles of the present blocks, then
  their Clebsch-Gordan doubles, both in block-index order (leg 0 fastest,
  su2struct.cpp:243-261) and each block column-major;
* ``enr-{ns}-{D}-{2S}-{ky}-{exci}.dat`` (save_enr, 1070-1098): one line per
  saved sweep, the Ritz value minus 1000 at 16 significant digits; read_enr
  (1100-1152) takes the last line and adds 1000 back;
* ``ww-{ns}-{D}-{2S}-{ky}-{exci}-{ns/2}.bin`` (save_ww, 1154-1186):
  max_dcut * phdim doubles of the last solve's ``wout`` (kept values sector by
  sector, then the cut ones).  The reference zeroes only the first max_dcut
  entries before that solve; the rest of its buffer keeps older values, which
  are written here as zeros.

A left site i (i < ns/2) is stored as the reference's three-leg
``uu[i]`` = [lb[i-1], site, lb[i]] with CG make_cgc(j_a, 1, j_b); a right
site as [site, rb[i+1], rb[i]] with make_cgc(1, j_r', j_r) (the form
tensor_su2::svd leaves, tensor_su2.cpp:1733-1743).  Its reduced block
(j_a, 1, j_b) is the projector's rows of the part j_a inside the fused sector
j_b.
"""
from __future__ import annotations

import os

import numpy as np

from .dmrg import PHYS, Enlarged
from .su2 import Bond, block_index, cgc, is_proportional_to, valid_blocks

SITE_BOND = Bond([PHYS], [1], 1)


def base_name(kind, ns, dcut, twoS, ky=0, exci=0):
    return f"{kind}-{ns}-{dcut}-{twoS}-{ky}-{exci}"


def write_tensor3(path, legs, blocks):
    """Write ``path``.dat/.bin for a three-leg, locspin-0 tensor.

    legs: three Bonds (direction kept); blocks: {(j0, j1, j2): ndarray}.
    """
    with open(path + ".dat", "w") as f:
        f.write(f"3\t0\n")
        for b in legs:
            f.write(f"{b.direction}\t{b.nsec}\n")
            f.write("".join(f"{j}\t" for j in b.twoj) + "\n")
            f.write("".join(f"{d}\t" for d in b.dims) + "\n")
    nten = legs[0].nsec * legs[1].nsec * legs[2].nsec
    flags = np.zeros(nten, dtype=np.int32)
    red, cg = [], []
    for labs in valid_blocks(legs, 0):
        if labs not in blocks:
            continue
        flags[block_index(legs, labs)] = 1
        red.append(np.asarray(blocks[labs], dtype=np.float64).ravel(order="F"))
        cg.append(np.asarray(cgc(*labs)).ravel(order="F"))
    with open(path + ".bin", "wb") as f:
        f.write(flags.tobytes())
        f.write(np.concatenate(red).tobytes() if red else b"")
        f.write(np.concatenate(cg).tobytes() if cg else b"")


def read_tensor3(path):
    """Inverse of write_tensor3: (legs, {labels: reduced block}).

    A stored block whose CG is r times make_cgc of its labels comes back as
    r * reduced (the reference's own readers keep the CG as stored;
    tensor::is_proportional_to, tensor.cpp:1231-1269, gives r).
    """
    with open(path + ".dat") as f:
        tok = f.read().split()
    pos = 0
    nb, loc = int(tok[0]), int(tok[1])
    pos = 2
    if nb != 3 or loc != 0:
        raise ValueError(f"{path}.dat: expected a three-leg locspin-0 tensor")
    legs = []
    for _ in range(3):
        d, n = int(tok[pos]), int(tok[pos + 1])
        pos += 2
        tj = [int(x) for x in tok[pos:pos + n]]
        pos += n
        dm = [int(x) for x in tok[pos:pos + n]]
        pos += n
        legs.append(Bond(tj, dm, d, sort=False))
    nten = legs[0].nsec * legs[1].nsec * legs[2].nsec
    raw = open(path + ".bin", "rb").read()
    flags = np.frombuffer(raw[:4 * nten], dtype=np.int32)
    present = [labs for labs in valid_blocks(legs, 0) if flags[block_index(legs, labs)]]
    dims = [tuple(b.dims[b.index(j)] for b, j in zip(legs, labs)) for labs in present]
    nred = sum(int(np.prod(d)) for d in dims)
    ncg = sum((a + 1) * (b + 1) * (c + 1) for a, b, c in present)
    body = np.frombuffer(raw[4 * nten:], dtype=np.float64)
    if body.size != nred + ncg:
        raise ValueError(f"{path}.bin: {body.size} doubles, expected {nred + ncg}")
    out, o, oc = {}, 0, nred
    for labs, d in zip(present, dims):
        n = int(np.prod(d))
        blk = body[o:o + n].reshape(d, order="F")
        o += n
        c = tuple(x + 1 for x in labs)
        nc = c[0] * c[1] * c[2]
        stored = body[oc:oc + nc].reshape(c, order="F")
        oc += nc
        r = is_proportional_to(stored, np.asarray(cgc(*labs)), tol=1e-10)
        if r is None:
            raise ValueError(f"{path}.bin: block {labs} CG is not a multiple of make_cgc")
        out[labs] = r * blk
    return legs, out


def left_tensor(lb_prev: Bond, lb: Bond, U):
    """Projector U {jb: [D_E(jb), d(jb)]} of a left site as uu[i] blocks."""
    E = Enlarged(lb_prev, left=True)
    legs = [Bond(lb_prev.twoj, lb_prev.dims, 1, sort=False), SITE_BOND,
            Bond(lb.twoj, lb.dims, -1, sort=False)]
    blocks = {}
    for jb in lb.twoj:
        for ja, o, d in E.parts(jb):
            blocks[(ja, PHYS, jb)] = U[jb][o:o + d, :].reshape(d, 1, -1)
    return legs, blocks


def right_tensor(rb_next: Bond, rb: Bond, V):
    """Projector V {jr: [D_F(jr), d(jr)]} of a right site as uu[i] blocks."""
    F = Enlarged(rb_next, left=False)
    legs = [SITE_BOND, Bond(rb_next.twoj, rb_next.dims, 1, sort=False),
            Bond(rb.twoj, rb.dims, -1, sort=False)]
    blocks = {}
    for jr in rb.twoj:
        for jq, o, d in F.parts(jr):
            blocks[(PHYS, jq, jr)] = V[jr][o:o + d, :].reshape(1, d, -1)
    return legs, blocks


def projector_from_left(legs, blocks):
    """(lb_prev, lb, U) from a left site's uu[i]."""
    lb_prev = Bond(legs[0].twoj, legs[0].dims, 1, sort=False)
    lb = Bond(legs[2].twoj, legs[2].dims, 1, sort=False)
    E = Enlarged(lb_prev, left=True)
    U = {}
    for jb, db in zip(lb.twoj, lb.dims):
        if jb not in E.dims:
            continue
        P = np.zeros((E.dims[jb], db))
        for ja, o, d in E.parts(jb):
            b = blocks.get((ja, PHYS, jb))
            if b is not None:
                P[o:o + d, :] = b.reshape(d, db)
        U[jb] = P
    return lb_prev, lb, U


def projector_from_right(legs, blocks):
    """(rb_next, rb, V) from a right site's uu[i]."""
    rb_next = Bond(legs[1].twoj, legs[1].dims, 1, sort=False)
    rb = Bond(legs[2].twoj, legs[2].dims, 1, sort=False)
    F = Enlarged(rb_next, left=False)
    V = {}
    for jr, dr in zip(rb.twoj, rb.dims):
        if jr not in F.dims:
            continue
        P = np.zeros((F.dims[jr], dr))
        for jq, o, d in F.parts(jr):
            b = blocks.get((PHYS, jq, jr))
            if b is not None:
                P[o:o + d, :] = b.reshape(d, dr)
        V[jr] = P
    return rb_next, rb, V


def save(dm, directory, ky=0, exci=0):
    """Checkpoint of an oracle DMRG after a sweep (save_mps2, save_ww, save_enr)."""
    ns, D, S = dm.ns, dm.D, dm.S
    base = os.path.join(directory, base_name("uu", ns, D, S, ky, exci))
    for i in range(ns):
        if i < ns // 2:
            legs, blocks = left_tensor(dm.lb[i - 1], dm.lb[i], dm.U[i])
        else:
            legs, blocks = right_tensor(dm.rb[i + 1], dm.rb[i], dm.V[i])
        write_tensor3(f"{base}-{i}", legs, blocks)
    ww = np.zeros(D * (PHYS + 1))
    w = np.asarray(dm.wout)[: ww.size]
    ww[: w.size] = w
    with open(os.path.join(directory, base_name("ww", ns, D, S, ky, exci) + f"-{ns // 2}.bin"),
              "wb") as f:
        f.write(ww.tobytes())
    with open(os.path.join(directory, base_name("enr", ns, D, S, ky, exci) + ".dat"), "a") as f:
        f.write(f"{dm.energies[-1] - 1000:.16g}\n")


def load(dm, directory, dcut=None, ky=0, exci=0):
    """prepare_sweep from files (read_mps + dmrg_su2_omp.cpp:160-193).

    Sets the projectors and bonds of every site and rebuilds the environments
    from both edges; dm.sweep() then continues from the centre.  Returns the
    saved energy (read_enr).
    """
    ns, S = dm.ns, dm.S
    D = dm.D if dcut is None else dcut
    base = os.path.join(directory, base_name("uu", ns, D, S, ky, exci))
    for i in range(ns // 2):
        legs, blocks = read_tensor3(f"{base}-{i}")
        lb_prev, lb, U = projector_from_left(legs, blocks)
        if lb_prev != dm.lb[i - 1]:
            raise ValueError(f"site {i}: left bond does not continue the chain")
        if i < ns // 2 - 1:
            dm.push_left(i, U)
        else:
            dm.U[i] = U
            dm.lb[i] = lb
    for i in range(ns - 1, ns // 2 - 1, -1):
        legs, blocks = read_tensor3(f"{base}-{i}")
        rb_next, rb, V = projector_from_right(legs, blocks)
        if rb_next != dm.rb[i + 1]:
            raise ValueError(f"site {i}: right bond does not continue the chain")
        dm.push_right(i, V)
    with open(os.path.join(directory, base_name("enr", ns, D, S, ky, exci) + ".dat")) as f:
        vals = [float(x) for x in f.read().split()]
    return vals[-1] + 1000


def lower_state(directory, ns, dcut, twoS, exci, ky=0):
    """A lower state read from its checkpoint files (read_mps into orth[exci],
    dmrg_su2_omp.cpp:944-1068; energy from read_enr, 1100-1152)."""
    from .dmrg import LowerState
    base = os.path.join(directory, base_name("uu", ns, dcut, twoS, ky, exci))
    beta, T = {}, {}
    for k in range(ns):
        legs, blocks = read_tensor3(f"{base}-{k}")
        T[k] = {}
        if k < ns // 2:
            beta[k - 1] = Bond(legs[0].twoj, legs[0].dims, 1, sort=False)
            beta[k] = Bond(legs[2].twoj, legs[2].dims, 1, sort=False)
            for (ja, _, jb), b in blocks.items():
                T[k][(ja, jb)] = b[:, 0, :]
        else:
            beta[k] = Bond(legs[1].twoj, legs[1].dims, 1, sort=False)
            beta[k - 1] = Bond(legs[2].twoj, legs[2].dims, 1, sort=False)
            for (_, jq, jr), b in blocks.items():
                T[k][(jr, jq)] = b[0, :, :].T
    with open(os.path.join(directory, base_name("enr", ns, dcut, twoS, ky, exci) + ".dat")) as f:
        energy = [float(x) for x in f.read().split()][-1] + 1000
    return LowerState(ns, beta, T, energy)
