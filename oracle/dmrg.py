"""Two-site SU(2) DMRG on square-lattice cylinders (oracle restatement).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Restates lingwangqs/SU2DMRG's driver for the J1-J2 family with Q = 0:
  * lattice tables hmap/coup/fll/frr            dmrg_su2.cpp:31-303
  * local spin operators sigma[0/1] = +-sqrt(3)/2 dmrg_su2.cpp:503-506
  * enlarged-basis environments                  dmrg_su2_omp.cpp:604-825
  * truncated-basis environments                 dmrg_su2_omp.cpp:319-601
  * H_eff matvec                                 dmrg_su2.cpp:543-664 (oracle.heff)
  * Lanczos (mlanc=5, full reorthogonalisation)  lanczos_su2.cpp:125-195, 237-338
  * truncated SVD across sectors                 tensor_su2.cpp:1619-1758, tensor.cpp:1472-1700
  * warm-up growth (do_idmrg) and sweeps         dmrg_su2_omp.cpp:20-157

Data layout: the MPS tensor of a left site is kept as its projector U from
the enlarged basis E_i = fuse(bond_{i-1}, site) to the new bond
(``proju`` of tensor_su2.cpp:1733-1739); the reference's three-leg uu[i] is
that projector contracted with the fusion isometry (tensor_su2.cpp:1740),
so every truncated-basis operator is U^T (enlarged operator) U.  The right
side mirrors this with F_i = fuse(site, bond_{i+1}).

The Lanczos start vector between steps is the previous wavefunction carried
to the new centre with the exact recoupling factor (oracle.recoupling.
left_to_right_psi); the reference reuses the reduced blocks without it
(tensor_su2.cpp:749-836), which only changes the start vector.

Pinned by exact diagonalisation of the same Hamiltonian (tests).
"""
from __future__ import annotations

import math

import numpy as np
from scipy.linalg import eigh_tridiagonal

from . import heff
from . import plaquette
from . import recoupling as rc
from .su2 import Bond, dsort2, fuse_bonds

PHYS = 1                      # spin-1/2 site, 2j = 1 (dmrg_su2.cpp:37)
OP = 2                        # spin-1 operator leg
SIGMA_LEFT = math.sqrt(3) / 2   # sigma[0] reduced element (dmrg_su2.cpp:504)
SIGMA_RIGHT = -math.sqrt(3) / 2  # sigma[1] (dmrg_su2.cpp:505)
NFREE = 2                     # dmrg_su2.cpp:54
MLANC = 5                     # dmrg_su2.cpp:530
SITE = Bond([PHYS], [1], 1)


# --------------------------------------------------------------------------
# lattice (dmrg_su2.cpp:97-303)
# --------------------------------------------------------------------------
class Lattice:
    def __init__(self, lx, ly, jcoup, qcoup=0.0, bdry=2):
        self.lx, self.ly, self.ns, self.bdry = lx, ly, lx * ly, bdry
        self.qcoup = qcoup
        ns = self.ns
        self.hmap = np.zeros((ns, ns), dtype=int)

        def site(x, y):
            return x * ly + (y if x % 2 == 0 else ly - 1 - y)

        for k in range(5):
            for i in range(ns):
                x1 = i // ly
                y1 = i % ly if x1 % 2 == 0 else ly - 1 - (i % ly)
                flag = True
                if bdry == 0:
                    if k >= 2:
                        continue
                    x2, y2 = ((x1 + 1) % lx, y1) if k == 0 else (x1, (y1 + 1) % ly)
                elif bdry == 2:
                    if k == 0:
                        x2, y2 = x1 + 1, y1
                    elif k == 1:
                        x2, y2 = x1, (y1 + 1) % ly
                    elif k == 2:
                        x2, y2 = x1 + 1, (y1 + 1) % ly
                    elif k == 3:
                        x2, y2 = x1 - 1, (y1 + 1) % ly
                    else:
                        flag = False
                elif bdry == 1:
                    if k == 0:
                        x2, y2 = x1 + 1, y1
                    elif k == 1:
                        x2, y2 = x1, (y1 + 1) % ly
                    elif k == 2:
                        x2, y2 = x1 + 2, y1
                    elif k == 3:
                        x2, y2 = x1, (y1 + 2) % ly
                    else:
                        flag = False
                elif bdry == 3:
                    # Shastry-Sutherland (dmrg_su2.cpp:154-170): square-lattice
                    # bonds plus one diagonal on every other plaquette, the
                    # diagonals of neighbouring plaquettes orthogonal.  The
                    # k = 4 case names a pair but no coupling class
                    # (dmrg_su2.cpp:180-191), so it adds no bond.
                    if k == 0:
                        x2, y2 = x1 + 1, y1
                    elif k == 1:
                        x2, y2 = x1, (y1 + 1) % ly
                    elif k == 2 and x1 % 2 == 0 and y1 % 2 == 0:
                        x2, y2 = x1 + 1, (y1 + 1) % ly
                    elif k == 3 and x1 % 2 == 1 and y1 % 2 == 0:
                        x2, y2 = x1 + 1, (y1 + ly - 1) % ly
                    else:
                        flag = False
                else:
                    raise NotImplementedError("boundary code")
                if not (flag and 0 <= x2 < lx and 0 <= y2 < ly):
                    continue
                j = site(x2, y2)
                # coupling class (dmrg_su2.cpp:180-197): 1 -> coup[1],
                # 2 -> coup[2], 3 -> coup[3]; on the Shastry-Sutherland
                # lattice the square bonds are class 2, the dimers class 1
                if bdry == 2 and k == 1 and (x1 == 0 or x1 == lx - 1):
                    v = 3
                elif (bdry == 3 and k < 2) or (bdry in (1, 2) and 2 <= k < 4):
                    v = 2
                elif (bdry == 3 and 2 <= k < 4) or (bdry != 3 and k < 2):
                    v = 1
                else:
                    continue
                self.hmap[i, j] = v
                self.hmap[j, i] = v
        # coupling constants (dmrg_su2.cpp:258-260)
        self.coup = [0.0, jcoup + qcoup / 2.0, 1.0 - abs(qcoup), jcoup + qcoup / 4.0]
        # plaquettes (x, y), (x, y+1), (x+1, y+1), (x+1, y) for x < lx-1 as
        # sorted snake positions (dmrg_su2.cpp:200-232); only with |Q| > 1e-5
        # (plqflg, dmrg_su2.cpp:263-281)
        self.plq = []
        if abs(qcoup) > 1e-5:
            for m in range(ns - ly):
                x1, y1 = m // ly, m % ly
                self.plq.append(tuple(sorted((site(x1, y1), site(x1, (y1 + 1) % ly),
                                              site(x1 + 1, (y1 + 1) % ly), site(x1 + 1, y1)))))
        # retention flags (dmrg_su2.cpp:284-303): S_j is kept while a bond
        # crosses i, S_p0 until p1 and S_p3 until p2 of a plaquette
        self.fll = np.zeros((ns, ns), dtype=int)
        self.frr = np.zeros((ns, ns), dtype=int)
        for i in range(ns):
            for j in range(i):
                if np.any(self.hmap[j, i + 1:] != 0):
                    self.fll[i, j] = 1
            for j in range(i + 1, ns):
                if np.any(self.hmap[j, :i] != 0):
                    self.frr[i, j] = 1
            for p0, p1, p2, p3 in self.plq:
                if p0 < i < p1:
                    self.fll[i, p0] = 1
                if p2 < i < p3:
                    self.frr[i, p3] = 1

    def J(self, i, j):
        return self.coup[self.hmap[i, j]] if self.hmap[i, j] else 0.0

    def bonds(self):
        return [(i, j, self.J(i, j)) for i in range(self.ns) for j in range(i + 1, self.ns)
                if self.hmap[i, j]]


# --------------------------------------------------------------------------
# fused bases
# --------------------------------------------------------------------------
class Enlarged:
    """A fused leg and where each (block sector) sits inside its sectors.

    Left (bond x site): tensor_su2::fuse(b1,b2) offsets (tensor_su2.cpp:876-897),
    b2 = site outer, b1 = bond inner -> ascending bond sector order.
    Right (site x bond): b1 = site, b2 = bond outer -> ascending bond order.
    """

    def __init__(self, block: Bond, left: bool):
        self.block = block
        self.bond = fuse_bonds(block, SITE, direction=-1) if left else \
            fuse_bonds(SITE, block, direction=-1)
        self.off = {}   # (j_block, j_fused) -> row offset inside the fused sector
        pos = {j: 0 for j in self.bond.twoj}
        for jb, db in zip(block.twoj, block.dims):
            for jf in range(abs(jb - PHYS), jb + PHYS + 1, 2):
                if jf in pos:
                    self.off[(jb, jf)] = pos[jf]
                    pos[jf] += db
        self.dims = dict(zip(self.bond.twoj, self.bond.dims))

    def parts(self, jf):
        """(j_block, offset, dim) of block sectors inside fused sector jf."""
        return [(jb, o, self.block.dims[self.block.index(jb)])
                for (jb, f), o in sorted(self.off.items(), key=lambda kv: kv[1]) if f == jf]


def _zeros_vec_op(basis_dims, transitions=None):
    out = {}
    for a, da in basis_dims.items():
        for b, db in basis_dims.items():
            if abs(a - b) <= OP <= a + b and (a + b) % 2 == 0:
                if transitions is None or (a, b) in transitions:
                    out[(a, b)] = np.zeros((da, db))
    return out


# --------------------------------------------------------------------------
# enlarged environments (dmrg_su2_omp.cpp:604-825 with uu1 = fusion isometry)
# --------------------------------------------------------------------------
def enlarge(E: Enlarged, H, S, pair_ops, left: bool):
    """Build H, carried S_k and the new-site S in the enlarged basis.

    H: scalar op on the block (or None), S: {k: vector op} to carry (k retained),
    pair_ops: list of (coef, op) summed into the bond operator that pairs with
    the new site (operator_pairup), or [] for none.
    Returns (Hhat, {k: Shat}, Snew).
    """
    fd = E.dims
    # H (x) 1 : overlap_transformation, factor 1 (tensor_su2_dmrg_src.cpp:531-688)
    Hh = {j: np.zeros((d, d)) for j, d in fd.items()}
    if H is not None:
        for jf in fd:
            for jb, o, d in E.parts(jf):
                if jb in H:
                    Hh[jf][o:o + d, o:o + d] += H[jb]
    # sum_k J S_k . s_site : operator_pairup with sigma[1] (left) / sigma[0] (right)
    if pair_ops:
        Q = heff.presum(pair_ops)
        sig = SIGMA_RIGHT if left else SIGMA_LEFT
        for jf in fd:
            for jb, o, d in E.parts(jf):
                for jbp, op_, dp in E.parts(jf):
                    blk = Q.get((jb, jbp))
                    if blk is None:
                        continue
                    f = (rc.pairup_left(jf, jb, jbp, OP, PHYS, PHYS) if left
                         else rc.pairup_right(jf, jb, jbp, OP, PHYS, PHYS))
                    if abs(f) < 1e-8:
                        continue
                    Hh[jf][o:o + d, op_:op_ + dp] += f * sig * blk
    # S_k (x) 1 : operator_transformation (tensor_su2_dmrg_src.cpp:221-402)
    Sh = {}
    for k, op in S.items():
        out = _zeros_vec_op(fd)
        for (jf, jfp), blk in out.items():
            for jb, o, d in E.parts(jf):
                for jbp, op_, dp in E.parts(jfp):
                    src = op.get((jb, jbp))
                    if src is None:
                        continue
                    f = (rc.op_transform_left(jb, jf, jbp, jfp, OP, PHYS) if left
                         else rc.op_transform_right(jb, jf, jbp, jfp, OP, PHYS))
                    if abs(f) < 1e-8:
                        continue
                    blk[o:o + d, op_:op_ + dp] += f * src
        Sh[k] = {key: v for key, v in out.items()}
    # 1 (x) s : operator_initial with sigma[0] (left) / sigma[1] (right)
    sig = SIGMA_LEFT if left else SIGMA_RIGHT
    Sn = _zeros_vec_op(fd)
    for (jf, jfp), blk in Sn.items():
        for jb, o, d in E.parts(jf):
            for jbp, op_, dp in E.parts(jfp):
                if jbp != jb:
                    continue
                f = (rc.op_initial_left(jb, jf, jfp, OP, PHYS, PHYS) if left
                     else rc.op_initial_right(jb, jf, jfp, OP, PHYS, PHYS))
                if abs(f) < 1e-8:
                    continue
                blk[o:o + d, op_:op_ + dp] += f * sig * np.eye(d)
    return Hh, Sh, Sn


def project(P, H, S):
    """Truncated-basis operators U^T O U (P: fused sector -> [D_E x D_new])."""
    Ht = None if H is None else {j: P[j].T @ H[j] @ P[j] for j in P}
    St = {}
    for k, op in S.items():
        St[k] = {(a, b): P[a].T @ blk @ P[b] for (a, b), blk in op.items() if a in P and b in P}
    return Ht, St


# --------------------------------------------------------------------------
# truncated SVD (tensor.cpp:1472-1700, tensor_su2.cpp:1619-1758)
# --------------------------------------------------------------------------
def fix_signs(v):
    """Make the first entry with |v_i| >= max|v|/2 of every column positive.
    dsyev_ leaves eigenvector signs to the LAPACK build; this convention
    (robust to the equal-magnitude entries symmetric lattices produce) makes
    the all-ones warm-up start vector (tensor_su2.cpp:2367) and hence the
    whole trajectory reproducible across eigensolvers."""
    if v.size == 0:
        return v
    a = np.abs(v)
    idx = np.argmax(a >= 0.5 * a.max(axis=0), axis=0)
    sg = np.sign(v[idx, np.arange(v.shape[1])])
    sg[sg == 0] = 1.0
    return v * sg


def _gram_eig(G):
    """Gram eigenpairs with the reference's refinement of small eigenvalues.

    tensor.cpp:1498-1529: eigenpairs with w/w_max < 0.1 are re-diagonalised in
    the projected Gram V^T G V, repeatedly, until one such pair remains.
    """
    n = G.shape[0]
    w, v = np.linalg.eigh(G)
    npp = 0
    # a zero Gram matrix gives 0/0: NaN compares false, as in the reference's
    # C loop, and no pair is refined
    with np.errstate(invalid="ignore", divide="ignore"):
        for i in range(n - 1, -1, -1):
            if w[i] / w[n - 1] < 0.1:
                npp = i + 1
                break
    if npp > 1:
        ap = G.copy()
        vp = v.copy()
        np_ = n
        while True:
            app = vp[:np_, :npp].T @ ap[:np_, :np_] @ vp[:np_, :npp]
            wsub, vpp = np.linalg.eigh(app)
            w[:npp] = wsub
            v[:, :npp] = v[:, :npp] @ vpp
            ap = app
            vp = vpp
            j = 0
            for i in range(npp - 1, -1, -1):
                if wsub[i] / wsub[npp - 1] < 0.1:
                    j = i + 1
                    break
            np_ = npp
            npp = j
            if npp <= 1:
                break
    return w, fix_signs(v)


def sector_svd(A, p1, p2, max_dcut, tol=1e-100):
    """tensor::svd: returns (Us^p1 [m x dc], Vs^p2 [n x dc], s ascending [dc])."""
    m, n = A.shape
    if m >= n:
        w, v = _gram_eig(A.T @ A)
        u = A @ v
        num = w[-1]
        dc = 0
        s = np.zeros(n)
        if num > tol:
            for i in range(n - 1, -1, -1):
                if w[i] / num > tol:
                    s[i] = math.sqrt(w[i])
                    dc += 1
        dc = min(dc, max_dcut, m)
        for i in range(n - dc, n):
            if s[i] > 0:
                u[:, i] *= s[i] ** (p1 - 1.0)
                v[:, i] *= s[i] ** p2
            else:
                u[:, i] = 0
                v[:, i] = 0
        return u[:, n - dc:], v[:, n - dc:], s[n - dc:]
    w, u = _gram_eig(A @ A.T)
    v = A.T @ u
    num = w[-1]
    dc = 0
    s = np.zeros(m)
    if num > tol:
        for i in range(m - 1, -1, -1):
            if w[i] / num > tol:
                s[i] = math.sqrt(w[i])
                dc += 1
    dc = min(dc, max_dcut, n)
    for i in range(m - dc, m):
        if s[i] > 0:
            v[:, i] *= s[i] ** (p2 - 1.0)
            u[:, i] *= s[i] ** p1
        else:
            v[:, i] = 0
            u[:, i] = 0
    return u[:, m - dc:], v[:, m - dc:], s[m - dc:]


def truncated_svd(psi, p1, p2, max_dcut, info=None):
    """tensor_su2::svd across sectors: global shell-sorted cut to max_dcut.

    Returns (U {j: [D_E x d_j]}, V {j: [D_F x d_j]}, new bond, kept values per
    sector ascending).  info (a dict, optional) receives "discarded", the sum
    of the squared singular values that the cut drops, and "wout", the
    reference's output array: the kept values sector by sector, then the cut
    ones sector by sector, each ascending (tensor_su2.cpp:1707-1731).
    """
    secs = sorted(psi)
    per = {}
    wsort, isort = [], []
    for i, j in enumerate(secs):
        u, v, s = sector_svd(psi[j], p1, p2, max_dcut)
        per[j] = (u, v, s)
        wsort += list(s)
        isort += [i] * len(s)
    w_sorted, i_sorted = dsort2(wsort, isort)
    n = len(w_sorted)
    k0 = 0 if max_dcut > n else n - max_dcut
    keep = {j: 0 for j in secs}
    for t in range(n - 1, k0 - 1, -1):
        keep[secs[i_sorted[t]]] += 1
    if info is not None:
        info["discarded"] = float(np.sum(np.square(w_sorted[:k0])))
    if info is not None:
        head, tail = [], []
        for j in secs:
            s_j = per[j][2]
            cut = len(s_j) - keep[j]
            head += list(s_j[cut:])
            tail += list(s_j[:cut])
        info["wout"] = np.array(head + tail)
    U, V, kept = {}, {}, {}
    tw, td = [], []
    for j in secs:
        u, v, s = per[j]
        d = keep[j]
        if d > 0:
            U[j] = u[:, len(s) - d:]
            V[j] = v[:, len(s) - d:]
            kept[j] = s[len(s) - d:]
            tw.append(j)
            td.append(d)
    bond = Bond(tw, td, 1, sort=False)
    return U, V, bond, kept


# --------------------------------------------------------------------------
# Lanczos (lanczos_su2.cpp:125-195, 237-338)
# --------------------------------------------------------------------------
def lanczos(apply, v0, mlanc=MLANC, tol=1e-6):
    """Returns (lowest Ritz value, normalised Ritz vector, number of matvecs)."""
    nmv = 0
    nrm = heff.norm(v0)
    ff = [heff.scal(1.0 / nrm, v0)]
    aal = [0.0] * mlanc
    nnl = [0.0] * (mlanc + 1)

    def hv(x):
        nonlocal nmv
        nmv += 1
        return apply(x)

    def ritz(m):
        d = np.array(aal[:m])
        e = np.array(nnl[1:m])
        w, z = eigh_tridiagonal(d, e) if m > 1 else (d.copy(), np.ones((1, 1)))
        if z[0, 0] < 0:          # Ritz vector with a positive overlap on the start vector
            z = -z
        vec = heff.scal(z[0, 0], ff[0])
        for j in range(1, m):
            vec = heff.axpy(z[j, 0], ff[j], vec)
        return w[0], vec

    def check(vec):
        nr = heff.norm(vec)
        vec = heff.scal(1.0 / nr, vec)
        hv_ = hv(vec)
        prod = heff.norm(hv_)
        ov = heff.inner(hv_, vec) / prod
        return vec, prod, abs(ov)

    f1 = hv(ff[0])
    aal[0] = heff.inner(ff[0], f1)
    f1 = heff.axpy(-aal[0], ff[0], f1)
    nnl[1] = heff.norm(f1)
    # an exactly invariant start vector (a 1-dim H_eff) leaves nnl = 0: the
    # unnormalised residual is then never read again
    ff.append(heff.scal(1.0 / nnl[1], f1) if nnl[1] > 0 else f1)
    if mlanc == 2:
        if nnl[1] < tol:
            e, vec = ritz(1)
            return e, vec, nmv
        f2 = hv(ff[1])
        aal[1] = heff.inner(ff[1], f2)
        e, vec = ritz(2)
        return e, heff.scal(1.0 / heff.norm(vec), vec), nmv
    for m in range(2, mlanc):
        fm = hv(ff[m - 1])
        aal[m - 1] = heff.inner(ff[m - 1], fm)
        fm = heff.axpy(-aal[m - 1], ff[m - 1], fm)
        fm = heff.axpy(-nnl[m - 1], ff[m - 2], fm)
        for j in range(m):
            q = heff.inner(fm, ff[j])
            fm = heff.axpy(-q, ff[j], fm)
        nnl[m] = heff.norm(fm)
        ff.append(heff.scal(1.0 / nnl[m], fm) if nnl[m] > 0 else fm)
        if nnl[m] < tol:
            e, vec = ritz(m)
            vec, _, _ = check(vec)
            return e, vec, nmv
        if m % 4 == 0:
            e, vec = ritz(m)
            vec, _, ov = check(vec)
            if abs(ov - 1.0) < 1e-8:
                return e, vec, nmv
    fm = hv(ff[mlanc - 1])
    aal[mlanc - 1] = heff.inner(ff[mlanc - 1], fm)
    e, vec = ritz(mlanc)
    vec, _, _ = check(vec)
    return e, vec, nmv


# --------------------------------------------------------------------------
# the DMRG driver
# --------------------------------------------------------------------------
def _ident(dims):
    return {j: np.eye(d) for j, d in dims.items()}


def _bond_of(P):
    return Bond(list(P), [P[j].shape[1] for j in P], 1, sort=False)


class LowerState:
    """A converged lower state for excited-state targeting (exci > 0).

    Its site tensors are kept once as blocks T[k][(j_a, j_b)] between the bond
    left of site k (beta[k-1]) and the bond right of it (beta[k]); the
    reference holds them as orth[j][k] and re-labels them between the left and
    right forms (tensor_su2.cpp:749-836), which for spin-1/2 sites is a block
    transpose (the re-coupling factor is 1).  `energy` is the Ritz value the
    penalty uses (gs_enr[j], read back with the 1000 added, dmrg_su2_omp.cpp:
    1100-1152).
    """

    def __init__(self, ns, beta, T, energy):
        self.ns = ns
        self.beta = beta          # k -> Bond, k = -1 .. ns-1
        self.T = T                # k -> {(ja, jb): [d(ja) x d(jb)]}
        self.energy = energy

    @classmethod
    def from_dmrg(cls, dm):
        """The state a finished oracle run holds (left sites U, right sites V)."""
        ns = dm.ns
        beta, T = {}, {}
        for k in range(-1, ns // 2):
            beta[k] = dm.lb[k]
        for k in range(ns // 2, ns + 1):
            beta[k - 1] = dm.rb[k]
        for k in range(ns):
            T[k] = {}
            if k < ns // 2:
                E = Enlarged(dm.lb[k - 1], left=True)
                for jb, P in dm.U[k].items():
                    for ja, o, d in E.parts(jb):
                        T[k][(ja, jb)] = P[o:o + d, :].copy()
            else:
                F = Enlarged(dm.rb[k + 1], left=False)
                for ja, P in dm.V[k].items():
                    for jb, o, d in F.parts(ja):
                        T[k][(ja, jb)] = P[o:o + d, :].T.copy()
        return cls(ns, beta, T, dm.energies[-1])

    def left_view(self, k):
        """Projector E = fuse(beta[k-1], site) -> beta[k] as {jb: [D_E x d]}."""
        E = Enlarged(self.beta[k - 1], left=True)
        out = {}
        for jb, db in zip(self.beta[k].twoj, self.beta[k].dims):
            if jb not in E.dims:
                continue
            P = np.zeros((E.dims[jb], db))
            for ja, o, d in E.parts(jb):
                blk = self.T[k].get((ja, jb))
                if blk is not None:
                    P[o:o + d, :] = blk
            out[jb] = P
        return E, out

    def right_view(self, k):
        """Projector F = fuse(site, beta[k]) -> beta[k-1] as {ja: [D_F x d]}."""
        F = Enlarged(self.beta[k], left=False)
        out = {}
        for ja, da in zip(self.beta[k - 1].twoj, self.beta[k - 1].dims):
            if ja not in F.dims:
                continue
            P = np.zeros((F.dims[ja], da))
            for jb, o, d in F.parts(ja):
                blk = self.T[k].get((ja, jb))
                if blk is not None:
                    P[o:o + d, :] = blk.T
            out[ja] = P
        return F, out


class DMRG:
    """MPS projectors + environments for one lattice and target spin 2S.

    lb[i]: bond right of site i (left part, lb[-1] = {0:1}); rb[i]: bond left
    of site i (right part, rb[ns] = {2S:1}, the auxiliary spin of
    dmrg_su2_omp.cpp:38-42).  U[i]: E_i -> lb[i]; V[i]: F_i -> rb[i].
    Lenv[i]/Renv[i]: (H, {k: S_k}, {(plaquette, channel): {K: op}}) in lb[i]/rb[i].
    """

    def __init__(self, lat: Lattice, totspin=0, max_dcut=200, verbose=False, lower=()):
        self.lat = lat
        self.ns = lat.ns
        self.S = totspin
        self.D = max_dcut
        self.verbose = verbose
        self.lb = {-1: Bond([0], [1], 1)}
        self.rb = {self.ns: Bond([totspin], [1], 1)}
        self.Lenv = {-1: (None, {}, {})}
        self.Renv = {self.ns: (None, {}, {})}
        self.U, self.V = {}, {}
        self.energies = []
        self.history = []          # per step: (kept multiplets, discarded weight, matvecs)
        self.nmatvec = 0
        self.psi = None
        self.center = None
        # excited states: overlap environments with each lower state
        # (ovlp[j][i] of dmrg_su2_omp.cpp:319-601), {j_lower x j_current}
        self.lower = list(lower)
        self.OL = [{-1: {0: np.ones((1, 1))}} for _ in self.lower]
        self.OR = [{self.ns: {totspin: np.ones((1, 1))}} for _ in self.lower]

    # ---- environments (dmrg_su2_omp.cpp:319-825) --------------------------
    def enlarged_left(self, i):
        E = Enlarged(self.lb[i - 1], left=True)
        H, S, P = self.Lenv[i - 1]
        keep = {k: S[k] for k in S if self.lat.fll[i, k]}
        pair = [(self.lat.J(i, k), S[k]) for k in range(i) if self.lat.hmap[i, k] and k in S]
        Hh, Sh, Sn = enlarge(E, H, keep, pair, left=True)
        Sh[i] = Sn
        Ph, hadd = plaquette.grow(self.lat, E, i, P, S, True, SIGMA_LEFT, SIGMA_RIGHT)
        if hadd:
            for (j, _), blk in hadd.items():
                Hh[j] = Hh[j] + blk
        return E, Hh, Sh, Ph

    def enlarged_right(self, i):
        E = Enlarged(self.rb[i + 1], left=False)
        H, S, P = self.Renv[i + 1]
        keep = {k: S[k] for k in S if self.lat.frr[i, k]}
        pair = [(self.lat.J(i, k), S[k]) for k in range(i + 1, self.ns)
                if self.lat.hmap[i, k] and k in S]
        Hh, Sh, Sn = enlarge(E, H, keep, pair, left=False)
        Sh[i] = Sn
        Ph, hadd = plaquette.grow(self.lat, E, i, P, S, False, SIGMA_RIGHT, SIGMA_LEFT)
        if hadd:
            for (j, _), blk in hadd.items():
                Hh[j] = Hh[j] + blk
        return E, Hh, Sh, Ph

    def push_left(self, i, U=None):
        """Truncated left environment after site i (prepare_site_operator_from_left(i,flag))."""
        E, Hh, Sh, Ph = self.enlarged_left(i)
        if U is None:
            U = _ident(E.dims)
        self.U[i] = U
        self.lb[i] = _bond_of(U)
        self.Lenv[i] = project(U, Hh, Sh) + (plaquette.project(U, Ph),)
        for j, st in enumerate(self.lower):
            # overlap_transformation (tensor_su2_dmrg_src.cpp:531-688): the
            # lower state's site tensor against this run's, through ovlp[i-1]
            Ej, Pj = st.left_view(i)
            prev = self.OL[j][i - 1]
            out = {}
            for jb, P in U.items():
                if jb not in Pj:
                    continue
                M = np.zeros((Pj[jb].shape[1], P.shape[1]))
                for ja, o, d in E.parts(jb):
                    if ja not in prev or (ja, jb) not in Ej.off:
                        continue
                    oj = Ej.off[(ja, jb)]
                    dj = prev[ja].shape[0]
                    M += Pj[jb][oj:oj + dj, :].T @ prev[ja] @ P[o:o + d, :]
                out[jb] = M
            self.OL[j][i] = out

    def push_right(self, i, V=None):
        E, Hh, Sh, Ph = self.enlarged_right(i)
        if V is None:
            V = _ident(E.dims)
        self.V[i] = V
        self.rb[i] = _bond_of(V)
        self.Renv[i] = project(V, Hh, Sh) + (plaquette.project(V, Ph),)
        for j, st in enumerate(self.lower):
            Fj, Pj = st.right_view(i)
            prev = self.OR[j][i + 1]
            out = {}
            for jr, P in V.items():
                if jr not in Pj:
                    continue
                M = np.zeros((Pj[jr].shape[1], P.shape[1]))
                for jq, o, d in E.parts(jr):
                    if jq not in prev or (jq, jr) not in Fj.off:
                        continue
                    oj = Fj.off[(jq, jr)]
                    dj = prev[jq].shape[0]
                    M += Pj[jr][oj:oj + dj, :].T @ prev[jq] @ P[o:o + d, :]
                out[jr] = M
            self.OR[j][i] = out

    # ---- one two-site solve ---------------------------------------------
    def terms(self, il, ir, Sl, Sr):
        """H_eff term list of dmrg_su2.cpp:571-610 (couplings presummed once)."""
        out = []
        if il < self.ns // 2:
            for j in range(il + 1):
                if j not in Sl:
                    continue
                right = [(self.lat.J(j, k), Sr[k]) for k in range(ir, self.ns)
                         if self.lat.hmap[j, k] and k in Sr]
                if right:
                    out.append((Sl[j], heff.presum(right)))
        else:
            for k in range(ir, self.ns):
                if k not in Sr:
                    continue
                left = [(self.lat.J(j, k), Sl[j]) for j in range(il + 1)
                        if self.lat.hmap[j, k] and j in Sl]
                if left:
                    out.append((heff.presum(left), Sr[k]))
        return out

    def step(self, il, ir, psi0, p1, p2):
        El, Hl, Sl, Pl = self.enlarged_left(il)
        Er, Hr, Sr, Pr = self.enlarged_right(ir)
        lay = {j: (El.dims[j], Er.dims[j]) for j in El.dims if j in Er.dims}
        if psi0 is None:
            psi0 = {j: np.ones(d) for j, d in lay.items()}
        else:
            # The vector carries every sector of the two-site space: a block the
            # carried-over start vector lacks is null there, and the reference's
            # operator+= (tensor_su2.cpp:157-190) fills it from H_eff's output
            # as soon as a term reaches it (tensor_su2_dmrg_src.cpp:921-928 builds
            # that output on the operators' full legs).
            psi0 = {j: psi0[j] if j in psi0 else np.zeros(d) for j, d in lay.items()}
        terms = self.terms(il, ir, Sl, Sr)
        terms += plaquette.heff_terms(self.lat, il, ir, Sl, Sr, Pl, Pr)
        self.last_terms = (Hl, Hr, terms)

        phis = self.lower_vectors(il, ir, El, Er, lay) if il + 1 == ir else []

        def apply(v):
            out = heff.heff_apply(v, Hl, Hr, terms)
            # - sum_j E_j <phi_j|v> phi_j (dmrg_su2.cpp:611-623)
            for e_j, phi in phis:
                c = -e_j * heff.inner(phi, v)
                out = heff.axpy(c, phi, out)
            return out

        nele = sum(a * b for a, b in lay.values())
        e, psi, nmv = lanczos(apply, psi0, mlanc=min(MLANC, nele + 1))
        self.nmatvec += nmv
        info = {}
        U, V, bond, kept = truncated_svd(psi, p1, p2, self.D, info)
        self.energies.append(e)
        self.history.append((sum(bond.dims), info["discarded"], nmv))
        self.wout = info["wout"]
        self.psi = psi
        self.center = (il, ir)
        self.kept = kept
        if self.verbose:
            print(f"step {il},{ir} E={e:.12f} D={sum(bond.dims)}")
        return e, U, V, bond, El, Er

    def lower_vectors(self, il, ir, El, Er, lay):
        """Each lower state in the current two-site basis (overlapvec of
        prepare_input_vector, dmrg_su2_omp.cpp:229-254): its tensors at il and
        ir mapped through the overlap environments, contracted on their bond."""
        out = []
        for j, st in enumerate(self.lower):
            Ej, A = st.left_view(il)
            Fj, B = st.right_view(ir)
            L, R = self.OL[j][il - 1], self.OR[j][ir + 1]
            phi = {}
            for jf, (da, db) in lay.items():
                if jf not in A or jf not in B:
                    phi[jf] = np.zeros((da, db))
                    continue
                At = np.zeros((da, A[jf].shape[1]))
                for ja, o, d in El.parts(jf):
                    if ja in L and (ja, jf) in Ej.off:
                        oj = Ej.off[(ja, jf)]
                        At[o:o + d, :] = L[ja].T @ A[jf][oj:oj + L[ja].shape[0], :]
                Bt = np.zeros((db, B[jf].shape[1]))
                for jq, o, d in Er.parts(jf):
                    if jq in R and (jq, jf) in Fj.off:
                        oj = Fj.off[(jq, jf)]
                        Bt[o:o + d, :] = R[jq].T @ B[jf][oj:oj + R[jq].shape[0], :]
                phi[jf] = At @ Bt.T
            out.append((st.energy, phi))
        return out

    # ---- Lanczos start vectors carried across one site --------------------
    def start_left(self, il, A, b):
        """psi0 at (il-1, il) from the left factor A (E_il x b) of (il, il+1)."""
        E = Enlarged(self.lb[il - 1], left=True)      # E_il parts: lb[il-1] sectors
        F = Enlarged(b, left=False)                   # F_il = site x b
        P = self.U[il - 1]                            # E_{il-1} -> lb[il-1]
        out = {}
        for jl in self.lb[il - 1].twoj:
            if jl not in F.dims or jl not in P:
                continue
            dl = self.lb[il - 1].dims[self.lb[il - 1].index(jl)]
            blk = np.zeros((dl, F.dims[jl]))
            for jb, o, d in F.parts(jl):
                if jb not in A or (jl, jb) not in E.off:
                    continue
                oe = E.off[(jl, jb)]
                f = rc.right_to_left_psi(jl, PHYS, jb)
                blk[:, o:o + d] = f * A[jb][oe:oe + dl, :]
            out[jl] = P[jl] @ blk
        return out

    def start_right(self, ir, B, b):
        """psi0 at (ir, ir+1) from the right factor B (F_ir x b) of (ir-1, ir)."""
        F = Enlarged(self.rb[ir + 1], left=False)     # F_ir parts: rb[ir+1] sectors
        E = Enlarged(b, left=True)                    # E_ir = b x site
        P = self.V[ir + 1]                            # F_{ir+1} -> rb[ir+1]
        out = {}
        for jr in self.rb[ir + 1].twoj:
            if jr not in E.dims or jr not in P:
                continue
            dr = self.rb[ir + 1].dims[self.rb[ir + 1].index(jr)]
            blk = np.zeros((E.dims[jr], dr))
            for jb, o, d in E.parts(jr):
                if jb not in B or (jr, jb) not in F.off:
                    continue
                of = F.off[(jr, jb)]
                f = rc.left_to_right_psi(jb, PHYS, jr)
                blk[o:o + d, :] = f * B[jb][of:of + dr, :].T
            out[jr] = blk @ P[jr].T
        return out

    # ---- warm-up (dmrg_su2_omp.cpp:20-95) -------------------------------
    def warmup(self):
        ns = self.ns
        last = None
        for i in range(0, ns // 2 - 1):
            if i not in self.Lenv:
                self.push_left(i)           # isometry: no solve reached site i yet
            r = ns - 1 - i
            if r not in self.Renv:
                self.push_right(r)
            il, ir = i + 1, ns - 2 - i
            solve = ((i + 2) * 2 * PHYS >= self.S and i >= 1) or i == ns // 2 - 2
            if not solve:
                continue
            final = i == ns // 2 - 2
            e, U, V, bond, El, Er = self.step(il, ir, None, 1 if final else 0, 0)
            if final:
                self.U[il] = U               # carries the singular values
                self.V[ir] = V
                self.lb[il] = bond
                self.rb[ir] = bond
            else:
                self.push_left(il, U)
                self.push_right(ir, V)
            last = e
        return last

    # ---- one sweep (dmrg_su2_omp.cpp:98-157) ----------------------------
    def sweep(self):
        ns = self.ns
        nf = NFREE
        # pass 1: centre -> left edge
        il, ir = ns // 2 - 1, ns // 2
        A = self.U[il]
        bond = self.lb[il]
        for i in range(ns // 2, nf + 1, -1):
            il, ir = i - 2, i - 1
            self.push_right(ir + 1, self.V[ir + 1])          # prepare_site_operator_from_right(i,1)
            psi0 = self.start_left(ir, A, bond)
            p1, p2 = (1, 0) if i > nf + 2 else (0, 1)
            e, U, V, bond, El, Er = self.step(il, ir, psi0, p1, p2)
            self.V[ir] = V
            self.rb[ir] = bond
            self.lb[il] = bond
            if p1 == 1:
                A = U
                self.U[il] = U
            else:
                self.push_left(il, U)
                B = V
        # pass 2: left edge -> right edge
        for i in range(nf, ns - nf - 2):
            il, ir = i + 1, i + 2
            self.push_left(il - 1, self.U[il - 1])             # prepare_site_operator_from_left(i,1)
            psi0 = self.start_right(il, B, bond)
            p1, p2 = (0, 1) if i < ns - nf - 3 else (1, 0)
            e, U, V, bond, El, Er = self.step(il, ir, psi0, p1, p2)
            self.U[il] = U
            self.lb[il] = bond
            self.rb[ir] = bond
            if p2 == 1:
                B = V
                self.V[ir] = V
            else:
                self.push_right(ir, V)
                A = U
        # pass 3: right edge -> centre
        for i in range(ns - nf - 1, ns // 2, -1):
            il, ir = i - 2, i - 1
            self.push_right(ir + 1, self.V[ir + 1])
            psi0 = self.start_left(ir, A, bond)
            e, U, V, bond, El, Er = self.step(il, ir, psi0, 1, 0)
            self.V[ir] = V
            self.rb[ir] = bond
            self.lb[il] = bond
            self.U[il] = U
            A = U
        return self.energies[-1]

    def run(self, nsweep=5, stop=1e-6):
        """Warm-up then sweeps; stops when |dE| < stop (main_omp.cpp:184-202)."""
        e_prev = self.warmup()
        per_sweep = []
        for _ in range(nsweep):
            e = self.sweep()
            per_sweep.append(e)
            if stop is not None and abs(e - e_prev) < stop:
                break
            e_prev = e
        return per_sweep
