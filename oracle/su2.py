"""SU(2) structure, Clebsch-Gordan builders and recoupling factors (oracle).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.

Spins are doubled everywhere (``twoj`` = 2j), as in the reference
(``tensor.cpp:1355``).  Dense tensors are numpy arrays whose axis ``i`` is
leg ``i``; the reference stores the same tensors column-major with leg 0
fastest (``tensor.cpp:1101-1119``), which only matters at the C boundary.

Pins (tests/test_oracle_su2.py):
  * ``isort2``/``dsort2``: against sorted() on random inputs plus the
    stability pattern a shell sort implies (ties keep a deterministic order
    that is checked against a literal trace of the gap sequence).
  * ``Bond.fuse``: closed-form angular momentum addition with multiplicity
    counting, brute force over m-states for tiny legs.
  * ``cgc``/``singlet``: sympy ``clebsch_gordan`` exact values (<=1e-14),
    orthonormality.
  * ``fac_hamilt_vec``: dense brute-force contraction of the full H_eff
    term (reduced blocks tensored with their CG tensors) projected onto the
    singlet, tests/test_oracle_heff.py.
"""
from __future__ import annotations

import math
from functools import lru_cache

import numpy as np

FUSE_CAP = 38      # su2bond.cpp:175-177 caps the fused 2j at 38
MAX_ANGM = 25      # dmrg_su2.cpp:46 (tables hold 2j < 25)


# --------------------------------------------------------------------------
# shell sorts (sort2.f)
# --------------------------------------------------------------------------
def isort2(a, b):
    """Shell sort of integer keys ``a`` carrying ``b`` (sort2.f:38-73).

    Gap sequence N/2, N/4, ..., 1 with a strict ``>`` exchange test.
    Returns new lists.
    """
    a = list(a)
    b = list(b)
    n = len(a)
    gap = n // 2
    while gap > 0:
        for i in range(gap, n):
            j = i - gap
            while True:
                if a[j] > a[j + gap]:
                    a[j], a[j + gap] = a[j + gap], a[j]
                    b[j], b[j + gap] = b[j + gap], b[j]
                else:
                    break
                if j >= gap:
                    j -= gap
                else:
                    break
        gap //= 2
    return a, b


def dsort2(w, idx):
    """Shell sort of doubles ``w`` carrying int ``idx`` (sort2.f:76-111).

    The order of equal keys decides which sector keeps a multiplet at the
    truncation cut (tensor_su2.cpp:1686-1695), so the exact gap sequence and
    the strict ``>`` test are kept.
    """
    w = [float(x) for x in w]
    idx = list(idx)
    n = len(w)
    gap = n // 2
    while gap > 0:
        for i in range(gap, n):
            j = i - gap
            while True:
                if w[j] > w[j + gap]:
                    w[j], w[j + gap] = w[j + gap], w[j]
                    idx[j], idx[j + gap] = idx[j + gap], idx[j]
                else:
                    break
                if j >= gap:
                    j -= gap
                else:
                    break
        gap //= 2
    return w, idx


# --------------------------------------------------------------------------
# legs (su2bond.cpp) and block structure (su2struct.cpp)
# --------------------------------------------------------------------------
class Bond:
    """One leg: sectors (2j, multiplicity) ascending in 2j, and a direction.

    Mirrors ``su2bond`` (su2bond.hpp:3-29); ``cgcdim`` is 2j+1.
    """

    __slots__ = ("twoj", "dims", "direction")

    def __init__(self, twoj, dims, direction=1, sort=True):
        if sort:  # su2bond::set_su2bond sorts with isort2 (su2bond.cpp:59-74)
            twoj, dims = isort2(twoj, dims)
        self.twoj = list(int(x) for x in twoj)
        self.dims = list(int(x) for x in dims)
        self.direction = int(direction)

    def __eq__(self, other):
        # su2bond::operator!= compares direction, labels and dims (91-101)
        return (self.direction == other.direction and self.twoj == other.twoj
                and self.dims == other.dims)

    def __repr__(self):
        return f"Bond(dir={self.direction}, {list(zip(self.twoj, self.dims))})"

    @property
    def nsec(self):
        return len(self.twoj)

    def index(self, twoj):
        """su2bond::get_angularmoment_index (su2bond.cpp:142-152)."""
        return self.twoj.index(twoj)

    def inverted(self):
        return Bond(self.twoj, self.dims, -self.direction, sort=False)

    def total(self):
        return sum(self.dims)

    def offsets(self):
        off, o = [], 0
        for d in self.dims:
            off.append(o)
            o += d
        return off


def fuse_bonds(b1: Bond, b2: Bond, direction=None) -> Bond:
    """su2bond::fuse(b1,b2) (su2bond.cpp:155-212) and fuse(b1,b2,dir) (215-260).

    Labels appear in first-seen order over (b1 sector, b2 sector, 2j from
    |j1-j2| to min(j1+j2, 38)), dims accumulate d1*d2, then isort2.  Without
    ``direction`` the two legs must share a direction and the result points
    the other way.
    """
    if direction is None:
        if b1.direction != b2.direction:
            raise ValueError("su2bond::fuse two bonds of different directions")
        direction = -b1.direction
    ang, dim = [], []
    for m1, d1 in zip(b1.twoj, b1.dims):
        for m2, d2 in zip(b2.twoj, b2.dims):
            lo = abs(m1 - m2)
            hi = min(m1 + m2, FUSE_CAP)
            for k in range(lo, hi + 1, 2):
                if k in ang:
                    dim[ang.index(k)] += d1 * d2
                else:
                    ang.append(k)
                    dim.append(d1 * d2)
    return Bond(ang, dim, direction, sort=True)


def check_angularmoments(labels, locspin=0):
    """su2struct::check_angularmoments (su2struct.cpp:264-289).

    A label tuple is allowed when sequential fusion (with the 38 cap) of its
    legs reaches ``locspin``; a single leg must equal it.
    """
    if len(labels) == 1:
        return labels[0] == locspin
    acc = Bond([labels[0]], [1], 1)
    for lab in labels[1:]:
        acc = fuse_bonds(acc, Bond([lab], [1], 1))
        acc.direction = 1
    return locspin in acc.twoj


def block_index(bonds, labels):
    """su2struct::get_tensor_index (su2struct.cpp:243-261): leg 0 fastest."""
    idx = 0
    for b, lab in zip(reversed(bonds), reversed(labels)):
        idx = idx * b.nsec + b.index(lab)
    return idx


def block_labels(bonds, idx):
    """su2struct::get_tensor_argument label decoding (su2struct.cpp:264-275)."""
    labs = []
    for b in bonds:
        labs.append(b.twoj[idx % b.nsec])
        idx //= b.nsec
    return tuple(labs)


def valid_blocks(bonds, locspin=0):
    """All allowed label tuples in block-index order."""
    n = 1
    for b in bonds:
        n *= b.nsec
    out = []
    for i in range(n):
        labs = block_labels(bonds, i)
        if check_angularmoments(labs, locspin):
            out.append(labs)
    return out


# --------------------------------------------------------------------------
# CG tensors (tensor.cpp)
# --------------------------------------------------------------------------
def _fact(n):
    return float(math.factorial(n))


@lru_cache(maxsize=None)
def cgc(j1, j2, j3):
    """<j1 m1; j2 m2 | j3 m3> as a dense [j1+1, j2+1, j3+1] array.

    Racah's closed form with doubled spins, index i <-> m = 2i - j
    (tensor.cpp:1353-1407).  Returns a read-only array.
    """
    t = np.zeros((j1 + 1, j2 + 1, j3 + 1))
    if not (abs(j1 - j2) <= j3 <= j1 + j2) or (j1 + j2 + j3) % 2:
        t.setflags(write=False)
        return t
    f1 = math.sqrt((j3 + 1) * _fact((j3 + j1 - j2) // 2) * _fact((j3 - j1 + j2) // 2)
                   * _fact((j1 + j2 - j3) // 2) / _fact((j1 + j2 + j3) // 2 + 1))
    for i in range(j1 + 1):
        m1 = 2 * i - j1
        for j in range(j2 + 1):
            m2 = 2 * j - j2
            m3 = m1 + m2
            if abs(m3) > j3:
                continue
            k = (m3 + j3) // 2
            f2 = math.sqrt(_fact((j1 + m1) // 2) * _fact((j1 - m1) // 2)
                           * _fact((j2 + m2) // 2) * _fact((j2 - m2) // 2)
                           * _fact((j3 + m3) // 2) * _fact((j3 - m3) // 2))
            k1 = (j1 + j2 - j3) // 2
            k2 = (j1 - m1) // 2
            k3 = (j2 + m2) // 2
            k4 = (j3 - j2 + m1) // 2
            k5 = (j3 - j1 - m2) // 2
            s = 0.0
            for z in range((j1 + j2 + j3) // 2 + 1):
                if k1 - z >= 0 and k2 - z >= 0 and k3 - z >= 0 and k4 + z >= 0 and k5 + z >= 0:
                    s += (-1.0) ** z / (_fact(z) * _fact(k1 - z) * _fact(k2 - z)
                                        * _fact(k3 - z) * _fact(k4 + z) * _fact(k5 + z))
            t[i, j, k] = f1 * f2 * s
    t.setflags(write=False)
    return t


@lru_cache(maxsize=None)
def singlet(j):
    """(-1)^((j-m)/2)/sqrt(j+1) at m1+m2=0 (tensor.cpp:1327-1350)."""
    t = np.zeros((j + 1, j + 1))
    for i in range(j + 1):
        m1 = 2 * i - j
        t[i, j - i] = (1.0 if ((j - m1) // 2) % 2 == 0 else -1.0) / math.sqrt(j + 1)
    t.setflags(write=False)
    return t


@lru_cache(maxsize=None)
def identity(j):
    """tensor.cpp:1310-1324."""
    t = np.eye(j + 1)
    t.setflags(write=False)
    return t


def contract(t1, i1, t2, i2):
    """Dense leg contraction with the reference's output leg order.

    Output legs are t1's legs after ``i1`` (cyclic) followed by t2's legs
    after ``i2`` (cyclic) -- tensor.cpp:521-583 (shift to front/back, one
    dgemm).
    """
    n1, n2 = t1.ndim, t2.ndim
    o1 = [(i1 + 1 + k) % n1 for k in range(n1 - 1)]
    o2 = [(i2 + 1 + k) % n2 for k in range(n2 - 1)]
    return np.tensordot(t1.transpose(o1 + [i1]), t2.transpose([i2] + o2),
                        axes=([n1 - 1], [0]))


def is_proportional_to(t, ref, tol=1e-12):
    """Ratio test of tensor.cpp:1231-1269; returns the ratio or None."""
    if t.shape != ref.shape:
        return None
    if np.all(np.abs(t) <= tol):
        return 0.0
    flat_r = ref.ravel(order="F")
    flat_t = t.ravel(order="F")
    pos = int(np.argmax(np.abs(flat_r)))   # first max in column-major order
    if abs(flat_r[pos]) <= tol:
        raise ValueError("is_proportional_to comparing to zero tensor")
    rat = flat_t[pos] / flat_r[pos]
    for a, b in zip(flat_t, flat_r):
        za, zb = abs(a) <= tol, abs(b) <= tol
        if za and zb:
            continue
        if za != zb:
            return None
        if abs(a / b - rat) > tol:
            return None
    return float(rat)


@lru_cache(maxsize=None)
def fac_hamilt_vec(i, j, k):
    """H_eff recoupling scalar for an (ket 2j=i, operator 2j=j, bra 2j=k) term.

    Restates dmrg_su2.cpp:759-775: every leg of CG(i,j,k) is turned with a
    singlet (contracting the CG leg with the singlet's second leg), and the
    factor is <CG, turned CG> * sqrt(j+1).  Zero outside the triangle or when
    i and k differ in parity.
    """
    if not (abs(i - j) <= k <= i + j) or (i % 2) != (k % 2):
        return 0.0
    c = cgc(i, j, k)
    t1 = contract(c, 0, singlet(i), 1)      # legs [j, k, i']
    t2 = contract(t1, 0, singlet(j), 1)     # legs [k, i', j']
    t1 = contract(t2, 0, singlet(k), 1)     # legs [i', j', k']
    return float(np.sum(c * t1) * math.sqrt(j + 1))
