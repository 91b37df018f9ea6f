"""ctypes binding of ``libsu2dmrg.so`` (the C ABI in ``include/su2dmrg.h``).

The library is the product: this module only loads it, declares the entry
points and converts the plain Python block layout used by the tests and the
benchmark (``2j -> ndarray`` wavefunctions, ``(2j_in, 2j_out) -> ndarray``
operators) to and from the C ABI's flat block-index order.  There is no
fallback: if the shared library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsu2dmrg.so")
HEADER = os.path.join(HERE, "..", "include", "su2dmrg.h")

if not os.path.exists(LIB_PATH):
    raise ImportError(f"{LIB_PATH} is not built; run `python -m su2dmrg_amd.build`")

lib = C.CDLL(LIB_PATH)

_p = C.c_void_p
_i = C.c_int
_ll = C.c_longlong
_d = C.c_double
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)
_llp = C.POINTER(C.c_longlong)

_SIG = {
    "su2_version": (C.c_char_p, []),
    "su2_ctx_create": (_i, [_i, C.POINTER(_p)]),
    "su2_ctx_destroy": (_i, [_p]),
    "su2_last_error": (C.c_char_p, [_p]),
    "su2_ctx_synchronize": (_i, [_p]),
    "su2_ctx_set_svd_tuning": (_i, [_p, _i, _i]),
    "su2_ctx_set_svd_precondition": (_i, [_p, _i]),
    "su2_ctx_memory": (_i, [_p, _llp, _llp, _llp, _i]),
    "su2_tensor_create": (_i, [_p, _i, _ip, _ip, _ip, _ip, _i, _ip, C.POINTER(_p)]),
    "su2_tensor_destroy": (_i, [_p, _p]),
    "su2_tensor_like": (_i, [_p, _p, C.POINTER(_p)]),
    "su2_tensor_size": (_i, [_p, _llp, _ip]),
    "su2_tensor_block": (_i, [_p, _i, _ip, _ip, _ip, _llp]),
    "su2_tensor_upload": (_i, [_p, _p, _dp, _ll]),
    "su2_tensor_download": (_i, [_p, _p, _dp, _ll]),
    "su2_tensor_copy": (_i, [_p, _p, _p]),
    "su2_tensor_axpy": (_i, [_p, _p, _d, _p]),
    "su2_tensor_scal": (_i, [_p, _p, _d]),
    "su2_tensor_dot": (_i, [_p, _p, _p, _dp]),
    "su2_tensor_normalize": (_i, [_p, _p, _dp]),
    "su2_tensor_dot_noalloc": (_i, [_p, _p, _p, _dp]),
    "su2_noalloc_scratch_bytes": (_i, [_llp]),
    "su2_heff_create": (_i, [_p, _p, _p, _p, _i, _ip, _ip, C.POINTER(_p), _dp, _ip,
                             C.POINTER(_p), _dp, C.POINTER(_p)]),
    "su2_env_transform": (_i, [_p, _i, _i, _i, _ip, _ip, _p, _p, _d, C.POINTER(_p)]),
    "su2_heff_apply": (_i, [_p, _p, _p, _p]),
    "su2_heff_flops": (_i, [_p, _dp, _dp]),
    "su2_heff_destroy": (_i, [_p, _p]),
    "su2_heff_time": (_i, [_p, _p, _p, _p, _i, _dp, _dp, _dp]),
    "su2_heff_autotune": (_i, [_p, _p, _i, _ip, C.POINTER(C.c_float), _ip]),
    "su2_heff_set_variants": (_i, [_p, _p, _i, _i]),
    "su2_heff_set_kernel_path": (_i, [_p, _p, _i, _p]),
    "su2_heff_set_precision": (_i, [_p, _p, _i]),
    "su2_svd_truncate": (_i, [_p, _p, _d, _d, _i, C.POINTER(_p), C.POINTER(_p), _dp, _i, _ip,
                              _dp]),
    "su2_recoupling": (_i, [_i, _ip, _dp]),
    "su2_dmrg_create": (_i, [_p, _i, _i, _d, _d, _i, _i, _i, C.POINTER(_p)]),
    "su2_dmrg_destroy": (_i, [_p, _p]),
    "su2_dmrg_warmup": (_i, [_p, _p, _dp]),
    "su2_dmrg_sweep": (_i, [_p, _p, _dp]),
    "su2_dmrg_set_max_dcut": (_i, [_p, _p, _i]),
    "su2_dmrg_set_heff_variants": (_i, [_p, _p, _i, _i]),
    "su2_dmrg_energies": (_i, [_p, _dp, _i, _ip]),
    "su2_dmrg_history": (_i, [_p, _ip, _dp, _ip, _i, _ip]),
    "su2_dmrg_set_verbose": (_i, [_p, _i]),
    "su2_dmrg_save": (_i, [_p, _p, C.c_char_p, _i, _i]),
    "su2_dmrg_load": (_i, [_p, _p, C.c_char_p, _i, _i, _i, _dp]),
    "su2_dmrg_add_lower_state": (_i, [_p, _p, C.c_char_p, _i, _i, _i]),
    "su2_dmrg_stats": (_i, [_p, _dp, _i]),
    "su2_dmrg_set_matvec_precision": (_i, [_p, _p, _i, _d]),
    "su2_dmrg_set_graph": (_i, [_p, _p, _i]),
    "su2_dmrg_centre": (_i, [_p, _p, C.POINTER(_p), C.POINTER(_p)]),
    "su2_batch_create": (_i, [_p, C.POINTER(_p)]),
    "su2_batch_destroy": (_i, [_p, _p]),
    "su2_batch_add": (_i, [_p, _p, _p, _i, _ip, _ip, _d, _d, _ip]),
    "su2_batch_tick": (_i, [_p, _p, _ip]),
    "su2_batch_run": (_i, [_p, _p]),
    "su2_batch_run_lanes": (_i, [C.POINTER(_p), C.POINTER(_p), _i]),
    "su2_batch_set_graph": (_i, [_p, _p, _i]),
    "su2_batch_energies": (_i, [_p, _i, _ip, _dp, _i, _ip]),
    "su2_batch_stats": (_i, [_p, _dp, _i]),
    "su2_server_create": (_i, [_p, _d, _i, C.c_char_p, C.POINTER(_p)]),
    "su2_server_destroy": (_i, [_p, _p]),
    "su2_server_submit": (_i, [_p, _p, _p, _d, _ip]),
    "su2_server_set_lanes": (_i, [_p, _p, _i]),
    "su2_server_run": (_i, [_p, _p, _i]),
    "su2_server_result": (_i, [_p, _i, _dp, _dp, _dp]),
    "su2_server_stats": (_i, [_p, _dp, _i]),
    "su2_job_footprint": (_i, [_p, _dp]),
}

for _name, (_res, _args) in _SIG.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def declared_symbols():
    """Entry points declared in include/su2dmrg.h."""
    with open(HEADER) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(su2_[a-z0-9_]+)\s*\(", text)))


class Job(C.Structure):
    """su2_job: one served DMRG job (the reference's run flags)."""
    _fields_ = [("lx", C.c_int), ("ly", C.c_int), ("jcoup", C.c_double), ("qcoup", C.c_double),
                ("bdry", C.c_int), ("totspin", C.c_int), ("max_dcut", C.c_int),
                ("exci", C.c_int), ("nsweep", C.c_int), ("stop", C.c_double)]

    def __init__(self, lx, ly, jcoup=2.0, qcoup=0.0, bdry=2, totspin=0, max_dcut=100, exci=0,
                 nsweep=2, stop=-1.0):
        super().__init__(lx, ly, float(jcoup), float(qcoup), bdry, totspin, max_dcut, exci,
                         nsweep, float(stop))

    def footprint(self):
        b = C.c_double()
        rc = lib.su2_job_footprint(C.byref(self), C.byref(b))
        if rc:
            raise Su2Error(rc, lib.su2_last_error(None).decode())
        return b.value


class Su2Error(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"su2 status {code}: {msg}")
        self.code = code


def _check_global(rc):
    """Status of a call that takes no context (its message is the thread's last)."""
    if rc:
        raise Su2Error(rc, lib.su2_last_error(None).decode())


def _ia(x):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int32))
    return a, a.ctypes.data_as(_ip)


class Context:
    """One HIP device + stream (su2_ctx)."""

    def __init__(self, device=0):
        h = _p()
        rc = lib.su2_ctx_create(device, C.byref(h))
        if rc:
            raise Su2Error(rc, lib.su2_last_error(None).decode())
        self.h = h
        self.device = device

    def check(self, rc):
        if rc:
            raise Su2Error(rc, lib.su2_last_error(self.h).decode())

    def sync(self):
        self.check(lib.su2_ctx_synchronize(self.h))

    def set_svd_precondition(self, min_k=256):
        """QR preconditioning of the Jacobi SVD for sector blocks with
        min(m, n) >= min_k (the context's default is 256; 0 or False: off;
        True: 256)."""
        if min_k is True:
            min_k = 256
        self.check(lib.su2_ctx_set_svd_precondition(self.h, int(min_k)))

    def set_svd_tuning(self, block_cols=32, inner_sweeps=1):
        """Column block width (16 or 32) and inner sweeps per pair visit (1..3)
        of the one-sided Jacobi SVD."""
        self.check(lib.su2_ctx_set_svd_tuning(self.h, int(block_cols), int(inner_sweeps)))

    def memory(self, reset_peak=False):
        """(live, peak, cached) device bytes of the context's cache."""
        a, b, c = _ll(), _ll(), _ll()
        self.check(lib.su2_ctx_memory(self.h, C.byref(a), C.byref(b), C.byref(c),
                                      int(bool(reset_peak))))
        return a.value, b.value, c.value

    def close(self):
        if self.h:
            lib.su2_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tensor:
    """Device block-sparse tensor (su2_tensor)."""

    def __init__(self, ctx, legs, locspin=0, present=None):
        """``legs`` = list of (twoj list, dims list, dir)."""
        self.ctx = ctx
        self.legs = [(list(t), list(d), int(s)) for t, d, s in legs]
        nsec = [len(t) for t, _, _ in self.legs]
        twoj = [x for t, _, _ in self.legs for x in t]
        dims = [x for _, d, _ in self.legs for x in d]
        dirs = [s for _, _, s in self.legs]
        a1, p1 = _ia(nsec)
        a2, p2 = _ia(twoj)
        a3, p3 = _ia(dims)
        a4, p4 = _ia(dirs)
        if present is not None:
            a5, p5 = _ia(present)
        else:
            p5 = None
        h = _p()
        ctx.check(lib.su2_tensor_create(ctx.h, len(self.legs), p1, p2, p3, p4, locspin, p5,
                                        C.byref(h)))
        self.h = h
        n = _ll()
        nb = _i()
        ctx.check(lib.su2_tensor_size(h, C.byref(n), C.byref(nb)))
        self.size = n.value
        self.blocks = []
        nl = len(self.legs)
        for i in range(nb.value):
            idx = _i()
            tj = (C.c_int * nl)()
            dm = (C.c_int * nl)()
            off = _ll()
            ctx.check(lib.su2_tensor_block(h, i, C.byref(idx), tj, dm, C.byref(off)))
            self.blocks.append((idx.value, tuple(tj), tuple(dm), off.value))

    def upload(self, flat):
        flat = np.ascontiguousarray(flat, dtype=np.float64)
        self.ctx.check(lib.su2_tensor_upload(self.ctx.h, self.h, flat.ctypes.data_as(_dp),
                                             flat.size))

    def download(self):
        out = np.empty(self.size, dtype=np.float64)
        self.ctx.check(lib.su2_tensor_download(self.ctx.h, self.h, out.ctypes.data_as(_dp),
                                               out.size))
        return out

    def close(self):
        if self.h:
            lib.su2_tensor_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def like(self, nleg=2):
        """A zeroed tensor of the same structure (su2_tensor_like)."""
        h = _p()
        self.ctx.check(lib.su2_tensor_like(self.ctx.h, self.h, C.byref(h)))
        return _wrap_tensor(self.ctx, h, nleg=nleg)

    # Lanczos algebra
    def copy_from(self, other):
        self.ctx.check(lib.su2_tensor_copy(self.ctx.h, self.h, other.h))

    def axpy(self, a, x):
        self.ctx.check(lib.su2_tensor_axpy(self.ctx.h, self.h, float(a), x.h))

    def scal(self, a):
        self.ctx.check(lib.su2_tensor_scal(self.ctx.h, self.h, float(a)))

    def dot(self, other):
        r = _d()
        self.ctx.check(lib.su2_tensor_dot(self.ctx.h, self.h, other.h, C.byref(r)))
        return r.value

    def dot_noalloc(self, other):
        """su2_tensor_dot_noalloc: the same dot with no allocation (fixed scratch)."""
        r = _d()
        self.ctx.check(lib.su2_tensor_dot_noalloc(self.ctx.h, self.h, other.h, C.byref(r)))
        return r.value

    def normalize(self):
        r = _d()
        self.ctx.check(lib.su2_tensor_normalize(self.ctx.h, self.h, C.byref(r)))
        return r.value


# --------------------------------------------------------------------------
# plain layout <-> flat block-index order
# --------------------------------------------------------------------------
def wavefunction(ctx, twoj, dims_a, dims_b, psi=None):
    t = Tensor(ctx, [(twoj, dims_a, 1), (twoj, dims_b, 1)], locspin=0)
    if psi is not None:
        t.upload(pack_wavefunction(t, psi))
    return t


def pack_wavefunction(t, psi):
    out = np.zeros(t.size)
    for _, tj, dm, off in t.blocks:
        blk = psi.get(tj[0])
        if blk is not None:
            out[off:off + dm[0] * dm[1]] = np.asarray(blk).ravel(order="F")
    return out


def unpack_wavefunction(t, flat):
    return {tj[0]: flat[off:off + dm[0] * dm[1]].reshape((dm[0], dm[1]), order="F")
            for _, tj, dm, off in t.blocks}


def scalar_operator(ctx, twoj, dims, h):
    t = Tensor(ctx, [(twoj, dims, 1), (twoj, dims, -1)], locspin=0)
    flat = np.zeros(t.size)
    for _, tj, dm, off in t.blocks:
        if tj[0] in h:
            flat[off:off + dm[0] * dm[1]] = np.asarray(h[tj[0]]).ravel(order="F")
    t.upload(flat)
    return t


def vector_operator(ctx, twoj, dims, op, rank=2):
    """Three-leg [ket, operator leg 2K = rank, bra] operator holding the blocks in ``op``."""
    n = len(twoj)
    present = np.zeros(n * n, dtype=np.int32)
    pos = {j: i for i, j in enumerate(twoj)}
    for (a, b) in op:
        present[pos[a] + n * pos[b]] = 1
    t = Tensor(ctx, [(twoj, dims, 1), ([rank], [1], 1), (twoj, dims, -1)], locspin=0,
               present=present)
    flat = np.zeros(t.size)
    for _, tj, dm, off in t.blocks:
        flat[off:off + dm[0] * dm[2]] = np.asarray(op[(tj[0], tj[2])]).ravel(order="F")
    t.upload(flat)
    return t


class HeffPlan:
    """su2_heff: a built H_eff matvec plan.

    ``terms`` = list of (left side, right side), each side a list of
    (coef, Tensor) summed once at plan creation.
    """

    def __init__(self, ctx, psi_layout, hl, hr, terms, ranks=None):
        """ranks: doubled operator-leg rank 2K of each term (default all 2)."""
        self.ctx = ctx
        lc, rc, lops, rops, lcoef, rcoef = [], [], [], [], [], []
        rk, _ = _ia(list(ranks) if ranks is not None else [2] * max(1, len(terms)))
        for left, right in terms:
            lc.append(len(left))
            rc.append(len(right))
            for c, t in left:
                lcoef.append(c)
                lops.append(t.h)
            for c, t in right:
                rcoef.append(c)
                rops.append(t.h)
        self._keep = (lops, rops)
        a1, p1 = _ia(lc if lc else [0])
        a2, p2 = _ia(rc if rc else [0])
        LO = (_p * max(1, len(lops)))(*lops)
        RO = (_p * max(1, len(rops)))(*rops)
        LC = (C.c_double * max(1, len(lcoef)))(*lcoef)
        RC = (C.c_double * max(1, len(rcoef)))(*rcoef)
        h = _p()
        ctx.check(lib.su2_heff_create(ctx.h, psi_layout.h, hl.h if hl else None,
                                      hr.h if hr else None, len(terms),
                                      rk.ctypes.data_as(_ip), p1, LO, LC, p2, RO, RC,
                                      C.byref(h)))
        self.h = h

    def apply(self, psi, out):
        self.ctx.check(lib.su2_heff_apply(self.ctx.h, self.h, psi.h, out.h))

    def flops(self):
        a, e = _d(), _d()
        _check_global(lib.su2_heff_flops(self.h, C.byref(a), C.byref(e)))
        return a.value, e.value

    def time(self, psi, out, reps):
        a, s1, s2 = _d(), _d(), _d()
        self.ctx.check(lib.su2_heff_time(self.ctx.h, self.h, psi.h, out.h, int(reps), C.byref(a),
                                         C.byref(s1), C.byref(s2)))
        return a.value, s1.value, s2.value

    def autotune(self, reps=3):
        """Time every tile variant for both launches; returns (chosen, times1, times2)."""
        ch = (C.c_int * 2)()
        nv = C.c_int()
        times = (C.c_float * 128)()
        self.ctx.check(lib.su2_heff_autotune(self.ctx.h, self.h, int(reps), ch, times,
                                             C.byref(nv)))
        n = nv.value
        return (ch[0], ch[1]), list(times[:n]), list(times[n:2 * n])

    def set_variants(self, v1, v2):
        self.ctx.check(lib.su2_heff_set_variants(self.ctx.h, self.h, int(v1), int(v2)))

    def set_precision(self, bits):
        """64: FP64 matvec (default); 32: FP32 MFMA with FP32 accumulation."""
        self.ctx.check(lib.su2_heff_set_precision(self.ctx.h, self.h, int(bits)))

    def set_kernel_path(self, uniform):
        """Straight-line kernel for single-layout launches (True) or the
        general kernel (False); returns the (stage 1, stage 2) layouts."""
        lay = (C.c_int * 2)()
        self.ctx.check(lib.su2_heff_set_kernel_path(self.ctx.h, self.h, int(bool(uniform)), lay))
        return lay[0], lay[1]

    def close(self):
        if self.h:
            lib.su2_heff_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def upload_workload(ctx, work):
    """Device tensors + plan for a synthetic workload (see ``synthetic``)."""
    twoj, da, db = work["twoj"], work["dims_a"], work["dims_b"]
    psi = wavefunction(ctx, twoj, da, db, work["psi"])
    out = wavefunction(ctx, twoj, da, db)
    hl = scalar_operator(ctx, twoj, da, work["hl"]) if work.get("hl") is not None else None
    hr = scalar_operator(ctx, twoj, db, work["hr"]) if work.get("hr") is not None else None
    dev_terms = []
    keep = []
    for L, R in work["terms"]:
        tl = vector_operator(ctx, twoj, da, L)
        tr = vector_operator(ctx, twoj, db, R)
        keep += [tl, tr]
        dev_terms.append(([(1.0, tl)], [(1.0, tr)]))
    plan = HeffPlan(ctx, psi, hl, hr, dev_terms)
    return {"psi": psi, "out": out, "hl": hl, "hr": hr, "ops": keep, "plan": plan}


# --------------------------------------------------------------------------
# recoupling tables, truncated SVD, DMRG driver
# --------------------------------------------------------------------------
RC_KINDS = {"op_transform_left": 0, "op_transform_right": 1, "op_initial_left": 2,
            "op_initial_right": 3, "pairup_left": 4, "pairup_right": 5,
            "psi_left_to_right": 6, "psi_right_to_left": 7, "hamilt_vec": 8,
            "couple_left": 9, "couple_right": 10, "reduce_left": 11, "reduce_right": 12}


def recoupling(kind, *labels):
    lab = (C.c_int * 6)(*(list(labels) + [0] * (6 - len(labels))))
    out = _d()
    rc = lib.su2_recoupling(RC_KINDS[kind], lab, C.byref(out))
    if rc:
        raise Su2Error(rc, lib.su2_last_error(None).decode())
    return out.value


ENV_KINDS = {"op_transform": 0, "overlap": 1, "pairup": 2, "initial": 3}


def env_transform(ctx, kind, left, block_twoj, block_dims, P=None, op=None, coef=1.0):
    """su2_env_transform: one environment kernel on one block operator (see the
    header); returns the new device tensor."""
    a1, p1 = _ia(block_twoj)
    a2, p2 = _ia(block_dims)
    h = _p()
    ctx.check(lib.su2_env_transform(ctx.h, ENV_KINDS[kind], 0 if left else 1, len(block_twoj),
                                    p1, p2, P.h if P is not None else None,
                                    op.h if op is not None else None, float(coef), C.byref(h)))
    return _wrap_tensor(ctx, h, nleg=2 if kind in ("overlap", "pairup") else 3)


def _wrap_tensor(ctx, h, nleg=2):
    t = Tensor.__new__(Tensor)
    t.ctx = ctx
    t.h = h
    n = _ll()
    nb = _i()
    ctx.check(lib.su2_tensor_size(h, C.byref(n), C.byref(nb)))
    t.size = n.value
    t.blocks = []
    for i in range(nb.value):
        idx = _i()
        tj = (C.c_int * 4)()
        dm = (C.c_int * 4)()
        off = _ll()
        ctx.check(lib.su2_tensor_block(h, i, C.byref(idx), tj, dm, C.byref(off)))
        t.blocks.append((idx.value, tuple(tj[:nleg]), tuple(dm[:nleg]), off.value))
    t.legs = None
    return t


def svd_truncate(ctx, psi, p1, p2, max_dcut):
    """Returns (U blocks {2j: ndarray}, V blocks, kept values, discarded weight)."""
    hu, hv = _p(), _p()
    cap = 1 << 16
    kept = (C.c_double * cap)()
    nk = _i()
    disc = _d()
    ctx.check(lib.su2_svd_truncate(ctx.h, psi.h, float(p1), float(p2), int(max_dcut),
                                   C.byref(hu), C.byref(hv), kept, cap, C.byref(nk),
                                   C.byref(disc)))
    U = _wrap_tensor(ctx, hu)
    V = _wrap_tensor(ctx, hv)
    Ub = unpack_wavefunction(U, U.download())
    Vb = unpack_wavefunction(V, V.download())
    return Ub, Vb, list(kept[:nk.value]), disc.value


class Dmrg:
    """su2_dmrg: two-site SU(2) DMRG resident on the device."""

    def __init__(self, ctx, lx, ly, jcoup, qcoup=0.0, bdry=2, totspin=0, max_dcut=200):
        self.ctx = ctx
        h = _p()
        ctx.check(lib.su2_dmrg_create(ctx.h, lx, ly, float(jcoup), float(qcoup), bdry, totspin,
                                      max_dcut, C.byref(h)))
        self.h = h

    def warmup(self):
        e = _d()
        self.ctx.check(lib.su2_dmrg_warmup(self.ctx.h, self.h, C.byref(e)))
        return e.value

    def sweep(self):
        e = _d()
        self.ctx.check(lib.su2_dmrg_sweep(self.ctx.h, self.h, C.byref(e)))
        return e.value

    def set_max_dcut(self, d):
        self.ctx.check(lib.su2_dmrg_set_max_dcut(self.ctx.h, self.h, int(d)))

    def save(self, directory, ky=0, exci=0):
        """Checkpoint files in the reference's format (uu-*, ww-*, enr-*)."""
        self.ctx.check(lib.su2_dmrg_save(self.ctx.h, self.h, os.fsencode(directory), ky, exci))

    def load(self, directory, file_dcut=0, ky=0, exci=0):
        """Restart from checkpoint files; returns the saved energy."""
        e = _d()
        self.ctx.check(lib.su2_dmrg_load(self.ctx.h, self.h, os.fsencode(directory),
                                         int(file_dcut), ky, exci, C.byref(e)))
        return e.value

    def add_lower_state(self, directory, file_dcut=0, ky=0, exci=0):
        """Target an excited state: keep the run orthogonal to the state saved
        in `directory` with index `exci` (call for 0 .. n-1 before warmup)."""
        self.ctx.check(lib.su2_dmrg_add_lower_state(self.ctx.h, self.h, os.fsencode(directory),
                                                    int(file_dcut), ky, exci))

    def set_verbose(self, level=1):
        _check_global(lib.su2_dmrg_set_verbose(self.h, int(level)))

    def set_heff_variants(self, v1, v2):
        self.ctx.check(lib.su2_dmrg_set_heff_variants(self.ctx.h, self.h, int(v1), int(v2)))

    def energies(self):
        n = _i()
        _check_global(lib.su2_dmrg_energies(self.h, None, 0, C.byref(n)))
        buf = (C.c_double * max(1, n.value))()
        _check_global(lib.su2_dmrg_energies(self.h, buf, n.value, C.byref(n)))
        return list(buf[:n.value])

    def history(self):
        """Per step: (kept multiplets, discarded weight, matvecs)."""
        n = _i()
        _check_global(lib.su2_dmrg_history(self.h, None, None, None, 0, C.byref(n)))
        m = max(1, n.value)
        kept, disc, mv = (_i * m)(), (C.c_double * m)(), (_i * m)()
        _check_global(lib.su2_dmrg_history(self.h, kept, disc, mv, n.value, C.byref(n)))
        return [(kept[i], disc[i], mv[i]) for i in range(n.value)]

    def set_graph(self, on=True):
        self.ctx.check(lib.su2_dmrg_set_graph(self.ctx.h, self.h, int(bool(on))))

    def centre(self):
        """(HeffPlan, wavefunction Tensor) of the centre bond of the current state."""
        hp, ht = _p(), _p()
        self.ctx.check(lib.su2_dmrg_centre(self.ctx.h, self.h, C.byref(hp), C.byref(ht)))
        plan = HeffPlan.__new__(HeffPlan)
        plan.ctx = self.ctx
        plan.h = hp
        plan._keep = ()
        psi = _wrap_tensor(self.ctx, ht, nleg=2)
        return plan, psi

    def set_matvec_precision(self, bits, min_flops=-1.0):
        """64: FP64 matvecs; 32: FP32 MFMA Krylov matvecs with FP64 checks, for
        solves whose matvec has at least `min_flops` FLOPs (< 0: default)."""
        self.ctx.check(lib.su2_dmrg_set_matvec_precision(self.ctx.h, self.h, int(bits),
                                                         float(min_flops)))

    def stats(self):
        out = (C.c_double * 17)()
        _check_global(lib.su2_dmrg_stats(self.h, out, 17))
        keys = ["t_env", "t_lanczos", "t_svd", "t_project", "t_start", "t_matvec",
                "heff_flops", "matvecs", "steps", "bond_dim", "t_jacobi", "jacobi_sweeps",
                "matvecs32", "t_matvec32", "heff_flops32", "max_fp64_correction", "svd_degraded"]
        return dict(zip(keys, list(out)))

    def run_schedule(self, schedule, stop=None, min_flops=-1.0):
        """Warm-up, then one sweep per (max_dcut, bits) entry of `schedule`
        (the warm-up uses the first entry); returns the energy after each
        sweep.  With `stop`, ends early once |dE| < stop (main_omp.cpp:197)."""
        d0, b0 = schedule[0]
        self.set_max_dcut(d0)
        self.set_matvec_precision(b0, min_flops)
        prev = self.warmup()
        out = []
        for d, bits in schedule:
            self.set_max_dcut(d)
            self.set_matvec_precision(bits, min_flops)
            e = self.sweep()
            out.append(e)
            if stop is not None and abs(e - prev) < stop:
                break
            prev = e
        return out

    def run(self, nsweep=5, stop=1e-6):
        """Warm-up, then sweeps until |dE| < stop (main_omp.cpp:184-202)."""
        prev = self.warmup()
        per = []
        for _ in range(nsweep):
            e = self.sweep()
            per.append(e)
            if stop is not None and abs(e - prev) < stop:
                break
            prev = e
        return per

    def close(self):
        if self.h:
            lib.su2_dmrg_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """su2_batch: several Dmrg runs of one context advanced in lockstep."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = _p()
        ctx.check(lib.su2_batch_create(ctx.h, C.byref(h)))
        self.h = h
        self._runs = []

    def add(self, dmrg, schedule, stop=-1.0, min_flops=-1.0):
        """schedule: list of (max_dcut, bits) sweeps; the warm-up uses entry 0."""
        n = len(schedule)
        d, _ = _ia([x[0] for x in schedule] or [0])
        b, _ = _ia([x[1] for x in schedule] or [64])
        rid = _i()
        self.ctx.check(lib.su2_batch_add(self.ctx.h, self.h, dmrg.h, n, d.ctypes.data_as(_ip),
                                         b.ctypes.data_as(_ip), float(min_flops), float(stop),
                                         C.byref(rid)))
        self._runs.append(dmrg)
        return rid.value

    def tick(self):
        n = _i()
        self.ctx.check(lib.su2_batch_tick(self.ctx.h, self.h, C.byref(n)))
        return n.value

    def run(self):
        self.ctx.check(lib.su2_batch_run(self.ctx.h, self.h))

    def set_graph(self, on=True):
        self.ctx.check(lib.su2_batch_set_graph(self.ctx.h, self.h, int(bool(on))))

    def energies(self, rid):
        """(finished, energy after each finished phase, warm-up first)."""
        f, n = _i(), _i()
        buf = (C.c_double * 256)()
        _check_global(lib.su2_batch_energies(self.h, int(rid), C.byref(f), buf, 256, C.byref(n)))
        return bool(f.value), list(buf[:min(n.value, 256)])

    def stats(self):
        out = (C.c_double * 8)()
        _check_global(lib.su2_batch_stats(self.h, out, 8))
        keys = ["ticks", "graphs_built", "graph_replays", "t_prepare", "t_lanczos", "t_matvec",
                "t_svd", "t_complete"]
        return dict(zip(keys, list(out)))

    def close(self):
        if self.h:
            lib.su2_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_lanes(batches):
    """su2_batch_run_lanes: runs the batches at once, one host thread each;
    every batch must be on its own Context (one stream per lane)."""
    n = len(batches)
    ctxs = (_p * n)(*[b.ctx.h for b in batches])
    hs = (_p * n)(*[b.h for b in batches])
    _check_global(lib.su2_batch_run_lanes(ctxs, hs, n))


class Server:
    """su2_server: request loop with continuous batching on one device; with
    lanes > 1 the slots are split over that many streams and host threads."""

    def __init__(self, ctx, budget_bytes, max_slots, workdir, lanes=1):
        self.ctx = ctx
        h = _p()
        ctx.check(lib.su2_server_create(ctx.h, float(budget_bytes), int(max_slots),
                                        os.fsencode(workdir), C.byref(h)))
        self.h = h
        self.n = 0
        if lanes != 1:
            ctx.check(lib.su2_server_set_lanes(ctx.h, self.h, int(lanes)))

    def submit(self, job, arrival_s=0.0):
        jid = _i()
        self.ctx.check(lib.su2_server_submit(self.ctx.h, self.h, C.byref(job), float(arrival_s),
                                             C.byref(jid)))
        self.n += 1
        return jid.value

    def run(self, verbose=0):
        self.ctx.check(lib.su2_server_run(self.ctx.h, self.h, int(verbose)))

    def result(self, jid):
        """dict(energy, t_arrive, t_admit, t_finish, latency, footprint)."""
        e, fp = _d(), _d()
        t = (C.c_double * 3)()
        _check_global(lib.su2_server_result(self.h, int(jid), C.byref(e), t, C.byref(fp)))
        return {"energy": e.value, "t_arrive": t[0], "t_admit": t[1], "t_finish": t[2],
                "latency": t[2] - t[0], "footprint": fp.value}

    def stats(self):
        out = (C.c_double * 3)()
        _check_global(lib.su2_server_stats(self.h, out, 3))
        return {"wall": out[0], "ticks": out[1], "max_concurrent": out[2]}

    def close(self):
        if self.h:
            lib.su2_server_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
