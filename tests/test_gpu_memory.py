"""Device memory cache of a context (GPU).

Blocks freed by tensors, GEMM plans and H_eff plans are held back until the
context stream has been synchronised and then reused; objects may also
outlive their context (Python collects them in any order).  These tests
rebuild plans of changing shapes on one context, so that freed blocks of
other sizes are handed out again, and check every result against the oracle;
and they destroy a context before the tensors, plans and DMRG driver made on
it.
"""
import sys

import numpy as np
import pytest

from oracle import heff as oheff
from su2dmrg_amd import synthetic

pytestmark = pytest.mark.gpu

RTOL = 1e-10


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _relerr(got, ref):
    num = sum(float(np.sum((got[j] - ref[j]) ** 2)) for j in ref)
    den = sum(float(np.sum(ref[j] ** 2)) for j in ref)
    return (num / den) ** 0.5


def test_reused_blocks_give_oracle_results():
    from su2dmrg_amd import lib as L
    _say("memory cache: plans of changing shapes on one context")
    ctx = L.Context(0)
    shapes = [((1, 3, 5), (120, 200, 60)), ((0, 2, 4, 6), (90, 150, 70, 9)),
              ((1, 3, 5), (120, 200, 60)), ((1, 3), (300, 40))]
    for i, (twoj, dims) in enumerate(shapes):
        work = synthetic.small(seed=30 + i, twoj=twoj, dims=dims, nterm=3)
        dev = L.upload_workload(ctx, work)
        ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], work["terms"])
        for _ in range(2):
            dev["plan"].apply(dev["psi"], dev["out"])
            got = L.unpack_wavefunction(dev["out"], dev["out"].download())
            assert _relerr(got, ref) < RTOL, (i, twoj, dims)
        for k in ("plan", "psi", "out", "hl", "hr"):
            if dev[k] is not None:
                dev[k].close()
        for t in dev["ops"]:
            t.close()
    ctx.close()


def test_objects_outlive_their_context():
    from su2dmrg_amd import lib as L
    _say("memory cache: context destroyed before its objects")
    ctx = L.Context(0)
    work = synthetic.small(seed=41, twoj=(1, 3), dims=(80, 50), nterm=2)
    dev = L.upload_workload(ctx, work)
    dev["plan"].apply(dev["psi"], dev["out"])
    got = L.unpack_wavefunction(dev["out"], dev["out"].download())
    ref = oheff.heff_apply(work["psi"], work["hl"], work["hr"], work["terms"])
    assert _relerr(got, ref) < RTOL
    d = L.Dmrg(ctx, 4, 2, 1.0, 0.0, 0, 0, 20)
    d.run(nsweep=1, stop=None)
    ctx.close()
    d.close()
    dev["plan"].close()
    for k in ("psi", "out", "hl", "hr"):
        if dev[k] is not None:
            dev[k].close()
    for t in dev["ops"]:
        t.close()
    # the device is still usable afterwards
    c2 = L.Context(0)
    dev2 = L.upload_workload(c2, work)
    dev2["plan"].apply(dev2["psi"], dev2["out"])
    got = L.unpack_wavefunction(dev2["out"], dev2["out"].download())
    assert _relerr(got, ref) < RTOL
    c2.close()
