"""The centre-bond H_eff of a run's own state (su2_dmrg_centre, GPU).

The plan is built from the run's environments exactly as each two-site solve
builds its own (dmrg_su2.cpp:543-664), and the wavefunction is the product of
the factors the last solve left (dmrg_su2_omp.cpp:135-150).  Checks: the
Rayleigh quotient of that wavefunction is the run's last Ritz value (exactly
where nothing was truncated, to the discarded weight otherwise), and the
plan is symmetric, <u|H v> = <H u|v>, on random vectors.
"""
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import ed

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from su2dmrg_amd import lib as L
    c = L.Context(0)
    yield c
    c.close()


def _rq_and_symmetry(ctx, g, seed=0):
    plan, psi = g.centre()
    out = psi.like()
    plan.apply(psi, out)
    rq = out.dot(psi) / psi.dot(psi)
    rng = np.random.default_rng(seed)
    u, v = psi.like(), psi.like()
    u.upload(rng.standard_normal(psi.size))
    v.upload(rng.standard_normal(psi.size))
    hu, hv = psi.like(), psi.like()
    plan.apply(u, hu)
    plan.apply(v, hv)
    a, b = u.dot(hv), hu.dot(v)
    for t in (out, u, v, hu, hv, psi):
        t.close()
    plan.close()
    return rq, a, b


def test_centre_of_an_exact_run(ctx):
    from su2dmrg_amd import lib as L
    _say("centre H_eff: 3x4 cylinder, D=100 (no truncation)")
    lat = od.Lattice(3, 4, jcoup=2.0, bdry=2)
    e_ed = ed.ground_energy(lat, 0)
    g = L.Dmrg(ctx, 3, 4, 2.0, 0.0, 2, 0, 100)
    g.run(nsweep=6, stop=None)
    rq, a, b = _rq_and_symmetry(ctx, g)
    e = g.energies()[-1]
    _say(f"  Rayleigh quotient {rq:.14f}, last Ritz {e:.14f}, ED {e_ed:.14f}")
    assert abs(rq - e) < 1e-10 * abs(e)
    assert abs(rq - e_ed) < 1e-10 * abs(e_ed)
    assert abs(a - b) < 1e-12 * max(abs(a), 1.0)
    g.close()


def test_centre_of_a_truncated_run(ctx):
    from su2dmrg_amd import lib as L
    _say("centre H_eff: 8x4 cylinder, D=120")
    g = L.Dmrg(ctx, 8, 4, 2.0, 0.0, 2, 0, 120)
    g.run(nsweep=2, stop=None)
    rq, a, b = _rq_and_symmetry(ctx, g, seed=1)
    e = g.energies()[-1]
    disc = g.history()[-1][1]
    _say(f"  Rayleigh quotient {rq:.12f}, last Ritz {e:.12f}, discarded {disc:.2e}")
    # the state lost the discarded weight of the last cut
    assert abs(rq - e) <= 10 * disc * abs(e) + 1e-10 * abs(e)
    assert abs(a - b) < 1e-12 * max(abs(a), 1.0)
    with pytest.raises(L.Su2Error):
        L.Dmrg(ctx, 4, 4, 2.0, 0.0, 2, 0, 50).centre()
    g.close()
