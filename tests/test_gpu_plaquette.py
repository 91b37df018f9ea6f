"""J-Q-J2 plaquette (Q) terms on the MI355X against the oracle and exact
diagonalisation (GPU).

Step-by-step parity is asserted on the trajectories that a 1e-15 relative
perturbation of every Gram matrix and of every wavefunction before its SVD
leaves unchanged to 3e-12 in the oracle.  On the symmetric cylinders the
warm-up keeps sectors whose weight is rounding noise (the 1e-100 rule of
tensor.cpp:1533-1568 counts them) and the kept direction is that noise, so
their trajectories depend on each implementation's rounding, the
reference's included; they are compared with exact diagonalisation of the
physical J1-J2-Q model once converged, as for the J1-J2 cylinders.
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


# (lx, ly, jcoup, qcoup, bdry, D, 2S)
STEP_CASES = [(3, 4, 0.8, 0.6, 0, 20, 0), (4, 3, 0.7, 0.3, 0, 16, 2)]


@pytest.mark.parametrize("lx,ly,jc,q,bdry,D,twoS", STEP_CASES)
def test_jq_steps_match_oracle(ctx, lx, ly, jc, q, bdry, D, twoS):
    from su2dmrg_amd import lib as L
    _say(f"plaquette step parity {lx}x{ly} J={jc} Q={q} bdry={bdry} D={D} 2S={twoS}")
    lat = od.Lattice(lx, ly, jcoup=jc, qcoup=q, bdry=bdry)
    o = od.DMRG(lat, twoS, D)
    o.run(nsweep=2, stop=None)
    g = L.Dmrg(ctx, lx, ly, jc, q, bdry, twoS, D)
    g.run(nsweep=2, stop=None)
    eg = np.array(g.energies())
    eo = np.array(o.energies)
    g.close()
    assert len(eg) == len(eo)
    np.testing.assert_allclose(eg, eo, rtol=1e-10, atol=1e-10)


# J-Q (J2 = 1 - |Q| = 0), J-Q-J2, negative Q, S > 0
ED_CASES = [(3, 4, 1.0, 0.5, 0), (4, 4, 1.0, 1.0, 0), (4, 3, 0.7, 0.3, 2), (3, 4, 1.2, -0.4, 4)]


@pytest.mark.parametrize("lx,ly,jc,q,twoS", ED_CASES)
def test_jq_cylinder_converges_to_ed(ctx, lx, ly, jc, q, twoS):
    # the reference's (jcoup, qcoup) Hamiltonian is the physical J1 = jcoup,
    # J2 = 1 - |Q|, Q model plus Q/8 per plaquette
    from su2dmrg_amd import lib as L
    _say(f"plaquette cylinder {lx}x{ly} J={jc} Q={q} 2S={twoS} vs ED")
    lat = od.Lattice(lx, ly, jcoup=jc, qcoup=q, bdry=2)
    e_ed = ed.jq_energies(lx, ly, jc, 1 - abs(q), q, twoS)[0] + q * len(lat.plq) / 8
    g = L.Dmrg(ctx, lx, ly, jc, q, 2, twoS, 200)
    g.run(nsweep=6, stop=None)
    e = g.energies()
    g.close()
    assert abs(e[-1] - e_ed) < 1e-10 * abs(e_ed)
    nstep = 2 * (lx * ly - 6)
    assert min(e[-nstep:]) > e_ed - 1e-10 * abs(e_ed)
