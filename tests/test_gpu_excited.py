"""Excited-state targeting on the MI355X (the reference's exci > 0 path)
against the oracle and exact diagonalisation (GPU).

The lower states are the device's own converged runs, saved in the
reference's file format and read back by both the device
(su2_dmrg_add_lower_state) and the oracle (oracle.mpsio.lower_state), so the
two excited runs see the same lower state to the last bit.
"""
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import ed, mpsio

pytestmark = pytest.mark.gpu

# 6 (open) x 2 (open) J1-J2 at J2/J1 = 0.5: the three lowest S = 0 levels are
# non-degenerate (-14.0265, -10.1350, -10.0104), and D = 100 is exact
LX, LY, JC, BD, D = 6, 2, 2.0, 0, 100


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from su2dmrg_amd import lib as L
    c = L.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def ladder(ctx, tmp_path_factory):
    from su2dmrg_amd import lib as L
    _say("excited: ground state 6x2 D=100 on the device")
    lat = od.Lattice(LX, LY, jcoup=JC, bdry=BD)
    levels = ed.sector_energies(lat, 0, 3)
    d = str(tmp_path_factory.mktemp("lower"))
    g = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 0, D)
    g.run(nsweep=8, stop=None)
    g.save(d, exci=0)
    e0 = g.energies()[-1]
    g.close()
    return lat, levels, d, e0


def test_ground_state_of_the_ladder(ladder):
    lat, levels, d, e0 = ladder
    assert abs(e0 - levels[0]) < 1e-10 * abs(levels[0])


def _ground_files(ctx, tmp_path_factory, lx, ly, dcut):
    from su2dmrg_amd import lib as L
    d = str(tmp_path_factory.mktemp(f"lower{lx}x{ly}d{dcut}"))
    g = L.Dmrg(ctx, lx, ly, JC, 0.0, BD, 0, dcut)
    g.run(nsweep=4, stop=None)
    g.save(d, exci=0)
    g.close()
    return d


# the exact ladder, and a 4 x 4 open lattice truncated to 12 multiplets (a
# trajectory that is well conditioned for the ground state as well)
@pytest.mark.parametrize("lx,ly,dcut", [(6, 2, 100), (4, 4, 12)])
def test_excited_steps_match_oracle(ctx, tmp_path_factory, lx, ly, dcut):
    from su2dmrg_amd import lib as L
    _say(f"excited: step parity with the oracle {lx}x{ly} D={dcut}, 3 sweeps")
    lat = od.Lattice(lx, ly, jcoup=JC, bdry=BD)
    d = _ground_files(ctx, tmp_path_factory, lx, ly, dcut)
    lower = mpsio.lower_state(d, lat.ns, dcut, 0, 0)
    o = od.DMRG(lat, 0, dcut, lower=[lower])
    o.run(nsweep=3, stop=None)
    g = L.Dmrg(ctx, lx, ly, JC, 0.0, BD, 0, dcut)
    g.add_lower_state(d, exci=0)
    g.run(nsweep=3, stop=None)
    eg = np.array(g.energies())
    eo = np.array(o.energies)
    g.close()
    assert len(eg) == len(eo)
    np.testing.assert_allclose(eg, eo, rtol=1e-10, atol=1e-10)
    # variational above the ground state's level of the same basis size
    assert eg[-1] > lower.energy


def test_first_and_second_excited_levels_match_ed(ctx, ladder):
    from su2dmrg_amd import lib as L
    lat, levels, d, e0 = ladder
    _say("excited: E1 and E2 against exact diagonalisation")
    x = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 0, D)
    x.add_lower_state(d, exci=0)
    x.run(nsweep=14, stop=None)
    e1 = x.energies()[-1]
    x.save(d, exci=1)
    x.close()
    assert abs(e1 - levels[1]) < 1e-10 * abs(levels[1])
    y = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 0, D)
    y.add_lower_state(d, exci=0)
    y.add_lower_state(d, exci=1)
    y.run(nsweep=14, stop=None)
    e2 = y.energies()[-1]
    y.close()
    assert abs(e2 - levels[2]) < 1e-10 * abs(levels[2])


def test_lower_state_after_a_solve_is_refused(ctx, ladder):
    from su2dmrg_amd import lib as L
    lat, levels, d, e0 = ladder
    g = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 0, D)
    g.warmup()
    with pytest.raises(L.Su2Error, match="before the first solve"):
        g.add_lower_state(d, exci=0)
    g.close()


def test_lower_state_of_another_spin_is_refused(ctx, ladder):
    # the files are named after 2S: a 2S = 2 run finds no 2S = 0 state
    from su2dmrg_amd import lib as L
    lat, levels, d, e0 = ladder
    g = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 2, D)
    with pytest.raises(L.Su2Error, match="cannot read"):
        g.add_lower_state(d, exci=0)
    g.close()
