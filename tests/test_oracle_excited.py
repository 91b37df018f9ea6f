"""Excited-state targeting of the oracle DMRG (the reference's exci > 0 path)
against exact diagonalisation (CPU only)."""
import numpy as np
import pytest

from oracle import dmrg as od
from oracle import ed, mpsio

# 6 (open) x 2 (open) J1-J2 at J2/J1 = 0.5: the three lowest S = 0 levels are
# non-degenerate (-14.0265, -10.1350, -10.0104), and D = 100 is exact
LAT = dict(lx=6, ly=2, jcoup=2.0, bdry=0)


@pytest.fixture(scope="module")
def ladder():
    lat = od.Lattice(LAT["lx"], LAT["ly"], jcoup=LAT["jcoup"], bdry=LAT["bdry"])
    levels = ed.sector_energies(lat, 0, 3)
    g = od.DMRG(lat, 0, 100)
    g.run(nsweep=8, stop=None)
    x = od.DMRG(lat, 0, 100, lower=[od.LowerState.from_dmrg(g)])
    x.run(nsweep=14, stop=None)
    return lat, levels, g, x


def test_first_and_second_excited_levels_match_ed(ladder):
    lat, levels, g, x = ladder
    assert abs(g.energies[-1] - levels[0]) < 1e-10 * abs(levels[0])
    assert abs(x.energies[-1] - levels[1]) < 1e-10 * abs(levels[1])
    y = od.DMRG(lat, 0, 100, lower=[od.LowerState.from_dmrg(g), od.LowerState.from_dmrg(x)])
    y.run(nsweep=14, stop=None)
    assert abs(y.energies[-1] - levels[2]) < 1e-10 * abs(levels[2])


def test_excited_state_is_orthogonal_to_the_lower_one(ladder):
    # the converged excited wavefunction has no weight on the ground state's
    # projection into the same two-site basis
    lat, levels, g, x = ladder
    il, ir = x.center
    El = od.Enlarged(x.lb[il - 1], left=True)
    Er = od.Enlarged(x.rb[ir + 1], left=False)
    lay = {j: x.psi[j].shape for j in x.psi}
    (e0, phi), = x.lower_vectors(il, ir, El, Er, lay)
    ov = sum(float(np.sum(phi[j] * x.psi[j])) for j in x.psi)
    nphi = np.sqrt(sum(float(np.sum(phi[j] ** 2)) for j in phi))
    assert abs(nphi - 1.0) < 1e-8          # the ground state lives in this basis
    assert abs(ov) < 1e-6


def test_lower_state_from_files_equals_the_run(ladder, tmp_path):
    lat, levels, g, x = ladder
    mpsio.save(g, str(tmp_path))
    a = od.LowerState.from_dmrg(g)
    b = mpsio.lower_state(str(tmp_path), g.ns, g.D, 0, 0)
    assert abs(a.energy - b.energy) < 1e-12
    for k in range(g.ns):
        assert a.beta[k] == b.beta[k]
        assert sorted(a.T[k]) == sorted(b.T[k])
        for key in a.T[k]:
            np.testing.assert_array_equal(a.T[k][key], b.T[k][key])
