"""Checkpoint files of the device DMRG against the oracle's (GPU)."""
import os
import sys

import numpy as np
import pytest

from oracle import dmrg as od
from oracle import mpsio

pytestmark = pytest.mark.gpu

# open 6x2 J1-J2 at D = 10: truncating, and well conditioned (its oracle
# trajectory moves by < 1e-11 under 1e-15 Gram perturbations)
LX, LY, JC, BD, D = 6, 2, 2.0, 0, 10


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
def states(ctx, tmp_path_factory):
    from su2dmrg_amd import lib as L
    lat = od.Lattice(LX, LY, jcoup=JC, bdry=BD)
    o = od.DMRG(lat, totspin=0, max_dcut=D)
    o.run(nsweep=1, stop=None)
    g = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 0, D)
    g.run(nsweep=1, stop=None)
    do = str(tmp_path_factory.mktemp("oracle"))
    dg = str(tmp_path_factory.mktemp("device"))
    mpsio.save(o, do)
    g.save(dg)
    yield lat, o, g, do, dg
    g.close()


def test_device_files_match_oracle_files(states):
    _say("checkpoint: device files vs oracle files")
    lat, o, g, do, dg = states
    ns = LX * LY
    assert sorted(os.listdir(do)) == sorted(os.listdir(dg))
    for i in range(ns):
        base = f"uu-{ns}-{D}-0-0-0-{i}"
        assert open(os.path.join(dg, base + ".dat")).read() == \
            open(os.path.join(do, base + ".dat")).read()
        lg, bg = mpsio.read_tensor3(os.path.join(dg, base))
        lo, bo = mpsio.read_tensor3(os.path.join(do, base))
        assert sorted(bg) == sorted(bo)
        for k in bo:
            np.testing.assert_allclose(bg[k], bo[k], rtol=0, atol=1e-9)
        # the CG doubles are make_cgc's to the last bit of a double
        rg = open(os.path.join(dg, base + ".bin"), "rb").read()
        ro = open(os.path.join(do, base + ".bin"), "rb").read()
        assert len(rg) == len(ro)
        nred = sum(b.size for b in bo.values())
        nten = lo[0].nsec * lo[1].nsec * lo[2].nsec
        assert rg[:4 * nten] == ro[:4 * nten]
        cg_g = np.frombuffer(rg[4 * nten:], dtype=np.float64)[nred:]
        cg_o = np.frombuffer(ro[4 * nten:], dtype=np.float64)[nred:]
        np.testing.assert_allclose(cg_g, cg_o, rtol=0, atol=1e-15)
    wg = np.fromfile(os.path.join(dg, f"ww-{ns}-{D}-0-0-0-{ns // 2}.bin"))
    wo = np.fromfile(os.path.join(do, f"ww-{ns}-{D}-0-0-0-{ns // 2}.bin"))
    np.testing.assert_allclose(wg, wo, rtol=0, atol=1e-10)
    eg = float(open(os.path.join(dg, f"enr-{ns}-{D}-0-0-0.dat")).read().split()[-1])
    eo = float(open(os.path.join(do, f"enr-{ns}-{D}-0-0-0.dat")).read().split()[-1])
    assert abs(eg - eo) < 1e-10


def test_device_round_trip_continues_the_trajectory(ctx, states):
    from su2dmrg_amd import lib as L
    _say("checkpoint: device save/load round trip")
    lat, o, g, do, dg = states
    full = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 0, D)
    full.run(nsweep=2, stop=None)
    ref = full.energies()
    full.close()
    r = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 0, D)
    e0 = r.load(dg)
    assert abs(e0 - g.energies()[-1]) < 1e-12
    r.sweep()
    got = r.energies()
    r.close()
    np.testing.assert_allclose(got, ref[-len(got):], rtol=0, atol=1e-12)


def test_device_continues_from_oracle_files(ctx, states):
    from su2dmrg_amd import lib as L
    _say("checkpoint: device loads the oracle's files")
    lat, o, g, do, dg = states
    o2 = od.DMRG(lat, totspin=0, max_dcut=D)
    mpsio.load(o2, do)
    o2.sweep()
    r = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 0, D)
    r.load(do)
    r.sweep()
    got = r.energies()
    r.close()
    np.testing.assert_allclose(got, o2.energies, rtol=1e-10, atol=1e-10)


def test_load_rejects_files_of_another_lattice(ctx, states, tmp_path):
    from su2dmrg_amd import lib as L
    lat, o, g, do, dg = states
    r = L.Dmrg(ctx, LX, LY, JC, 0.0, BD, 0, D)
    with pytest.raises(RuntimeError):
        r.load(str(tmp_path))          # no files at all
    r.close()
