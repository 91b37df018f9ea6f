"""Mixed-precision DMRG (FP32 MFMA Krylov matvecs, FP64 checks) on the MI355X.

The schedule of a mixed run: early sweeps with FP32 Krylov matvecs, whose
Ritz-vector check runs in FP64 and gives the step energy as an FP64 Rayleigh
quotient; final sweeps entirely FP64.  The final energy must equal exact
diagonalisation to 1e-10 relative, as the plain FP64 run does.
"""
import sys

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


@pytest.mark.parametrize("lx,ly,jcoup,S,D,n64", [
    (4, 4, 2.0, 0, 200, 2),   # config 1: J2/J1 = 0.5 on 4 (open) x 4 (periodic)
    (3, 4, 2.0, 0, 100, 3),
])
def test_mixed_schedule_matches_ed(ctx, lx, ly, jcoup, S, D, n64):
    from su2dmrg_amd import lib as L
    _say(f"mixed dmrg {lx}x{ly} jcoup={jcoup} 2S={S} D={D}: 3 sweeps FP32 + {n64} FP64")
    lat = od.Lattice(lx, ly, jcoup=jcoup, bdry=2)
    e_ed = ed.ground_energy(lat, S)
    g = L.Dmrg(ctx, lx, ly, jcoup, 0.0, 2, S, D)
    # every solve mixed (min_flops = 0), however small its matvec
    per = g.run_schedule([(D, 32)] * 3 + [(D, 64)] * n64, min_flops=0.0)
    st = g.stats()
    _say(f"  sweeps {per}, ED {e_ed}, fp32 matvecs {st['matvecs32']:.0f} of "
         f"{st['matvecs']:.0f}, largest FP64 correction {st['max_fp64_correction']:.2e}")
    assert st["matvecs32"] > 0
    # the FP32 Krylov space shows in the Ritz values, the FP64 check corrects
    assert 0 < st["max_fp64_correction"] < 1e-4 * abs(e_ed)
    # after the FP32 sweeps the energy is already close; the FP64 sweeps finish
    assert abs(per[2] - e_ed) < 1e-6 * abs(e_ed)
    assert abs(per[-1] - e_ed) < 1e-10 * abs(e_ed)
    g.close()


def test_mixed_and_fp64_runs_agree_on_16x8_at_d500(ctx):
    # the headline lattice at a bond dimension the box runs in seconds:
    # 2 sweeps FP32 then 2 FP64, against 4 FP64 sweeps; both runs' energies
    # per site are printed, and agree to the truncation's precision
    from su2dmrg_amd import lib as L
    _say("mixed vs fp64 16x8 D=500, 4 sweeps")
    res = {}
    for name, sched in (("fp64", [(500, 64)] * 4), ("mixed", [(500, 32)] * 2 + [(500, 64)] * 2)):
        g = L.Dmrg(ctx, 16, 8, 2.0, 0.0, 2, 0, 500)
        per = g.run_schedule(sched)
        st = g.stats()
        res[name] = (per, st)
        _say(f"  {name}: E/site {[0.5 * e / 128 for e in per]}, fp32 matvecs "
             f"{st['matvecs32']:.0f} of {st['matvecs']:.0f}")
        g.close()
    assert res["mixed"][1]["matvecs32"] > 0
    assert res["fp64"][1]["matvecs32"] == 0
    e64 = res["fp64"][0][-1]
    emx = res["mixed"][0][-1]
    assert abs(emx - e64) < 1e-4 * abs(e64)
