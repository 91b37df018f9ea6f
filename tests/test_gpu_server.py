"""Served stream of DMRG jobs with continuous batching (su2_server, GPU).

Jobs of different lattices, couplings, spin sectors, bond dimensions and
target levels arrive at Poisson times; the server admits them by estimated
device footprint into a fixed number of slots and advances the admitted ones
together.  Each job's result must be bit for bit what it gives served alone,
and excited levels lie above the ground state of their sector.
"""
import sys
import tempfile

import numpy as np
import pytest

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


def _jobs(L):
    return [
        L.Job(3, 4, jcoup=2.0, max_dcut=20, nsweep=2),
        L.Job(4, 3, jcoup=0.7, qcoup=0.3, max_dcut=16, nsweep=2),
        L.Job(6, 2, jcoup=2.0, max_dcut=40, exci=1, nsweep=2),
        L.Job(4, 4, jcoup=2.0, totspin=2, max_dcut=24, nsweep=2),
        L.Job(6, 2, jcoup=2.0, max_dcut=40, exci=2, nsweep=2),
        L.Job(3, 4, jcoup=1.0, max_dcut=30, nsweep=3, stop=1e-6),
    ]


def test_served_jobs_equal_jobs_alone(ctx):
    from su2dmrg_amd import lib as L
    rng = np.random.default_rng(7)
    jobs = _jobs(L)
    arrivals = np.cumsum(rng.exponential(0.05, len(jobs)))
    _say(f"serving {len(jobs)} jobs, arrivals {np.round(arrivals, 3).tolist()}")
    with tempfile.TemporaryDirectory() as d:
        srv = L.Server(ctx, 8e9, 4, d)
        ids = [srv.submit(j, float(t)) for j, t in zip(jobs, arrivals)]
        srv.run()
        mixed = [srv.result(i) for i in ids]
        st = srv.stats()
        srv.close()
    _say(f"  wall {st['wall']:.2f} s, {st['ticks']:.0f} ticks, "
         f"up to {st['max_concurrent']:.0f} jobs at once")
    assert st["max_concurrent"] >= 2
    lat = np.array([r["latency"] for r in mixed])
    _say(f"  latency p50 {np.percentile(lat, 50):.3f} s, p99 {np.percentile(lat, 99):.3f} s")
    assert np.all(lat > 0)
    for r in mixed:
        assert r["t_arrive"] <= r["t_admit"] + 1e-9 and r["t_admit"] <= r["t_finish"]
    alone = []
    for j in jobs:
        with tempfile.TemporaryDirectory() as d:
            srv = L.Server(ctx, 8e9, 1, d)
            i = srv.submit(j, 0.0)
            srv.run()
            alone.append(srv.result(i))
            srv.close()
    for k, (a, b) in enumerate(zip(alone, mixed)):
        _say(f"  job {k}: alone {a['energy']:.14f} served {b['energy']:.14f}")
        assert a["energy"] == b["energy"]
    # the excited levels of the 6x2 ladder lie above its ground level
    with tempfile.TemporaryDirectory() as d:
        srv = L.Server(ctx, 8e9, 1, d)
        i = srv.submit(L.Job(6, 2, jcoup=2.0, max_dcut=40, nsweep=2), 0.0)
        srv.run()
        e0 = srv.result(i)["energy"]
        srv.close()
    assert e0 < mixed[2]["energy"] < mixed[4]["energy"]


def test_lanes_serve_the_same_energies(ctx):
    # the same stream of jobs over three lanes (contexts, host threads)
    from su2dmrg_amd import lib as L
    rng = np.random.default_rng(7)
    jobs = _jobs(L)
    arrivals = np.cumsum(rng.exponential(0.05, len(jobs)))
    out = {}
    for lanes in (1, 3):
        _say(f"serving {len(jobs)} jobs in {lanes} lane(s)")
        with tempfile.TemporaryDirectory() as d:
            srv = L.Server(ctx, 8e9, 6, d, lanes=lanes)
            ids = [srv.submit(j, float(t)) for j, t in zip(jobs, arrivals)]
            srv.run()
            out[lanes] = [srv.result(i) for i in ids]
            st = srv.stats()
            srv.close()
        _say(f"  wall {st['wall']:.2f} s, {st['ticks']:.0f} ticks, "
             f"up to {st['max_concurrent']:.0f} jobs at once")
    for a, b in zip(out[1], out[3]):
        assert a["energy"] == b["energy"]
        assert b["t_arrive"] <= b["t_admit"] + 1e-9 and b["t_admit"] <= b["t_finish"]


def test_admission_respects_the_budget(ctx):
    from su2dmrg_amd import lib as L
    j = L.Job(4, 4, jcoup=2.0, max_dcut=30, nsweep=1)
    fp = j.footprint()
    with tempfile.TemporaryDirectory() as d:
        # room for two jobs at a time out of four slots
        srv = L.Server(ctx, 2.5 * fp, 4, d)
        ids = [srv.submit(j, 0.0) for _ in range(4)]
        srv.run()
        st = srv.stats()
        res = [srv.result(i) for i in ids]
        srv.close()
    assert st["max_concurrent"] == 2
    assert len(set(r["energy"] for r in res)) == 1
    # the same budget over two lanes: still two jobs at a time
    with tempfile.TemporaryDirectory() as d:
        srv = L.Server(ctx, 2.5 * fp, 4, d, lanes=2)
        ids = [srv.submit(j, 0.0) for _ in range(4)]
        srv.run()
        st = srv.stats()
        res2 = [srv.result(i) for i in ids]
        srv.close()
    assert st["max_concurrent"] == 2
    assert all(r["energy"] == res[0]["energy"] for r in res2)
    with pytest.raises(L.Su2Error):
        srv2 = L.Server(ctx, 0.5 * fp, 1, "/tmp")
        srv2.submit(j, 0.0)
