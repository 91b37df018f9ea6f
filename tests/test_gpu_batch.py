"""Several DMRG runs in lockstep (su2_batch) against the same runs alone (GPU).

A batch puts the H_eff applies of all runs' Lanczos solves into one grouped
launch per GEMM stage (replayed from a HIP graph every round) and all their
wavefunctions into one Jacobi SVD.  Every problem keeps its own K split and
every block its own slicing, so each run's step energies must equal, bit for
bit, those it gives alone -- whatever else shares the launches.
"""
import sys

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


# (lx, ly, jcoup, qcoup, 2S, schedule of (D, bits) sweeps): plain, a higher
# spin sector, a J-Q lattice with plaquette terms, a mixed-precision run
RUNS = [
    (3, 4, 2.0, 0.0, 0, [(20, 64), (20, 64)]),
    (4, 4, 2.0, 0.0, 2, [(24, 64), (30, 64)]),
    (4, 3, 0.7, 0.3, 0, [(16, 64), (16, 64)]),
    (6, 2, 2.0, 0.0, 0, [(12, 32), (12, 32), (12, 64)]),
]


def _alone(L, ctx, spec, graph=True):
    lx, ly, j, q, S, sched = spec
    g = L.Dmrg(ctx, lx, ly, j, q, 2, S, sched[0][0])
    g.set_graph(graph)
    d0, b0 = sched[0]
    g.set_max_dcut(d0)
    g.set_matvec_precision(b0, 0.0)
    g.warmup()
    for d, b in sched:
        g.set_max_dcut(d)
        g.set_matvec_precision(b, 0.0)
        g.sweep()
    out = (np.array(g.energies()), g.history())
    g.close()
    return out


def _batched(L, ctx, specs, graph=True):
    b = L.Batch(ctx)
    b.set_graph(graph)
    runs = []
    for lx, ly, j, q, S, sched in specs:
        g = L.Dmrg(ctx, lx, ly, j, q, 2, S, sched[0][0])
        # every solve of a 32-bit sweep mixed, however small (as in _alone)
        runs.append((g, b.add(g, sched, min_flops=0.0)))
    ticks = 0
    while b.tick() > 0:
        ticks += 1
    st = b.stats()
    out = []
    for g, rid in runs:
        fin, ph = b.energies(rid)
        assert fin
        out.append((np.array(g.energies()), g.history(), ph))
    b.close()
    for g, _ in runs:
        g.close()
    return out, ticks, st


def test_batched_runs_equal_runs_alone(ctx):
    from su2dmrg_amd import lib as L
    _say("batch of 4 runs vs each alone")
    alone = [_alone(L, ctx, s) for s in RUNS]
    got, ticks, st = _batched(L, ctx, RUNS)
    _say(f"  {ticks} ticks, {st['graphs_built']:.0f} graphs built, "
         f"{st['graph_replays']:.0f} replays")
    assert ticks == max(len(a[0]) for a in alone)
    assert st["graph_replays"] > st["graphs_built"] > 0
    for (ea, ha), (eb, hb, ph) in zip(alone, got):
        assert len(ea) == len(eb)
        assert np.array_equal(ea, eb), np.max(np.abs(ea - eb))
        assert ha == hb
        assert ph[-1] == eb[-1]


def test_graph_replays_equal_direct_launches(ctx):
    # every round replays a recorded launch sequence with new inputs, with
    # the runs' own vector algebra and host waits between replays
    from su2dmrg_amd import lib as L
    _say("batch with and without HIP graphs")
    a, _, sa = _batched(L, ctx, RUNS[:3], graph=True)
    b, _, sb = _batched(L, ctx, RUNS[:3], graph=False)
    assert sb["graphs_built"] == 0 and sa["graphs_built"] > 0
    for (ea, _, _), (eb, _, _) in zip(a, b):
        assert np.array_equal(ea, eb)


def test_single_run_graph_and_direct_agree(ctx):
    from su2dmrg_amd import lib as L
    _say("single run with and without HIP graphs")
    ea, _ = _alone(L, ctx, RUNS[1], graph=True)
    eb, _ = _alone(L, ctx, RUNS[1], graph=False)
    assert np.array_equal(ea, eb)


def test_lanes_equal_runs_alone(ctx):
    # three lanes, each a batch on its own context (stream) and host thread;
    # the runs are spread over them round robin and run at once
    from su2dmrg_amd import lib as L
    _say("runs in 3 lanes vs each alone")
    alone = [_alone(L, ctx, s) for s in RUNS]
    ctxs = [ctx, L.Context(0), L.Context(0)]
    lanes = [L.Batch(c) for c in ctxs]
    runs = []
    for i, (lx, ly, j, q, S, sched) in enumerate(RUNS):
        k = i % 3
        g = L.Dmrg(ctxs[k], lx, ly, j, q, 2, S, sched[0][0])
        runs.append((k, g, lanes[k].add(g, sched, min_flops=0.0)))
    L.run_lanes(lanes)
    for (ea, ha), (k, g, rid) in zip(alone, runs):
        fin, ph = lanes[k].energies(rid)
        assert fin
        eb = np.array(g.energies())
        assert np.array_equal(ea, eb), np.max(np.abs(ea - eb))
        assert ha == g.history()
        assert ph[-1] == eb[-1]
    for _, g, _ in runs:
        g.close()
    for b in lanes:
        b.close()
    for c in ctxs[1:]:
        c.close()


def test_lanes_reject_a_shared_context(ctx):
    from su2dmrg_amd import lib as L
    a, b = L.Batch(ctx), L.Batch(ctx)
    with pytest.raises(L.Su2Error, match="one context"):
        L.run_lanes([a, b])
    a.close()
    b.close()
