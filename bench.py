"""Benchmark of the SU(2) DMRG hot path on one MI355X.

A "step" is one H_eff matvec (dmrg_su2::hamiltonian_vector_multiplication_
idmrg) over the centre bond of a 16x8 J1-J2 cylinder at D=2000 SU(2)
multiplets: enlarged sectors 2j=1,3,5,7,9 with 1362/1362/629/124/9
multiplets per leg, 8 spin-1 term pairs plus both block Hamiltonians,
synthetic FP64 N(0,1) blocks from seed 0 (``su2dmrg_amd.synthetic``).
Inputs are resident in HBM before timing starts.

Prints one JSON line: matvec TFLOP/s (algorithmic FLOPs in the reference's
GEMM order / wall time), the roofline of the dominant kernel (the grouped
FP64 MFMA GEMM, HIP-event timed on the library stream) and a CPU baseline
(the numpy restatement on this host's cores, on a bounded sample).  Further
legs of the same line (each can be switched off):

* fp32: the same matvec on FP32 MFMA with FP32 accumulation;
* sweep: the real 16x8 J1-J2 (J2/J1 = 0.5) cylinder at D=2000 -- the FP64
  run (warm-up + sweeps) and the mixed-precision run (FP32 Krylov matvecs in
  the early sweeps with D ramping 500 -> 1000 -> 2000, FP64 for the last
  two), with each run's sweep times and the mixed run's |dE/site| against
  the FP64 run;
* scan: independent J-Q-J2 DMRG jobs on 8x4 and 12x6 cylinders run one after
  another, then all together in lockstep (su2_batch: one grouped launch per
  GEMM stage per Lanczos round, HIP-graph replayed), then dealt over
  --lanes lockstep batches, each on its own stream and host thread
  (su2_batch_run_lanes); jobs/hour of each;
* serve: a Poisson stream of mixed jobs (lattices, couplings, bond
  dimensions, ground and excited levels) through the continuous-batching
  server with 16 slots over --lanes lanes, jobs/hour and p50/p99 latency
  against the same jobs served one at a time (and, with --serve-one-lane,
  with the 16 slots in one lane); 64 jobs by default, so that the start of
  the stream and the drain of its last jobs are a small part of the wall
  time;
* cpu_sweep: the oracle's CPU sweep of a 16x8 cylinder at a small D beside
  the device's sweep of the same lattice and D.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 MFMA
PEAK_FP32_TFLOPS = 157.3    # MI355X dense FP32 MFMA
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def measured_traffic(variants):
    """HBM bytes of one matvec (stage 1 + stage 2 launches) from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary, when it was taken with the
    same GEMM kernel and tile variants; None otherwise."""
    path = os.path.join(ROOT, "profiles", "heff_matvec_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    same = (list(t.get("gemm_variants", [])) == list(variants)
            and t.get("kernel") == "grouped_gemm_f64_uniform")
    return t["bytes_per_matvec"] if same else None


def cpu_baseline(work, budget_s=20.0, repeats=3):
    """Oracle matvec (numpy/OpenBLAS) on the host cores: one untimed pass on a
    bounded sample of terms, then `repeats` timed passes of that sample."""
    import numpy as np  # noqa: F401
    from oracle import heff as oheff
    from su2dmrg_amd import synthetic
    model, ncpu = cpu_info()
    threads = int(os.environ.get("OMP_NUM_THREADS", ncpu))
    sub = dict(work)
    sub["terms"] = work["terms"][:1]
    t0 = time.perf_counter()
    oheff.heff_apply(work["psi"], None, None, sub["terms"])
    t1 = time.perf_counter() - t0
    per_pass = budget_s / max(1, repeats + 1)
    nterm = max(1, min(len(work["terms"]), int(per_pass / max(t1, 1e-3))))
    sub["terms"] = work["terms"][:nterm]
    oheff.heff_apply(work["psi"], work["hl"], work["hr"], sub["terms"])   # warm-up
    times = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        oheff.heff_apply(work["psi"], work["hl"], work["hr"], sub["terms"])
        times.append(time.perf_counter() - t0)
    dt = min(times)
    f = synthetic.flops(sub)
    return {"value": f / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "cpu_model": model, "nproc": ncpu,
            "sample": f"oracle numpy matvec with {nterm} of {len(work['terms'])} term pairs "
                      f"plus both block Hamiltonians: one warm-up pass, best of {repeats} "
                      f"timed passes ({', '.join(f'{t:.2f}' for t in times)} s)"}


def centre_matvec(L, ctx, dm, reps=10):
    """The FP64 matvec of the run's own centre-bond H_eff (su2_dmrg_centre),
    HIP-event timed: its sectors and TFLOP/s beside the synthetic workload's."""
    plan, psi = dm.centre()
    out = psi.like()
    alg, _ = plan.flops()
    ms, ms1, ms2 = plan.time(psi, out, reps)
    dims = sorted(((b[1][0], b[2][0], b[2][1]) for b in psi.blocks), key=lambda x: -x[1] * x[2])
    rec = {"sectors": len(psi.blocks),
           "largest_blocks_2j_rows_cols": [list(d) for d in dims[:5]],
           "psi_doubles": psi.size, "gflop": round(alg / 1e9, 3),
           "ms_per_matvec": round(ms, 4), "stage_ms": [round(ms1, 4), round(ms2, 4)],
           "tflops": round(alg / (ms * 1e-3) / 1e12, 3)}
    plan.close()
    out.close()
    psi.close()
    return rec


def run_schedule(L, ctx, lx, ly, schedule, label, centre=False, graph=True):
    """Warm-up then one sweep per (D, bits) entry; per-sweep wall times (device
    synchronised on both sides), energies and the device-side breakdown."""
    dm = L.Dmrg(ctx, lx, ly, 2.0, 0.0, 2, 0, schedule[0][0])
    dm.set_graph(graph)
    d0, b0 = schedule[0]
    dm.set_max_dcut(d0)
    dm.set_matvec_precision(b0)
    ns = lx * ly
    t0 = time.perf_counter()
    e_w = dm.warmup()
    ctx.sync()
    t_warm = time.perf_counter() - t0
    log(f"{label}: warm-up {t_warm:.1f} s, E/site {0.5 * e_w / ns:.10f}")
    prev = dm.stats()
    sweeps = []
    for k, (d, bits) in enumerate(schedule):
        dm.set_max_dcut(d)
        dm.set_matvec_precision(bits)
        t0 = time.perf_counter()
        e = dm.sweep()
        ctx.sync()
        ts = time.perf_counter() - t0
        st = dm.stats()
        dd = {x: st[x] - prev[x] for x in st}
        prev = st
        nstep = int(dd["steps"])
        rec = {"max_dcut": d, "bits": bits, "time_s": round(ts, 3),
               "energy_per_site": 0.5 * e / ns,
               "matvecs": int(dd["matvecs"]), "fp32_matvecs": int(dd["matvecs32"]),
               "matvec_tflops": round(dd["heff_flops"] / max(dd["t_matvec"], 1e-12) / 1e12, 3),
               "breakdown_s": {x: round(dd[x], 3) for x in
                               ("t_env", "t_lanczos", "t_matvec", "t_matvec32", "t_svd",
                                "t_jacobi", "t_project", "t_start")},
               "jacobi_sweeps_per_step": round(dd["jacobi_sweeps"] / max(nstep, 1), 2)}
        sweeps.append(rec)
        log(f"{label}: sweep {k + 1} D={d} {bits}-bit {ts:.1f} s, E/site {rec['energy_per_site']:.12f}"
            f", SVD {dd['t_svd']:.1f} s, matvec {dd['t_matvec']:.1f} s")
    hist = dm.history()
    st = dm.stats()
    out = {"lattice": f"{lx}x{ly} cylinder, J2/J1=0.5, S=0", "schedule": schedule,
           "warmup_time_s": round(t_warm, 3), "sweeps": sweeps,
           "total_sweep_time_s": round(sum(s["time_s"] for s in sweeps), 3),
           "final_energy_per_site": sweeps[-1]["energy_per_site"],
           "max_kept": max(h[0] for h in hist),
           "max_fp64_correction": st["max_fp64_correction"]}
    if centre:
        out["centre_bond_matvec"] = centre_matvec(L, ctx, dm)
        c = out["centre_bond_matvec"]
        log(f"{label}: centre bond {c['sectors']} sectors, {c['gflop']} GFLOP, "
            f"{c['ms_per_matvec']} ms, {c['tflops']} TFLOP/s")
    dm.close()
    return out


def sweep_legs(L, ctx, args):
    lx, ly = (int(x) for x in args.sweep_lattice.split("x"))
    D = args.sweep_d
    n = args.sweeps
    g = not args.no_graph
    fp64 = run_schedule(L, ctx, lx, ly, [(D, 64)] * n, "fp64 run", centre=True, graph=g)
    ramp = [d for d in (500, 1000, 2000) if d < D] + [D]
    ramp = (ramp + [D] * 3)[:max(1, n - 2)]
    mixed_sched = [(d, 32) for d in ramp] + [(D, 64)] * min(2, n)
    mixed = run_schedule(L, ctx, lx, ly, mixed_sched, "mixed run", graph=g)
    d_e = abs(mixed["final_energy_per_site"] - fp64["final_energy_per_site"])
    return {"fp64": fp64, "mixed": mixed, "mixed_vs_fp64_abs_de_per_site": d_e,
            "note": "both runs start from the reference's all-ones warm-up vector; with "
                    f"{n} sweeps neither is converged, so |dE/site| measures the two "
                    "schedules' different trajectories, not FP32 rounding"}


def scan_jobs(njobs, dvals, seed=1):
    """J-Q-J2 jobs of the parameter scan: J2/J1 uniform on [0.3, 0.7], Q/J1 on
    [0, 2], s = J2 + Q, -jcoup 1/s -qcoup Q/s (dmrg_su2.cpp:258-260); 8x4 and
    12x6 cylinders alternately, 2S from {0, 2, 4}, D from dvals."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for k in range(njobs):
        j2 = rng.uniform(0.3, 0.7)
        q = rng.uniform(0.0, 2.0)
        s = j2 + q
        lx, ly = ((8, 4), (12, 6))[k % 2]
        out.append(dict(lx=lx, ly=ly, jcoup=1.0 / s, qcoup=q / s,
                        totspin=int(rng.choice([0, 2, 4])), D=int(rng.choice(dvals))))
    return out


def scan_leg(L, ctx, njobs, nsweep, dvals, lanes=1):
    jobs = scan_jobs(njobs, dvals)
    sched = lambda D: [(D, 64)] * nsweep  # noqa: E731

    def make(j):
        return L.Dmrg(ctx, j["lx"], j["ly"], j["jcoup"], j["qcoup"], 2, j["totspin"], j["D"])

    # one job after another (each alone on the device)
    ctx.sync()
    t0 = time.perf_counter()
    seq = []
    for j in jobs:
        g = make(j)
        seq.append(g.run_schedule(sched(j["D"]))[-1])
        g.close()
    ctx.sync()
    t_seq = time.perf_counter() - t0
    # all together in lockstep
    b = L.Batch(ctx)
    runs = []
    t0 = time.perf_counter()
    for j in jobs:
        g = make(j)
        runs.append((g, b.add(g, sched(j["D"]))))
    b.run()
    ctx.sync()
    t_bat = time.perf_counter() - t0
    same = all(b.energies(rid)[1][-1] == e for (_, rid), e in zip(runs, seq))
    st = b.stats()
    b.close()
    for g, _ in runs:
        g.close()
    out = {"jobs": njobs, "lattices": "8x4 and 12x6", "D": list(dvals),
           "sweeps_per_job": nsweep,
           "sequential_s": round(t_seq, 3), "batched_s": round(t_bat, 3),
           "jobs_per_hour_sequential": round(3600 * njobs / t_seq, 1),
           "jobs_per_hour_batched": round(3600 * njobs / t_bat, 1),
           "speedup": round(t_seq / t_bat, 3),
           "energies_bitwise_equal": same,
           "batch_stats": {k: round(v, 3) for k, v in st.items()}}
    if lanes > 1:
        # the same jobs over several lanes: a batch per context (stream), run
        # at once from one host thread each, jobs dealt round robin
        ctxs = [ctx] + [L.Context(ctx.device) for _ in range(lanes - 1)]
        bs = [L.Batch(c) for c in ctxs]
        runs = []
        t0 = time.perf_counter()
        for i, j in enumerate(jobs):
            k = i % lanes
            g = L.Dmrg(ctxs[k], j["lx"], j["ly"], j["jcoup"], j["qcoup"], 2, j["totspin"], j["D"])
            runs.append((k, g, bs[k].add(g, sched(j["D"]))))
        L.run_lanes(bs)
        for c in ctxs:
            c.sync()
        t_lan = time.perf_counter() - t0
        same_l = all(bs[k].energies(rid)[1][-1] == e for (k, _, rid), e in zip(runs, seq))
        for _, g, _ in runs:
            g.close()
        for b in bs:
            b.close()
        for c in ctxs[1:]:
            c.close()
        out.update({"lanes": lanes, "lanes_s": round(t_lan, 3),
                    "jobs_per_hour_lanes": round(3600 * njobs / t_lan, 1),
                    "speedup_lanes": round(t_seq / t_lan, 3),
                    "energies_bitwise_equal_lanes": same_l})
    return out


def serve_leg(L, ctx, njobs, rate_per_s, lattices, dvals, seed=3, lanes=1, one_lane=False):
    import numpy as np
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(njobs):
        lx, ly = lattices[int(rng.integers(len(lattices)))]
        j2 = rng.uniform(0.3, 0.7)
        q = rng.uniform(0.0, 2.0)
        s = j2 + q
        D = int(rng.choice(dvals))
        exci = int(rng.choice((1, 2))) if k % 2 else 0
        jobs.append(L.Job(lx, ly, jcoup=1.0 / s, qcoup=q / s, max_dcut=D, exci=exci, nsweep=1))
    arrivals = np.cumsum(rng.exponential(1.0 / rate_per_s, njobs))
    budget = 0.8 * 288e9

    def serve(slots, nl):
        with tempfile.TemporaryDirectory() as d:
            srv = L.Server(ctx, budget, slots, d, lanes=nl)
            ids = [srv.submit(j, float(t)) for j, t in zip(jobs, arrivals)]
            srv.run()
            res = [srv.result(i) for i in ids]
            st = srv.stats()
            srv.close()
        lat = np.array([r["latency"] for r in res])
        return res, {"wall_s": round(st["wall"], 3), "max_concurrent": int(st["max_concurrent"]),
                     "jobs_per_hour": round(3600 * njobs / st["wall"], 1),
                     "p50_latency_s": round(float(np.percentile(lat, 50)), 3),
                     "p99_latency_s": round(float(np.percentile(lat, 99)), 3)}

    res, main = serve(16, lanes)            # 16 slots over the lanes
    res1, seq = serve(1, 1)                 # the same stream one job at a time
    out = {"jobs": njobs, "poisson_rate_per_s": rate_per_s,
           "lattices": [f"{a}x{b}" for a, b in lattices], "D": list(dvals),
           "excited_share": 0.5, "slots": 16, "lanes": lanes, "budget_bytes": budget}
    out.update(main)
    out["sequential"] = seq
    if lanes > 1 and one_lane:
        resb, one = serve(16, 1)            # the 16 slots in one lane
        out["one_lane"] = one
        out["one_lane"]["throughput_ratio"] = round(seq["wall_s"] / one["wall_s"], 3)
        res1 = res1 if all(a["energy"] == b["energy"] for a, b in zip(resb, res1)) else None
    out["throughput_ratio"] = round(seq["wall_s"] / main["wall_s"], 3)
    out["energies_bitwise_equal_to_sequential"] = res1 is not None and all(
        a["energy"] == b["energy"] for a, b in zip(res, res1))
    return out


def cpu_sweep_leg(L, ctx, lx, ly, D):
    """The oracle's CPU sweep beside the device's, same lattice, D and start."""
    from oracle import dmrg as od
    lat = od.Lattice(lx, ly, jcoup=2.0, bdry=2)
    o = od.DMRG(lat, totspin=0, max_dcut=D)
    o.warmup()
    t0 = time.perf_counter()
    e_cpu = o.sweep()
    t_cpu = time.perf_counter() - t0
    g = L.Dmrg(ctx, lx, ly, 2.0, 0.0, 2, 0, D)
    g.warmup()
    ctx.sync()
    t0 = time.perf_counter()
    e_gpu = g.sweep()
    ctx.sync()
    t_gpu = time.perf_counter() - t0
    g.close()
    model, ncpu = cpu_info()
    return {"lattice": f"{lx}x{ly}", "max_dcut": D, "cpu_sweep_s": round(t_cpu, 3),
            "gpu_sweep_s": round(t_gpu, 3), "ratio": round(t_cpu / t_gpu, 2),
            "cpu_energy_per_site": 0.5 * e_cpu / (lx * ly),
            "gpu_energy_per_site": 0.5 * e_gpu / (lx * ly),
            "energy_note": "a timing comparison: on this symmetric cylinder the warm-up keeps "
                           "sectors of rounding-noise weight, so the two trajectories part after "
                           "it; energy parity is tested on well-conditioned lattices and against "
                           "exact diagonalisation",
            "cpu": f"oracle (numpy/OpenBLAS, {model}, {ncpu} cores), one sweep after its own "
                   "warm-up"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cyl16x8_d2000",
                    choices=["cyl16x8_d2000", "cyl8x4_d1000"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--variants", default="",
                    help="pin the GEMM tile variants 'stage1,stage2' instead of autotuning")
    ap.add_argument("--legs", default="fp32,sweep,scan,serve,cpu_sweep",
                    help="comma list of extra legs (fp32, sweep, scan, serve, cpu_sweep) or 'none'")
    ap.add_argument("--sweep-lattice", default="16x8")
    ap.add_argument("--sweep-d", type=int, default=2000)
    ap.add_argument("--sweeps", type=int, default=3,
                    help="sweeps of each config-3 run (the mixed run: ramp, then two FP64)")
    ap.add_argument("--lanes", type=int, default=6,
                    help="lanes (contexts with their own stream and host thread) of the scan "
                         "and serve legs")
    ap.add_argument("--scan-jobs", type=int, default=32)
    ap.add_argument("--scan-sweeps", type=int, default=1)
    ap.add_argument("--scan-d", default="300,500", help="bond dimensions drawn by the scan jobs")
    ap.add_argument("--serve-jobs", type=int, default=64)
    ap.add_argument("--serve-one-lane", action="store_true",
                    help="also serve the stream with the 16 slots in one lane")
    ap.add_argument("--serve-rate", type=float, default=8.0)
    ap.add_argument("--serve-lattices", default="6x4,8x4,8x6,12x6")
    ap.add_argument("--serve-d", default="64,100,150")
    ap.add_argument("--cpu-sweep", default="16x8:64")
    ap.add_argument("--no-graph", action="store_true",
                    help="sweep legs: launch each Lanczos matvec directly instead of replaying "
                         "its HIP graph (same kernels and bits; for kernel tracers)")
    ap.add_argument("--svd-block", type=int, default=32, choices=[16, 32])
    ap.add_argument("--svd-inner", type=int, default=1, choices=[1, 2, 3])
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding of the matvec is designed but not built", file=sys.stderr)
        sys.exit(2)
    legs = set() if args.legs == "none" else set(args.legs.split(","))

    from su2dmrg_amd import lib as L
    from su2dmrg_amd import synthetic

    log(f"workload {args.workload}: generating synthetic blocks")
    work = getattr(synthetic, args.workload)(seed=0)
    flops = synthetic.flops(work)
    ctx = L.Context(0)
    ctx.set_svd_tuning(args.svd_block, args.svd_inner)
    dev = L.upload_workload(ctx, work)
    plan = dev["plan"]
    alg, executed = plan.flops()
    assert abs(alg - flops) < 1e-6 * flops
    log(f"plan built: {alg / 1e9:.1f} GFLOP algorithmic, {executed / 1e9:.1f} GFLOP executed")

    plan.apply(dev["psi"], dev["out"])
    if args.variants:
        chosen = tuple(int(x) for x in args.variants.split(","))
        plan.set_variants(*chosen)
        log(f"tile variants pinned to {chosen}")
    else:
        chosen, t1, t2 = plan.autotune(reps=3)
        log(f"autotune: stage1 variant {chosen[0]} ({' '.join(f'{t:.2f}' for t in t1)} ms), "
            f"stage2 variant {chosen[1]} ({' '.join(f'{t:.2f}' for t in t2)} ms)")
    for _ in range(args.warmup):
        plan.apply(dev["psi"], dev["out"])
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.apply(dev["psi"], dev["out"])
    ctx.sync()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3
    value = alg * args.steps / wall / 1e12
    log(f"timed {args.steps} steps: {ms_step:.3f} ms/step, {value:.2f} TFLOP/s")

    # kernel-level timing with HIP events on the library stream
    ms_apply, ms1, ms2 = plan.time(dev["psi"], dev["out"], max(3, min(args.steps, 10)))
    gemm_ms = ms1 + ms2
    achieved = alg / (gemm_ms * 1e-3) / 1e12
    log(f"events: apply {ms_apply:.3f} ms, stage1 {ms1:.3f} ms, stage2 {ms2:.3f} ms")

    fp32 = None
    if "fp32" in legs:
        plan.set_precision(32)
        ch32, _, _ = plan.autotune(reps=3)
        a32, s1, s2 = plan.time(dev["psi"], dev["out"], 10)
        fp32 = {"variants": list(ch32), "ms_apply": round(a32, 4),
                "ms_stage1": round(s1, 4), "ms_stage2": round(s2, 4),
                "gemm_tflops": round(alg / ((s1 + s2) * 1e-3) / 1e12, 3),
                "pct_fp32_peak": round(100 * alg / ((s1 + s2) * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 2),
                "kernel": "grouped_gemm_f64_uniform<float,...> (v_mfma_f32_16x16x4_f32, FP32 "
                          "accumulation)"}
        log(f"fp32 matvec: {fp32['gemm_tflops']} TFLOP/s ({fp32['pct_fp32_peak']}% of FP32 peak)")
        plan.set_precision(64)
    for k in ("psi", "out", "hl", "hr", "plan"):
        if dev.get(k) is not None:
            dev[k].close()
    for t in dev["ops"]:
        t.close()

    sweep = None
    if "sweep" in legs:
        try:
            sweep = sweep_legs(L, ctx, args)
        except L.Su2Error as exc:
            log(f"sweep legs failed: {exc}")
            sweep = {"error": str(exc)}
    scan = None
    if "scan" in legs:
        log(f"scan: {args.scan_jobs} jobs, sequential then batched")
        scan = scan_leg(L, ctx, args.scan_jobs, args.scan_sweeps,
                        [int(x) for x in args.scan_d.split(",")], lanes=args.lanes)
        log(f"scan: {scan['jobs_per_hour_sequential']} -> {scan['jobs_per_hour_batched']} jobs/h "
            f"({scan['speedup']}x), bitwise equal: {scan['energies_bitwise_equal']}")
        if args.lanes > 1:
            log(f"scan: {args.lanes} lanes {scan['jobs_per_hour_lanes']} jobs/h "
                f"({scan['speedup_lanes']}x), bitwise equal: {scan['energies_bitwise_equal_lanes']}")
    serve = None
    if "serve" in legs:
        log(f"serve: {args.serve_jobs} jobs at {args.serve_rate}/s")
        serve = serve_leg(L, ctx, args.serve_jobs, args.serve_rate,
                          [tuple(int(y) for y in x.split("x")) for x in args.serve_lattices.split(",")],
                          [int(x) for x in args.serve_d.split(",")], lanes=args.lanes,
                          one_lane=args.serve_one_lane)
        log(f"serve: {serve['jobs_per_hour']} jobs/h ({args.lanes} lanes), p50 "
            f"{serve['p50_latency_s']} s, p99 {serve['p99_latency_s']} s; sequential "
            f"{serve['sequential']['jobs_per_hour']} jobs/h"
            + (f"; one lane {serve['one_lane']['jobs_per_hour']} jobs/h" if "one_lane" in serve else "")
            + f"; ratio {serve['throughput_ratio']}")

    cpu = None
    cpu_sweep = None
    if not args.no_cpu:
        log("cpu baseline (oracle matvec) ...")
        cpu = cpu_baseline(work, args.cpu_budget)
        log(f"cpu baseline {cpu['value']:.3f} TFLOP/s on {cpu['cores']} threads")
        if "cpu_sweep" in legs:
            lat, dcut = args.cpu_sweep.split(":")
            lx, ly = (int(x) for x in lat.split("x"))
            log(f"cpu sweep: oracle vs device, {lat} at D={dcut}")
            cpu_sweep = cpu_sweep_leg(L, ctx, lx, ly, int(dcut))
            log(f"cpu sweep {cpu_sweep['cpu_sweep_s']} s vs device {cpu_sweep['gpu_sweep_s']} s")
        if sweep and "fp64" in sweep:
            # the headline sweep on the CPU cannot finish in the bench; its GEMM
            # work alone, at the measured CPU matvec rate, bounds it from below
            s0 = sweep["fp64"]["sweeps"][0]
            tflop = s0["matvec_tflops"] * s0["breakdown_s"]["t_matvec"]
            cpu["sweep_lower_bound_s"] = round(tflop / cpu["value"], 1)
            cpu["sweep_lower_bound_note"] = (
                "the first FP64 16x8 D=2000 sweep's algorithmic matvec FLOPs divided by the "
                "CPU matvec rate above (matvec GEMMs only; no SVD, environments or copies)")

    line = {
        "metric": "Heff matvec TFLOP/s (% FP64 MFMA peak) + sweep time, 16x8 J1-J2 D=2000",
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "pct_fp64_peak": round(100.0 * value / PEAK_FP64_TFLOPS, 2),
        "config": {
            "workload": f"{args.workload}: H_eff matvec, centre bond, "
                        f"{len(work['terms'])} spin-1 term pairs + hL + hR, "
                        f"sectors 2j={work['twoj']} dims={work['dims_a']}",
            "algorithmic_gflop_per_step": round(alg / 1e9, 3),
            "sweep_time_s": None if not sweep or "fp64" not in sweep
            else sweep["fp64"]["sweeps"][0]["time_s"],
            "gemm_variants": list(chosen),
            "svd_block_cols": args.svd_block,
            "svd_inner_sweeps": args.svd_inner,
        },
        "fp32_matvec": fp32,
        "sweep": sweep,
        "scan": scan,
        "serve": serve,
        "cpu_sweep": cpu_sweep,
        "roofline": {
            "bound": "mfma",
            "achieved": round(achieved, 4),
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP64_TFLOPS, 4),
            "traffic": measured_traffic(chosen),
            "kernel": "grouped_gemm_f64_uniform (stage 1 + stage 2 launches of one matvec; "
                      "traffic in bytes per matvec from profiles/heff_matvec_traffic.json)",
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
