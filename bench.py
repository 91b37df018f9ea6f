"""Benchmark of the SU(2) H_eff matvec on one MI355X.

A "step" is one H_eff matvec (dmrg_su2::hamiltonian_vector_multiplication_
idmrg) over the centre bond of a 16x8 J1-J2 cylinder at D=2000 SU(2)
multiplets: enlarged sectors 2j=1,3,5,7,9 with 1362/1362/629/124/9
multiplets per leg, 8 spin-1 term pairs plus both block Hamiltonians,
synthetic FP64 N(0,1) blocks from seed 0 (``su2dmrg_amd.synthetic``).
Inputs are resident in HBM before timing starts.

Prints one JSON line: matvec TFLOP/s (algorithmic FLOPs in the reference's
GEMM order / wall time), the roofline of the dominant kernel (the grouped
FP64 MFMA GEMM, HIP-event timed on the library stream) and a CPU baseline
(the numpy restatement on this host's cores, on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6     # MI355X dense FP64 MFMA
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def cpu_baseline(work, budget_s=20.0):
    """Time the oracle matvec (numpy/OpenBLAS) on a bounded sample of terms."""
    import numpy as np
    from oracle import heff as oheff
    from su2dmrg_amd import synthetic
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # one term first to size the sample
    sub = dict(work)
    sub["terms"] = work["terms"][:1]
    t0 = time.perf_counter()
    oheff.heff_apply(work["psi"], None, None, sub["terms"])
    t1 = time.perf_counter() - t0
    nterm = max(1, min(len(work["terms"]), int(budget_s / max(t1, 1e-3))))
    sub["terms"] = work["terms"][:nterm]
    t0 = time.perf_counter()
    oheff.heff_apply(work["psi"], work["hl"], work["hr"], sub["terms"])
    dt = time.perf_counter() - t0
    f = synthetic.flops(sub)
    return {"value": f / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"oracle numpy matvec with {nterm} of {len(work['terms'])} term pairs "
                      f"plus both block Hamiltonians, one pass ({dt:.2f} s)"}


def sweep_leg(L, ctx, lx, ly, dcut, variants):
    """Time one full sweep of the real 16x8 J1-J2 (J2/J1 = 0.5) cylinder.

    The warm-up (do_idmrg) grows the blocks at the same D and is untimed; the
    sweep (dmrg_su2::sweep, three passes, nfree = 2) is timed end to end on
    the host clock, after a device synchronisation on both sides.
    """
    dm = L.Dmrg(ctx, lx, ly, 2.0, 0.0, 2, 0, dcut)
    dm.set_verbose(1)
    if variants:
        dm.set_heff_variants(*variants)
    t0 = time.perf_counter()
    e_w = dm.warmup()
    ctx.sync()
    t_warm = time.perf_counter() - t0
    s0 = dm.stats()
    log(f"sweep leg: warm-up {t_warm:.1f} s, E = {e_w:.10f}")
    t0 = time.perf_counter()
    e = dm.sweep()
    ctx.sync()
    t_sweep = time.perf_counter() - t0
    s1 = dm.stats()
    d = {k: s1[k] - s0[k] for k in s1}
    hist = dm.history()
    nstep = int(d["steps"])
    out = {
        "lattice": f"{lx}x{ly} cylinder, J2/J1=0.5, S=0",
        "max_dcut": dcut,
        "sweep_time_s": round(t_sweep, 3),
        "warmup_time_s": round(t_warm, 3),
        "steps": nstep,
        "energy": e,
        "energy_per_site": 0.5 * e / (lx * ly),
        "max_kept": max(h[0] for h in hist[-nstep:]),
        "matvecs": int(d["matvecs"]),
        "matvec_tflops": round(d["heff_flops"] / max(d["t_matvec"], 1e-12) / 1e12, 3),
        "breakdown_s": {k: round(d[k], 3) for k in
                        ("t_env", "t_lanczos", "t_matvec", "t_svd", "t_jacobi", "t_project",
                         "t_start")},
        "jacobi_sweeps_per_step": round(d["jacobi_sweeps"] / max(nstep, 1), 2),
    }
    dm.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="cyl16x8_d2000",
                    choices=["cyl16x8_d2000", "cyl8x4_d1000"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--variants", default="",
                    help="pin the GEMM tile variants 'stage1,stage2' instead of autotuning")
    ap.add_argument("--sweep-lattice", default="16x8",
                    help="lattice of the timed sweep, or 'none' to skip the sweep leg")
    ap.add_argument("--sweep-d", type=int, default=2000, help="max multiplets of the sweep leg")
    ap.add_argument("--svd-block", type=int, default=32, choices=[16, 32],
                    help="column block width of the Jacobi SVD")
    ap.add_argument("--svd-inner", type=int, default=1, choices=[1, 2, 3],
                    help="inner Jacobi sweeps per visit of a column pair")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding of the matvec is designed but not built", file=sys.stderr)
        sys.exit(2)

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

    sweep = None
    if args.sweep_lattice != "none":
        lx, ly = (int(x) for x in args.sweep_lattice.split("x"))
        log(f"sweep leg: {lx}x{ly} at D={args.sweep_d}")
        try:
            sweep = sweep_leg(L, ctx, lx, ly, args.sweep_d, None)
            log(f"sweep leg: {sweep['sweep_time_s']:.1f} s, E/site {sweep['energy_per_site']:.10f}")
        except L.Su2Error as exc:
            # the matvec figures above stand on their own; record why the sweep did not
            log(f"sweep leg failed: {exc}")
            sweep = {"lattice": f"{lx}x{ly}", "max_dcut": args.sweep_d, "error": str(exc)}

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (oracle) ...")
        cpu = cpu_baseline(work, args.cpu_budget)
        log(f"cpu baseline {cpu['value']:.3f} TFLOP/s on {cpu['cores']} threads")

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
            "sweep_time_s": None if sweep is None else sweep.get("sweep_time_s"),
            "sweep": sweep,
            "gemm_variants": list(chosen),
            "svd_block_cols": args.svd_block,
            "svd_inner_sweeps": args.svd_inner,
        },
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
