#!/usr/bin/env python3
"""Throughput of the node kernel on one MI355X.

Workload (one step = one pass of the node kernel over the lattice): the d3q19
MRT porous channel, 256^3 nodes, periodic in x and z, bounce-back walls at
y = 0 and 255, body force g_x = 1e-6, random Solid spheres of radius 4-12 up
to 10% solid (splitmix64 seed 2), tau = 0.6 with s_e = 1.19, s_eps = 1.4,
s_q = 1.2.  `value` is MLUPS (every node counted, as the reference counts
them, src/main.cpp:101-107) with FP64 storage; the other storage formats
are reported beside it.  Inputs are resident in HBM when the timed region
starts; every step runs the full kernel (the last one of the timed run also
computes the Globals).

Prints ONE JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS & HBM-roofline %, d3q19 FP64 vs FP16-store; batched solves/s, p99 latency"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def traffic_table():
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), scaled by the copy-kernel calibration; see profiles/."""
    p = os.path.join(ROOT, "profiles", "traffic_d3q19_porous256.json")
    if not os.path.exists(p):
        return {}
    with open(p) as f:
        return json.load(f).get("per_launch_bytes", {})


def cpu_baseline(n, seconds, threads):
    """The oracle (a C restatement of the reference CPU build's node kernel,
    FP64) on a bounded sample of the same workload."""
    from oracle import oracle as O
    from tclb_amd import cases

    case = cases.d3q19_porous(n=n)
    L = O.lattice_for_case(case, "f64")
    O.set_threads(threads)
    t0 = time.perf_counter()
    L.iterate(1)
    t1 = time.perf_counter() - t0
    steps = max(1, min(20000, int(seconds / max(t1, 1e-6))))
    t0 = time.perf_counter()
    L.iterate(steps)
    dt = time.perf_counter() - t0
    return {
        "value": case.nodes * steps / dt / 1e6,
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "sample": f"same generator at {n}^3 (seed 2, 10% solid), {steps} steps, FP64, "
                  f"OpenMP over (y,z) rows, {threads} threads, MPI ranks = 1, {dt:.1f} s",
    }


def batched_section(steps, seq_cases):
    """The 256-case d2q9 cylinder sweep (d2q9 MRT with Drag/Lift) in one
    batch (FP16-shift storage, FP32 collision, autotuned tiles) against the
    same cases solved one at a time as FP64 lattices (a sample of
    `seq_cases`, scaled by nodes)."""
    import tclb_amd
    from tclb_amd import cases
    model = "d2q9_drag"
    sweep = cases.d2q9_cylinder_sweep(ncases=256, steps=steps, model=model)
    nodes = sum(c.nodes for c in sweep)
    probe = tclb_amd.Batch(model, "f16s", 1 << 20)
    cap = sum(probe.case_bytes(c.nx, c.ny) for c in sweep) + (16 << 20)
    probe.close()
    b = tclb_amd.Batch(model, "f16s", capacity_bytes=cap)
    for c in sweep:
        b.add(c)
    tune = b.autotune(steps=5)
    t0 = time.perf_counter()
    b.iterate(steps)
    wall = time.perf_counter() - t0
    b.close()
    seq_nodes, seq_t = 0, 0.0
    for c in sweep[:seq_cases]:
        lat = c.build("f64")
        t0 = time.perf_counter()
        lat.iterate(steps)
        seq_t += time.perf_counter() - t0
        seq_nodes += c.nodes
        lat.close()
    seq_mlups = seq_nodes * steps / seq_t / 1e6
    batch_mlups = nodes * steps / wall / 1e6
    return {
        "workload": f"256 d2q9 (MRT) cylinder cases (nx 128-1024, ny 64-256, Re 20-200), {nodes} nodes, "
                    f"{steps} steps timed, FP16-shift storage",
        "MLUPS": round(batch_mlups, 1),
        "solves_per_s_8000_steps": round(256 / (wall * 8000 / steps), 2),
        "tuning": tune,
        "sequential_f64_MLUPS": round(seq_mlups, 1),
        "sequential_f64_sample_cases": seq_cases,
        "speedup_solves_per_s": round(batch_mlups / seq_mlups, 2),
    }


def serving_section(rate, window, sample):
    from tclb_amd import serve
    r = serve.run(rate=rate, window=window, storage="f16s", budget_gb=240.0, sample=sample)
    keep = ["rate_per_s", "window_s", "requests", "completed", "stopped_by_rule", "solves_per_s",
            "latency_p50_s", "latency_p99_s", "queue_wait_p50_s", "MLUPS", "MLUPS_while_busy",
            "sequential_f64_MLUPS", "sequential_f64_solves_per_s", "speedup_solves_per_s"]
    out = {k: (round(r[k], 4) if isinstance(r.get(k), float) else r.get(k)) for k in keep}
    out["workload"] = ("Poisson stream of d3q19 porous boxes (sides log-U[48,256], porosity U[0,25]%, "
                       "g_x U[1e-7,1e-5], tau U[0.52,0.9], budget U[1k,20k] steps or VOLvx relative change "
                       "< 1e-6 over 500 steps), FP16-shift storage, 240 GB HBM budget")
    return out


def cylinder_section(steps):
    """The 1024x128 d2q9 cylinder in FP64 with the drag/lift Globals on the
    last step of each 1000-step segment."""
    import tclb_amd
    from tclb_amd import cases
    c = cases.d2q9_cylinder(steps=steps)
    lat = c.build("f64")
    lat.iterate(100)
    t0 = time.perf_counter()
    for _ in range(steps // 1000):
        lat.iterate(1000)
    wall = time.perf_counter() - t0
    g = lat.global_dict()
    return {"workload": "d2q9 cylinder 1024x128, FP64, Globals every 1000 steps", "steps": steps,
            "MLUPS": round(c.nodes * steps / wall / 1e6, 1), "drag": g["Drag"], "lift": g["Lift"]}


def cavity_section(n, steps):
    """The 640^3 d3q19 lid-driven cavity in every storage format: MLUPS,
    and rho/u/lid-torque errors against the FP64 run after `steps` steps."""
    from tclb_amd import cavity
    r = cavity.run(n=n, steps=steps, chunk=steps, log=log)
    for v in r["formats"].values():
        for k in list(v):
            if isinstance(v[k], float):
                v[k] = round(v[k], 6) if abs(v[k]) >= 1e-3 else float(f"{v[k]:.4g}")
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--n", type=int, default=256, help="lattice side")
    ap.add_argument("--formats", default="f64,f32s,f16s,bf16s")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-n", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched, serving, cylinder and cavity sections")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=2.0)
    ap.add_argument("--serve-window", type=float, default=15.0)
    ap.add_argument("--cavity-n", type=int, default=640)
    ap.add_argument("--cavity-steps", type=int, default=300)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs one GPU only: the node kernel is measured on one MI355X "
              f"(one subdomain); --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import tclb_amd
    from tclb_amd import cases

    log(f"building the porous case {args.n}^3")
    case = cases.d3q19_porous(n=args.n)
    N = case.nodes
    copy_gbs, _ = tclb_amd.copy_bandwidth(1 << 31, 10)
    log(f"device copy ceiling {copy_gbs:.0f} GB/s")
    traffic = traffic_table()
    results = {}
    for fmt in args.formats.split(","):
        lat = case.build(fmt)
        lat.iterate(args.warmup)
        t0 = time.perf_counter()
        lat.iterate(args.steps, tclb_amd.ITER_NORM | tclb_amd.ITER_LASTGLOB)
        wall = time.perf_counter() - t0
        ev_ms, _ = lat.timing()
        bpn = lat.bytes_per_node()
        per_launch_ms = ev_ms / args.steps
        achieved = bpn * N / (per_launch_ms * 1e-3) / 1e9
        g = lat.global_dict()
        results[fmt] = {
            "MLUPS": N * args.steps / wall / 1e6,
            "ms_per_step": wall * 1e3 / args.steps,
            "kernel_ms": per_launch_ms,
            "bytes_per_node": bpn,
            "achieved_GBs": achieved,
            "roofline_frac": achieved / HBM_PEAK_GBS,
            "mean_ux_lattice": g["VOLvx"] / g["VOLvolume"] / 1.9,
        }
        log(f"{fmt}: {results[fmt]['MLUPS']:.0f} MLUPS, {achieved:.0f} GB/s "
            f"({100 * achieved / HBM_PEAK_GBS:.1f}% of 8 TB/s), {per_launch_ms:.3f} ms/step")
        lat.close()

    main_fmt = "f64" if "f64" in results else list(results)[0]
    r = results[main_fmt]
    t = traffic.get(main_fmt)
    out = {
        "metric": METRIC,
        "value": round(r["MLUPS"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"d3q19 MRT porous channel {args.n}^3, FP64 storage, 10% random spheres r 4-12, "
                        f"tau 0.6, s_e 1.19, s_eps 1.4, s_q 1.2, g_x 1e-6",
            "storage": main_fmt,
            "nodes": N,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(r["achieved_GBs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(r["roofline_frac"], 4),
            "traffic": t,
        },
        "formats": {k: {kk: (round(vv, 4) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in results.items()},
        "copy_ceiling_GBs": round(copy_gbs, 1),
    }
    if "f16s" in results and "f64" in results:
        out["f16s_over_f64"] = round(results["f16s"]["MLUPS"] / results["f64"]["MLUPS"], 3)
    if not args.no_extra:
        log("batched sweep")
        out["batched"] = batched_section(args.batch_steps, 16)
        log(f"batched: {out['batched']['MLUPS']} MLUPS, x{out['batched']['speedup_solves_per_s']} vs sequential FP64")
        log("serving")
        out["serving"] = serving_section(args.serve_rate, args.serve_window, 2)
        log(f"serving: p50 {out['serving']['latency_p50_s']} s, p99 {out['serving']['latency_p99_s']} s")
        log("cylinder 1024x128")
        out["cylinder"] = cylinder_section(5000)
        log(f"cavity {args.cavity_n}^3")
        out["cavity"] = cavity_section(args.cavity_n, args.cavity_steps)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        log(f"CPU baseline: oracle on {args.cpu_n}^3, {threads} threads")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_seconds, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
