#!/usr/bin/env python3
"""Throughput of the node kernel on one MI355X.

Workload (one step = one pass of the node kernel over the lattice): the d3q19
MRT porous channel, 256^3 nodes, periodic in x and z, bounce-back walls at
y = 0 and 255, body force g_x = 1e-6, random Solid spheres of radius 4-12 up
to 10% solid (splitmix64 seed 2), tau = 0.6 with s_e = 1.19, s_eps = 1.4,
s_q = 1.2.  `value` is MLUPS (every node counted, as the reference counts
them, src/main.cpp:101-107) with FP64 storage; the other storage formats
are reported beside it ("f16s+sr": FP16-shift with stochastic rounding of
its stores).  Inputs are resident in HBM when the timed region
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


def _oracle_mlups(n, seconds, threads):
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
    return case.nodes * steps / dt / 1e6, steps, dt


def cpu_share():
    """Host threads this process may use: the OpenMP setting of the host
    (the GPU box gives one GPU's process a share of its cores), else every
    core."""
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        return int(env)
    return os.cpu_count() or 1


def cpu_baseline(n, seconds, threads):
    """The oracle (a C restatement of the reference CPU build's node kernel,
    FP64, gcc -O3) on a bounded sample of the same workload: OpenMP over the
    host threads given to this process, and one thread (one MPI rank of the
    reference, no OpenMP)."""
    mt, steps, dt = _oracle_mlups(n, seconds, threads)
    st, steps1, dt1 = _oracle_mlups(n, seconds / 3, 1)
    return {
        "value": round(mt, 2),
        "unit": "MLUPS",
        "cores": threads,
        "kind": "port",
        "single_thread_value": round(st, 3),
        "nproc": os.cpu_count(),
        "sample": f"same generator at {n}^3 (seed 2, 10% solid), FP64, oracle built gcc -O3; "
                  f"{steps} steps on {threads} OpenMP threads over (y,z) rows ({dt:.1f} s; the host's "
                  f"OMP_NUM_THREADS share of its {os.cpu_count()} logical cores) and {steps1} steps on "
                  f"1 thread ({dt1:.1f} s); MPI ranks = 1",
    }


def batched_section(steps, seq_cases):
    """The 256-case d2q9 cylinder sweep (d2q9 MRT with Drag/Lift) in one
    batch (FP16-shift storage, FP32 collision, autotuned tiles) against the
    same cases solved one at a time as plain lattices, in the same storage
    format (the gain of batching) and in FP64 (a sample of `seq_cases`,
    scaled by nodes).  The lone lattices take their fastest step, the
    temporally blocked one where their timing picks it."""
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
    b.prepare()
    t0 = time.perf_counter()
    b.iterate(steps)
    wall = time.perf_counter() - t0
    b.close()
    batch_mlups = nodes * steps / wall / 1e6
    out = {
        "workload": f"256 d2q9 (MRT) cylinder cases (nx 128-1024, ny 64-256, Re 20-200), {nodes} nodes, "
                    f"{steps} steps timed, FP16-shift storage",
        "MLUPS": round(batch_mlups, 1),
        "solves_per_s_8000_steps": round(256 / (wall * 8000 / steps), 2),
        "tuning": tune,
        "sequential_sample_cases": seq_cases,
    }
    for st in ("f16s", "f64"):
        seq_nodes, seq_t = 0, 0.0
        for c in sweep[:seq_cases]:
            lat = c.build(st)
            lat.prepare()
            t0 = time.perf_counter()
            lat.iterate(steps)
            seq_t += time.perf_counter() - t0
            seq_nodes += c.nodes
            lat.close()
        seq_mlups = seq_nodes * steps / seq_t / 1e6
        out[f"sequential_{st}_MLUPS"] = round(seq_mlups, 1)
        out[f"speedup_vs_sequential_{st}"] = round(batch_mlups / seq_mlups, 2)
    return out


SERVE_KEEP = ["rate_per_s", "window_s", "arrival_span_s", "requests", "completed", "stopped_by_rule", "rejected",
              "makespan_s", "solves_per_s", "latency_p50_s", "latency_p99_s", "queue_wait_p50_s", "MLUPS",
              "MLUPS_while_busy", "mean_nodes", "mean_steps_run",
              "sequential_f16s_MLUPS", "sequential_f16s_solves_per_s", "speedup_vs_sequential_f16s",
              "sequential_f64_MLUPS", "sequential_f64_solves_per_s", "speedup_vs_sequential_f64",
              "sequential_sample"]


def _serve_keep(r):
    return {k: (round(r[k], 4) if isinstance(r[k], float) else r[k]) for k in SERVE_KEEP if k in r}


def serving_section(rate, window, sample, nsat):
    """The d3q19 serving stream: at the stated 50 req/s (backlogged: the
    first `nsat` requests arrive within half a second, far faster than they
    are served, so solves/s is the server's capacity) and at a rate below
    capacity (latency without queueing)."""
    from tclb_amd import serve
    out = {"workload": ("Poisson stream of d3q19 porous boxes (sides log-U[48,256], porosity U[0,25]%, "
                        "g_x U[1e-7,1e-5], tau U[0.52,0.9], budget U[1k,20k] steps or VOLvx relative change "
                        "< 1e-6 over 500 steps), FP16-shift storage, 240 GB HBM budget")}
    out["saturated"] = _serve_keep(serve.saturated(nreq=nsat, rate=50.0, storage="f16s", budget_gb=240.0,
                                                   sample=sample, log=log))
    out["below_capacity"] = _serve_keep(serve.run(rate=rate, window=window, storage="f16s", budget_gb=240.0,
                                                  sample=sample, log=log))
    return out


def cylinder_section(steps):
    """The 1024x128 d2q9 cylinder in FP64 with the drag/lift Globals on the
    last step of each 1000-step segment; plain steps as the lattice chooses
    by timing (`temporal_blocking`: steps per launch, 0 for single steps)."""
    import tclb_amd
    from tclb_amd import cases
    c = cases.d2q9_cylinder(steps=steps)
    lat = c.build("f64")
    lat.prepare()
    lat.iterate(400)   # its first 270 steps time the temporal blocking
    t0 = time.perf_counter()
    for _ in range(steps // 1000):
        lat.iterate(1000)
    wall = time.perf_counter() - t0
    g = lat.global_dict()
    return {"workload": "d2q9 cylinder 1024x128, FP64, Globals every 1000 steps", "steps": steps,
            "MLUPS": round(c.nodes * steps / wall / 1e6, 1), "temporal_blocking": lat.temporal_blocking,
            "drag": g["Drag"], "lift": g["Lift"]}


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
    ap.add_argument("--formats", default="f64,f32s,f16s,bf16s,f16s+sr")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-n", type=int, default=64)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched, serving, cylinder and cavity sections")
    ap.add_argument("--batch-steps", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=2.0)
    ap.add_argument("--serve-window", type=float, default=15.0)
    ap.add_argument("--serve-saturated", type=int, default=24,
                    help="requests of the 50 req/s stream played backlogged")
    ap.add_argument("--perm-steps", type=int, default=5000, help="FP64 steps before the permeability is read")
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
    nu = (0.6 - 0.5) / 3     # tau = 0.6
    g_x = 1e-6
    permeability = None
    for fmt in args.formats.split(","):
        # "<f16s|bf16s>+sr": that format with stochastic rounding of its stores
        storage, sr = (fmt[:-3], True) if fmt.endswith("+sr") else (fmt, False)
        lat = case.build(storage)
        if sr:
            lat.set_stochastic_rounding(True, seed=1)
        # the rows-per-tile timing and the graph capture happen here, not in
        # the timed steps
        lat.prepare()
        lat.iterate(args.warmup)
        t0 = time.perf_counter()
        lat.iterate(args.steps, tclb_amd.ITER_NORM | tclb_amd.ITER_LASTGLOB)
        wall = time.perf_counter() - t0
        ev_ms, _ = lat.timing()
        bpn = lat.bytes_per_node()
        per_launch_ms = ev_ms / args.steps
        achieved = bpn * N / (per_launch_ms * 1e-3) / 1e9
        g = lat.global_dict()
        # VOLvx sums the fluid nodes' velocity scaled by 1.9 (the reference
        # computes its Globals after R /= Mw, Dynamics.c.Rt:262-289)
        ux_fluid = g["VOLvx"] / g["VOLvolume"] / 1.9
        results[fmt] = {
            "MLUPS": N * args.steps / wall / 1e6,
            "ms_per_step": wall * 1e3 / args.steps,
            "kernel_ms": per_launch_ms,
            "bytes_per_node": bpn,
            "achieved_GBs": achieved,
            "roofline_frac": achieved / HBM_PEAK_GBS,
            "steps_run": args.warmup + args.steps,
            "mean_ux_fluid": float(f"{ux_fluid:.6g}"),
        }
        log(f"{fmt}: {results[fmt]['MLUPS']:.0f} MLUPS, {achieved:.0f} GB/s "
            f"({100 * achieved / HBM_PEAK_GBS:.1f}% of 8 TB/s), {per_launch_ms:.3f} ms/step "
            f"(wall {results[fmt]['ms_per_step']:.3f} ms/step)")
        if fmt == "f64" and args.perm_steps > 0:
            # the stated 5,000 steps, untimed, for the permeability
            more = args.perm_steps - (args.warmup + args.steps)
            if more > 0:
                lat.iterate(more, tclb_amd.ITER_NORM | tclb_amd.ITER_LASTGLOB)
            g = lat.global_dict()
            ux_f = g["VOLvx"] / g["VOLvolume"] / 1.9
            ux_d = g["VOLvx"] / N / 1.9
            permeability = {
                "steps": lat.iteration,
                "mean_ux_fluid": float(f"{ux_f:.6g}"),
                "darcy_velocity": float(f"{ux_d:.6g}"),
                "porosity": float(f"{g['VOLvolume'] / N:.6g}"),
                "k_lattice_units": float(f"{nu * ux_d / g_x:.6g}"),
                "note": "k = nu <u_x>/g_x with <u_x> the Darcy (superficial) velocity, sum of u_x over all nodes; "
                        "not converged at this step count: the viscous time of the 256-node channel, "
                        "H^2/nu, is about 2e6 steps",
            }
            log(f"permeability after {lat.iteration} steps: k = {permeability['k_lattice_units']}")
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
        "formats": {k: {kk: (round(vv, 4) if isinstance(vv, float) and kk != "mean_ux_fluid" else vv)
                        for kk, vv in v.items()}
                    for k, v in results.items()},
        "copy_ceiling_GBs": round(copy_gbs, 1),
    }
    if permeability:
        out["permeability"] = permeability
    if "f16s" in results and "f64" in results:
        out["f16s_over_f64"] = round(results["f16s"]["MLUPS"] / results["f64"]["MLUPS"], 3)
    if "f16s+sr" in results and "f64" in results:
        out["f16s_sr_over_f64"] = round(results["f16s+sr"]["MLUPS"] / results["f64"]["MLUPS"], 3)
    if not args.no_extra:
        log("batched sweep")
        out["batched"] = batched_section(args.batch_steps, 16)
        bt = out["batched"]
        log(f"batched: {bt['MLUPS']} MLUPS, x{bt['speedup_vs_sequential_f16s']} vs sequential f16s, "
            f"x{bt['speedup_vs_sequential_f64']} vs sequential FP64")
        log("serving")
        out["serving"] = serving_section(args.serve_rate, args.serve_window, 2, args.serve_saturated)
        sv = out["serving"]
        log(f"serving: saturated {sv['saturated']['solves_per_s']} solves/s "
            f"(x{sv['saturated'].get('speedup_vs_sequential_f16s')} vs f16s, "
            f"x{sv['saturated'].get('speedup_vs_sequential_f64')} vs FP64); below capacity p50 "
            f"{sv['below_capacity']['latency_p50_s']} s, p99 {sv['below_capacity']['latency_p99_s']} s")
        log("cylinder 1024x128")
        out["cylinder"] = cylinder_section(5000)
        log(f"cavity {args.cavity_n}^3")
        out["cavity"] = cavity_section(args.cavity_n, args.cavity_steps)
    if not args.no_cpu:
        threads = cpu_share()
        log(f"CPU baseline: oracle on {args.cpu_n}^3, {threads} threads and 1 thread")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_seconds, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
