"""A Poisson stream of d3q19 porous-box solves played into the
continuous-batching server.

    python -m tclb_amd.serve --rate 2 --window 30 --storage f16s --budget-gb 240

Requests (tclb_amd.cases.d3q19_serving_requests) are generated before the
clock starts and submitted at their arrival times; each runs to its step
budget or stops when the relative change of its VOLvx Global over 500 steps
is below 1e-6 (the reference's <Stop VOLvxPercentChange="1e-6"
Iterations="500"/>).  Reports latency (finish - submit) percentiles,
solves/s over the makespan and aggregate MLUPS, and the solves/s of the same
requests solved one at a time as plain lattices in the same storage format
and in FP64, measured on a sample and scaled by work (the first ratio is the
gain of batching alone, the second includes the storage format's).  With
--saturated N the first N requests of the stream are played at --rate (50/s
is far above capacity), so the server runs backlogged.  Prints one JSON
object.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import ITER_LASTGLOB, ITER_NORM, Server
from . import cases as C

STOP = [("VOLvx", "PercentChange", 1e-6)]
INTERVAL = 500


def sequential_solve(case, max_steps, storage="f64"):
    """One request alone on a Lattice with the same budget and stop rule;
    returns (steps, seconds)."""
    lat = case.build(storage)
    lat.prepare()  # launch-shape timing and graph capture, outside the clock
    g_id = lat.global_names.index("VOLvx")
    old, steps = -12341234.0, 0
    t0 = time.perf_counter()
    while steps < max_steps:
        n = min(INTERVAL, max_steps - steps)
        lat.iterate(n, ITER_NORM | ITER_LASTGLOB)
        steps += n
        if n == INTERVAL:
            v = lat.globals[g_id]
            if abs((old - v) / v) <= STOP[0][2]:
                break
            old = v
    dt = time.perf_counter() - t0
    lat.close()
    return steps, dt


def _sequential(pairs, storages, sample):
    """Sequential solves of the first `sample` (case, budget) pairs, one
    Lattice at a time, per storage format: {storage: MLUPS}."""
    out = {}
    for st in storages:
        t, work = 0.0, 0
        for case, steps in pairs[:sample]:
            s, dt = sequential_solve(case, steps, st)
            t += dt
            work += s * case.nodes
        out[st] = work / t / 1e6
    return out


def _play(stream, storage, budget_gb, max_chunk, say):
    """Submit every (arrival, case, budget) at its arrival time into a new
    server, wait for all; returns (results, cases, makespan, stats)."""
    srv = Server("d3q19", storage, hbm_budget=int(budget_gb * (1 << 30)), max_chunk=max_chunk)
    t_start = time.perf_counter()
    ids = []
    for arrival, case, steps in stream:
        delay = arrival - (time.perf_counter() - t_start)
        if delay > 0:
            time.sleep(delay)
        ids.append((srv.submit(case, steps, stop=STOP, interval=INTERVAL), case))
    srv.wait_all(3600)
    makespan = time.perf_counter() - t_start
    res = [srv.poll(i) for i, _ in ids]
    st = srv.stats()
    srv.close()
    return res, [c for _, c in ids], makespan, st


def _summary(res, cs, makespan, st):
    done = [r for r in res if r["state"].startswith("done")]
    lat = np.array([r["t_finish"] - r["t_submit"] for r in done])
    wait = np.array([r["t_admit"] - r["t_submit"] for r in done])
    node_steps = sum(r["steps"] * c.nodes for r, c in zip(res, cs))
    return {
        "requests": len(res), "completed": len(done),
        "stopped_by_rule": int(np.sum([r["state"] == "done_stop" for r in res])),
        "rejected": int(np.sum([r["state"] == "rejected" for r in res])),
        "makespan_s": makespan, "solves_per_s": len(done) / makespan,
        "latency_p50_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "latency_p99_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "queue_wait_p50_s": float(np.percentile(wait, 50)) if len(wait) else None,
        "MLUPS": node_steps / makespan / 1e6,
        "MLUPS_while_busy": st["node_steps"] / max(st["busy_s"], 1e-9) / 1e6,
        "mean_nodes": float(np.mean([c.nodes for c in cs])),
        "mean_steps_run": float(np.mean([r["steps"] for r in res])),
        "node_steps": node_steps,
    }


def _versus(out, seq, node_steps, completed):
    """Solves/s of the same requests solved one at a time at each storage
    format (their MLUPS on the sample applied to the work the server did),
    and the server's ratio to each."""
    for st, mlups in seq.items():
        t_all = node_steps / (mlups * 1e6)
        sps = completed / t_all if t_all > 0 else None
        out[f"sequential_{st}_MLUPS"] = mlups
        out[f"sequential_{st}_solves_per_s"] = sps
        out[f"speedup_vs_sequential_{st}"] = out["solves_per_s"] / sps if sps else None


def run(rate=2.0, window=30.0, storage="f16s", budget_gb=240.0, seed=5, max_chunk=500, sample=4, log=None,
        seq_storages=("f16s", "f64")):
    """Requests of the stream that arrive within `window` s at `rate`/s."""
    say = log or (lambda m: None)
    n_guess = max(1, int(rate * window * 1.5) + 10)
    stream = [r for r in C.d3q19_serving_requests(n_guess, rate=rate, seed=seed) if r[0] < window]
    say(f"{len(stream)} requests arrive in {window:.0f} s at {rate}/s")
    res, cs, makespan, st = _play(stream, storage, budget_gb, max_chunk, say)
    out = {"storage": storage, "rate_per_s": rate, "window_s": window}
    out.update(_summary(res, cs, makespan, st))
    if sample > 0:
        seq = _sequential([(c, r["steps"]) for c, r in zip(cs, res)], seq_storages, sample)
        _versus(out, seq, out["node_steps"], out["completed"])
        out["sequential_sample"] = sample
    return out


def saturated(nreq=24, rate=50.0, storage="f16s", budget_gb=240.0, seed=5, max_chunk=500, sample=3, log=None,
              seq_storages=("f16s", "f64")):
    """The first `nreq` requests of the stated 50/s stream: they arrive far
    faster than they are served, so the server runs backlogged and its
    solves/s is its capacity on this request mix."""
    say = log or (lambda m: None)
    stream = C.d3q19_serving_requests(nreq, rate=rate, seed=seed)
    say(f"{nreq} requests at {rate}/s (backlogged)")
    res, cs, makespan, st = _play(stream, storage, budget_gb, max_chunk, say)
    out = {"storage": storage, "rate_per_s": rate, "arrival_span_s": stream[-1][0]}
    out.update(_summary(res, cs, makespan, st))
    if sample > 0:
        seq = _sequential([(c, r["steps"]) for c, r in zip(cs, res)], seq_storages, sample)
        _versus(out, seq, out["node_steps"], out["completed"])
        out["sequential_sample"] = sample
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--rate", type=float, default=2.0)
    ap.add_argument("--window", type=float, default=30.0)
    ap.add_argument("--storage", default="f16s")
    ap.add_argument("--budget-gb", type=float, default=240.0)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--chunk", type=int, default=500)
    ap.add_argument("--sample", type=int, default=4)
    ap.add_argument("--saturated", type=int, default=0, help="play the first N requests at --rate (backlogged)")
    a = ap.parse_args()
    say = lambda m: print(m, file=sys.stderr, flush=True)
    if a.saturated:
        r = saturated(a.saturated, a.rate, a.storage, a.budget_gb, a.seed, a.chunk, a.sample, log=say)
    else:
        r = run(a.rate, a.window, a.storage, a.budget_gb, a.seed, a.chunk, a.sample, log=say)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
