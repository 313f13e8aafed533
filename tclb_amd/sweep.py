"""Batched sweep of independent cylinder cases (one persistent launch per
step over a ragged arena) against the same cases solved one at a time.

    python -m tclb_amd.sweep --cases 256 --steps 8000 --storage f16s

Reports the batch's aggregate MLUPS and solves/s, the solves/s of the cases
solved sequentially as single FP64 lattices, and per case the relative
error of the drag against that FP64 solve: of the drag averaged over the
last `window` steps (sampled every `every` steps; shedding cases oscillate,
so the last-step value also carries the phase of the oscillation) and of
the last step's drag.  --tuning FLAT,BLOCK,ROWS_PER_TILE,GRID,ORDER runs
that tile shape instead of the autotuner's pick (for profiled runs, whose
timing misleads the autotuner).  Prints one JSON object.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import Batch, ITER_LASTGLOB, ITER_NORM, Model
from . import cases as C


def _advance(it, steps, window, every, gd):
    """steps in all; the drag (Globals index gd) sampled every `every`
    steps over the last `window`: returns the samples [n_samples, ...]"""
    window = min(window, steps)
    it(steps - window, ITER_NORM)
    samples = []
    for _ in range(window // every):
        samples.append(it(every, ITER_NORM | ITER_LASTGLOB))
    return np.array(samples)


def run(ncases=256, steps=8000, storage="f16s", seed=4, reference=True, autotune_steps=5, log=None,
        model="d2q9_drag", window=2000, every=100, tuning=None):
    """`tuning`: a tile shape {flat, block, rows_per_tile, grid, order} to
    run instead of the autotuner's (a profiler run perturbs the autotuner's
    timing, so its pick differs from an unprofiled run's)"""
    say = log or (lambda m: None)
    sweep = C.d2q9_cylinder_sweep(ncases=ncases, seed=seed, steps=steps, model=model)
    nodes = sum(c.nodes for c in sweep)
    probe = Batch(model, storage, 1 << 20)
    cap = sum(probe.case_bytes(c.nx, c.ny) for c in sweep) + (16 << 20)
    probe.close()
    b = Batch(model, storage, capacity_bytes=cap)
    ids = [b.add(c) for c in sweep]
    gd, gl = Model(model).global_id("Drag"), Model(model).global_id("Lift")
    if tuning is None:
        tune = b.autotune(steps=autotune_steps)
    else:
        tune = dict(tuning)
        b.set_tuning(block=tune["block"], rows_per_tile=tune["rows_per_tile"], grid=tune["grid"],
                     flat=bool(tune["flat"]), order=tune["order"])
    say(f"batch of {ncases} cases, {nodes} nodes, tuning {tune}")
    def bat(n, flags):
        b.iterate(n, flags)
        if flags & ITER_LASTGLOB:
            return [b.globals(i)[gd] for i in ids]
        return None
    t0 = time.perf_counter()
    samples = _advance(bat, steps, window, every, gd)
    wall = time.perf_counter() - t0
    drag = samples[-1]
    drag_mean = samples.mean(axis=0)
    lift = np.array([b.globals(i)[gl] for i in ids])
    out = {
        "model": model, "cases": ncases, "steps": steps, "storage": storage, "nodes": nodes,
        "batch_seconds": wall, "batch_MLUPS": nodes * steps / wall / 1e6,
        "batch_solves_per_s": ncases / wall, "tuning": tune,
        "finite": bool(np.all(np.isfinite(drag))),
    }
    b.close()
    if reference:
        ref_drag = np.empty(ncases)
        ref_mean = np.empty(ncases)
        t_seq = 0.0
        for k, c in enumerate(sweep):
            lat = c.build("f64")

            def one(n, flags):
                lat.iterate(n, flags)
                return lat.globals[gd] if flags & ITER_LASTGLOB else None
            t0 = time.perf_counter()
            s = _advance(one, steps, window, every, gd)
            t_seq += time.perf_counter() - t0
            ref_drag[k], ref_mean[k] = s[-1], s.mean()
            lat.close()
        # errors are taken over the cases whose FP64 solve stays finite (all
        # of them with the MRT model; BGK with the Zou-He velocity inlet of
        # d2q9_SRT is unstable below tau ~ 0.55 in every format, FP64
        # included), and the unstable ones are counted
        ok = np.isfinite(ref_drag) & np.isfinite(drag) & np.isfinite(ref_mean) & np.isfinite(drag_mean)
        err = np.abs(drag[ok] - ref_drag[ok]) / np.maximum(np.abs(ref_drag[ok]), 1e-300)
        errm = np.abs(drag_mean[ok] - ref_mean[ok]) / np.maximum(np.abs(ref_mean[ok]), 1e-300)
        out.update({
            "sequential_f64_seconds": t_seq,
            "sequential_f64_MLUPS": nodes * steps / t_seq / 1e6,
            "sequential_f64_solves_per_s": ncases / t_seq,
            "speedup_solves_per_s": t_seq / wall,
            "unstable_f64": int(np.sum(~np.isfinite(ref_drag))),
            "unstable_batch": int(np.sum(~np.isfinite(drag))),
            "unstable_batch_only": int(np.sum(~np.isfinite(drag) & np.isfinite(ref_drag))),
            "compared_cases": int(np.sum(ok)),
            "mean_drag_window": [min(window, steps), every],
            "mean_drag_rel_err_max": float(np.max(errm)) if len(errm) else None,
            "mean_drag_rel_err_median": float(np.median(errm)) if len(errm) else None,
            "mean_drag_rel_err_p90": float(np.percentile(errm, 90)) if len(errm) else None,
            "drag_rel_err_max": float(np.max(err)) if len(err) else None,
            "drag_rel_err_median": float(np.median(err)) if len(err) else None,
            "drag_rel_err_p90": float(np.percentile(err, 90)) if len(err) else None,
        })
    out["drag_median"] = float(np.median(drag[np.isfinite(drag)])) if np.any(np.isfinite(drag)) else None
    out["lift_abs_max"] = float(np.max(np.abs(lift[np.isfinite(lift)]))) if np.any(np.isfinite(lift)) else None
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--cases", type=int, default=256)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--storage", default="f16s")
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--no-reference", action="store_true")
    ap.add_argument("--model", default="d2q9_drag", help="d2q9_drag (MRT) or d2q9_SRT_drag (BGK)")
    ap.add_argument("--window", type=int, default=2000, help="steps at the end over which drag is averaged")
    ap.add_argument("--every", type=int, default=100, help="drag sampling interval in that window")
    ap.add_argument("--tuning", default="", help="FLAT,BLOCK,ROWS_PER_TILE,GRID,ORDER: this tile shape, no autotuner")
    a = ap.parse_args()
    tuning = None
    if a.tuning:
        v = [int(x) for x in a.tuning.split(",")]
        tuning = dict(zip(("flat", "block", "rows_per_tile", "grid", "order"), v))
    r = run(a.cases, a.steps, a.storage, a.seed, not a.no_reference,
            log=lambda m: print(m, file=sys.stderr, flush=True), model=a.model, window=a.window, every=a.every,
            tuning=tuning)
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
