"""The 2D cylinder channel (tclb_amd.cases.d2q9_cylinder, 1024x128, BGK with
Drag and Lift) stepped in one storage format.

    python -m tclb_amd.cylinder --steps 8192 --storage f64 [--temporal-blocking K]

With --temporal-blocking K the plain steps run K per launch on LDS tiles
(bit-identical to single steps; 0: single steps; -1, the default: the
lattice times 1, 3 and 5 per launch on the first 270 warm-up steps and
keeps the fastest).  `prepare()` (rows-per-tile timing, graph capture) and
512 warm-up steps, then `steps` timed steps: wall-clock and HIP-event microseconds per step,
MLUPS, and the Drag and Lift of a final Globals step.  The environment
variable TCLBX_GRAPH_STEPS sets the length of the captured step loops (0:
every step launched directly).  Prints one JSON object.
"""
import argparse
import json
import os
import sys
import time

from . import ITER_LASTGLOB, ITER_NORM
from . import cases as C


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=8192)
    ap.add_argument("--storage", default="f64")
    ap.add_argument("--model", default="d2q9_SRT_drag")
    ap.add_argument("--temporal-blocking", type=int, default=-1)
    a = ap.parse_args(argv)
    c = C.d2q9_cylinder(model=a.model, steps=0)
    lat = c.build(a.storage)
    try:
        lat.temporal_blocking = a.temporal_blocking
        lat.prepare()
        lat.iterate(512, ITER_NORM)
        t0 = time.perf_counter()
        lat.iterate(a.steps, ITER_NORM)
        wall = time.perf_counter() - t0
        ms, _ = lat.timing()
        lat.iterate(1, ITER_NORM | ITER_LASTGLOB)
        g = lat.global_dict()
        print(json.dumps({"workload": f"{a.model} cylinder {c.nx}x{c.ny}", "storage": a.storage, "steps": a.steps,
                          "graph_steps": os.environ.get("TCLBX_GRAPH_STEPS", "32"),
                          "temporal_blocking": lat.temporal_blocking,
                          "us_per_step_wall": round(wall / a.steps * 1e6, 3),
                          "us_per_step_events": round(ms / a.steps * 1e3, 3),
                          "MLUPS": round(c.nodes * a.steps / wall / 1e6, 1),
                          "drag": g.get("Drag"), "lift": g.get("Lift")}), flush=True)
    finally:
        lat.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
