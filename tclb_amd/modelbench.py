"""Step throughput of each model on a periodic box of collision nodes.

    python -m tclb_amd.modelbench --n 256 --steps 200 --models d3q19,d3q27,d3q27_cumulant --formats f64,f16s
    python -m tclb_amd.modelbench --models d3q19 --formats f64 --shapes 256x256x256,256x256x8

Every node of an n^3 (2D models: n^2) periodic box carries the model's
collision NodeType and nothing else, so each step is the plain pull +
collision of every node; a small body force keeps the flow moving.  For
each model and storage format: `warmup` steps (2D models: at least 300,
whose first 270 time the choice of the temporal blocking, reported as
`temporal_blocking`), then `steps` timed steps
(HIP events around the step launches), MLUPS and the algorithmic HBM
bytes per node step (read and write Q stored values, the 2-byte NodeType,
the FP64 averages of the AVG builds read and written in place, the FP64
parameter densities of the bc builds read).  --shapes runs boxes of the
given sizes instead of cubes (with the size of the two snapshots beside
them: a box whose snapshots fit the 256 MB Infinity Cache steps from the
cache).  The launch-shape timing and the graph capture happen before the
timed steps.  Prints one JSON object per line.
"""
import argparse
import json
import sys

import numpy as np

from . import Lattice, Model

# the collision NodeType and the settings of each model's box
COLLISION = {"d2q9_SRT": "BGK", "d2q9_SRT_drag": "BGK"}
SETTINGS = {
    "d2q9_SRT": [("nu", 0.02)], "d2q9_SRT_drag": [("nu", 0.02)],
    "d3q19": [("nu", 0.02), ("ForceX", 1e-6)], "d3q19_lid": [("nu", 0.02), ("ForceX", 1e-6)],
    "d3q27": [("nu", 0.02), ("ForceX", 1e-6)],
}


def settings_of(model):
    if model in SETTINGS:
        return SETTINGS[model]
    if model.startswith("d3q27_cumulant"):
        return [("nu", 0.02), ("ForceX", 1e-6)] + ([("Smag", 0.16)] if model.endswith("SMAG") else [])
    return [("Viscosity", 0.02), ("GravitationX", 1e-6)]      # d2q9 (MRT) and its builds


def run(model, fmt, n, steps, warmup, shape=None):
    m = Model(model)
    if shape is None:
        shape = (n, n, 1 if m.Q == 9 else n)        # the d2q9 models are 2D
    nx, ny, nz = shape
    lat = Lattice(model, nx, ny, nz, storage=fmt)
    try:
        lat.set_flags(np.full((nz, ny, nx), m.flag(COLLISION.get(model, "MRT")), np.uint16))
        for k, v in settings_of(model):
            lat.set_setting(k, v)
        lat.init()
        lat.prepare()
        # the 2D lattices time their temporal blocking on the first 270
        # steps of an Iterate: those stay in the warm-up
        lat.iterate(max(warmup, 300) if m.Q == 9 else warmup)
        lat.iterate(steps)
        ms, _ = lat.timing()
        nodes = nx * ny * nz
        mlups = nodes * steps / (ms * 1e-3) / 1e6
        bpn = lat.bytes_per_node()
        return {"model": model, "storage": fmt, "shape": f"{nx}x{ny}x{nz}", "nodes": nodes, "steps": steps,
                "MLUPS": round(mlups, 1), "bytes_per_node": bpn, "GBs": round(mlups * 1e6 * bpn / 1e9, 1),
                "snapshots_MB": round(nodes * bpn / 1e6, 1), "temporal_blocking": lat.temporal_blocking,
                "finite": bool(np.isfinite(lat.get_f()[:, :1]).all())}
    finally:
        lat.close()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=256, help="box side (3D: n^3 nodes, 2D: n^2)")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--models", default="d3q19,d3q27,d3q27_cumulant")
    ap.add_argument("--formats", default="f64,f16s")
    ap.add_argument("--shapes", default="", help="boxes NXxNYxNZ,... instead of cubes of side --n")
    a = ap.parse_args(argv)
    shapes = [tuple(int(v) for v in s.split("x")) for s in a.shapes.split(",") if s] or [None]
    for model in a.models.split(","):
        for fmt in a.formats.split(","):
            for shape in shapes:
                steps = a.steps
                if shape is not None:   # about 2e9 node updates per shape, whole graph chunks
                    steps = max(a.steps, int(2e9 / (shape[0] * shape[1] * shape[2])) // 32 * 32)
                print(json.dumps(run(model, fmt, a.n, steps, a.warmup, shape)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
