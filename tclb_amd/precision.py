"""Accuracy of the reduced storage formats, model by model, against FP64.

    python -m tclb_amd.precision --models d2q9_SRT_drag,d2q9_drag,d3q19,d3q27,d3q27_cumulant --steps 4000

The storage formats store f - w_i (f32s, f16s, bf16s) or f (f32) and
collide in float; how the float collision meets the stored values differs
by model family: d3q19 and the 2D models collide on the deviations f - w,
the 27-velocity models on f = stored + w as the reference's float build
does.  This runs one flow per model in FP64 and in each format and reports
the relative L2 and max-norm of the velocity (over the fluid nodes,
max-norm relative to the largest FP64 speed), the relative L2 of the
density, and the relative error of one Global:

- 2D models: the 1024x128 cylinder channel (Zou-He velocity inlet with a
  parabolic profile, peak 0.01, pressure outlet, walls, a disc of diameter
  32; nu 0.02), Global Drag;
- 3D models: a 128^3 porous channel (10 % Solid or Wall spheres of radius
  4-12, walls at y = 0 and 127, periodic in x and z, force 2e-6 along x,
  nu 0.02), Global the x flux (VOLvx for d3q19, XFlux for d3q27, Flux for
  the cumulant models).

Prints one JSON object per model and format.
"""
import argparse
import json
import sys

import numpy as np

from . import ITER_LASTGLOB, ITER_NORM, Model
from . import cases as C

GLOBAL = {"d3q19": "VOLvx", "d3q27": "XFlux"}


def flow(model, n3=128):
    """(case, fluid mask, Global name) of the model's flow."""
    m = Model(model)
    if m.Q == 9:
        c = C.d2q9_cylinder(model=model, steps=0)
        coll = m.flag("BGK") if model.startswith("d2q9_SRT") else m.flag("MRT")
        fluid = (c.flags & ((1 << m.group("SETTINGZONE")[1]) - 1)) == coll
        return c, fluid, "Drag" if "Drag" in m.global_names else m.global_names[0]
    base = C.d3q19_porous(n=n3, force=2e-6, steps=0)
    solid = base.flags != 0x20
    wall = m.flag("Wall")
    flags = np.where(solid, wall, m.flag("MRT")).astype(np.uint16)
    st = [("nu", 0.02), ("ForceX", 2e-6)] + ([("Smag", 0.16)] if model.endswith("SMAG") else [])
    if model.startswith("d3q19"):
        st = [("omega", 1 / 0.56), ("ForceX", 2e-6)]
    c = C.Case(f"porous_{n3}", model, n3, n3, n3, flags, st)
    name = GLOBAL.get(model, "Flux" if model.startswith("d3q27_cumulant") else m.global_names[0])
    return c, ~solid, name


def run(model, formats, steps):
    case, fluid, gname = flow(model)
    out = []
    ref = None
    for fmt in ["f64"] + [f for f in formats if f != "f64"]:
        lat = case.build(fmt)
        lat.prepare()
        lat.iterate(steps, ITER_NORM | ITER_LASTGLOB)
        u = lat.quantity("U")[fluid]
        # density: Rho, or 1 + 3 P (P = (rho - 1)/3 in the 3D models)
        rho = lat.quantity("Rho")[fluid] if "Rho" in lat.spec.quantities else 1 + 3 * lat.quantity("P")[fluid]
        g = lat.global_dict()[gname]
        finite = bool(np.all(np.isfinite(u)))
        lat.close()
        if ref is None:
            ref = (u, rho, g)
            continue
        u0, rho0, g0 = ref
        r = {"model": model, "storage": fmt, "workload": case.name, "steps": steps, "finite": finite,
             "u_rel_L2": float(np.linalg.norm(u - u0) / np.linalg.norm(u0)),
             "u_max_over_umax": float(np.max(np.abs(u - u0)) / np.max(np.abs(u0))),
             "rho_rel_L2": float(np.linalg.norm(rho - rho0) / np.linalg.norm(rho0)),
             "global": gname, "global_rel_err": float(abs(g - g0) / abs(g0)) if g0 != 0 else None}
        out.append(r)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", default="d2q9_SRT_drag,d2q9_drag,d3q19,d3q27,d3q27_cumulant")
    ap.add_argument("--formats", default="f32,f32s,f16s,bf16s")
    ap.add_argument("--steps", type=int, default=4000)
    a = ap.parse_args(argv)
    for model in a.models.split(","):
        for r in run(model, a.formats.split(","), a.steps):
            print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
