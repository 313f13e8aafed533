"""Precision sweep of the d3q19 lid-driven cavity: the same case run with
every storage format, errors of rho, u and the lid torque against the FP64
run.

    python -m tclb_amd.cavity --n 640 --steps 20000 --formats f64,f32s,f16s,bf16s,f16s+sr,bf16s+sr

The case (tclb_amd.cases.d3q19_cavity): n^3 nodes, the top face a MovingWall
at u_x = 0.05, bounce-back Wall on the other five faces, MRT at tau = 0.52
with the reference's rates.  Each format runs `steps` steps (Globals on the
last one); errors are taken over the MRT (fluid) nodes against the first
format of the list: relative L2 and max-norm of u (max-norm relative to the
lid speed) and of rho, and the relative error of MovingWallTorque.  A
format "<f16s|bf16s>+sr" runs that format with stochastic rounding.  Prints
one JSON object.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import ITER_LASTGLOB, ITER_NORM
from . import cases as C


def run(n=640, steps=20000, formats=("f64", "f32s", "f16s", "bf16s"), u_lid=0.05, tau=0.52, chunk=1000,
        log=None):
    say = log or (lambda m: None)
    case = C.d3q19_cavity(n=n, u_lid=u_lid, tau=tau, steps=steps)
    fluid = (case.flags & 0x20) != 0
    out = {"workload": f"d3q19_lid cavity {n}^3, lid u_x {u_lid}, tau {tau}", "nodes": case.nodes,
           "steps": steps, "formats": {}}
    ref = None
    for fmt in formats:
        # "<storage>+sr": the 16-bit format with stochastic rounding
        storage, sr = (fmt[:-3], True) if fmt.endswith("+sr") else (fmt, False)
        lat = case.build(storage)
        if sr:
            lat.set_stochastic_rounding(True, seed=1)
        lat.prepare()
        done = 0
        t0 = time.perf_counter()
        while done < steps:
            k = min(chunk, steps - done)
            last = done + k == steps
            lat.iterate(k, ITER_NORM | (ITER_LASTGLOB if last else 0))
            done += k
            say(f"{fmt}: {done}/{steps} steps, {time.perf_counter() - t0:.0f} s")
        wall = time.perf_counter() - t0
        g = lat.global_dict()
        u = lat.quantity("U")[fluid]
        rho = 1.0 + 3.0 * lat.quantity("P")[fluid]
        lat.close()
        r = {"MLUPS": case.nodes * steps / wall / 1e6, "seconds": wall,
             "lid_torque": g["MovingWallTorque"], "lid_Fx": g["MovingWallFx"],
             "mass_drift": float(np.sum(rho - 1.0) / rho.size), "finite": bool(np.all(np.isfinite(u)))}
        if ref is None:
            ref = (u, rho, g["MovingWallTorque"])
        else:
            u0, rho0, t0_ = ref
            du, drho = u - u0, rho - rho0
            r.update({
                "u_rel_L2": float(np.linalg.norm(du) / np.linalg.norm(u0)),
                "u_max_over_lid": float(np.max(np.abs(du)) / u_lid),
                "rho_rel_L2": float(np.linalg.norm(drho) / np.linalg.norm(rho0)),
                "rho_max": float(np.max(np.abs(drho))),
                "lid_torque_rel": float(abs(g["MovingWallTorque"] - t0_) / abs(t0_)) if t0_ != 0 else None,
            })
        del u, rho
        out["formats"][fmt] = r
        say(f"{fmt}: {r['MLUPS']:.0f} MLUPS" + (f", u L2 {r['u_rel_L2']:.2e}" if "u_rel_L2" in r else ""))
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=640)
    ap.add_argument("--steps", type=int, default=20000)
    ap.add_argument("--formats", default="f64,f32s,f16s,bf16s")
    ap.add_argument("--u-lid", type=float, default=0.05)
    ap.add_argument("--tau", type=float, default=0.52)
    ap.add_argument("--chunk", type=int, default=1000)
    a = ap.parse_args()
    r = run(a.n, a.steps, a.formats.split(","), a.u_lid, a.tau, a.chunk,
            log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(r), flush=True)


if __name__ == "__main__":
    main()
