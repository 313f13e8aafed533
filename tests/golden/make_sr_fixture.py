"""Writes tests/golden/sr_conversions.npz: gfx950's stochastic-rounding
conversions of float to binary16 and bfloat16 on (x, r) pairs, the
vectors the oracle's orc_f32_to_f16_sr / orc_f32_to_bf16_sr are checked
against (tests/test_oracle_pins.py).  Needs an MI355X:

    python tests/golden/make_sr_fixture.py [--out DIR]

builds sr_probe.hip with hipcc (into --out, default a temporary folder),
runs it and stores x, r and the two results as uint32/uint16 arrays.
"""
import argparse
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="folder for the probe binary and its raw output")
    ap.add_argument("--fixture", default=os.path.join(HERE, "sr_conversions.npz"))
    a = ap.parse_args(argv)
    work = a.out or tempfile.mkdtemp()
    os.makedirs(work, exist_ok=True)
    exe, raw = os.path.join(work, "sr_probe"), os.path.join(work, "sr_probe.bin")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", os.path.join(HERE, "sr_probe.hip"), "-o", exe], check=True)
    subprocess.run([exe, raw], check=True)
    d = open(raw, "rb").read()
    n = int(np.frombuffer(d[:4], np.int32)[0])
    o = 4
    x = np.frombuffer(d, np.uint32, n, o); o += 4 * n
    r = np.frombuffer(d, np.uint32, n, o); o += 4 * n
    h = np.frombuffer(d, np.uint16, n, o); o += 2 * n
    b = np.frombuffer(d, np.uint16, n, o)
    np.savez_compressed(a.fixture, x=x, r=r, f16=h, bf16=b)
    print(f"{n} conversions -> {a.fixture}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
