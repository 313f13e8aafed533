"""Writes the small data fixtures under tests/golden/.

d3q19_mrtmat.json: the 19x19 integer matrix MRTMAT as listed in
models/flow/d3q19/Dynamics.R:1-21 (and src/lib/d3q19.R:1-21), read from the
reference's text.  R's matrix(c(...), 19, 19) fills column-major, so the
k-th listed row is column k (moment k); the file stores it as listed,
`rows[k][i]` = moment k of direction i.

Run from the repository root:  python tests/golden/make_fixtures.py
"""
import json
import os
import re

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def mrtmat():
    src = open(os.path.join(REF, "models/flow/d3q19/Dynamics.R")).read()
    body = src[src.index("matrix(c(") + len("matrix(c("): src.index("),19,19)")]
    nums = [int(v) for v in re.findall(r"-?\d+", body)]
    assert len(nums) == 361
    return [nums[19 * k: 19 * k + 19] for k in range(19)]


if __name__ == "__main__":
    with open(os.path.join(HERE, "d3q19_mrtmat.json"), "w") as f:
        rows = mrtmat()
        f.write('{"source": "models/flow/d3q19/Dynamics.R:1-21",\n')
        f.write(' "layout": "rows[k][i] = moment k of direction i",\n "rows": [\n')
        f.write(",\n".join("  " + json.dumps(r) for r in rows))
        f.write("\n]}\n")
    print("wrote d3q19_mrtmat.json")
