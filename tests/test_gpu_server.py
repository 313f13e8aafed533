"""The continuous-batching server: every request gives the same result alone
and in a mix, stop rules follow the reference's <Stop> handler, admission
respects the HBM budget, oversized requests are rejected."""
import time

import numpy as np
import pytest

import tclb_amd
from tclb_amd import cases

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def small_requests(n, seed=7):
    return cases.d3q19_serving_requests(n, rate=100.0, seed=seed, side=(12, 40), budget=(60, 400))


def test_same_result_alone_and_in_mix():
    say("server: alone vs mix")
    reqs = small_requests(6)
    mix = tclb_amd.Server("d3q19", "f16s", hbm_budget=256 << 20, max_chunk=50)
    ids = []
    for k, (_, case, steps) in enumerate(reqs):
        ids.append(mix.submit(case, steps))
        if k == 2:
            time.sleep(0.05)  # later requests join while the first ones run
    assert mix.wait_all(120)
    for (_, case, steps), rid in zip(reqs, ids):
        r = mix.poll(rid)
        assert r["state"] == "done_budget" and r["steps"] == steps
        alone = tclb_amd.Server("d3q19", "f16s", hbm_budget=64 << 20, max_chunk=7)
        a = alone.submit(case, steps)
        assert alone.wait_all(60)
        ra = alone.poll(a)
        assert ra["steps"] == steps
        assert np.array_equal(r["globals"], ra["globals"])
        # and the same as a Lattice run for the same steps
        lat = case.build("f16s")
        lat.iterate(steps)
        assert np.array_equal(lat.globals, r["globals"])
        alone.close()


def test_stop_rule_semantics():
    say("server: stop rules")
    _, case, _ = small_requests(1, seed=11)[0]
    s = tclb_amd.Server("d3q19", "f64", hbm_budget=64 << 20, max_chunk=1000)
    # VOLvolume is constant: a Change criterion passes from the second check on
    r1 = s.submit(case, 10000, stop=[("VOLvolume", "Change", 0.5)], interval=20, times=1)
    # Above with an unreachable limit never passes: runs to the budget
    r2 = s.submit(case, 90, stop=[("VOLvx", "Above", 1e9)], interval=10, times=1)
    # three consecutive passing checks
    r3 = s.submit(case, 10000, stop=[("VOLvolume", "Change", 0.5)], interval=15, times=3)
    assert s.wait_all(120)
    a, b, c = s.poll(r1), s.poll(r2), s.poll(r3)
    assert a["state"] == "done_stop" and a["steps"] == 40
    assert b["state"] == "done_budget" and b["steps"] == 90
    assert c["state"] == "done_stop" and c["steps"] == 60


def test_admission_by_budget_and_rejection():
    say("server: admission")
    reqs = small_requests(5, seed=13)
    sizes = [tclb_amd.Batch("d3q19", "f16s", 1 << 20).case_bytes(c.nx, c.ny, c.nz) for _, c, _ in reqs]
    budget = int(max(sizes) * 1.5)   # at most one of the larger ones at a time
    s = tclb_amd.Server("d3q19", "f16s", hbm_budget=budget, max_chunk=25)
    ids = [s.submit(c, st) for _, c, st in reqs]
    big = cases.d3q19_serving_requests(1, seed=3, side=(64, 64), budget=(10, 10))[0][1]
    rbig = s.submit(big, 10)
    assert s.wait_all(120)
    res = [s.poll(i) for i in ids]
    assert all(r["state"] == "done_budget" for r in res)
    assert s.poll(rbig)["state"] == "rejected"
    st = s.stats()
    assert st["completed"] == 5 and st["rejected"] == 1 and st["active"] == 0
    # some request had to wait for memory
    assert max(r["t_admit"] - r["t_submit"] for r in res) > 0
