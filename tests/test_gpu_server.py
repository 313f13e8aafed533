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


def _stated_requests():
    """Requests of the serving stream as stated (seed 5; sides log-uniform in
    [48, 256], porosity 0-25%, g_x 1e-7..1e-5, tau 0.52..0.9), with short
    step budgets so the oracle can follow them."""
    reqs = cases.d3q19_serving_requests(12)
    pick = [(1, 20), (2, 24), (3, 30), (4, 36), (9, 40), (6, 12)]
    out = [(reqs[k][1], steps) for k, steps in pick]
    sides = [d for c, _ in out for d in (c.nx, c.ny, c.nz)]
    assert 48 <= min(sides) and max(sides) <= 256 and max(sides) >= 240
    return out


def test_served_requests_match_the_oracle(oracle_mod):
    """Six requests at the serving workload's stated sizes (up to 4.2M
    nodes, sides 49-248), served in FP64, three of them joining while the
    first three run and each retiring at its own budget: every request's
    Globals at its end within 1e-10 of the oracle's run of the same case
    (the bar of the reference's harness, tools/csvdiff:7)."""
    say("server: stated request sizes vs the oracle")
    reqs = _stated_requests()
    s = tclb_amd.Server("d3q19", "f64", hbm_budget=8 << 30, max_chunk=8)
    ids = [s.submit(c, st) for c, st in reqs[:3]]
    t0 = time.time()
    while s.poll(ids[0])["state"] == "queued" and time.time() - t0 < 60:
        time.sleep(0.005)
    ids += [s.submit(c, st) for c, st in reqs[3:]]
    assert s.wait_all(300)
    oracle_mod.set_threads(16)
    try:
        for (case, steps), rid in zip(reqs, ids):
            r = s.poll(rid)
            assert r["state"] == "done_budget" and r["steps"] == steps, (case.name, r)
            g = oracle_mod.lattice_for_case(case, "f64").iterate(steps)
            say(f"  {case.name}: {steps} steps, VOLvx {r['globals'][16]:.9e} (oracle {g[16]:.9e})")
            for i, (a, b) in enumerate(zip(r["globals"], g)):
                scale = max(abs(a), abs(b))
                assert scale < 1e-10 or abs(a - b) <= 1e-10 * scale, (case.name, i, a, b)
    finally:
        oracle_mod.set_threads(1)
    s.close()


def test_served_f16s_request_tracks_the_oracle(oracle_mod):
    """One request of the stated stream (61x71x84, 8% solid, 40 steps)
    served in FP16-shift storage.  Serving adds nothing: its Globals are
    those of the same case run as a plain f16s lattice, bit for bit.  Against
    the oracle's FP64 run, each volume Global stays within what storing
    every population in half precision can move it: per node and step at
    most half the format's spacing on each population that enters the sum
    (the spacing of the largest stored deviation |f - w| of the FP64 state,
    subnormal here: 2^-24), accumulated over all fluid nodes and steps.  The
    oracle's own f16s emulation, which forms f = stored + w where the kernel
    collides on the deviations, is printed beside it: at these velocities
    (u ~ 2e-4, deviations in half's subnormal range, every node starting
    from the same state, so rounding errors are correlated) the two float
    forms round to different sides and differ by up to a few percent."""
    say("server: f16s request vs a plain lattice and the oracle")
    case, steps = _stated_requests()[4]
    s = tclb_amd.Server("d3q19", "f16s", hbm_budget=1 << 30, max_chunk=500)
    rid = s.submit(case, steps)
    assert s.wait_all(120)
    r = s.poll(rid)
    assert r["state"] == "done_budget" and r["steps"] == steps
    s.close()
    lat = case.build("f16s")
    lat.iterate(steps, tclb_amd.ITER_NORM | tclb_amd.ITER_LASTGLOB)
    assert np.array_equal(r["globals"], lat.globals)
    lat.close()
    oracle_mod.set_threads(16)
    try:
        o64 = oracle_mod.lattice_for_case(case, "f64")
        g64 = o64.iterate(steps)
        g16 = oracle_mod.lattice_for_case(case, "f16s").iterate(steps)
        f64 = o64.get_f()
    finally:
        oracle_mod.set_threads(1)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    maxd = float(np.max(np.abs(f64 - w[:, None, None, None])))
    spacing = max(2.0 ** -24, 2.0 ** (np.floor(np.log2(maxd)) - 10))
    names = tclb_amd.Model("d3q19").global_names
    fluid = g64[names.index("VOLvolume")]
    assert r["globals"][names.index("VOLvolume")] == fluid
    # populations entering each sum (10 with e_x != 0, all 19 for rho) and
    # the sum's scale (VOLv* = 1.9 sum u, VOLp* = sum J / 10, VOLrho =
    # sum rho / 19: the Globals are taken after R /= Mw)
    per = {"VOLvx": 10 * 1.9, "VOLvy": 10 * 1.9, "VOLvz": 10 * 1.9,
           "VOLpx": 10 / 10, "VOLpy": 10 / 10, "VOLpz": 10 / 10, "VOLrho": 19 / 19}
    for name, k in per.items():
        i = names.index(name)
        a, b, c = r["globals"][i], g16[i], g64[i]
        bound = k * fluid * steps * spacing / 2
        say(f"  {name}: served {a:.6e}  oracle f16s {b:.6e}  oracle f64 {c:.6e}  bound {bound:.2e}")
        assert abs(a - c) <= bound, (name, a, c, bound)


def test_served_request_with_stochastic_rounding_matches_a_lattice():
    """With stochastic rounding on, a served f16s request ends with the
    populations and Globals of the same case run as a lone Lattice with the
    same seed, bit for bit, whatever else shares the server's batch."""
    say("server: stochastic rounding vs a plain lattice")
    reqs = _stated_requests()
    (case, steps), (other, osteps) = reqs[4], reqs[2]
    s = tclb_amd.Server("d3q19", "f16s", hbm_budget=1 << 30, max_chunk=17)
    s.set_stochastic_rounding(True, seed=5)
    rid = s.submit(case, steps)
    s.submit(other, min(osteps, 3 * steps))
    assert s.wait_all(120)
    r = s.poll(rid)
    assert r["state"] == "done_budget" and r["steps"] == steps
    s.close()
    lat = case.build("f16s")
    lat.set_stochastic_rounding(True, seed=5)
    lat.iterate(steps, tclb_amd.ITER_NORM | tclb_amd.ITER_LASTGLOB)
    assert np.array_equal(r["globals"], lat.globals)
    plain = case.build("f16s")
    plain.iterate(steps, tclb_amd.ITER_NORM | tclb_amd.ITER_LASTGLOB)
    assert not np.array_equal(plain.globals, lat.globals)
    lat.close()
    plain.close()


def test_malformed_request_fails_at_submit():
    """A request the batch cannot hold (a 2D model with nz = 2) fails in
    submit() with a message; the server keeps serving the valid request
    submitted beside it."""
    import dataclasses
    say("server: malformed request")
    good = cases.d2q9_cylinder(nx=64, ny=32, diameter=8, cx=16, cy=16, steps=50)
    bad = dataclasses.replace(good, nz=2, flags=np.concatenate([good.flags, good.flags]))
    s = tclb_amd.Server("d2q9_SRT", "f64", hbm_budget=64 << 20, max_chunk=20)
    r0 = s.submit(good, 50)
    with pytest.raises(tclb_amd.TclbxError, match="nz = 1"):
        s.submit(bad, 50)
    r1 = s.submit(good, 30)
    assert s.wait_all(60)
    assert s.poll(r0)["state"] == "done_budget" and s.poll(r1)["state"] == "done_budget"
    assert s.stats()["completed"] == 2
    s.close()
