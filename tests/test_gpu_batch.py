"""The batch (ragged arena, persistent tile kernel) against single lattices.

A case in a batch runs the same node arithmetic over its own tiles, and its
Globals are reduced over its own tiles only, so the populations and the
Globals of a case are bit-identical whether it runs as a Lattice (with the
same tile width), alone in a batch, or among other cases; the oracle checks
the mixed batch in FP64.
"""
import numpy as np
import pytest

from conftest import oracle_for_case

import tclb_amd
from tclb_amd import cases
from test_gpu_parity import glob_close, random_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def mixed_cases(model, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(5):
        if model.startswith("d2q9"):
            nx, ny, nz = int(rng.integers(9, 90)), int(rng.integers(5, 40)), 1
        else:
            nx, ny, nz = int(rng.integers(5, 30)), int(rng.integers(4, 12)), int(rng.integers(3, 9))
        out.append(random_case(model, nx, ny, nz, seed=seed * 10 + k))
    return out


@pytest.mark.parametrize("model,storage", [("d2q9_SRT_drag", "f64"), ("d3q19", "f64"), ("d3q19", "f16s"),
                                           ("d2q9_SRT", "bf16s")])
def test_batch_matches_single_lattices(model, storage):
    say(f"batch vs lattices {model} {storage}")
    cs = mixed_cases(model, 3)
    b = tclb_amd.Batch(model, storage, capacity_bytes=64 << 20)
    b.set_tuning(block=64, rows_per_tile=2, grid=17)
    ids = [b.add(c) for c in cs]
    solo = []
    for c in cs:
        lat = c.build(storage)
        lat.set_block(64)
        solo.append(lat)
    for n in (7, 1, 12):
        b.iterate(n)
        for cid, lat in zip(ids, solo):
            lat.iterate(n)
            assert np.array_equal(b.get_f(cid), lat.get_f(), equal_nan=True)
            assert np.array_equal(b.globals(cid), lat.globals, equal_nan=True)
            assert b.iteration(cid) == lat.iteration


@pytest.mark.parametrize("model,storage", [("d3q19", "f16s"), ("d2q9", "bf16s"), ("d3q27_cumulant", "f16s")])
def test_batch_stochastic_rounding_matches_single_lattices(model, storage):
    """With stochastic rounding on, every case of a batch gives bit for bit
    what a lone Lattice with the same seed gives, also a case that joins
    after the others have run: each store's random bits are keyed by the
    case's own node index and iteration.  A different seed changes the
    result."""
    say(f"batch stochastic rounding {model} {storage}")
    cs = mixed_cases(model, 5)
    b = tclb_amd.Batch(model, storage, capacity_bytes=64 << 20)
    b.set_tuning(block=64, rows_per_tile=2, grid=17)
    b.set_stochastic_rounding(True, seed=99)
    ids = [b.add(c) for c in cs[:-1]]
    solo = []
    for c in cs:
        lat = c.build(storage)
        lat.set_block(64)
        lat.set_stochastic_rounding(True, seed=99)
        solo.append(lat)
    b.iterate(6)
    for lat in solo[:-1]:
        lat.iterate(6)
    ids.append(b.add(cs[-1]))
    for n in (5, 1, 9):
        b.iterate(n)
        for cid, lat in zip(ids, solo):
            lat.iterate(n)
            assert b.iteration(cid) == lat.iteration
            assert np.array_equal(b.get_f(cid), lat.get_f(), equal_nan=True)
            assert np.array_equal(b.globals(cid), lat.globals, equal_nan=True)
    other = cs[0].build(storage)
    other.set_block(64)
    other.set_stochastic_rounding(True, seed=100)
    other.iterate(6 + 15)
    assert not np.array_equal(other.get_f(), solo[0].get_f())


@pytest.mark.parametrize("flat", [False, True])
def test_flat_tiles_match_row_tiles(flat):
    """Flat tiles (runs of the node index) compute the same populations as
    row tiles; their Globals agree to the reduction's rounding."""
    say(f"flat tiles {flat}")
    cs = mixed_cases("d3q19", 17)
    a = tclb_amd.Batch("d3q19", "f64", capacity_bytes=64 << 20)
    b = tclb_amd.Batch("d3q19", "f64", capacity_bytes=64 << 20)
    a.set_tuning(block=128, rows_per_tile=3, grid=0, flat=flat)
    b.set_tuning(block=64, rows_per_tile=1, grid=5, flat=False)
    ia = [a.add(c) for c in cs]
    ib = [b.add(c) for c in cs]
    a.iterate(9)
    b.iterate(9)
    for x, y in zip(ia, ib):
        assert np.array_equal(a.get_f(x), b.get_f(y), equal_nan=True)
        glob_close(a.globals(x), b.globals(y), 1e-13)


@pytest.mark.parametrize("grid", [0, 17, 2048])
@pytest.mark.parametrize("flat", [False, True])
def test_xcd_tile_order_changes_nothing(grid, flat):
    """Walking the tiles in XCD-contiguous order (each of the eight XCDs
    takes a contiguous run of the tile list) updates every tile exactly
    once, in one block per tile and in persistent grids of 17 and 2048
    blocks: distributions and Globals bit for bit as in list order.  The
    13 cases of random sizes make a tile count that is in general not a
    multiple of 8."""
    say(f"xcd tile order grid={grid} flat={flat}")
    cs = mixed_cases("d2q9_SRT_drag", 13)
    a = tclb_amd.Batch("d2q9_SRT_drag", "f32s", capacity_bytes=64 << 20)
    b = tclb_amd.Batch("d2q9_SRT_drag", "f32s", capacity_bytes=64 << 20)
    a.set_tuning(block=64, rows_per_tile=1, grid=grid, flat=flat, order=1)
    b.set_tuning(block=64, rows_per_tile=1, grid=grid, flat=flat, order=0)
    assert a.tile_order() == 1 and b.tile_order() == 0
    ia = [a.add(c) for c in cs]
    ib = [b.add(c) for c in cs]
    for n in (5, 33):
        a.iterate(n)
        b.iterate(n)
        for x, y in zip(ia, ib):
            assert np.array_equal(a.get_f(x), b.get_f(y), equal_nan=True)
            assert np.array_equal(a.globals(x), b.globals(y), equal_nan=True)


def test_batch_case_alone_and_in_mix_and_oracle(oracle_mod):
    say("case alone vs in a mix vs oracle")
    cs = mixed_cases("d3q19", 5)
    mix = tclb_amd.Batch("d3q19", "f64", capacity_bytes=64 << 20)
    ids = [mix.add(c) for c in cs]
    alone = tclb_amd.Batch("d3q19", "f64", capacity_bytes=16 << 20)
    a_id = alone.add(cs[2])
    mix.iterate(15)
    alone.iterate(15)
    assert np.array_equal(mix.get_f(ids[2]), alone.get_f(a_id))
    assert np.array_equal(mix.globals(ids[2]), alone.globals(a_id))
    for c, cid in zip(cs, ids):
        orc = oracle_for_case(oracle_mod, c, "f64")
        g = orc.iterate(15)
        assert np.array_equal(mix.get_f(cid), orc.get_f())
        glob_close(mix.globals(cid), g[:mix.nglobals], 1e-10)


def test_autotune_keeps_state():
    say("autotune keeps the state")
    cs = mixed_cases("d2q9_SRT", 7)
    a = tclb_amd.Batch("d2q9_SRT", "f32s", capacity_bytes=32 << 20)
    b = tclb_amd.Batch("d2q9_SRT", "f32s", capacity_bytes=32 << 20)
    ia = [a.add(c) for c in cs]
    ib = [b.add(c) for c in cs]
    a.iterate(5)
    b.iterate(5)
    t = a.autotune(steps=2)
    assert t["block"] in (64, 128, 256) and t["rows_per_tile"] in (1, 2, 4) and t["us_per_step"] > 0
    assert t["flat"] in (0, 1)
    a.iterate(8)
    b.iterate(8)
    for x, y in zip(ia, ib):
        assert np.array_equal(a.get_f(x), b.get_f(y), equal_nan=True)


def test_admission_by_capacity_and_reuse():
    say("arena admission")
    need = tclb_amd.Batch("d2q9_SRT", "f16s", capacity_bytes=1 << 20).case_bytes(64, 32)
    b = tclb_amd.Batch("d2q9_SRT", "f16s", capacity_bytes=3 * need)
    ids = [b.add_case(64, 32) for _ in range(3)]
    assert None not in ids
    assert b.add_case(64, 32) is None          # full
    b.remove_case(ids[1])
    mid = b.add_case(64, 32)                     # the freed slot is reused
    assert mid is not None
    assert b.info()["cases"] == 3
    assert b.add_case(64, 64) is None
    b.remove_case(ids[0])
    b.remove_case(mid)                           # two adjacent free blocks coalesce
    assert b.add_case(64, 64) is not None        # 64x64 needs exactly two 64x32 blocks


def test_graph_replay_batch():
    say("batch graph replay")
    cs = mixed_cases("d3q19", 9)
    a = tclb_amd.Batch("d3q19", "f64", capacity_bytes=64 << 20)
    b = tclb_amd.Batch("d3q19", "f64", capacity_bytes=64 << 20)
    ia = [a.add(c) for c in cs]
    ib = [b.add(c) for c in cs]
    a.set_graph_steps(4)
    b.set_graph_steps(0)
    for n in (3, 17, 1):
        a.iterate(n)
        b.iterate(n)
    for x, y in zip(ia, ib):
        assert np.array_equal(a.get_f(x), b.get_f(y), equal_nan=True)
        assert np.array_equal(a.globals(x), b.globals(y), equal_nan=True)


def test_late_admission_joins_mid_run():
    """A case admitted while others are half-way through gives the same
    result as when it starts alone."""
    say("late admission")
    cs = mixed_cases("d2q9_SRT_drag", 11)
    b = tclb_amd.Batch("d2q9_SRT_drag", "f64", capacity_bytes=64 << 20)
    first = [b.add(c) for c in cs[:3]]
    b.iterate(5)
    late = b.add(cs[3])
    b.iterate(12)
    solo = tclb_amd.Batch("d2q9_SRT_drag", "f64", capacity_bytes=16 << 20)
    s = solo.add(cs[3])
    solo.iterate(12)
    assert b.iteration(late) == 12 and b.iteration(first[0]) == 17
    assert np.array_equal(b.get_f(late), solo.get_f(s), equal_nan=True)
    assert np.array_equal(b.globals(late), solo.globals(s), equal_nan=True)


def test_sweep_generator_cases_run():
    """A few cases of the 256-case cylinder sweep generator in one batch."""
    say("cylinder sweep cases")
    sweep = cases.d2q9_cylinder_sweep(ncases=6, steps=50)
    b = tclb_amd.Batch(sweep[0].model, "f16s", capacity_bytes=256 << 20)
    ids = [b.add(c) for c in sweep]
    b.iterate(50)
    gd = tclb_amd.Model(sweep[0].model).global_id("Drag")
    for cid in ids:
        assert np.all(np.isfinite(b.globals(cid)))
        assert b.globals(cid)[gd] > 0          # the flow pushes every disc downstream
        assert np.all(np.isfinite(b.quantity(cid, "U")))


def test_full_sweep_batch_case_by_case(oracle_mod):
    """The cylinder sweep at its stated size: all 256 cases (24.9M nodes)
    in one f16s batch, autotuned, 1,000 steps.  Every case stays finite, the
    drag of most cases points downstream, and a sample of eight
    cases (the four smallest, four at random) equals, bit for bit, its own
    solve in a batch of one with the same tile shape.  In FP64 the three smallest cases of the sweep, run
    as one batch, match the oracle after 1,000 steps (populations to
    1e-12, Drag/Lift to 1e-10)."""
    say("full 256-case sweep, case by case")
    steps = 1000
    sweep = cases.d2q9_cylinder_sweep(ncases=256, steps=steps)
    model = sweep[0].model
    probe = tclb_amd.Batch(model, "f16s", 1 << 20)
    cap = sum(probe.case_bytes(c.nx, c.ny) for c in sweep) + (16 << 20)
    probe.close()
    b = tclb_amd.Batch(model, "f16s", capacity_bytes=cap)
    ids = [b.add(c) for c in sweep]
    assert b.info()["nodes"] == sum(c.nodes for c in sweep)
    t = b.autotune(steps=2)
    b.iterate(steps)
    gd = tclb_amd.Model(model).global_id("Drag")
    drags = []
    for cid in ids:
        g = b.globals(cid)
        assert np.all(np.isfinite(g))
        drags.append(g[gd])
    # the start-up pressure waves still swing some discs' drag below zero;
    # the flow pushes the discs downstream as a whole
    say(f"  drag > 0 in {np.mean(np.array(drags) > 0):.0%} of the cases")
    assert np.median(drags) > 0 and np.mean(np.array(drags) > 0) > 0.5
    order = sorted(range(256), key=lambda k: sweep[k].nodes)
    sample = order[:4] + list(np.random.default_rng(3).choice(order[4:], 4, replace=False))
    for k in sample:
        one = tclb_amd.Batch(model, "f16s", capacity_bytes=sweep[k].nodes * 64 + (4 << 20))
        one.set_tuning(block=t["block"], rows_per_tile=t["rows_per_tile"], grid=t["grid"], flat=t["flat"],
                       order=t["order"])
        c1 = one.add(sweep[k])
        one.iterate(steps)
        assert np.array_equal(b.get_f(ids[k]), one.get_f(c1)), sweep[k].name
        assert np.array_equal(b.globals(ids[k]), one.globals(c1)), sweep[k].name
        one.close()
    b.close()
    small = [sweep[k] for k in order[:3]]
    f = tclb_amd.Batch(model, "f64", capacity_bytes=64 << 20)
    fid = [f.add(c) for c in small]
    f.iterate(steps)
    for c, cid in zip(small, fid):
        orc = oracle_for_case(oracle_mod, c, "f64")
        g = orc.iterate(steps)
        a, o = f.get_f(cid), orc.get_f()
        assert np.max(np.abs(a - o)) <= 1e-12 * np.max(np.abs(o)), c.name
        glob_close(f.globals(cid), g[:f.nglobals], 1e-10)
