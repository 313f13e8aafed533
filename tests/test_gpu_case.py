"""Case files run end to end on the GPU (tclbx_case_run): the handler order
and scheduling of the reference (Model, Param, Solve with Log / Stop /
Failcheck / VTK / SaveBinary callbacks, LoadBinary, Units, deprecated
Params), the Log CSV and VTI outputs, and parity of the lattice a case
produces with the oracle run from the oracle's own rasterization of the
same file."""
import base64
import csv
import os
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from conftest import oracle_for_case

import tclb_amd
from tclb_amd import cases
from oracle import geometry as G

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def run_case(model, xml, tmp_path, name="case.xml", storage="f64"):
    c = tclb_amd.Case(model, xml=xml, name=name, storage=storage, output_root=str(tmp_path) + "/")
    c.run()
    return c


CHANNEL = """<CLBConfig output="out/">
  <Geometry nx="64" ny="24">
    <BGK><Box/></BGK>
    <WVelocity name="Inlet"><Inlet/></WVelocity>
    <EPressure name="Outlet"><Outlet/></EPressure>
    <Wall mask="ALL"><Channel/></Wall>
    <Solid name="body"><OffgridSphere x="16" y="12" R="4.3"/></Solid>
  </Geometry>
  <Model>
    <Param name="nu" value="0.05"/>
    <Param name="Velocity_x" value="0.02" zone="Inlet"/>
    <Param name="DragInObj" value="0.5" zone="body"/>
  </Model>
  <Log Iterations="50"/>
  <Solve Iterations="200"/>
</CLBConfig>"""


def read_log(path):
    with open(path) as f:
        rows = list(csv.reader(f, skipinitialspace=True))
    return rows[0], [[float(v) for v in r] for r in rows[1:]]


def test_channel_case_matches_oracle_and_logs(oracle_mod, tmp_path):
    """The d2q9_SRT_drag channel case: the final rho/u within 1e-12 of the
    oracle run from the oracle's rasterization, the Log rows (every 50
    steps) carrying the oracle's Drag/Lift/Objective of those steps."""
    say("case: d2q9 channel with Log")
    c = run_case("d2q9_SRT_drag", CHANNEL, tmp_path, "chan.xml")
    assert c.iteration == 200
    flags, zones, _ = G.read_case(CHANNEL, "d2q9_SRT_drag", is_text=True)
    assert np.array_equal(c.lattice.get_flags(), flags)
    oc = cases.Case("chan", "d2q9_SRT_drag", 64, 24, 1, flags, settings=[("nu", 0.05)],
                    zonal=[("Velocity_x", zones["Inlet"], 0.02), ("DragInObj", zones["body"], 0.5)])
    orc = oracle_for_case(oracle_mod, oc, "f64")
    head, rows = read_log(tmp_path / "out" / "chan_Log_P00_00000000.csv")
    assert [r[0] for r in rows] == [50, 100, 150, 200]
    for r in rows:
        g = orc.iterate(50)
        for name, want in zip(["Drag", "Lift", "Objective"], g):
            got = r[head.index(name)]
            assert abs(got - want) <= 1e-10 * max(abs(want), 1e-300) or abs(want) < 1e-10, (name, got, want)
        assert r[head.index("nu")] == 0.05 and r[head.index("Velocity_x-Inlet")] == 0.02
    u, uo = c.lattice.quantity("U"), orc.quantity("U")
    assert np.max(np.abs(u - uo)) <= 1e-12 * np.max(np.abs(uo))


def test_log_header_is_the_reference_layout(tmp_path):
    """d2q9_SRT: Iteration, Time_si, Walltime, Optimization, then value and
    _si of every setting, of every zonal setting per zone (zones sorted by
    name), of every Global, then the dx/dt/dm scales (src/Solver.cpp.Rt:
    126-172)."""
    xml = CHANNEL.replace('<Param name="DragInObj" value="0.5" zone="body"/>', "")
    c = run_case("d2q9_SRT", xml, tmp_path, "hdr.xml")
    head, rows = read_log(tmp_path / "out" / "hdr_Log_P00_00000000.csv")
    want = ["Iteration", "Time_si", "Walltime", "Optimization"]
    for s in ["omega", "nu", "Density", "Threshold"]:
        want += [s, s + "_si"]
    for s in ["Velocity", "Velocity_x", "Velocity_y", "GravitationX", "GravitationY"]:
        for z in ["DefaultZone", "Inlet", "Outlet", "body"]:
            want += [f"{s}-{z}", f"{s}-{z}_si"]
    want += ["Objective", "Objective_si", "dx_si", "dt_si", "dm_si"]
    assert head == want
    assert len(rows) == 4 and all(len(r) == len(want) for r in rows)
    assert rows[0][1] == 50.0 and rows[0][-3:] == [1.0, 1.0, 1.0]
    text = open(tmp_path / "out" / "hdr_Log_P00_00000000.csv").read().splitlines()[1]
    assert text.startswith("50, 5.0000000000000e+01, ")   # %d, %.13le, ...


def test_stop_ends_the_solve(tmp_path):
    """<Stop VOLvolumeChange=..>: the fluid volume is constant, so the
    second check (and with Times=3 the fourth) passes and stops the Solve."""
    say("case: Stop")
    base = """<CLBConfig>
      <Geometry nx="16" ny="12" nz="10"><MRT><Box/></MRT><Wall mask="ALL"><Channel/></Wall></Geometry>
      <Model><Param name="nu" value="0.1"/><Param name="ForceX" value="1e-5"/></Model>
      {stop}
      <Solve Iterations="10000"/>
    </CLBConfig>"""
    c = run_case("d3q19", base.format(stop='<Stop VOLvolumeChange="0.5" Iterations="100"/>'), tmp_path, "s1.xml")
    assert c.iteration == 200
    c = run_case("d3q19", base.format(stop='<Stop VOLvolumeChange="0.5" Times="3" Iterations="100"/>'), tmp_path,
                 "s2.xml")
    assert c.iteration == 400
    c = run_case("d3q19", base.format(stop='<Stop VOLvxAbove="1e9" Iterations="100"/>'), tmp_path, "s3.xml")
    assert c.iteration == 10000


def test_save_load_and_failcheck(tmp_path):
    """SaveBinary writes <output>_<file>_0.pri; a NaN planted in it and read
    back with LoadBinary is found by Failcheck at its first check, which
    stops the Solve there."""
    say("case: SaveBinary, LoadBinary, Failcheck")
    geom = '<Geometry nx="32" ny="16"><BGK><Box/></BGK><Wall mask="ALL"><Channel/></Wall></Geometry>'
    a = run_case("d2q9_SRT", f"""<CLBConfig output="out/">{geom}
        <Model><Param name="nu" value="0.1"/><Param name="GravitationX" value="1e-5"/></Model>
        <SaveBinary file="state" Iterations="20"/><Solve Iterations="20"/></CLBConfig>""", tmp_path, "a.xml")
    pri = tmp_path / "out" / "a_state_0.pri"
    # the reference's layout (margin blocks, no footer)
    data = np.fromfile(pri, dtype=np.uint8)
    assert data.size == tclb_amd.pri_bytes("d2q9_SRT", "f64", 32, 16, 1)
    f, _ = tclb_amd.pri_unpack("d2q9_SRT", "f64", 32, 16, 1, data, 9)
    f = f.reshape(9, 16, 32)
    assert np.array_equal(f, a.lattice.get_f().reshape(9, 16, 32))
    b = run_case("d2q9_SRT", f"""<CLBConfig output="out/">{geom}
        <Model><Param name="nu" value="0.1"/><Param name="GravitationX" value="1e-5"/></Model>
        <LoadBinary file="{tmp_path}/out/a_state"/><Solve Iterations="30"/></CLBConfig>""", tmp_path, "b.xml")
    a2 = run_case("d2q9_SRT", f"""<CLBConfig output="out/">{geom}
        <Model><Param name="nu" value="0.1"/><Param name="GravitationX" value="1e-5"/></Model>
        <Solve Iterations="50"/></CLBConfig>""", tmp_path, "a2.xml")
    assert np.array_equal(b.lattice.get_f(), a2.lattice.get_f())   # 20 + 30 steps = 50 steps
    f[3, 8, 10] = np.nan
    tclb_amd.pri_pack("d2q9_SRT", "f64", f.reshape(9, 1, 16, 32)).tofile(tmp_path / "out" / "bad_0.pri")
    c = run_case("d2q9_SRT", f"""<CLBConfig output="out/">{geom}
        <Model><Param name="nu" value="0.1"/></Model>
        <LoadBinary file="{tmp_path}/out/bad"/>
        <Failcheck Iterations="10" nx="32" ny="16"/><Solve Iterations="100"/></CLBConfig>""", tmp_path, "c.xml")
    assert c.iteration == 10
    assert any("NaN" in m for m in c.messages)


def _vti_arrays(path):
    root = ET.parse(path).getroot()
    out = {}
    for da in root.iter("DataArray"):
        text = da.text.strip()
        n = struct.unpack("<i", base64.b64decode(text[:8]))[0]
        raw = base64.b64decode(text[8:])
        assert len(raw) == n
        dt = {"Float64": np.float64, "UInt8": np.uint8, "UInt16": np.uint16}[da.get("type")]
        out[da.get("Name")] = (np.frombuffer(raw, dtype=dt), int(da.get("NumberOfComponents")))
    return root, out


def test_vtk_output(tmp_path):
    """The VTI file of a <VTK> callback: extent, the NodeType groups as
    UInt8 fields, P and U as Float64 fields equal to the lattice's
    quantities, and the .pvti that names it."""
    say("case: VTK")
    xml = """<CLBConfig output="out/">
      <Geometry nx="16" ny="12" nz="10"><MRT><Box/></MRT><Wall mask="ALL"><Channel/></Wall>
        <XYslice><Box dz="5" nz="1"/></XYslice></Geometry>
      <Model><Param name="nu" value="0.1"/><Param name="ForceX" value="1e-5"/></Model>
      <VTK Iterations="40"/><Solve Iterations="40"/></CLBConfig>"""
    c = run_case("d3q19", xml, tmp_path, "v.xml")
    root, arr = _vti_arrays(tmp_path / "out" / "v_VTK_P00_00000040.vti")
    assert root.find("ImageData").get("WholeExtent") == "0 16 0 12 0 10"
    assert set(arr) == {"ADDITIONALS", "BOUNDARY", "COLLISION", "SETTINGZONE", "P", "U"}
    fl = c.lattice.get_flags().reshape(-1)
    assert np.array_equal(arr["BOUNDARY"][0], ((fl & 0x1C) >> 2).astype(np.uint8))
    assert np.array_equal(arr["ADDITIONALS"][0], (fl & 0x3).astype(np.uint8))
    assert np.array_equal(arr["P"][0], c.lattice.quantity("P").reshape(-1))
    assert arr["U"][1] == 3 and np.array_equal(arr["U"][0], c.lattice.quantity("U").reshape(-1))
    p = ET.parse(tmp_path / "out" / "v_VTK_P00_00000040.pvti").getroot()
    assert p.find("PImageData/Piece").get("Source") == "v_VTK_P00_00000040.vti"


def test_units_and_deprecated_params(tmp_path):
    """Units gauges convert Solve Iterations ("0.05s" = 50 steps at 1000
    steps per second) and Param values; a deprecated <Params a="v"/>
    becomes <Param name="a" value="v"/> and makes the case permissive."""
    say("case: Units, Params")
    xml = """<CLBConfig output="out/">
      <Units><Param name="size" value="1mm" gauge="1"/><Param name="time" value="1s" gauge="1000"/></Units>
      <Geometry nx="20mm" ny="8mm"><BGK><Box/></BGK><Wall mask="ALL"><Channel/></Wall></Geometry>
      <Model><Params nu="1e-4m2/s" GravitationX="1e-6" NotASetting="3"/></Model>
      <Solve Iterations="0.05s"/></CLBConfig>"""
    c = run_case("d2q9_SRT", xml, tmp_path, "u.xml")
    assert (c.nx, c.ny) == (20, 8)
    assert c.iteration == 50
    assert c.lattice.get_setting("nu") == pytest.approx(1e-4 * 1e6 / 1000, rel=1e-12)
    assert any("Unknown setting NotASetting" in m for m in c.messages)


def test_unknown_setting_fails_when_not_permissive(tmp_path):
    xml = """<CLBConfig><Geometry nx="8" ny="4"><BGK><Box/></BGK></Geometry>
      <Model><Param name="NoSuch" value="1"/></Model></CLBConfig>"""
    with pytest.raises(tclb_amd.TclbxError):
        run_case("d2q9_SRT", xml, tmp_path, "bad.xml")


def test_d2q9_mrt_channel_case(oracle_mod, tmp_path):
    """The d2q9 (MRT) channel case file: the reference's Inlet/Outlet
    objective nodes and pressure-loss/flux Globals, Log rows against the
    oracle, final u within 1e-12."""
    say("case: d2q9 MRT channel")
    path = os.path.join(os.path.dirname(__file__), "golden", "cases", "channel_d2q9.xml")
    c = tclb_amd.Case("d2q9", path, output_root=str(tmp_path) + "/")
    c.run()
    assert c.iteration == 300
    flags, zones, _ = G.read_case(path, "d2q9")
    assert np.array_equal(c.lattice.get_flags(), flags)
    oc = cases.Case("chan", "d2q9", 160, 40, 1, flags, settings=[("Viscosity", 0.02), ("S3", 0.2)],
                    zonal=[("VelocityX", zones["Inlet"], 0.02), ("PressureLossInObj", -1, 1.0)])
    orc = oracle_for_case(oracle_mod, oc, "f64")
    head, rows = read_log(tmp_path / "output" / "channel_d2q9_Log_P00_00000000.csv")
    assert head[4:8] == ["RelaxationRate", "RelaxationRate_si", "Viscosity", "Viscosity_si"]
    assert [r[0] for r in rows] == [100, 200, 300]
    for r in rows:
        g = orc.iterate(100)
        for name, want in zip(["PressureLoss", "OutletFlux", "InletFlux", "Objective"], g):
            got = r[head.index(name)]
            assert abs(got - want) <= 1e-10 * abs(want) or abs(want) < 1e-10, (name, got, want)
        assert r[head.index("S2")] == pytest.approx(1 - 1 / (3 * 0.02 + 0.5), rel=1e-15)
    u, uo = c.lattice.quantity("U"), orc.quantity("U")
    assert np.max(np.abs(u - uo)) <= 1e-12 * np.max(np.abs(uo))
    assert rows[-1][head.index("InletFlux")] > 0


def test_save_checkpoint_keeps_the_last_and_restarts(tmp_path):
    """SaveCheckpoint writes <output>_checkpoint_<iter>_0.pri and
    <output>_restart_<iter>.xml, keeping the last `keep` pairs; the restart
    case is the case file with a LoadBinary of the checkpoint before the
    first Solve, and running it continues the flow exactly (the kernels are
    deterministic): checkpoint at 40 + 60 steps = 100 steps."""
    say("case: SaveCheckpoint and restart")
    geom = '<Geometry nx="32" ny="16"><BGK><Box/></BGK><Wall mask="ALL"><Channel/></Wall></Geometry>'
    model = '<Model><Param name="nu" value="0.1"/><Param name="GravitationX" value="1e-5"/></Model>'
    a = run_case("d2q9_SRT", f"""<CLBConfig output="out/">{geom}{model}
        <Solve Iterations="60"><SaveCheckpoint Iterations="20" keep="2"/></Solve></CLBConfig>""", tmp_path, "ck.xml")
    out = tmp_path / "out"
    names = sorted(p.name for p in out.iterdir())
    assert names == ["ck_checkpoint_00000040_0.pri", "ck_checkpoint_00000060_0.pri",
                     "ck_restart_00000040.xml", "ck_restart_00000060.xml"], names
    data = np.fromfile(out / "ck_checkpoint_00000060_0.pri", dtype=np.uint8)
    pri, _ = tclb_amd.pri_unpack("d2q9_SRT", "f64", 32, 16, 1, data, 9)
    assert np.array_equal(pri.reshape(9, 16, 32), a.lattice.get_f().reshape(9, 16, 32))
    text = (out / "ck_restart_00000040.xml").read_text()
    assert text.startswith('<?xml version="1.0"?>')
    root = ET.fromstring(text)
    kids = [c.tag for c in root]
    assert kids.index("LoadBinary") == kids.index("Solve") - 1
    assert root.find("LoadBinary").get("file") == str(out / "ck_checkpoint_00000040")
    r = tclb_amd.Case("d2q9_SRT", str(out / "ck_restart_00000040.xml"), output_root=str(tmp_path / "r") + "/")
    r.run()
    b = run_case("d2q9_SRT", f"""<CLBConfig output="out2/">{geom}{model}
        <Solve Iterations="100"/></CLBConfig>""", tmp_path, "ref.xml")
    assert np.array_equal(r.lattice.get_f(), b.lattice.get_f())
    # keep="all" keeps every checkpoint
    run_case("d2q9_SRT", f"""<CLBConfig output="all/">{geom}{model}
        <Solve Iterations="60"><SaveCheckpoint Iterations="20" keep="all"/></Solve></CLBConfig>""", tmp_path, "k.xml")
    assert len(list((tmp_path / "all").glob("k_checkpoint_*_0.pri"))) == 3


@pytest.mark.parametrize("model,fname,settings,zonal", [
    ("d3q27", "channel_d3q27.xml", [("nu", 0.04), ("Smag", 0.16)],
     [("Velocity", "Inlet", 0.03), ("XDragForceInObj", -1, 1.0)]),
    ("d3q27_cumulant", "channel_cumulant.xml", [("nu", 0.03), ("nubuffer", 0.1), ("Omega", 0.8)],
     [("Velocity", "Inlet", 0.03), ("DragInObj", -1, 1.0)]),
])
def test_d3q27_channel_cases(oracle_mod, tmp_path, model, fname, settings, zonal):
    """The 3D channel case files of the 27-velocity models: the NodeType map
    equals the oracle's rasterization, each Log row's Globals are within
    1e-10 of the oracle run on that map, and the final u within 1e-12."""
    say(f"case: {model} channel")
    path = os.path.join(os.path.dirname(__file__), "golden", "cases", fname)
    c = tclb_amd.Case(model, path, output_root=str(tmp_path) + "/")
    c.run()
    assert c.iteration == 300
    flags, zones, _ = G.read_case(path, model)
    assert np.array_equal(c.lattice.get_flags(), flags)
    nz, ny, nx = flags.shape
    zon = [(n, zones[z] if isinstance(z, str) else z, v) for n, z, v in zonal]
    oc = cases.Case("chan", model, nx, ny, nz, flags, settings=settings, zonal=zon)
    orc = oracle_for_case(oracle_mod, oc, "f64")
    head, rows = read_log(tmp_path / "output" / (fname[:-4] + "_Log_P00_00000000.csv"))
    assert [r[0] for r in rows] == [100, 200, 300]
    names = tclb_amd.Model(model).global_names
    for r in rows:
        g = orc.iterate(100)
        for name, want in zip(names, g):
            got = r[head.index(name)]
            assert abs(got - want) <= 1e-10 * abs(want) or abs(want) < 1e-10, (name, got, want)
    u, uo = c.lattice.quantity("U"), orc.quantity("U")
    assert np.max(np.abs(u - uo)) <= 1e-12 * np.max(np.abs(uo))
    # the drag Globals of the body see the flow (their sign follows each
    # model's own formula: d3q27 sums 2 f e after the swap, the cumulant
    # model the momentum change)
    drag = rows[-1][head.index("XDragForce" if model == "d3q27" else "Drag")]
    assert abs(drag) > 1e-6
    if model == "d3q27":   # a 3D VTI: every quantity of the model, x fastest
        vti = tmp_path / "output" / "channel_d3q27_VTK_P00_00000300.vti"
        root, arr = _vti_arrays(vti)
        assert root.find("ImageData").get("WholeExtent") == f"0 {nx} 0 {ny} 0 {nz}"
        for q, ncomp in (("P", 1), ("U", 3), ("Fd", 3)):
            v, k = arr[q]
            assert k == ncomp and v.size == nx * ny * nz * ncomp
        assert np.allclose(arr["U"][0].reshape(nz, ny, nx, 3), u, rtol=0, atol=0)


IB_CHANNEL = """<CLBConfig output="output/">
  <Geometry nx="40" ny="18" nz="14">
    <MRT><Box/></MRT>
    <WVelocity name="Inlet"><Inlet/></WVelocity>
    <EPressure name="Outlet"><Outlet/></EPressure>
    <IB><STL file="{stl}" side="surface"/></IB>
    <Body><STL file="{stl}" side="surface"/></Body>
  </Geometry>
  <Model>
    <Param name="nu" value="0.03"/>
    <Param name="Velocity" value="0.03" zone="Inlet"/>
    <Param name="DragInObj" value="1"/>
  </Model>
  <Log Iterations="100"/>
  <Solve Iterations="300"/>
</CLBConfig>"""


@pytest.mark.parametrize("model", ["d3q27_cumulant_IB", "d3q27_cumulant_IB_SMAG"])
def test_ib_sphere_case(oracle_mod, tmp_path, model):
    """A sphere given as an STL surface in a channel, the IB builds of the
    cumulant model: the case's NodeType map and cut distances equal the
    oracle's rasterization, the lattice holds those cuts, each Log row's
    Globals (Drag of the Body nodes among them) are within 1e-10 of the
    oracle run on that map, and the final u within 1e-12."""
    from test_geometry import icosphere, write_stl
    say(f"case: {model} STL sphere with interpolated bounce-back")
    stl = write_stl(tmp_path / "sphere.stl", icosphere((12.3, 9.1, 6.8), 4.2))
    xml = IB_CHANNEL.format(stl=stl)
    c = tclb_amd.Case(model, xml=xml, name="ib.xml", output_root=str(tmp_path) + "/")
    c.run()
    assert c.iteration == 300
    flags, zones, _, cuts = G.read_case(xml, model, is_text=True, cuts=True)
    assert cuts is not None and (cuts != 65535).sum() > 500
    assert np.array_equal(c.lattice.get_flags(), flags)
    assert np.array_equal(c.lattice.get_cuts(), cuts)
    nz, ny, nx = flags.shape
    settings = [("nu", 0.03)] + ([("Smag", 0.0)] if model.endswith("SMAG") else [])
    zon = [("Velocity", zones["Inlet"], 0.03), ("DragInObj", -1, 1.0)]
    oc = cases.Case("ib", model, nx, ny, nz, flags, settings=settings, zonal=zon, cuts=cuts)
    orc = oracle_for_case(oracle_mod, oc, "f64")
    head, rows = read_log(tmp_path / "output" / "ib_Log_P00_00000000.csv")
    assert [r[0] for r in rows] == [100, 200, 300]
    names = tclb_amd.Model(model).global_names
    for r in rows:
        g = orc.iterate(100)
        for name, want in zip(names, g):
            got = r[head.index(name)]
            assert abs(got - want) <= 1e-10 * abs(want) or abs(want) < 1e-10, (name, got, want)
    assert abs(rows[-1][head.index("Drag")]) > 1e-6
    u, uo = c.lattice.quantity("U"), orc.quantity("U")
    assert np.max(np.abs(u - uo)) <= 1e-12 * np.max(np.abs(uo))


AVG_CHANNEL = """<CLBConfig output="output/">
  <Geometry nx="32" ny="14" nz="10">
    <MRT><Box/></MRT>
    <WVelocity name="Inlet"><Inlet/></WVelocity>
    <EPressure name="Outlet"><Outlet/></EPressure>
    <Wall mask="ALL"><OffgridSphere x="10" y="7" z="5" R="2.6"/></Wall>
  </Geometry>
  <Model>
    <Param name="nu" value="0.02"/>
    <Param name="Velocity" value="0.04" zone="Inlet"/>
  </Model>
  <VTK Iterations="300" what="avgU,varU,KinE,averageP,U"/>
  <Average Iterations="100"/>
  <Solve Iterations="300"/>
</CLBConfig>"""


def test_average_handler_case(oracle_mod, tmp_path):
    """<Average Iterations=100> resets the averages when set up and every 100
    iterations; the VTK callback listed before it writes, at iteration 300,
    the averages of iterations 201-300, equal to the oracle's after the same
    resets."""
    model = "d3q27_cumulant_AVG"
    say("case: d3q27_cumulant_AVG with <Average>")
    c = tclb_amd.Case(model, xml=AVG_CHANNEL, name="avg.xml", output_root=str(tmp_path) + "/")
    c.run()
    assert c.iteration == 300
    flags, zones, _ = G.read_case(AVG_CHANNEL, model, is_text=True)
    nz, ny, nx = flags.shape
    oc = cases.Case("avg", model, nx, ny, nz, flags, settings=[("nu", 0.02)],
                    zonal=[("Velocity", zones["Inlet"], 0.04)])
    orc = oracle_for_case(oracle_mod, oc, "f64")
    orc.iterate(200, False)
    orc.reset_average()
    orc.iterate(100, False)
    root, arr = _vti_arrays(tmp_path / "output" / "avg_VTK_P00_00000300.vti")
    for q, ncomp in (("avgU", 3), ("varU", 3), ("KinE", 1), ("averageP", 1), ("U", 3)):
        v, k = arr[q]
        assert k == ncomp
        want = orc.quantity(q)
        got = v.reshape(want.shape)
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want)), q
    mu = arr["avgU"][0].reshape(nz, ny, nx, 3)
    assert mu[..., 0].max() > 0.01          # the mean flow is there


def _read_txt(path, nx, comp):
    """A TXT quantity file: one row of nx values (comp numbers each) per line."""
    if str(path).endswith(".gz"):
        import gzip
        with gzip.open(path, "rt") as f:
            lines = f.read().splitlines()
    else:
        with open(path) as f:
            lines = f.read().splitlines()
    rows = [np.array([float(v) for v in ln.split()]) for ln in lines]
    assert all(r.size == nx * comp for r in rows)
    return np.concatenate(rows)


def test_txt_container_repeat(tmp_path):
    """<Repeat Times="3"> runs its Solve and TXT three times; the TXT files
    (src/vtkLattice.cpp:97-160) hold the info header and the quantities at
    16 digits, a row of nx values per line, plain and gzipped.  A <Log>
    stacked inside a <Container> comes off when the container ends, so the
    Solve after it adds no rows."""
    say("case: TXT, Container, Repeat")
    geo = """<Geometry nx="16" ny="12" nz="4"><MRT><Box/></MRT><Wall mask="ALL"><Channel/></Wall></Geometry>
      <Model><Param name="nu" value="0.1"/><Param name="ForceX" value="1e-5"/></Model>"""
    xml = f"""<CLBConfig output="out/">{geo}
      <Repeat Times="3"><Solve Iterations="10"/><TXT name="T" what="P,U"/></Repeat>
      <Container><Log Iterations="5"/><Solve Iterations="10"/></Container>
      <Solve Iterations="10"/>
      <TXT name="Z" gzip="true"/></CLBConfig>"""
    c = run_case("d3q19", xml, tmp_path, "t.xml")
    assert c.iteration == 50
    out = tmp_path / "out"
    for it in (10, 20, 30):
        assert (out / f"t_T_P00_{it:08d}_P.txt").exists() and (out / f"t_T_P00_{it:08d}_U.txt").exists()
    assert not (out / "t_T_P00_00000010_Solid.txt").exists()
    info = (out / "t_T_P00_00000030_info.txt").read_text().splitlines()
    assert info[4:] == ["size: 768", "NX: 16", "NY: 12", "NZ: 4"] and info[0] == "dx: 1"
    # the state after 30 steps, from a case that only solves 30
    ref = run_case("d3q19", f"""<CLBConfig output="ref/">{geo}<Solve Iterations="30"/></CLBConfig>""", tmp_path, "r.xml")
    np.testing.assert_allclose(_read_txt(out / "t_T_P00_00000030_P.txt", 16, 1),
                               ref.lattice.quantity("P").reshape(-1), rtol=1e-15, atol=0)
    np.testing.assert_allclose(_read_txt(out / "t_T_P00_00000030_U.txt", 16, 3),
                               ref.lattice.quantity("U").reshape(-1), rtol=1e-15, atol=1e-300)
    np.testing.assert_allclose(_read_txt(out / "t_Z_P00_00000050_U.txt.gz", 16, 3),
                               c.lattice.quantity("U").reshape(-1), rtol=1e-15, atol=1e-300)
    logs = sorted(out.glob("t_Log*.csv"))
    assert len(logs) == 1
    _, rows = read_log(logs[0])
    assert [r[0] for r in rows] == [35.0, 40.0]
