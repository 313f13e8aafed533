"""The C-ABI library without a GPU: it loads, exports every entry point
include/tclbx.h declares, and its model registry holds the reference's
NodeType bit layout (src/conf.R:600-699), Globals order (src/conf.R:740-761)
and settings."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, gpu_available

import tclb_amd


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "tclbx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(tclbx_\w+)\s*\(", text, re.M)))


def test_every_declared_symbol_is_exported():
    lib = tclb_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(tclb_amd.SYMBOLS) == syms


def test_models():
    assert tclb_amd.models() == ["d2q9_SRT", "d2q9_SRT_drag", "d3q19", "d3q19_lid", "d2q9", "d2q9_drag", "d3q27",
                                 "d3q27_cumulant", "d2q9_autosym", "d2q9_autosym2", "d3q27_cumulant_SMAG",
                                 "d3q27_cumulant_IB", "d3q27_cumulant_IB_SMAG", "d3q27_cumulant_AVG",
                                 "d3q27_cumulant_AVG_IB", "d3q27_cumulant_AVG_SMAG", "d3q27_cumulant_AVG_IB_SMAG",
                                 "d2q9_bc", "d2q9_bc_autosym", "d2q9_bc_autosym2"]


def test_d2q9_mrt_contract():
    """models/flow/d2q9/Dynamics.R: BOUNDARY has 10 types (4 bits), then
    COLLISION (MRT), then OBJECTIVE (Inlet, Outlet); zone from bit 7."""
    m = tclb_amd.Model("d2q9")
    order = ["EPressure", "WPressure", "NVelocity", "SVelocity", "WVelocity", "EVelocity", "NSymmetry",
             "SSymmetry", "Solid", "Wall"]
    for k, n in enumerate(order):
        assert m.nodetype(n) == (k + 1, 0xF)
    assert m.nodetype("MRT") == (0x10, 0x10)
    assert m.nodetype("Inlet") == (0x20, 0x60) and m.nodetype("Outlet") == (0x40, 0x60)
    assert m.group("SETTINGZONE") == (0xFF80, 7)
    assert m.settings == ["RelaxationRate", "Viscosity", "GravitationX", "GravitationY", "S2", "S3", "S4",
                          "Threshold"]
    assert m.zonal_settings == ["VelocityX", "VelocityY", "Pressure", "PressureLossInObj", "OutletFluxInObj",
                                "InletFluxInObj"]
    assert m.global_names == ["PressureLoss", "OutletFlux", "InletFlux", "Objective"]
    assert m.quantities == ["Rho", "U"]


def test_d2q9_srt_contract():
    m = tclb_amd.Model("d2q9_SRT")
    # NodeType values (BOUNDARY group, 3 bits; COLLISION 1 bit; zone from bit 4)
    want = {"EPressure": 1, "WPressure": 2, "WVelocity": 3, "EVelocity": 4, "Solid": 5, "Wall": 6}
    for n, v in want.items():
        assert m.nodetype(n) == (v, 0x7)
    assert m.nodetype("BGK") == (0x8, 0x8)
    assert m.group("SETTINGZONE") == (0xFFF0, 4)
    assert m.nodetype("None") == (0, 0)
    assert m.nodetype("Clear") == (0, 0xFFFF)
    assert m.global_names == ["Objective"]
    assert m.quantities == ["U", "Rho"]
    assert m.settings[:3] == ["omega", "nu", "Density"]
    assert m.zonal_settings == ["Velocity", "Velocity_x", "Velocity_y", "GravitationX", "GravitationY"]
    assert m.setting_default("Density") == 1.0
    assert m.setting_id("nonexistent") == -1


def test_d3q19_contract():
    m = tclb_amd.Model("d3q19")
    assert [m.nodetype(n)[0] for n in ("XYslice", "XZslice", "YZslice")] == [1, 2, 3]
    assert m.group("ADDITIONALS") == (0x3, 0)
    want = {"EPressure": 0x4, "Solid": 0x8, "Wall": 0xC, "WPressure": 0x10, "WPressureL": 0x14, "WVelocity": 0x18}
    for n, v in want.items():
        assert m.nodetype(n) == (v, 0x1C)
    assert m.nodetype("MRT") == (0x20, 0x20)
    assert m.group("SETTINGZONE") == (0xFFC0, 6)
    g = m.global_names
    assert len(g) == 26 and g[0] == "Flux" and g[24] == "Objective" and g[25] == "MaxV"
    assert g[16:24] == ["VOLvx", "VOLvy", "VOLvz", "VOLpx", "VOLpy", "VOLpz", "VOLrho", "VOLvolume"]
    assert len(m.zonal_settings) == 24 and m.zonal_settings[0] == "FluxInObj"
    assert m.quantities == ["P", "U"]
    assert m.quantity_id("U") == (1, True)


def test_lid_model_extends_d3q19():
    """d3q19_lid keeps every d3q19 bit, setting and Global at the same
    place, adds MovingWall as the seventh BOUNDARY type (0x1C) and the wall
    force/torque Globals before Objective."""
    a, b = tclb_amd.Model("d3q19"), tclb_amd.Model("d3q19_lid")
    for n in ("XYslice", "EPressure", "Solid", "Wall", "WPressure", "WPressureL", "WVelocity", "MRT"):
        assert a.nodetype(n) == b.nodetype(n)
    assert b.nodetype("MovingWall") == (0x1C, 0x1C)
    assert a.group("SETTINGZONE") == b.group("SETTINGZONE")
    assert b.settings[:13] == a.settings[:13]
    assert b.settings[13:] == ["MovingWallVelocityX", "MovingWallVelocityY", "MovingWallVelocityZ", "Threshold"]
    ga, gb = a.global_names, b.global_names
    assert gb[:24] == ga[:24]
    assert gb[24:] == ["MovingWallFx", "MovingWallFy", "MovingWallFz", "MovingWallTorque", "Objective", "MaxV"]
    assert b.zonal_settings[24:] == ["MovingWallFxInObj", "MovingWallFyInObj", "MovingWallFzInObj",
                                     "MovingWallTorqueInObj"]


def test_d2q9_drag_extends_d2q9():
    a, b = tclb_amd.Model("d2q9"), tclb_amd.Model("d2q9_drag")
    assert b.global_names == ["PressureLoss", "OutletFlux", "InletFlux", "Drag", "Lift", "Objective"]
    assert b.zonal_settings == a.zonal_settings + ["DragInObj", "LiftInObj"]
    assert b.settings == a.settings and b.group("SETTINGZONE") == a.group("SETTINGZONE")


def test_drag_model_globals():
    m = tclb_amd.Model("d2q9_SRT_drag")
    assert m.global_names == ["Drag", "Lift", "Objective"]
    assert m.zonal_settings[-2:] == ["DragInObj", "LiftInObj"]


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_lattice_creation_fails_loudly_without_gpu():
    with pytest.raises(tclb_amd.TclbxError):
        tclb_amd.Lattice("d3q19", 8, 8, 8)


def test_unknown_model_is_an_error():
    with pytest.raises(tclb_amd.TclbxError):
        tclb_amd.Lattice("d3q27_cumulant_part", 8, 8, 8)


def test_d3q27_contract():
    """models/flow/d3q27/Dynamics.R: groups BOUNDARY (8 types, 4 bits),
    COLLISION, ENTROPIC, LES in alphabetical order, zone from bit 7."""
    m = tclb_amd.Model("d3q27")
    assert m.Q == 27
    order = ["NSymmetry", "SSymmetry", "ISymmetry", "EPressure", "Solid", "Wall", "WPressure", "WVelocity"]
    for k, n in enumerate(order):
        assert m.nodetype(n) == (k + 1, 0xF)
    assert m.nodetype("MRT") == (0x10, 0x10)
    assert m.nodetype("Stab") == (0x20, 0x20)
    assert m.nodetype("Smagorinsky") == (0x40, 0x40)
    assert m.group("SETTINGZONE") == (0xFF80, 7)
    assert m.settings == ["omega", "nu", "Smag", "ForceX", "ForceY", "ForceZ", "Threshold"]
    assert m.zonal_settings == ["Velocity", "Pressure", "Turbulence", "XFluxInObj", "YFluxInObj", "ZFluxInObj",
                                "XDragForceInObj", "YDragForceInObj", "ZDragForceInObj"]
    assert m.global_names == ["XFlux", "YFlux", "ZFlux", "XDragForce", "YDragForce", "ZDragForce", "Objective"]
    assert m.quantities == ["P", "U", "Fd"]
    assert m.quantity_id("Fd") == (2, True)


def test_d3q27_cumulant_contract():
    """models/flow/d3q27_cumulant/Dynamics.R (no build options): groups
    ADDITIONALS, BODY, BOUNDARY (12 types, 4 bits), COLLISION, zone from bit 8."""
    m = tclb_amd.Model("d3q27_cumulant")
    assert m.nodetype("NSymmetry") == (1, 3) and m.nodetype("SSymmetry") == (2, 3)
    assert m.nodetype("Body") == (4, 4)
    order = ["Buffer", "WVelocityTurbulent", "NVelocity", "SVelocity", "NPressure", "SPressure", "EPressure",
             "EVelocity", "Solid", "Wall", "WPressure", "WVelocity"]
    for k, n in enumerate(order):
        assert m.nodetype(n) == ((k + 1) << 3, 0x78)
    assert m.nodetype("MRT") == (0x80, 0x80)
    assert m.group("SETTINGZONE") == (0xFF00, 8)
    assert m.settings == ["nu", "nubuffer", "GalileanCorrection", "ForceX", "ForceY", "ForceZ", "Omega", "Threshold"]
    assert m.zonal_settings[:3] == ["Velocity", "Pressure", "Turbulence"]
    assert m.global_names == ["Density", "Flux", "Drag", "Lift", "Lateral", "Mass", "XMomentum", "YMomentum",
                              "ZMomentum", "Objective"]
    assert m.quantities == ["P", "U", "Solid"]
    s = tclb_amd.Model("d3q27_cumulant_SMAG")
    assert s.settings == m.settings[:-1] + ["Smag", "Threshold"]
    assert s.global_names == m.global_names and s.group("SETTINGZONE") == m.group("SETTINGZONE")


def test_modelbench_covers_every_model():
    """python -m tclb_amd.modelbench can set up a box of every model: its
    collision NodeType and settings exist in each model's contract."""
    from tclb_amd import modelbench
    for name in tclb_amd.models():
        m = tclb_amd.Model(name)
        assert m.flag(modelbench.COLLISION.get(name, "MRT")) != 0, name
        for k, _ in modelbench.settings_of(name):
            assert m.setting_id(k) >= 0, (name, k)


@pytest.mark.parametrize("model", ["d3q27_cumulant_AVG", "d2q9_bc_autosym"])
def test_batch_refuses_models_with_extra_densities(model):
    """A batch's arena holds distributions, flags and zone tables only; the
    AVG and bc builds keep per-node densities that are not streamed, so a
    batch (and the server built on it) refuses them before allocating."""
    with pytest.raises(tclb_amd.TclbxError, match="use a lattice"):
        tclb_amd.Batch(model, "f64", 1 << 20)


@pytest.mark.parametrize("model,n,storage", [("d3q19", 813, "f64"), ("d3q19", 1025, "f32"), ("d3q19", 1291, "f16s"),
                                             ("d3q27", 1025, "f32s")])
def test_density_plane_limit(model, n, storage):
    """The step kernel addresses a density plane with 32-bit byte offsets, so
    a lattice whose plane passes 4 GiB is refused before anything is
    allocated: FP64 past 2^29 nodes (813^3), FP32 past 2^30 (1025^3), 16-bit
    formats past 2^31 (1291^3)."""
    with pytest.raises(tclb_amd.TclbxError, match="4 GiB per plane"):
        tclb_amd.Lattice(model, n, n, n, storage)


def test_density_plane_limit_boundary():
    """812^3 FP64 (535,387,328 nodes, a plane of 4,283,098,624 bytes) is
    within the limit: creation gets past the size check (without a GPU it
    then fails on the device, not on the size)."""
    if gpu_available():
        pytest.skip("would allocate 163 GB")
    with pytest.raises(tclb_amd.TclbxError) as e:
        tclb_amd.Lattice("d3q19", 812, 812, 812, "f64")
    assert "4 GiB" not in str(e.value)



def test_temporal_blocking_entry_points_refuse_a_null_lattice():
    """tclbx_set_temporal_blocking and tclbx_get_temporal_blocking check the
    lattice handle before anything else (no GPU needed): an error with a
    message, and the output left untouched."""
    lib = tclb_amd.lib()
    assert lib.tclbx_set_temporal_blocking(None, 3) != 0
    assert "null lattice handle" in lib.tclbx_last_error().decode()
    k = ctypes.c_int(-7)
    assert lib.tclbx_get_temporal_blocking(None, ctypes.byref(k)) != 0
    assert k.value == -7
