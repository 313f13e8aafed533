"""Case files on the host: the unit system and the NodeType rasterizer of
the engine (libtclbx.so through tclbx_case_*), pinned by hand-derived
answers and checked bit for bit against the oracle's independent
restatement (oracle/geometry.py) on every primitive, and on the
reference's own example case files when the reference tree is present."""
import glob
import math
import os
import re

import numpy as np
import pytest

import tclb_amd
from oracle import geometry as G

HERE = os.path.dirname(os.path.abspath(__file__))
CASES = os.path.join(HERE, "golden", "cases")
REF_EXAMPLES = "/root/reference/example/flow"


def case(model, body, units="", size='nx="10" ny="6"'):
    xml = f'<CLBConfig>{units}<Geometry {size}>{body}</Geometry></CLBConfig>'
    return tclb_amd.Case(model, xml=xml, name="t.xml"), xml


def flags(model, body, units="", size='nx="10" ny="6"'):
    c, xml = case(model, body, units, size)
    a = c.rasterize()
    b, z, _ = G.read_case(xml, model, is_text=True)
    assert np.array_equal(a, b)
    assert c.zones == z
    return a


# ------------------------------------------------------------- units ---
GAUGES = ('<Units><Param name="size" value="1cm" gauge="4"/>'
          '<Param name="nu" value="1e-6m2/s" gauge="0.1"/></Units>')


def test_units_without_gauges_are_plain_numbers():
    c, _ = case("d2q9_SRT", "")
    assert c.units("1e-6") == 1e-6
    assert c.units("0.02") == 0.02
    assert c.units("-3") == -3.0
    assert c.units("10+2-0.5") == 11.5
    assert c.units("2.5e+2") == 250.0
    assert c.units("3m") == 3.0            # every base unit is 1 without gauges
    assert c.units("45d") == pytest.approx(math.pi / 4, rel=1e-15)


def test_units_gauges_closed_forms():
    """1 cm = 4 lattice lengths and nu = 1e-6 m2/s = 0.1 fix the metre
    (400) and the second (400^2/0.1 * 1e-6 = 1.6 per... s = 1/1.6e-3)."""
    c, _ = case("d2q9_SRT", "", units=GAUGES)
    m = 400.0
    s = m * m * 1e-6 / 0.1          # lattice time steps per second
    assert c.units("1m") == pytest.approx(m, rel=1e-12)
    assert c.units("1cm") == pytest.approx(4, rel=1e-12)
    assert c.units("10cm+2") == pytest.approx(42, rel=1e-12)
    assert c.units("1e-6m2/s") == pytest.approx(0.1, rel=1e-12)
    assert c.units("1s") == pytest.approx(s, rel=1e-12)
    assert c.units("2mm") == pytest.approx(0.8, rel=1e-12)     # m + m -> milli
    assert c.units("1kgm/s2") == pytest.approx(m / s ** 2, rel=1e-12)
    assert c.units("3N") == pytest.approx(3 * m / s ** 2, rel=1e-12)


@pytest.mark.parametrize("expr", ["1", "0.001sPa", "10cm+2", "-5cm", "1e-3m/s", "2.5kg/m3", "0.5Pa", "7us",
                                  "1km-3m", "1m2/s", "3e+1mm", "1.5h", "90d", "1N", "2J", "1W"])
def test_units_bit_exact_with_oracle(expr):
    units = ('<Units><Param name="nu" value="0.001sPa" gauge="0.015"/>'
             '<Param name="u" value=".5cm/s" gauge="0.1"/><Param name="rho" value="1000kg/m3" gauge="1"/></Units>')
    c, xml = case("d2q9_SRT", "", units=units)
    _, _, u = G.read_case(xml, "d2q9_SRT", is_text=True)
    assert c.units(expr) == u.alt(expr)


def test_bad_units_raise():
    c, _ = case("d2q9_SRT", "")
    for bad in ("1m//s", "1m*s"):
        with pytest.raises(tclb_amd.TclbxError):
            c.units(bad)


# -------------------------------------------------------- rasterizer ---
def test_box_regions_by_hand():
    """dx/dy offsets, nx/ny sizes, negative offsets from the end, fx/fy last
    indices (negative from the end), '<' and '>' offset prefixes, nesting
    relative to the parent element (10 x 6 lattice, d2q9_SRT Wall = 6)."""
    W = 6
    f = flags("d2q9_SRT", '<Wall><Box dx="2" nx="3" dy="1" ny="2"/></Wall>')[0]
    want = np.zeros((6, 10), np.uint16)
    want[1:3, 2:5] = W
    assert np.array_equal(f, want)
    f = flags("d2q9_SRT", '<Wall><Box dx="-1"/></Wall>')[0]           # last column
    want = np.zeros((6, 10), np.uint16); want[:, 9] = W
    assert np.array_equal(f, want)
    f = flags("d2q9_SRT", '<Wall><Box fx="1" fy="-2"/></Wall>')[0]     # x 0..1, y 0..4
    want = np.zeros((6, 10), np.uint16); want[0:5, 0:2] = W
    assert np.array_equal(f, want)
    f = flags("d2q9_SRT", '<Wall><Box dx="&lt;-3" nx="2"/></Wall>')[0]  # from the end: x 7, 8
    want = np.zeros((6, 10), np.uint16); want[:, 7:9] = W
    assert np.array_equal(f, want)
    f = flags("d2q9_SRT", '<Wall><Box dx="&gt;-2" nx="3"/></Wall>')[0]  # '>': -2 taken as is, x -2..0
    want = np.zeros((6, 10), np.uint16); want[:, 0] = W
    assert np.array_equal(f, want)
    f = flags("d2q9_SRT", '<Wall dx="4" dy="2"><Box dx="1" nx="2" ny="1"/></Wall>')[0]  # nested
    want = np.zeros((6, 10), np.uint16); want[2, 5:7] = W
    assert np.array_equal(f, want)


def test_sphere_and_wedge_counts_by_hand():
    """A Sphere in a 4x4 box covers the 12 nodes whose centres
    ((2i+1)/4 - 1/2)*2 = -3/4, -1/4, 1/4, 3/4 satisfy a^2 + b^2 < 1 (all but
    the 4 corners, 9/16 + 9/16 > 1); a Wedge on n x n keeps x <= y: n(n+1)/2."""
    f = flags("d2q9_SRT", '<Solid><Sphere dx="3" nx="4" dy="1" ny="4"/></Solid>')[0]
    assert np.count_nonzero(f) == 12 and np.count_nonzero(f[1:5, 3:7]) == 12
    assert f[1, 3] == 0 and f[1, 4] == 5
    f = flags("d2q9_SRT", '<Solid><Wedge nx="5" ny="5"/></Solid>')[0]
    assert np.count_nonzero(f) == 15
    ys, xs = np.nonzero(f)
    assert np.all(xs <= ys)
    f = flags("d2q9_SRT", '<Solid><Wedge nx="5" ny="5" direction="LowerRight"/></Solid>')[0]
    ys, xs = np.nonzero(f)
    assert len(xs) == 15 and np.all((4 - xs) <= (4 - ys))


def test_offgrid_sphere_brute_force():
    """OffgridSphere marks exactly the nodes whose centre (x+1/2, y+1/2,
    z+1/2) lies inside the ellipsoid, in the whole lattice."""
    f = flags("d3q19", '<Solid><OffgridSphere x="7.3" y="5.6" z="4.2" Rx="3.1" Ry="2.4" Rz="2.9"/></Solid>',
              size='nx="16" ny="12" nz="10"')
    z, y, x = np.meshgrid(np.arange(10), np.arange(12), np.arange(16), indexing="ij")
    inside = ((x + .5 - 7.3) / 3.1) ** 2 + ((y + .5 - 5.6) / 2.4) ** 2 + ((z + .5 - 4.2) / 2.9) ** 2 < 1
    # a node within 1e-12 of the surface could differ by rounding; there is none here
    assert np.array_equal(f == 8, inside)


def test_modes_masks_and_zones():
    """overwrite replaces the masked bits; fill only paints nodes with none
    of the mask's bits set; change only nodes with some; name= numbers zones
    in order of first use and widens the mask by the zone bits."""
    BGK, WALL, SOLID = 8, 6, 5
    base = '<BGK><Box/></BGK><Wall><Box nx="2"/></Wall><Solid mode="fill"><Box nx="4"/></Solid>'
    f = flags("d2q9_SRT", base)[0]
    assert np.all(f[:, 0:2] == BGK | WALL)      # fill skips nodes that have BOUNDARY bits
    assert np.all(f[:, 2:4] == BGK | SOLID)
    assert np.all(f[:, 4:] == BGK)
    f = flags("d2q9_SRT", base + '<WVelocity mode="change"><Box dx="1" nx="4"/></WVelocity>')[0]
    assert np.all(f[:, 0] == BGK | WALL)        # outside the box
    assert np.all(f[:, 1:4] == BGK | 3)         # change paints nodes that have BOUNDARY bits
    assert np.all(f[:, 4:] == BGK)              # x = 4 has none
    f = flags("d2q9_SRT", '<Wall mask="ALL"><Box nx="1"/></Wall><BGK><Box/></BGK>')[0]
    assert np.all(f[:, 0] == WALL | BGK) and np.all(f[:, 1:] == BGK)
    f = flags("d2q9_SRT", '<Wall name="b"><Box nx="1"/></Wall><Solid name="a"><Box dx="1" nx="1"/></Solid>'
                          '<BGK name="b"><Box dx="2" nx="1"/></BGK>')[0]
    assert np.all(f[:, 0] == WALL | (1 << 4))
    assert np.all(f[:, 1] == SOLID | (2 << 4))
    assert np.all(f[:, 2] == BGK | (1 << 4))


def test_attribute_order_matters_like_the_reference():
    """name= widens the mask by the zone bits, mask= replaces the mask: with
    mask first the zone bits are cleared and set, with name first the later
    mask= drops them from the mask, so the zone is OR-ed in."""
    base = '<Wall name="z2"><Box dx="9"/></Wall><Wall name="z1"><Box/></Wall>'
    f1 = flags("d2q9_SRT", base + '<Solid mask="BOUNDARY" name="z2"><Box nx="1"/></Solid>')[0]
    f2 = flags("d2q9_SRT", base + '<Solid name="z2" mask="BOUNDARY"><Box nx="1"/></Solid>')[0]
    assert np.all(f1[:, 0] == 5 | (1 << 4))        # zone replaced: z2 = 1
    assert np.all(f2[:, 0] == 5 | (2 << 4) | (1 << 4))  # zone OR-ed: z1 | z2


def test_predefined_zones():
    f = flags("d3q19", '<Wall><Inlet/></Wall><Solid><Outlet/></Solid><WVelocity><Channel/></WVelocity>',
              size='nx="6" ny="5" nz="4"')
    assert np.all(f[:, 1:4, 0] == 0xC) and np.all(f[:, 1:4, 5] == 0x8)
    assert np.all(f[:, 0, :] == 0x18) and np.all(f[:, 4, :] == 0x18)
    f = flags("d3q19", '<Wall><Tunnel/></Wall>', size='nx="6" ny="5" nz="4"')
    inner = f[1:3, 1:4, :]
    assert np.all(inner == 0) and np.count_nonzero(f) == f.size - inner.size


def test_text_primitive(tmp_path):
    """Text reads one integer per node of its box, z fastest, and draws the
    nonzero ones inside the parent's region."""
    vals = np.zeros((3, 2, 2), int)   # [x][y][z]
    vals[0, 0, 0] = 1; vals[2, 1, 1] = 7; vals[1, 1, 0] = -1
    p = tmp_path / "pattern.txt"
    p.write_text(" ".join(str(v) for v in vals.reshape(-1)))
    f = flags("d3q19", f'<Solid><Text file="{p}" dx="1" nx="3" ny="2" nz="2"/></Solid>', size='nx="5" ny="3" nz="2"')
    want = np.zeros((2, 3, 5), np.uint16)
    want[0, 0, 1] = want[1, 1, 3] = want[0, 1, 2] = 8
    assert np.array_equal(f, want)


@pytest.mark.parametrize("body,msg", [
    ("<Nothing><Box/></Nothing>", "Unknown flag (in xml): Nothing"),
    ('<Wall mask="SOME"><Box/></Wall>', "Unknown mask (in xml): SOME"),
    ('<Wall mode="paint"><Box/></Wall>', "Unknown mode (in xml): paint"),
    ("<Wall><Blob/></Wall>", "Unknown geometry element: Blob"),
])
def test_errors_name_the_problem(body, msg):
    c, _ = case("d2q9_SRT", body)
    with pytest.raises(tclb_amd.TclbxError, match=re.escape(msg)):
        c.rasterize()


@pytest.mark.parametrize("model,path", [
    ("d2q9_SRT", "geometry_2d.xml"), ("d2q9_SRT_drag", "geometry_2d.xml"),
    ("d3q19", "geometry_3d.xml"), ("d3q19_lid", "geometry_3d.xml")])
def test_every_primitive_matches_oracle(model, path):
    f = os.path.join(CASES, path)
    c = tclb_amd.Case(model, f)
    a = c.rasterize()
    b, z, _ = G.read_case(f, model)
    assert np.array_equal(a, b)
    assert c.zones == z
    assert len(np.unique(a)) > 12


def _examples():
    if not os.path.isdir(REF_EXAMPLES):
        return []
    return sorted(glob.glob(os.path.join(REF_EXAMPLES, "2d", "*.xml")) +
                  glob.glob(os.path.join(REF_EXAMPLES, "3d", "*.xml")))


@pytest.mark.skipif(not _examples(), reason="the reference's example case files are not present")
@pytest.mark.parametrize("model", ["d2q9_SRT", "d2q9", "d3q19"])
def test_reference_examples(model):
    """Each example case file of the reference (example/flow/2d, 3d) either
    rasterizes bit-identically in the engine and the oracle, or fails in
    both with the same message (an element the model lacks)."""
    done = 0
    for f in _examples():
        try:
            b, z, _ = G.read_case(f, model)
            err = None
        except KeyError as e:
            err = e.args[0]
            if "STL" in err or "Sweep" in err:   # surface meshes: not restated here
                continue
        except Exception:   # not a case this restatement reads (e.g. STL)
            continue
        try:
            c = tclb_amd.Case(model, f)
            if model.startswith("d2q9") and c.nz != 1:
                continue
            a = c.rasterize()
        except tclb_amd.TclbxError as e:
            if "2D model" in str(e):
                continue
            assert err is not None and err in str(e), (f, err, str(e))
            continue
        assert err is None, (f, err)
        assert np.array_equal(a, b), f
        assert c.zones == z, f
        done += 1
    assert done >= 1
