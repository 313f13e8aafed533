"""tclbx_get_quantity_noalloc, the entry point that allocates nothing (its
kernel stages into a device array the library reserves per GPU), against
tclbx_get_quantity on the same lattice: bit-identical values, on two calls in
a row (the second over a smaller region, so stale contents of the shared
array would show), for a scalar and a vector quantity, and a region beyond
the array's size refused with an error."""
import ctypes

import numpy as np
import pytest

import tclb_amd
from tclb_amd import cases

pytestmark = pytest.mark.gpu


def _noalloc(lat, name, region, scale=1.0):
    q, vec = lat.spec.quantity_id(name)
    dx, dy, dz, nx, ny, nz = region
    out = np.full((nz, ny, nx) + ((3,) if vec else ()), np.nan)
    rc = tclb_amd.lib().tclbx_get_quantity_noalloc(lat._h, q, dx, dy, dz, nx, ny, nz,
                                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), scale)
    if rc != 0:
        raise tclb_amd.TclbxError(tclb_amd.lib().tclbx_last_error().decode())
    return out


def test_noalloc_quantity_matches_get_quantity():
    print("no-alloc quantity vs get_quantity", flush=True)
    case = cases.d3q19_porous(n=48, steps=30)
    lat = case.build("f64")
    lat.iterate(30)
    for name in ("U", "P"):
        for region, scale in (((0, 0, 0, 48, 48, 48), 1.0), ((5, 7, 3, 20, 11, 9), 2.5)):
            want = lat.quantity(name, region=region, scale=scale)
            got = _noalloc(lat, name, region, scale)
            assert np.array_equal(got.reshape(want.shape), want), (name, region)
    lat.close()
    # a second lattice, model and storage: the shared array holds the last
    # call's values, which every call overwrites
    c2 = cases.d2q9_cylinder(nx=96, ny=40, diameter=10, cx=24, cy=20, steps=20)
    l2 = c2.build("f16s")
    l2.iterate(20)
    for name, region in (("U", (0, 0, 0, 96, 40, 1)), ("Rho", (10, 3, 0, 7, 5, 1))):
        want = l2.quantity(name, region=region)
        got = _noalloc(l2, name, region)
        assert np.array_equal(got.reshape(want.shape), want), (name, region)
    l2.close()


def test_noalloc_quantity_refuses_a_region_beyond_its_array():
    print("no-alloc quantity: size limit", flush=True)
    lat = tclb_amd.Lattice("d3q19", 160, 160, 120)   # 3,072,000 nodes: U needs 9,216,000 doubles
    lat.init()
    with pytest.raises(tclb_amd.TclbxError, match="staging"):
        _noalloc(lat, "U", (0, 0, 0, 160, 160, 120))
    # the scalar P of the same region fits (3,072,000 doubles)
    p = _noalloc(lat, "P", (0, 0, 0, 160, 160, 120))
    assert np.array_equal(p, lat.quantity("P"))
    lat.close()
