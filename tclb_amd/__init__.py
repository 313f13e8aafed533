"""tclb_amd -- MI355X-native lattice Boltzmann node-kernel engine.

Python mirror of the C ABI in include/tclbx.h (libtclbx.so, HIP kernels for
gfx950).  The Lattice class follows the reference's Lattice interface
(src/Lattice.h.Rt:39-191): SetSetting, FlagOverwrite, Init, Iterate,
globals, GetQuantity, Get/Set densities, save/load.

The library is loaded from this directory; there is no fallback: if it is
missing or fails to load, every call raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# TCLBX_LIB selects another build of the library (an A/B comparison of two
# builds); by default the in-tree build is loaded
LIBPATH = os.environ.get("TCLBX_LIB") or os.path.join(HERE, "libtclbx.so")

ITER_NORM = 0x001
ITER_GLOBS = 0x010
ITER_LASTGLOB = 0x080

# exported symbols declared in include/tclbx.h
SYMBOLS = [
    "tclbx_last_error", "tclbx_version", "tclbx_model_count", "tclbx_model_name",
    "tclbx_lattice_create", "tclbx_lattice_destroy", "tclbx_model_info", "tclbx_setting_default",
    "tclbx_setting_id", "tclbx_zone_setting_id", "tclbx_global_id", "tclbx_global_name",
    "tclbx_setting_name", "tclbx_zone_setting_name", "tclbx_quantity_name", "tclbx_quantity_id",
    "tclbx_nodetype", "tclbx_nodetype_group", "tclbx_set_setting", "tclbx_get_setting",
    "tclbx_set_zone_setting", "tclbx_flag_overwrite", "tclbx_get_flags", "tclbx_cuts_overwrite",
    "tclbx_get_cuts", "tclbx_init",
    "tclbx_iterate", "tclbx_get_globals", "tclbx_get_iteration", "tclbx_get_quantity",
    "tclbx_get_quantity_noalloc",
    "tclbx_get_density", "tclbx_set_density", "tclbx_extra_density_name", "tclbx_reset_average", "tclbx_save",
    "tclbx_load", "tclbx_save_layout", "tclbx_pri_bytes", "tclbx_pri_blocks", "tclbx_pri_pack", "tclbx_pri_unpack",
    "tclbx_get_timing",
    "tclbx_bytes_per_node", "tclbx_set_block", "tclbx_set_graph_steps", "tclbx_copy_bandwidth",
    "tclbx_set_rows_per_tile", "tclbx_get_rows_per_tile", "tclbx_prepare",
    "tclbx_set_sorted_dispatch", "tclbx_get_sorted_dispatch", "tclbx_set_stochastic_rounding",
    "tclbx_set_temporal_blocking", "tclbx_get_temporal_blocking",
    "tclbx_batch_set_stochastic_rounding", "tclbx_server_set_stochastic_rounding",
    "tclbx_batch_create", "tclbx_batch_destroy", "tclbx_batch_case_bytes", "tclbx_batch_add_case",
    "tclbx_batch_remove_case", "tclbx_batch_set_setting", "tclbx_batch_set_zone_setting",
    "tclbx_batch_flag_overwrite", "tclbx_batch_init_case", "tclbx_batch_iterate", "tclbx_batch_get_globals",
    "tclbx_batch_get_iteration", "tclbx_batch_get_quantity", "tclbx_batch_get_density",
    "tclbx_batch_set_density", "tclbx_batch_autotune", "tclbx_batch_set_tuning", "tclbx_batch_get_timing",
    "tclbx_batch_info", "tclbx_batch_tile_order", "tclbx_batch_set_graph_steps", "tclbx_batch_prepare",
    "tclbx_server_create", "tclbx_server_destroy", "tclbx_server_submit", "tclbx_server_poll",
    "tclbx_server_wait_all", "tclbx_server_stats", "tclbx_server_set_tuning",
    "tclbx_case_load", "tclbx_case_load_string", "tclbx_case_destroy", "tclbx_case_size", "tclbx_case_units",
    "tclbx_case_rasterize", "tclbx_case_cuts", "tclbx_case_zone_count", "tclbx_case_zone_name", "tclbx_case_set_output_root",
    "tclbx_case_run", "tclbx_case_check", "tclbx_case_lattice", "tclbx_case_iteration", "tclbx_case_message_count",
    "tclbx_case_message",
]


class TclbxError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libtclbx.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise TclbxError(f"{LIBPATH} is not built; run __graft_entry__.build() or make -C tclb_amd/csrc")
    L = ctypes.CDLL(LIBPATH)
    c_int, c_double, c_char_p, vp = ctypes.c_int, ctypes.c_double, ctypes.c_char_p, ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    u16p = ctypes.POINTER(ctypes.c_uint16)
    ip = ctypes.POINTER(ctypes.c_int)
    L.tclbx_last_error.restype = c_char_p
    L.tclbx_model_name.restype = c_char_p
    L.tclbx_model_name.argtypes = [c_int]
    L.tclbx_lattice_create.argtypes = [c_char_p, c_int, c_int, c_int, c_char_p, ctypes.POINTER(vp)]
    L.tclbx_lattice_destroy.argtypes = [vp]
    L.tclbx_model_info.argtypes = [c_char_p, ip, ip, ip, ip, ip]
    for n in ["tclbx_setting_id", "tclbx_zone_setting_id", "tclbx_global_id"]:
        getattr(L, n).argtypes = [c_char_p, c_char_p]
    for n in ["tclbx_global_name", "tclbx_setting_name", "tclbx_zone_setting_name", "tclbx_quantity_name"]:
        getattr(L, n).argtypes = [c_char_p, c_int]
        getattr(L, n).restype = c_char_p
    L.tclbx_setting_default.argtypes = [c_char_p, c_int, dp]
    L.tclbx_quantity_id.argtypes = [c_char_p, c_char_p, ip]
    L.tclbx_nodetype.argtypes = [c_char_p, c_char_p, u16p, u16p]
    L.tclbx_nodetype_group.argtypes = [c_char_p, c_char_p, u16p, ip]
    L.tclbx_set_setting.argtypes = [vp, c_int, c_double]
    L.tclbx_get_setting.argtypes = [vp, c_int, dp]
    L.tclbx_set_zone_setting.argtypes = [vp, c_int, c_int, c_double]
    L.tclbx_flag_overwrite.argtypes = [vp, u16p] + [c_int] * 6
    L.tclbx_get_flags.argtypes = [vp] + [c_int] * 6 + [u16p]
    L.tclbx_cuts_overwrite.argtypes = [vp, u16p] + [c_int] * 6
    L.tclbx_get_cuts.argtypes = [vp] + [c_int] * 6 + [u16p]
    L.tclbx_init.argtypes = [vp]
    L.tclbx_iterate.argtypes = [vp, c_int, c_int]
    L.tclbx_get_globals.argtypes = [vp, dp, c_int]
    L.tclbx_get_iteration.argtypes = [vp, ip]
    L.tclbx_get_quantity.argtypes = [vp, c_int] + [c_int] * 6 + [dp, c_double]
    L.tclbx_get_quantity_noalloc.argtypes = [vp, c_int] + [c_int] * 6 + [dp, c_double]
    L.tclbx_get_density.argtypes = [vp, c_int, dp]
    L.tclbx_reset_average.argtypes = [vp]
    L.tclbx_extra_density_name.argtypes = [c_char_p, c_int, ip]
    L.tclbx_extra_density_name.restype = c_char_p
    L.tclbx_set_density.argtypes = [vp, c_int, dp]
    L.tclbx_save.argtypes = [vp, c_char_p]
    L.tclbx_load.argtypes = [vp, c_char_p]
    L.tclbx_save_layout.argtypes = [vp, c_char_p, c_int]
    szp = ctypes.POINTER(ctypes.c_size_t)
    L.tclbx_pri_bytes.argtypes = [c_char_p, c_char_p, c_int, c_int, c_int, szp]
    L.tclbx_pri_blocks.argtypes = [c_char_p, c_int, c_int, c_int, ip, szp, c_int, ip]
    L.tclbx_pri_pack.argtypes = [c_char_p, c_char_p, c_int, c_int, c_int, vp, dp, vp]
    L.tclbx_pri_unpack.argtypes = [c_char_p, c_char_p, c_int, c_int, c_int, vp, vp, dp]
    L.tclbx_get_timing.argtypes = [vp, dp, dp]
    L.tclbx_bytes_per_node.argtypes = [vp, dp]
    L.tclbx_set_block.argtypes = [vp, c_int]
    L.tclbx_set_graph_steps.argtypes = [vp, c_int]
    L.tclbx_set_rows_per_tile.argtypes = [vp, c_int]
    L.tclbx_get_rows_per_tile.argtypes = [vp, ip]
    L.tclbx_prepare.argtypes = [vp]
    L.tclbx_set_sorted_dispatch.argtypes = [vp, c_int]
    L.tclbx_get_sorted_dispatch.argtypes = [vp, ip, ctypes.POINTER(ctypes.c_longlong)]
    L.tclbx_set_stochastic_rounding.argtypes = [vp, c_int, ctypes.c_uint]
    L.tclbx_set_temporal_blocking.argtypes = [vp, c_int]
    L.tclbx_get_temporal_blocking.argtypes = [vp, ip]
    L.tclbx_batch_set_stochastic_rounding.argtypes = [vp, c_int, ctypes.c_uint]
    L.tclbx_server_set_stochastic_rounding.argtypes = [vp, c_int, ctypes.c_uint]
    L.tclbx_copy_bandwidth.argtypes = [ctypes.c_longlong, c_int, dp, dp]
    ll, llp = ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong)
    L.tclbx_batch_create.argtypes = [c_char_p, c_char_p, ll, ctypes.POINTER(vp)]
    L.tclbx_batch_destroy.argtypes = [vp]
    L.tclbx_batch_case_bytes.argtypes = [vp, c_int, c_int, c_int, llp]
    L.tclbx_batch_add_case.argtypes = [vp, c_int, c_int, c_int, ip]
    L.tclbx_batch_remove_case.argtypes = [vp, c_int]
    L.tclbx_batch_set_setting.argtypes = [vp, c_int, c_int, c_double]
    L.tclbx_batch_set_zone_setting.argtypes = [vp, c_int, c_int, c_int, c_double]
    L.tclbx_batch_flag_overwrite.argtypes = [vp, c_int, u16p]
    L.tclbx_batch_init_case.argtypes = [vp, c_int]
    L.tclbx_batch_iterate.argtypes = [vp, c_int, c_int]
    L.tclbx_batch_get_globals.argtypes = [vp, c_int, dp, c_int]
    L.tclbx_batch_get_iteration.argtypes = [vp, c_int, llp]
    L.tclbx_batch_get_quantity.argtypes = [vp, c_int, c_int, dp, c_double]
    L.tclbx_batch_get_density.argtypes = [vp, c_int, c_int, dp]
    L.tclbx_batch_set_density.argtypes = [vp, c_int, c_int, dp]
    L.tclbx_batch_autotune.argtypes = [vp, c_int, ip, ip, ip, ip, dp]
    L.tclbx_batch_set_tuning.argtypes = [vp, c_int, c_int, c_int, c_int]
    L.tclbx_batch_get_timing.argtypes = [vp, dp, dp]
    L.tclbx_batch_info.argtypes = [vp, ip, llp, llp, llp]
    L.tclbx_batch_tile_order.argtypes = [vp, c_int, ip]
    L.tclbx_batch_set_graph_steps.argtypes = [vp, c_int]
    L.tclbx_batch_prepare.argtypes = [vp]
    L.tclbx_server_create.argtypes = [c_char_p, c_char_p, ll, c_int, ctypes.POINTER(vp)]
    L.tclbx_server_destroy.argtypes = [vp]
    L.tclbx_server_submit.argtypes = [vp, c_int, c_int, c_int, u16p, c_int, ip, dp, c_int, ip, ip, dp, ll,
                                      c_int, ip, ip, dp, c_int, c_int, ip]
    L.tclbx_server_poll.argtypes = [vp, c_int, ip, llp, dp, c_int, dp, dp, dp]
    L.tclbx_server_wait_all.argtypes = [vp, c_double, ip]
    L.tclbx_server_stats.argtypes = [vp, llp, llp, llp, dp, ip, ip, dp]
    L.tclbx_server_set_tuning.argtypes = [vp, c_int, c_int, c_int, c_int]
    L.tclbx_case_load.argtypes = [c_char_p, c_char_p, c_char_p, ctypes.POINTER(vp)]
    L.tclbx_case_load_string.argtypes = [c_char_p, c_char_p, c_char_p, c_char_p, ctypes.POINTER(vp)]
    L.tclbx_case_destroy.argtypes = [vp]
    L.tclbx_case_size.argtypes = [vp, ip, ip, ip]
    L.tclbx_case_units.argtypes = [vp, c_char_p, dp]
    L.tclbx_case_rasterize.argtypes = [vp, u16p, ll]
    L.tclbx_case_cuts.argtypes = [vp, u16p, ll, ip]
    L.tclbx_case_zone_count.argtypes = [vp]
    L.tclbx_case_zone_name.argtypes = [vp, c_int]
    L.tclbx_case_zone_name.restype = c_char_p
    L.tclbx_case_set_output_root.argtypes = [vp, c_char_p]
    L.tclbx_case_run.argtypes = [vp]
    L.tclbx_case_check.argtypes = [vp]
    L.tclbx_case_lattice.argtypes = [vp, ctypes.POINTER(vp)]
    L.tclbx_case_iteration.argtypes = [vp]
    L.tclbx_case_message_count.argtypes = [vp]
    L.tclbx_case_message.argtypes = [vp, c_int]
    L.tclbx_case_message.restype = c_char_p
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        raise TclbxError(lib().tclbx_last_error().decode())
    return rc


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """(GB/s counting read + write, ms per copy) of a plain device copy."""
    g, t = ctypes.c_double(), ctypes.c_double()
    _check(lib().tclbx_copy_bandwidth(int(nbytes), int(reps), ctypes.byref(g), ctypes.byref(t)))
    return g.value, t.value


STORAGE_DTYPE = {"f64": np.float64, "f32": np.float32, "f32s": np.float32, "f16s": np.float16, "bf16s": np.uint16}


def pri_bytes(model, storage, nx, ny, nz=1):
    """Size of the reference's <file>_0.pri for this lattice (host only)."""
    n = ctypes.c_size_t()
    _check(lib().tclbx_pri_bytes(model.encode(), storage.encode(), nx, ny, nz, ctypes.byref(n)))
    return n.value


def pri_blocks(model, nx, ny, nz=1):
    """[(block number 1..27, values)] of the reference's snapshot, in file order."""
    idx = (ctypes.c_int * 27)()
    val = (ctypes.c_size_t * 27)()
    cnt = ctypes.c_int()
    _check(lib().tclbx_pri_blocks(model.encode(), nx, ny, nz, idx, val, 27, ctypes.byref(cnt)))
    return [(idx[i], val[i]) for i in range(cnt.value)]


def pri_pack(model, storage, planes, extra=None):
    """The reference's file bytes of a snapshot given as stored values
    [Q, nz, ny, nx] (the storage format's dtype) and extra densities
    [nextra, nz, ny, nx] (float64)."""
    planes = np.ascontiguousarray(planes, dtype=STORAGE_DTYPE[storage])
    nz, ny, nx = planes.shape[1:]
    out = np.zeros(pri_bytes(model, storage, nx, ny, nz), dtype=np.uint8)
    ex = None if extra is None else np.ascontiguousarray(extra, dtype=np.float64)
    _check(lib().tclbx_pri_pack(model.encode(), storage.encode(), nx, ny, nz, planes.ctypes.data,
                                None if ex is None else _dp(ex), out.ctypes.data))
    return out


def pri_unpack(model, storage, nx, ny, nz, data, q, nextra=0):
    """(planes [Q, nz, ny, nx] stored values, extra [nextra, nz, ny, nx])
    of the reference's file bytes."""
    data = np.ascontiguousarray(np.frombuffer(bytes(data), dtype=np.uint8) if not isinstance(data, np.ndarray)
                                else data.view(np.uint8))
    planes = np.empty((q, nz, ny, nx), dtype=STORAGE_DTYPE[storage])
    extra = np.empty((nextra, nz, ny, nx), dtype=np.float64)
    _check(lib().tclbx_pri_unpack(model.encode(), storage.encode(), nx, ny, nz, data.ctypes.data, planes.ctypes.data,
                                  _dp(extra) if nextra else None))
    return planes, extra


def models():
    L = lib()
    return [L.tclbx_model_name(i).decode() for i in range(L.tclbx_model_count())]


class Model:
    """The model contract (Model_m, src/Lists.h:13-165) of one compiled-in
    model: settings, zonal settings, Globals, Quantities, NodeType flags.
    Needs no GPU."""

    def __init__(self, name):
        L = lib()
        self.name = name
        b = name.encode()
        vals = [ctypes.c_int() for _ in range(5)]
        _check(L.tclbx_model_info(b, *[ctypes.byref(v) for v in vals]))
        self.Q, self.nglobals, self.nsettings, self.nzonal, self.nquantities = [v.value for v in vals]
        self.settings = [L.tclbx_setting_name(b, i).decode() for i in range(self.nsettings)]
        self.zonal_settings = [L.tclbx_zone_setting_name(b, i).decode() for i in range(self.nzonal)]
        self.global_names = [L.tclbx_global_name(b, i).decode() for i in range(self.nglobals)]
        self.quantities = [L.tclbx_quantity_name(b, i).decode() for i in range(self.nquantities)]
        # densities that are not streamed (density Q + k): averages, then parameters
        self.extras, self.averages, self.params = [], [], []
        isavg = ctypes.c_int()
        while True:
            nm = L.tclbx_extra_density_name(b, len(self.extras), ctypes.byref(isavg))
            if not nm:
                break
            self.extras.append(nm.decode())
            (self.averages if isavg.value else self.params).append(nm.decode())

    def setting_id(self, name):
        return lib().tclbx_setting_id(self.name.encode(), name.encode())

    def zone_setting_id(self, name):
        return lib().tclbx_zone_setting_id(self.name.encode(), name.encode())

    def global_id(self, name):
        return lib().tclbx_global_id(self.name.encode(), name.encode())

    def quantity_id(self, name):
        isv = ctypes.c_int()
        q = lib().tclbx_quantity_id(self.name.encode(), name.encode(), ctypes.byref(isv))
        return q, bool(isv.value)

    def setting_default(self, name):
        v = ctypes.c_double()
        _check(lib().tclbx_setting_default(self.name.encode(), self.setting_id(name), ctypes.byref(v)))
        return v.value

    def nodetype(self, name):
        """(value, group mask) of NODE_<name>."""
        v, m = ctypes.c_uint16(), ctypes.c_uint16()
        _check(lib().tclbx_nodetype(self.name.encode(), name.encode(), ctypes.byref(v), ctypes.byref(m)))
        return v.value, m.value

    def group(self, name):
        """(mask, shift) of NodeType group `name`."""
        m, s = ctypes.c_uint16(), ctypes.c_int()
        _check(lib().tclbx_nodetype_group(self.name.encode(), name.encode(), ctypes.byref(m), ctypes.byref(s)))
        return m.value, s.value

    def flag(self, *names, zone=0):
        """NodeType value of the OR of `names` plus a zone number."""
        v = 0
        for n in names:
            v |= self.nodetype(n)[0]
        return v | (zone << self.group("SETTINGZONE")[1])


class Lattice:
    """One lattice on the GPU (one subdomain, periodic unless NodeTypes say
    otherwise), mirroring the reference Lattice (src/Lattice.h.Rt:39-191)."""

    def __init__(self, model, nx, ny, nz=1, storage="f64", _handle=None):
        L = lib()
        if _handle is None:
            h = ctypes.c_void_p()
            _check(L.tclbx_lattice_create(model.encode(), nx, ny, nz, storage.encode(), ctypes.byref(h)))
            self._owned = True
        else:
            h = _handle  # a case's lattice: the case frees it
            self._owned = False
        self._h = h
        self.model, self.storage = model, storage
        self.nx, self.ny, self.nz = nx, ny, nz
        m = Model(model)
        self.spec = m
        self.Q, self.nglobals = m.Q, m.nglobals
        self.settings, self.zonal_settings = m.settings, m.zonal_settings
        self.global_names, self.quantities = m.global_names, m.quantities

    def close(self):
        if getattr(self, "_h", None):
            if self._owned:
                lib().tclbx_lattice_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    # --- model registry ---------------------------------------------------
    def nodetype(self, name):
        return self.spec.nodetype(name)

    def group(self, name):
        return self.spec.group(name)

    # --- settings ---------------------------------------------------------
    def set_setting(self, name, value):
        i = self.spec.setting_id(name)
        if i < 0:
            raise TclbxError(f"unknown setting {name}")
        _check(lib().tclbx_set_setting(self._h, i, float(value)))

    def get_setting(self, name):
        i = self.spec.setting_id(name)
        if i < 0:
            raise TclbxError(f"unknown setting {name}")
        v = ctypes.c_double()
        _check(lib().tclbx_get_setting(self._h, i, ctypes.byref(v)))
        return v.value

    def set_zone_setting(self, name, value, zone=-1):
        i = self.spec.zone_setting_id(name)
        if i < 0:
            raise TclbxError(f"unknown zonal setting {name}")
        _check(lib().tclbx_set_zone_setting(self._h, i, int(zone), float(value)))

    # --- flags ------------------------------------------------------------
    def set_flags(self, flags, region=None):
        f = np.ascontiguousarray(flags, dtype=np.uint16)
        if region is None:
            region = (0, 0, 0, self.nx, self.ny, self.nz)
        dx, dy, dz, nx, ny, nz = region
        assert f.size == nx * ny * nz
        _check(lib().tclbx_flag_overwrite(self._h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                          dx, dy, dz, nx, ny, nz))

    def set_cuts(self, cuts, region=None):
        """Lattice::CutsOverwrite: uint16 [26, nz, ny, nx] cut distances of
        the 26 links (direction - 1 in the d3q27 order), NO_CUT = 65535."""
        c = np.ascontiguousarray(cuts, dtype=np.uint16)
        if region is None:
            region = (0, 0, 0, self.nx, self.ny, self.nz)
        dx, dy, dz, nx, ny, nz = region
        assert c.size == 26 * nx * ny * nz
        _check(lib().tclbx_cuts_overwrite(self._h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
                                          dx, dy, dz, nx, ny, nz))

    def get_cuts(self, region=None):
        if region is None:
            region = (0, 0, 0, self.nx, self.ny, self.nz)
        dx, dy, dz, nx, ny, nz = region
        out = np.zeros((26, nz, ny, nx), dtype=np.uint16)
        _check(lib().tclbx_get_cuts(self._h, dx, dy, dz, nx, ny, nz,
                                    out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))))
        return out

    def get_flags(self, region=None):
        if region is None:
            region = (0, 0, 0, self.nx, self.ny, self.nz)
        dx, dy, dz, nx, ny, nz = region
        out = np.empty((nz, ny, nx), dtype=np.uint16)
        _check(lib().tclbx_get_flags(self._h, dx, dy, dz, nx, ny, nz,
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))))
        return out

    # --- run --------------------------------------------------------------
    def init(self):
        _check(lib().tclbx_init(self._h))

    def iterate(self, n, iter_type=ITER_NORM | ITER_LASTGLOB):
        _check(lib().tclbx_iterate(self._h, int(n), int(iter_type)))

    @property
    def globals(self):
        out = np.zeros(self.nglobals)
        _check(lib().tclbx_get_globals(self._h, _dp(out), self.nglobals))
        return out

    def global_dict(self):
        return dict(zip(self.global_names, self.globals))

    @property
    def iteration(self):
        v = ctypes.c_int()
        _check(lib().tclbx_get_iteration(self._h, ctypes.byref(v)))
        return v.value

    def quantity(self, name, region=None, scale=1.0):
        q, isv = self.spec.quantity_id(name)
        if q < 0:
            raise TclbxError(f"unknown quantity {name}")
        if region is None:
            region = (0, 0, 0, self.nx, self.ny, self.nz)
        dx, dy, dz, nx, ny, nz = region
        shape = (nz, ny, nx) + ((3,) if isv else ())
        out = np.empty(shape, dtype=np.float64)
        _check(lib().tclbx_get_quantity(self._h, q, dx, dy, dz, nx, ny, nz, _dp(out), float(scale)))
        return out

    def get_f(self):
        out = np.empty((self.Q,) + self.shape, dtype=np.float64)
        for q in range(self.Q):
            _check(lib().tclbx_get_density(self._h, q, _dp(out[q])))
        return out

    def set_f(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape((self.Q,) + self.shape)
        for q in range(self.Q):
            plane = np.ascontiguousarray(a[q])
            _check(lib().tclbx_set_density(self._h, q, _dp(plane)))

    def get_extra(self, name):
        """A density that is not streamed (an average, BC[0], ...), [nz, ny, nx]."""
        a = np.empty(self.shape, dtype=np.float64)
        _check(lib().tclbx_get_density(self._h, self.Q + self.spec.extras.index(name), _dp(a)))
        return a

    def set_extra(self, name, values):
        a = np.ascontiguousarray(np.broadcast_to(values, self.shape), dtype=np.float64)
        _check(lib().tclbx_set_density(self._h, self.Q + self.spec.extras.index(name), _dp(a)))

    def get_averages(self):
        """The averaged densities (AVG builds), dict name -> [nz, ny, nx]."""
        return {n: self.get_extra(n) for n in self.spec.averages}

    def reset_average(self):
        """Lattice::resetAverage: the averages to zero, counted from now."""
        _check(lib().tclbx_reset_average(self._h))

    def save(self, path, layout="engine"):
        """layout "engine" (density planes + footer) or "reference" (the
        reference's margin blocks, tclbx_save_layout)."""
        if layout == "engine":
            _check(lib().tclbx_save(self._h, str(path).encode()))
        else:
            _check(lib().tclbx_save_layout(self._h, str(path).encode(), {"reference": 1}[layout]))

    def load(self, path):
        _check(lib().tclbx_load(self._h, str(path).encode()))

    def timing(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().tclbx_get_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def bytes_per_node(self):
        v = ctypes.c_double()
        _check(lib().tclbx_bytes_per_node(self._h, ctypes.byref(v)))
        return v.value

    def set_block(self, block):
        _check(lib().tclbx_set_block(self._h, int(block)))

    def set_graph_steps(self, steps):
        _check(lib().tclbx_set_graph_steps(self._h, int(steps)))

    def prepare(self):
        """Time the launch shape and capture the step loop now (no state
        change), so that the next iterate() runs steps only."""
        _check(lib().tclbx_prepare(self._h))

    @property
    def sorted_dispatch(self):
        """NodeType-sorted dispatch of the plain steps (bulk collision pass,
        then a pass over the other nodes sorted by NodeType class)."""
        on = ctypes.c_int()
        _check(lib().tclbx_get_sorted_dispatch(self._h, ctypes.byref(on), None))
        return bool(on.value)

    @sorted_dispatch.setter
    def sorted_dispatch(self, on):
        _check(lib().tclbx_set_sorted_dispatch(self._h, int(bool(on))))

    def set_stochastic_rounding(self, on=True, seed=0):
        """Stochastic rounding of the stores of the 16-bit formats."""
        _check(lib().tclbx_set_stochastic_rounding(self._h, int(bool(on)), int(seed) & 0xFFFFFFFF))

    @property
    def temporal_blocking(self):
        """Plain steps per launch of the temporally blocked 2D step in use (0:
        single steps).  Set k >= 2 to fix it, 0 to turn it off, -1 to let
        the lattice time 1, 3 and 5 steps per launch on the first 270 plain
        steps of its first Iterate that runs that many and keep the fastest
        (the default of d2q9_SRT, d2q9_SRT_drag, d2q9 and d2q9_drag)."""
        k = ctypes.c_int()
        _check(lib().tclbx_get_temporal_blocking(self._h, ctypes.byref(k)))
        return k.value

    @temporal_blocking.setter
    def temporal_blocking(self, steps):
        _check(lib().tclbx_set_temporal_blocking(self._h, int(steps)))

    def sorted_list_nodes(self):
        n = ctypes.c_longlong()
        _check(lib().tclbx_get_sorted_dispatch(self._h, None, ctypes.byref(n)))
        return n.value

    @property
    def rows_per_tile(self):
        """Rows walked by one block of a plain step (tuned on the first long
        iterate unless set)."""
        v = ctypes.c_int()
        _check(lib().tclbx_get_rows_per_tile(self._h, ctypes.byref(v)))
        return v.value

    @rows_per_tile.setter
    def rows_per_tile(self, rows):
        _check(lib().tclbx_set_rows_per_tile(self._h, int(rows)))


class Batch:
    """Many independent lattices ("cases") of one model and storage format in
    one ragged device arena, advanced together by one persistent kernel per
    step.  Each per-case method mirrors the Lattice method of the same name;
    a case computes exactly what the same case computes alone."""

    def __init__(self, model, storage="f64", capacity_bytes=1 << 30):
        h = ctypes.c_void_p()
        _check(lib().tclbx_batch_create(model.encode(), storage.encode(), int(capacity_bytes), ctypes.byref(h)))
        self._h = h
        self.model, self.storage = model, storage
        self.spec = Model(model)
        self.Q, self.nglobals = self.spec.Q, self.spec.nglobals
        self.global_names = self.spec.global_names
        self.shapes = {}

    def close(self):
        if getattr(self, "_h", None):
            lib().tclbx_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def case_bytes(self, nx, ny, nz=1):
        v = ctypes.c_longlong()
        _check(lib().tclbx_batch_case_bytes(self._h, nx, ny, nz, ctypes.byref(v)))
        return v.value

    def add_case(self, nx, ny, nz=1):
        """New case id, or None when the arena has no room for it."""
        i = ctypes.c_int()
        _check(lib().tclbx_batch_add_case(self._h, nx, ny, nz, ctypes.byref(i)))
        if i.value < 0:
            return None
        self.shapes[i.value] = (nz, ny, nx)
        return i.value

    def remove_case(self, cid):
        _check(lib().tclbx_batch_remove_case(self._h, cid))
        self.shapes.pop(cid, None)

    def set_setting(self, cid, name, value):
        i = self.spec.setting_id(name)
        if i < 0:
            raise TclbxError(f"unknown setting {name}")
        _check(lib().tclbx_batch_set_setting(self._h, cid, i, float(value)))

    def set_zone_setting(self, cid, name, value, zone=-1):
        i = self.spec.zone_setting_id(name)
        if i < 0:
            raise TclbxError(f"unknown zonal setting {name}")
        _check(lib().tclbx_batch_set_zone_setting(self._h, cid, i, int(zone), float(value)))

    def set_flags(self, cid, flags):
        f = np.ascontiguousarray(flags, dtype=np.uint16)
        assert f.size == int(np.prod(self.shapes[cid]))
        _check(lib().tclbx_batch_flag_overwrite(self._h, cid, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))))

    def init_case(self, cid):
        _check(lib().tclbx_batch_init_case(self._h, cid))

    def iterate(self, n, iter_type=ITER_NORM | ITER_LASTGLOB):
        _check(lib().tclbx_batch_iterate(self._h, int(n), int(iter_type)))

    def globals(self, cid):
        out = np.zeros(self.nglobals)
        _check(lib().tclbx_batch_get_globals(self._h, cid, _dp(out), self.nglobals))
        return out

    def iteration(self, cid):
        v = ctypes.c_longlong()
        _check(lib().tclbx_batch_get_iteration(self._h, cid, ctypes.byref(v)))
        return v.value

    def quantity(self, cid, name, scale=1.0):
        q, isv = self.spec.quantity_id(name)
        if q < 0:
            raise TclbxError(f"unknown quantity {name}")
        out = np.empty(self.shapes[cid] + ((3,) if isv else ()), dtype=np.float64)
        _check(lib().tclbx_batch_get_quantity(self._h, cid, q, _dp(out), float(scale)))
        return out

    def get_f(self, cid):
        out = np.empty((self.Q,) + self.shapes[cid], dtype=np.float64)
        for q in range(self.Q):
            _check(lib().tclbx_batch_get_density(self._h, cid, q, _dp(out[q])))
        return out

    def set_f(self, cid, f):
        a = np.ascontiguousarray(f, dtype=np.float64).reshape((self.Q,) + self.shapes[cid])
        for q in range(self.Q):
            plane = np.ascontiguousarray(a[q])
            _check(lib().tclbx_batch_set_density(self._h, cid, q, _dp(plane)))

    def add(self, case):
        """Admit a tclb_amd.cases.Case: flags, settings, zonal settings,
        Init, optional initial state.  Returns the id or None (no room)."""
        if case.model != self.model:
            raise TclbxError(f"case model {case.model} != batch model {self.model}")
        cid = self.add_case(case.nx, case.ny, case.nz)
        if cid is None:
            return None
        self.set_flags(cid, case.flags)
        for n, v in case.settings:
            self.set_setting(cid, n, v)
        for n, z, v in case.zonal:
            self.set_zone_setting(cid, n, v, z)
        self.init_case(cid)
        if case.init_f is not None:
            self.set_f(cid, case.init_f)
        return cid

    def autotune(self, steps=10):
        fl, b, r, g, us = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        _check(lib().tclbx_batch_autotune(self._h, int(steps), ctypes.byref(fl), ctypes.byref(b), ctypes.byref(r),
                                          ctypes.byref(g), ctypes.byref(us)))
        return {"flat": fl.value, "block": b.value, "rows_per_tile": r.value, "grid": g.value,
                "order": self.tile_order(), "us_per_step": us.value}

    def set_tuning(self, block=256, rows_per_tile=1, grid=0, flat=False, order=0):
        _check(lib().tclbx_batch_set_tuning(self._h, int(bool(flat)), int(block), int(rows_per_tile), int(grid)))
        self.tile_order(order)

    def set_stochastic_rounding(self, on=True, seed=0):
        """Stochastic rounding of every case's stores (16-bit formats);
        a case is keyed as a lone Lattice with the same seed."""
        _check(lib().tclbx_batch_set_stochastic_rounding(self._h, int(bool(on)), int(seed) & 0xFFFFFFFF))

    def tile_order(self, set=None):
        """The tile order of the step (0: list order, 1: XCD-contiguous);
        `set` changes it."""
        o = ctypes.c_int()
        _check(lib().tclbx_batch_tile_order(self._h, -1 if set is None else int(set), ctypes.byref(o)))
        return o.value

    def timing(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(lib().tclbx_batch_get_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def info(self):
        n, nodes, used, cap = ctypes.c_int(), ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        _check(lib().tclbx_batch_info(self._h, ctypes.byref(n), ctypes.byref(nodes), ctypes.byref(used),
                                      ctypes.byref(cap)))
        return {"cases": n.value, "nodes": nodes.value, "arena_used": used.value, "arena_capacity": cap.value}

    def set_graph_steps(self, steps):
        _check(lib().tclbx_batch_set_graph_steps(self._h, int(steps)))

    def prepare(self):
        """Build the device tables and capture the step loop now."""
        _check(lib().tclbx_batch_prepare(self._h))


STOP_TYPES = {"Change": 0, "PercentChange": 1, "Above": 2, "Below": 3}
REQUEST_STATES = {-1: "rejected", 0: "queued", 1: "running", 2: "done_budget", 3: "done_stop"}


def _iarr(v):
    return (ctypes.c_int * max(1, len(v)))(*[int(x) for x in v])


def _darr(v):
    return (ctypes.c_double * max(1, len(v)))(*[float(x) for x in v])


class Server:
    """Continuous-batching solve server (native worker thread over one
    Batch): submit() from any thread; requests are admitted while the HBM
    budget has room, run together, and retire on their step budget or stop
    rule (the reference's <Stop>: [(Global, "Change"|"PercentChange"|
    "Above"|"Below", limit)] checked every `interval` steps, `times`
    consecutive passes)."""

    def __init__(self, model, storage="f16s", hbm_budget=1 << 30, max_chunk=500):
        h = ctypes.c_void_p()
        _check(lib().tclbx_server_create(model.encode(), storage.encode(), int(hbm_budget), int(max_chunk),
                                         ctypes.byref(h)))
        self._h = h
        self.model, self.storage = model, storage
        self.spec = Model(model)

    def close(self):
        if getattr(self, "_h", None):
            lib().tclbx_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def submit(self, case, max_steps, stop=(), interval=500, times=1):
        f = np.ascontiguousarray(case.flags, dtype=np.uint16).reshape(-1)
        sids = [self.spec.setting_id(n) for n, _ in case.settings]
        zids = [self.spec.zone_setting_id(n) for n, _, _ in case.zonal]
        if min(sids + zids + [0]) < 0:
            raise TclbxError("unknown setting in the case")
        gids = [self.spec.global_id(g) for g, _, _ in stop]
        if min(gids + [0]) < 0:
            raise TclbxError("unknown Global in the stop rule")
        rid = ctypes.c_int()
        _check(lib().tclbx_server_submit(
            self._h, case.nx, case.ny, case.nz, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
            len(sids), _iarr(sids), _darr([v for _, v in case.settings]),
            len(zids), _iarr(zids), _iarr([z for _, z, _ in case.zonal]), _darr([v for _, _, v in case.zonal]),
            int(max_steps), len(gids), _iarr(gids), _iarr([STOP_TYPES[t] for _, t, _ in stop]),
            _darr([lim for _, _, lim in stop]), int(interval), int(times), ctypes.byref(rid)))
        return rid.value

    def poll(self, rid):
        st, steps = ctypes.c_int(), ctypes.c_longlong()
        g = np.zeros(self.spec.nglobals)
        ts, ta, tf = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(lib().tclbx_server_poll(self._h, rid, ctypes.byref(st), ctypes.byref(steps), _dp(g), len(g),
                                       ctypes.byref(ts), ctypes.byref(ta), ctypes.byref(tf)))
        return {"state": REQUEST_STATES[st.value], "steps": steps.value, "globals": g,
                "t_submit": ts.value, "t_admit": ta.value, "t_finish": tf.value}

    def wait_all(self, timeout=3600.0):
        idle = ctypes.c_int()
        _check(lib().tclbx_server_wait_all(self._h, float(timeout), ctypes.byref(idle)))
        return bool(idle.value)

    def stats(self):
        c, r, ns = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        busy, now = ctypes.c_double(), ctypes.c_double()
        a, q = ctypes.c_int(), ctypes.c_int()
        _check(lib().tclbx_server_stats(self._h, ctypes.byref(c), ctypes.byref(r), ctypes.byref(ns),
                                        ctypes.byref(busy), ctypes.byref(a), ctypes.byref(q), ctypes.byref(now)))
        return {"completed": c.value, "rejected": r.value, "node_steps": ns.value, "busy_s": busy.value,
                "active": a.value, "queued": q.value, "now": now.value}

    def set_tuning(self, block=256, rows_per_tile=1, grid=0, flat=True):
        _check(lib().tclbx_server_set_tuning(self._h, int(bool(flat)), int(block), int(rows_per_tile), int(grid)))

    def set_stochastic_rounding(self, on=True, seed=0):
        """Stochastic rounding of the served requests' stores (16-bit formats)."""
        _check(lib().tclbx_server_set_stochastic_rounding(self._h, int(bool(on)), int(seed) & 0xFFFFFFFF))


class Case:
    """A TCLB case file run on the engine, as CLB/<model>/main <case.xml>
    runs it (src/main.cpp:171-420): Units, the Geometry size, then every
    element of <CLBConfig> as a handler.  `path` is a case file, or pass
    `xml=` text (with `name` standing for the file name, which names the
    outputs)."""

    def __init__(self, model, path=None, storage="f64", xml=None, name="case.xml", output_root=None):
        L = lib()
        h = ctypes.c_void_p()
        if xml is not None:
            _check(L.tclbx_case_load_string(model.encode(), storage.encode(), xml.encode(), name.encode(),
                                            ctypes.byref(h)))
        else:
            _check(L.tclbx_case_load(model.encode(), storage.encode(), os.fspath(path).encode(), ctypes.byref(h)))
        self._h = h
        self.model, self.storage = model, storage
        if output_root is not None:
            _check(L.tclbx_case_set_output_root(h, os.fspath(output_root).encode()))
        nx, ny, nz = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(L.tclbx_case_size(h, ctypes.byref(nx), ctypes.byref(ny), ctypes.byref(nz)))
        self.nx, self.ny, self.nz = nx.value, ny.value, nz.value
        self._lat = None

    def close(self):
        if getattr(self, "_h", None):
            if self._lat is not None:
                self._lat.close()
            lib().tclbx_case_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def units(self, text):
        """Text with units -> lattice units (UnitEnv::alt)."""
        v = ctypes.c_double()
        _check(lib().tclbx_case_units(self._h, text.encode(), ctypes.byref(v)))
        return v.value

    def rasterize(self):
        """The NodeType map of the case's Geometry, uint16 [nz, ny, nx] (host only)."""
        out = np.zeros(self.shape, dtype=np.uint16)
        _check(lib().tclbx_case_rasterize(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), out.size))
        return out

    def cuts(self):
        """Cut distances a side="surface" STL drew, uint16 [26, nz, ny, nx],
        or None when the geometry has none (call rasterize() first)."""
        out = np.zeros((26,) + tuple(self.shape), dtype=np.uint16)
        present = ctypes.c_int(0)
        _check(lib().tclbx_case_cuts(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), out.size,
                                     ctypes.byref(present)))
        return out if present.value else None

    @property
    def zones(self):
        """Setting zones, name -> number (DefaultZone = 0)."""
        L = lib()
        out = {}
        for k in range(L.tclbx_case_zone_count(self._h)):
            n = L.tclbx_case_zone_name(self._h, k)
            if n is not None:
                out[n.decode()] = k
        return out

    def run(self):
        _check(lib().tclbx_case_run(self._h))
        return self

    def check(self):
        """tclbx_case_check: every handler made and stepped through with no
        lattice (no GPU); raises on a failing handler.  The messages hold
        the notices and errors."""
        _check(lib().tclbx_case_check(self._h))
        return self

    @property
    def lattice(self):
        if self._lat is None:
            h = ctypes.c_void_p()
            _check(lib().tclbx_case_lattice(self._h, ctypes.byref(h)))
            self._lat = Lattice(self.model, self.nx, self.ny, self.nz, self.storage, _handle=h)
        return self._lat

    @property
    def iteration(self):
        return lib().tclbx_case_iteration(self._h)

    @property
    def messages(self):
        L = lib()
        return [L.tclbx_case_message(self._h, i).decode() for i in range(L.tclbx_case_message_count(self._h))]
