"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference CPU build's node
kernel (see oracle.h for the reference files it follows) and is the checker
of the HIP path.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

D2Q9_SRT, D2Q9_SRT_DRAG, D3Q19, D3Q19_LID, D2Q9, D2Q9_DRAG, D3Q27, D3Q27_CUMULANT = 0, 1, 2, 3, 4, 5, 6, 7
D2Q9_AUTOSYM, D2Q9_AUTOSYM2, D3Q27_CUMULANT_SMAG, D3Q27_CUMULANT_IB, D3Q27_CUMULANT_IB_SMAG = 8, 9, 10, 11, 12
D3Q27_CUMULANT_AVG, D3Q27_CUMULANT_AVG_IB, D3Q27_CUMULANT_AVG_SMAG, D3Q27_CUMULANT_AVG_IB_SMAG = 13, 14, 15, 16
D2Q9_BC, D2Q9_BC_AUTOSYM, D2Q9_BC_AUTOSYM2 = 17, 18, 19
F64, F32, F32S, F16S, BF16S = 0, 1, 2, 3, 4
MODELS = {"d2q9_SRT": D2Q9_SRT, "d2q9_SRT_drag": D2Q9_SRT_DRAG, "d3q19": D3Q19, "d3q19_lid": D3Q19_LID,
          "d2q9": D2Q9, "d2q9_drag": D2Q9_DRAG, "d3q27": D3Q27, "d3q27_cumulant": D3Q27_CUMULANT,
          "d2q9_autosym": D2Q9_AUTOSYM, "d2q9_autosym2": D2Q9_AUTOSYM2, "d3q27_cumulant_SMAG": D3Q27_CUMULANT_SMAG,
          "d3q27_cumulant_IB": D3Q27_CUMULANT_IB, "d3q27_cumulant_IB_SMAG": D3Q27_CUMULANT_IB_SMAG,
          "d3q27_cumulant_AVG": D3Q27_CUMULANT_AVG, "d3q27_cumulant_AVG_IB": D3Q27_CUMULANT_AVG_IB,
          "d3q27_cumulant_AVG_SMAG": D3Q27_CUMULANT_AVG_SMAG,
          "d3q27_cumulant_AVG_IB_SMAG": D3Q27_CUMULANT_AVG_IB_SMAG, "d2q9_bc": D2Q9_BC,
          "d2q9_bc_autosym": D2Q9_BC_AUTOSYM, "d2q9_bc_autosym2": D2Q9_BC_AUTOSYM2}
AVERAGES = ["avgP", "varUX", "varUY", "varUZ", "varUXUY", "varUXUZ", "varUYUZ", "avgdxu2", "avgdyv2", "avgdzw2",
            "avgUX", "avgUY", "avgUZ"]
STORAGE = {"f64": F64, "f32": F32, "f32s": F32S, "f16s": F16S, "bf16s": BF16S}

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ip, ip, ip, ip, ip]
        L.orc_destroy.argtypes = [vp]
        L.orc_q.argtypes = [vp]
        L.orc_nglobals.argtypes = [vp]
        L.orc_set_params.argtypes = [vp, dp, ip]
        L.orc_set_zone_table.argtypes = [vp, dp, ip, ip]
        L.orc_set_flags.argtypes = [vp, ctypes.POINTER(ctypes.c_uint16)]
        L.orc_set_cuts.argtypes = [vp, ctypes.POINTER(ctypes.c_uint16)]
        L.orc_reset_average.argtypes = [vp]
        L.orc_nextra.argtypes = [vp]
        L.orc_get_extra.argtypes = [vp, dp]
        L.orc_set_extra.argtypes = [vp, dp]
        L.orc_cum_ib.argtypes = [dp, ctypes.c_uint16, ctypes.POINTER(ctypes.c_uint16), dp, dp, dp]
        L.orc_init.argtypes = [vp]
        L.orc_iterate.argtypes = [vp, ip, ip, dp]
        L.orc_get_f.argtypes = [vp, dp]
        L.orc_set_f.argtypes = [vp, dp]
        L.orc_get_quantity.argtypes = [vp, ip, dp]
        L.orc_f32_to_f16.restype = ctypes.c_uint16
        L.orc_f32_to_f16.argtypes = [ctypes.c_float]
        L.orc_f16_to_f32.restype = ctypes.c_float
        L.orc_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_f32_to_bf16.restype = ctypes.c_uint16
        L.orc_f32_to_bf16.argtypes = [ctypes.c_float]
        L.orc_bf16_to_f32.restype = ctypes.c_float
        L.orc_bf16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_d2q9_feq.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.orc_d3q19_feq.argtypes = [dp] + [ctypes.c_double] * 4
        L.orc_d3q19_collide.argtypes = [dp, dp]
        L.orc_d3q27_feq.argtypes = [dp] + [ctypes.c_double] * 4
        L.orc_d3q27_node.argtypes = [dp, ctypes.c_uint16, dp, dp, dp]
        L.orc_cum_node.argtypes = [dp, ctypes.c_uint16, dp, dp, dp, ip]
        L.orc_d3q27_basis.argtypes = [dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_set_threads.argtypes = [ip]
        u32 = ctypes.c_uint32
        L.orc_sr_bits.restype = u32
        L.orc_sr_bits.argtypes = [u32, u32, u32, u32]
        L.orc_f32_to_f16_sr.restype = ctypes.c_uint16
        L.orc_f32_to_f16_sr.argtypes = [ctypes.c_float, u32]
        L.orc_f32_to_bf16_sr.restype = ctypes.c_uint16
        L.orc_f32_to_bf16_sr.argtypes = [ctypes.c_float, u32]
        L.orc_set_stochastic.argtypes = [vp, ctypes.c_int, u32]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleLattice:
    """One lattice of the oracle; layout of f is [q][z][y][x]."""

    def __init__(self, model, nx, ny, nz=1, storage="f64"):
        self.model = MODELS[model] if isinstance(model, str) else model
        self.storage = STORAGE[storage] if isinstance(storage, str) else storage
        self.nx, self.ny, self.nz = nx, ny, nz
        self.h = lib().orc_create(self.model, nx, ny, nz, self.storage)
        self.Q = lib().orc_q(self.h)
        self.nglob = lib().orc_nglobals(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        return (self.nz, self.ny, self.nx)

    def set_params(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        lib().orc_set_params(self.h, _dp(p), len(p))

    def set_zone_table(self, table):
        t = np.ascontiguousarray(np.atleast_2d(table), dtype=np.float64)
        lib().orc_set_zone_table(self.h, _dp(t), t.shape[0], t.shape[1])

    def set_flags(self, flags):
        f = np.ascontiguousarray(flags, dtype=np.uint16).reshape(-1)
        assert f.size == self.nx * self.ny * self.nz
        lib().orc_set_flags(self.h, f.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))

    def set_cuts(self, cuts):
        c = np.ascontiguousarray(cuts, dtype=np.uint16).reshape(-1)
        assert c.size == 26 * self.nx * self.ny * self.nz
        lib().orc_set_cuts(self.h, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)))

    def init(self):
        lib().orc_init(self.h)

    def iterate(self, n, globals_last=True):
        g = np.zeros(max(self.nglob, 1), dtype=np.float64)
        lib().orc_iterate(self.h, n, 1 if globals_last else 0, _dp(g))
        return g

    def get_f(self):
        out = np.empty((self.Q,) + self.shape, dtype=np.float64)
        lib().orc_get_f(self.h, _dp(out))
        return out

    def set_f(self, f):
        a = np.ascontiguousarray(f, dtype=np.float64)
        assert a.size == self.Q * self.nx * self.ny * self.nz
        lib().orc_set_f(self.h, _dp(a))

    def reset_average(self):
        lib().orc_reset_average(self.h)

    def set_stochastic(self, on=True, seed=0):
        """Stochastic rounding of the 16-bit formats' stores in steps."""
        lib().orc_set_stochastic(self.h, 1 if on else 0, int(seed) & 0xFFFFFFFF)

    def extra(self):
        """[k, nz, ny, nx] of the densities that are not streamed"""
        a = np.empty((lib().orc_nextra(self.h),) + self.shape)
        lib().orc_get_extra(self.h, _dp(a))
        return a

    def set_extra(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.size == lib().orc_nextra(self.h) * self.nx * self.ny * self.nz
        lib().orc_set_extra(self.h, _dp(a))

    def averages(self):
        """dict name -> [nz, ny, nx] of the AVG builds' averaged densities"""
        return dict(zip(AVERAGES, self.extra()))

    def quantity(self, which):
        w = {"Rho": 0, "U": 1, "P": 2, "Fd": 3, "Solid": 4, "KinE": 5, "ReStr": 6, "Dissipation": 7, "avgU": 8,
             "varU": 9, "averageP": 10}[which] if isinstance(which, str) else which
        n = self.nx * self.ny * self.nz
        vec = w in (1, 3, 6, 8, 9)
        out = np.empty(n * (3 if vec else 1), dtype=np.float64)
        lib().orc_get_quantity(self.h, w, _dp(out))
        return out.reshape(self.shape + ((3,) if vec else ()))


def set_threads(n):
    """Threads of the oracle's step loop (OpenMP over (y,z) rows)."""
    lib().orc_set_threads(int(n))


def d2q9_feq(d, ux, uy):
    out = np.empty(9)
    lib().orc_d2q9_feq(_dp(out), d, ux, uy)
    return out


def d3q19_feq(rho, jx, jy, jz):
    out = np.empty(19)
    lib().orc_d3q19_feq(_dp(out), rho, jx, jy, jz)
    return out


def d3q19_collide(f, params):
    a = np.array(f, dtype=np.float64)
    p = np.zeros(16)
    p[: len(params)] = params
    lib().orc_d3q19_collide(_dp(a), _dp(p))
    return a


def d3q27_feq(rho, jx, jy, jz):
    out = np.empty(27)
    lib().orc_d3q27_feq(_dp(out), rho, jx, jy, jz)
    return out


def d3q27_node(f, nt, params, zone=None):
    """one d3q27 node update; returns (f, Globals[7])"""
    a = np.array(f, dtype=np.float64)
    p = np.zeros(16)
    p[: len(params)] = params
    z = np.zeros(32)
    if zone is not None:
        z[: len(zone)] = zone
    g = np.zeros(7)
    lib().orc_d3q27_node(_dp(a), nt, _dp(p), _dp(z), _dp(g))
    return a, g


def cum_node(f, nt, params, zone=None, smag=False):
    """one d3q27_cumulant node update (smag: the SMAG build); returns (f, Globals[10])"""
    a = np.array(f, dtype=np.float64)
    p = np.zeros(16)
    p[: len(params)] = params
    z = np.zeros(32)
    if zone is not None:
        z[: len(zone)] = zone
    g = np.zeros(10)
    lib().orc_cum_node(_dp(a), nt, _dp(p), _dp(z), _dp(g), 1 if smag else 0)
    return a, g


def cum_ib(f, nt, cut, old, zone=None):
    """InterpolatedBounceBack of the IB builds on one node: f pulled, cut[26]
    its links' cut distances, old[27] its populations in the read snapshot;
    returns (f, Globals[10])"""
    a = np.array(f, dtype=np.float64)
    c = np.ascontiguousarray(cut, dtype=np.uint16)
    o = np.ascontiguousarray(old, dtype=np.float64)
    z = np.zeros(32)
    if zone is not None:
        z[: len(zone)] = zone
    g = np.zeros(10)
    lib().orc_cum_ib(_dp(a), nt, c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)), _dp(o), _dp(z), _dp(g))
    return a, g


def d3q27_basis():
    """(M [direction, moment], order [moment], powers [moment, axis])"""
    M = np.empty(27 * 27)
    o = np.empty(27, dtype=np.int32)
    pw = np.empty(81, dtype=np.int32)
    lib().orc_d3q27_basis(_dp(M), o.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                          pw.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return M.reshape(27, 27), o, pw.reshape(27, 3)


D3Q19_PARAMS = ["omega", "InletVelocity", "InletDensity", "ForceX", "ForceY", "ForceZ",
                "MRT_Se", "MRT_Seps", "MRT_Sq", "MRT_Spi", "MRT_Sm"]
D3Q19_LID_PARAMS = D3Q19_PARAMS + ["MovingWallVelocityX", "MovingWallVelocityY", "MovingWallVelocityZ"]
D2Q9_PARAMS = ["S2", "S3", "S4", "GravitationX", "GravitationY"]
D3Q27_PARAMS = ["omega", "Smag", "ForceX", "ForceY", "ForceZ"]
CUMULANT_PARAMS = ["nu", "nubuffer", "GalileanCorrection", "ForceX", "ForceY", "ForceZ", "Omega"]
CUMULANT_SMAG_PARAMS = CUMULANT_PARAMS + ["Smag"]

# derived settings of the models' Dynamics.R (AddSetting(name, <target>=expr))
DERIVED = {
    "nu": ("omega", lambda v: 1.0 / (3 * v + 0.5)),
    "InletPressure": ("InletDensity", lambda v: 1.0 + v * 3),
    "Viscosity": ("RelaxationRate", lambda v: 1.0 / (3 * v + 0.5)),
    "RelaxationRate": ("S2", lambda v: 1 - v),
}


def resolved_settings(model, settings):
    """Setting values after applying `settings` in order the way
    Lattice::SetSetting does (src/Lattice.cu.Rt:1350-1358): defaults of the
    non-deriving settings, then each value and its derived target."""
    from tclb_amd import Model
    m = Model(model)
    vals = {n: 0.0 for n in m.settings}
    for n in m.settings:
        if n not in DERIVED:
            vals[n] = m.setting_default(n)
    for n, v in settings:
        while True:
            vals[n] = v
            if n not in DERIVED or DERIVED[n][0] not in vals:
                break
            n, v = DERIVED[n][0], DERIVED[n][1](v)
    return vals


def lattice_for_case(case, storage="f64"):
    """OracleLattice holding the same case as case.build(storage)."""
    from tclb_amd import Model
    m = Model(case.model)
    L = OracleLattice(case.model, case.nx, case.ny, case.nz, storage)
    vals = resolved_settings(case.model, case.settings)
    if case.model == "d3q19":
        L.set_params([vals[n] for n in D3Q19_PARAMS])
    elif case.model == "d3q19_lid":
        L.set_params([vals[n] for n in D3Q19_LID_PARAMS])
    elif case.model in ("d2q9", "d2q9_drag", "d2q9_autosym", "d2q9_autosym2", "d2q9_bc", "d2q9_bc_autosym",
                        "d2q9_bc_autosym2"):
        L.set_params([vals[n] for n in D2Q9_PARAMS])
    elif case.model == "d3q27":
        L.set_params([vals[n] for n in D3Q27_PARAMS])
    elif case.model.startswith("d3q27_cumulant"):
        L.set_params([vals[n] for n in (CUMULANT_SMAG_PARAMS if case.model.endswith("SMAG") else CUMULANT_PARAMS)])
    else:
        L.set_params([vals["omega"], vals["Density"]])
    zshift = m.group("SETTINGZONE")[1]
    used = int((np.asarray(case.flags) >> zshift).max())
    rows = max(used, max([z for _, z, _ in case.zonal], default=0)) + 1
    table = np.zeros((rows, m.nzonal))
    for n, z, v in case.zonal:
        c = m.zonal_settings.index(n)
        if z < 0:
            table[:, c] = v
        else:
            table[z, c] = v
    L.set_zone_table(table)
    L.set_flags(case.flags)
    if getattr(case, "cuts", None) is not None:
        L.set_cuts(case.cuts)
    L.init()
    if case.init_f is not None:
        L.set_f(case.init_f)
    if getattr(case, "extra", None):
        e = L.extra()
        for n, v in case.extra.items():
            e[m.extras.index(n)] = v
        L.set_extra(e)
    return L
