"""Synthetic case builders for the node-kernel workloads.

A Case holds what a TCLB case file would produce before the first
iteration: the NodeType map (as Geometry rasterizes it, src/Geometry.cpp.Rt:
141-319), Settings and zonal Settings (acParam, src/Handlers/acParam.cpp:
38-56), and optionally an initial state that replaces Init() (the
reference's LoadBinary route, src/Handlers/acLoadBinary.cpp:5-24).

Every random draw comes from an in-repo splitmix64 stream with a stated seed,
so a case is the same on every machine.
"""
from dataclasses import dataclass, field

import numpy as np

from . import Lattice, Model

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(seed, n):
    """n outputs of splitmix64 started at `seed` (Steele, Lea, Flood 2014)."""
    with np.errstate(over="ignore"):
        k = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, n, lo=0.0, hi=1.0):
    """n doubles uniform in [lo, hi): the top 53 bits of splitmix64."""
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


class Stream:
    """Sequential draws from one splitmix64 stream."""

    def __init__(self, seed):
        self.seed = np.uint64(seed)
        self.pos = 0

    def uniform(self, n=None, lo=0.0, hi=1.0):
        m = 1 if n is None else n
        with np.errstate(over="ignore"):
            start = self.seed + np.uint64(self.pos) * np.uint64(0x9E3779B97F4A7C15)
        v = uniform(start, m, lo, hi)
        self.pos += m
        return float(v[0]) if n is None else v


@dataclass
class Case:
    name: str
    model: str
    nx: int
    ny: int
    nz: int
    flags: np.ndarray                      # uint16 [nz, ny, nx]
    settings: list = field(default_factory=list)   # [(name, value)] in order
    zonal: list = field(default_factory=list)      # [(name, zone, value)]
    init_f: np.ndarray = None              # optional [Q, nz, ny, nx] float64
    steps: int = 1
    cuts: np.ndarray = None                # optional uint16 [26, nz, ny, nx] cut distances
    extra: dict = None                     # optional {extra density name: [nz, ny, nx]}, set after Init

    @property
    def nodes(self):
        return self.nx * self.ny * self.nz

    def build(self, storage="f64"):
        """A Lattice on the GPU holding this case, initialised."""
        lat = Lattice(self.model, self.nx, self.ny, self.nz, storage)
        self.apply(lat)
        return lat

    def apply(self, lat):
        lat.set_flags(self.flags)
        if self.cuts is not None:
            lat.set_cuts(self.cuts)
        for n, v in self.settings:
            lat.set_setting(n, v)
        for n, z, v in self.zonal:
            lat.set_zone_setting(n, v, z)
        lat.init()
        if self.init_f is not None:
            lat.set_f(self.init_f)
        for n, v in (self.extra or {}).items():
            lat.set_extra(n, v)
        return lat


# ----------------------------------------------------------- equilibria ---
def d2q9_feq(d, ux, uy):
    """SetEquilibrium of models/flow/d2q9_SRT/Dynamics.c:120-130 (float64)."""
    d, ux, uy = np.broadcast_arrays(np.asarray(d, float), np.asarray(ux, float), np.asarray(uy, float))
    f = np.empty((9,) + d.shape)
    f[0] = (2. + (-uy*uy - ux*ux)*3.)*d*2./9.
    f[1] = (2. + (-uy*uy + (1 + ux)*ux*2.)*3.)*d/18.
    f[2] = (2. + (-ux*ux + (1 + uy)*uy*2.)*3.)*d/18.
    f[3] = (2. + (-uy*uy + (-1 + ux)*ux*2.)*3.)*d/18.
    f[4] = (2. + (-ux*ux + (-1 + uy)*uy*2.)*3.)*d/18.
    f[5] = (1. + ((1 + uy)*uy + (1 + ux + uy*3.)*ux)*3.)*d/36.
    f[6] = (1. + ((1 + uy)*uy + (-1 + ux - uy*3.)*ux)*3.)*d/36.
    f[7] = (1. + ((-1 + uy)*uy + (-1 + ux + uy*3.)*ux)*3.)*d/36.
    f[8] = (1. + ((-1 + uy)*uy + (1 + ux - uy*3.)*ux)*3.)*d/36.
    return f


# ---------------------------------------------------------- workloads ---
def parabola(y, lo, hi, umax):
    """Poiseuille profile between walls at y = lo and y = hi, peak umax."""
    h = hi - lo
    return umax * 4.0 * (y - lo) * (hi - y) / (h * h)


def d2q9_cylinder(nx=1024, ny=128, diameter=32, cx=256, cy=64, umax=0.01, nu=0.02, steps=10000,
                  seed=1, perturb=1e-4, model="d2q9_SRT_drag"):
    """2D channel with a disc: BGK (d2q9_SRT models) or MRT (d2q9 models)
    everywhere in the fluid, no-slip bounce-back walls (mask ALL, no
    collision) on rows 0 and ny-1, Zou-He velocity inlet at x = 0 with a
    parabolic profile (one zone per row carrying Velocity_x, or VelocityX
    for d2q9), Zou-He pressure outlet (rho = 1) at x = nx-1, and a Solid
    disc (mask ALL) of the given diameter, rasterized as OffgridSphere does:
    (x+.5-cx)^2 + (y+.5-cy)^2 < r^2 (src/Geometry.cpp.Rt:793-834).  Initial
    state: equilibrium at rho = 1, u = the inlet profile, u_y perturbed
    uniformly in +-perturb (splitmix64, `seed`); walls and disc at rest.
    With the d2q9 models the viscosity sets S2 through RelaxationRate and
    the 3rd/4th order moments relax fully (S3 = S4 = 0, the model's
    defaults)."""
    m = Model(model)
    mrt = not model.startswith("d2q9_SRT")
    BGK = m.nodetype("MRT" if mrt else "BGK")[0]
    WALL, SOLID = m.nodetype("Wall")[0], m.nodetype("Solid")[0]
    WVEL, EPRES = m.nodetype("WVelocity")[0], m.nodetype("EPressure")[0]
    zshift = m.group("SETTINGZONE")[1]
    flags = np.full((1, ny, nx), BGK, dtype=np.uint16)
    y = np.arange(ny)
    zonal = []
    for j in range(1, ny - 1):
        flags[0, j, 0] = BGK | WVEL | (j << zshift)
        zonal.append(("VelocityX" if mrt else "Velocity_x", j, parabola(j, 0.5, ny - 1.5, umax)))
        flags[0, j, nx - 1] = BGK | EPRES
    X, Y = np.meshgrid(np.arange(nx), y)
    r = diameter / 2.0
    disc = (X + 0.5 - cx) ** 2 + (Y + 0.5 - cy) ** 2 < r * r
    flags[0][disc] = SOLID
    flags[0, 0, :] = WALL
    flags[0, ny - 1, :] = WALL
    ux = np.broadcast_to(parabola(y, 0.5, ny - 1.5, umax)[:, None], (ny, nx)).copy()
    uy = uniform(seed, nx * ny, -perturb, perturb).reshape(ny, nx)
    solid = (flags[0] == WALL) | (flags[0] == SOLID)
    ux[solid] = 0.0
    uy[solid] = 0.0
    init_f = d2q9_feq(1.0, ux, uy)[:, None]
    settings = [("Viscosity", nu)] if mrt else [("nu", nu), ("Density", 1.0)]
    return Case("d2q9_cylinder", model, nx, ny, 1, flags, settings=settings, zonal=zonal, init_f=init_f,
                steps=steps)


def random_spheres(nx, ny, nz, fraction, rmin, rmax, seed, periodic=(True, False, True)):
    """Boolean solid map of spheres with radius ~ U[rmin, rmax], centres
    x, z ~ U[0, n) and y ~ U[r, ny-1-r], added until the solid fraction is at
    least `fraction` (counted exactly); a node is inside when
    sum((i + .5 - c)^2) < r^2, with periodic images along the periodic axes."""
    solid = np.zeros((nz, ny, nx), dtype=bool)
    st = Stream(seed)
    target = fraction * nx * ny * nz
    count = 0
    while count < target:
        r = st.uniform(None, rmin, rmax)
        cx = st.uniform(None, 0, nx)
        cy = st.uniform(None, r, ny - 1 - r)
        cz = st.uniform(None, 0, nz)
        ri = int(np.ceil(r)) + 1
        xs = np.arange(int(np.floor(cx)) - ri, int(np.floor(cx)) + ri + 1)
        ys = np.arange(max(0, int(np.floor(cy)) - ri), min(ny, int(np.floor(cy)) + ri + 1))
        zs = np.arange(int(np.floor(cz)) - ri, int(np.floor(cz)) + ri + 1)
        dx2 = (xs + 0.5 - cx) ** 2
        dy2 = (ys + 0.5 - cy) ** 2
        dz2 = (zs + 0.5 - cz) ** 2
        inside = dz2[:, None, None] + dy2[None, :, None] + dx2[None, None, :] < r * r
        xi = np.mod(xs, nx) if periodic[0] else xs
        zi = np.mod(zs, nz) if periodic[2] else zs
        sel = np.ix_(zi, ys, xi)
        before = solid[sel].sum()
        solid[sel] |= inside
        count += int(solid[sel].sum() - before)
    return solid


def d3q19_porous(n=256, fraction=0.10, rmin=4, rmax=12, force=1e-6, tau=0.6, se=1.19, seps=1.4, sq=1.2,
                 steps=5000, seed=2):
    """d3q19 MRT porous channel: periodic in x and z, bounce-back walls
    (mask ALL) at y = 0 and y = n-1, body force ForceX, Solid spheres (mask
    ALL) up to the solid fraction, rho = 1 and u = 0 initially (Init)."""
    m = Model("d3q19")
    MRT = m.nodetype("MRT")[0]
    WALL, SOLID = m.nodetype("Wall")[0], m.nodetype("Solid")[0]
    flags = np.full((n, n, n), MRT, dtype=np.uint16)
    solid = random_spheres(n, n, n, fraction, rmin, rmax, seed)
    flags[solid] = SOLID
    flags[:, 0, :] = WALL
    flags[:, n - 1, :] = WALL
    settings = [("omega", 1.0 / tau), ("ForceX", force), ("MRT_Se", se), ("MRT_Seps", seps), ("MRT_Sq", sq)]
    return Case("d3q19_porous", "d3q19", n, n, n, flags, settings=settings, steps=steps)


def d3q19_cavity(n=640, u_lid=0.05, tau=0.52, steps=20000, ny=None, nz=None):
    """d3q19_lid lid-driven cavity: the top face y = ny-1 is a MovingWall
    (mask ALL) moving at (u_lid, 0, 0); the other five faces are Wall (mask
    ALL) bounce-back; MRT inside with omega = 1/tau and the reference's
    rates; rho = 1 and u = 0 initially.  The lid covers the whole top face,
    its edges included."""
    ny = n if ny is None else ny
    nz = n if nz is None else nz
    m = Model("d3q19_lid")
    MRT = m.nodetype("MRT")[0]
    WALL, LID = m.nodetype("Wall")[0], m.nodetype("MovingWall")[0]
    flags = np.full((nz, ny, n), MRT, dtype=np.uint16)
    flags[:, :, 0] = WALL
    flags[:, :, n - 1] = WALL
    flags[:, 0, :] = WALL
    flags[0, :, :] = WALL
    flags[nz - 1, :, :] = WALL
    flags[:, ny - 1, :] = LID
    settings = [("omega", 1.0 / tau), ("MovingWallVelocityX", u_lid)]
    return Case(f"d3q19_cavity_{n}x{ny}x{nz}", "d3q19_lid", n, ny, nz, flags, settings=settings, steps=steps)


def d2q9_cylinder_sweep(ncases=256, seed=4, steps=8000, umax=0.02, model="d2q9_drag"):
    """Independent cylinder cases of uneven size: nx uniform in
    {128, 144, ..., 1024}, ny uniform in {64, 80, ..., 256}, a Solid disc of
    diameter ny/4 at (nx/4, ny/2), Reynolds number uniform in [20, 200] with
    nu = umax * D / Re (umax the peak of the parabolic inlet); case k draws
    from splitmix64 stream seed*1000 + k.  The default model is the MRT
    d2q9 (with Drag/Lift): BGK (d2q9_SRT_drag) with the Zou-He velocity
    inlet is unstable below tau of about 0.55, which most of this Re range
    reaches."""
    out = []
    for k in range(ncases):
        st = Stream(seed * 1000 + k)
        nx = 128 + 16 * int(st.uniform(None, 0, 57))
        ny = 64 + 16 * int(st.uniform(None, 0, 13))
        re = st.uniform(None, 20.0, 200.0)
        d = ny / 4.0
        nu = umax * d / re
        c = d2q9_cylinder(nx=nx, ny=ny, diameter=d, cx=nx / 4.0, cy=ny / 2.0, umax=umax, nu=nu, steps=steps,
                          seed=seed * 1000 + k, model=model)
        c.name = f"cylinder_{k}_{nx}x{ny}_Re{re:.1f}"
        out.append(c)
    return out


def d3q19_serving_requests(nreq, rate=50.0, seed=5, side=(48, 256), porosity=(0.0, 0.25), force=(1e-7, 1e-5),
                           tau=(0.52, 0.9), budget=(1000, 20000)):
    """A Poisson request stream of d3q19 porous boxes: arrival gaps
    exponential with `rate` per second; each side length log-uniform in
    `side`; Solid spheres (radius 4-12) up to a solid fraction uniform in
    `porosity`; bounce-back walls at y = 0 and ny-1, periodic in x and z;
    ForceX uniform in `force`; tau uniform in `tau` (omega = 1/tau, reference
    MRT rates); a step budget uniform in `budget`.  Returns
    [(arrival_s, Case, max_steps)], request k drawing from splitmix64
    stream seed*100003 + k."""
    m = Model("d3q19")
    MRT = m.nodetype("MRT")[0]
    WALL, SOLID = m.nodetype("Wall")[0], m.nodetype("Solid")[0]
    out = []
    t = 0.0
    arrivals = Stream(seed)
    for k in range(nreq):
        t += -np.log(1.0 - arrivals.uniform()) / rate
        st = Stream(seed * 100003 + k)
        lo, hi = np.log(side[0]), np.log(side[1])
        nx, ny, nz = [int(round(np.exp(st.uniform(None, lo, hi)))) for _ in range(3)]
        frac = st.uniform(None, *porosity)
        g = st.uniform(None, *force)
        ta = st.uniform(None, *tau)
        steps = int(st.uniform(None, budget[0], budget[1] + 1))
        flags = np.full((nz, ny, nx), MRT, dtype=np.uint16)
        if frac > 0:
            flags[random_spheres(nx, ny, nz, frac, 4, 12, seed * 100003 + k)] = SOLID
        flags[:, 0, :] = WALL
        flags[:, ny - 1, :] = WALL
        c = Case(f"request_{k}_{nx}x{ny}x{nz}", "d3q19", nx, ny, nz, flags,
                 settings=[("omega", 1.0 / ta), ("ForceX", g)], steps=steps)
        out.append((t, c, steps))
    return out
