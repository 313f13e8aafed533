"""Pins of the CPU oracle (oracle/) against facts that do not come from the
oracle itself.

The reference's own golden data for these models lives in the TCLB_tests
submodule (tests/external, .gitmodules:1-3), which is absent, so the oracle
is pinned here by:
  * a sympy derivation of the d3q19 equilibrium moments following the
    algorithm of src/lib/d3q19.R:24-62 (product-form feq, truncation to
    second order in J) on the MRTMAT data of models/flow/d3q19/Dynamics.R;
  * what the mathematics fixes: the BGK fixed point, conservation of mass and
    momentum, the momentum gain of a body force, MRT with equal rates being
    BGK, the relaxation factor of every moment group, Zou-He invariants,
    bounce-back reversal, one-hot streaming with periodic wrap, and the exact
    Poiseuille solution that halfway bounce-back gives with the reference's
    q-moment rate 8(2-w)/(8-w);
  * numpy's IEEE float16 rounding and a brute-force bfloat16 rounding.
Each check is chosen so that a plausible slip (a wrong sign, index, rate
group, missing term or transposed matrix) fails it.
"""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def mrtmat():
    rows = json.load(open(os.path.join(GOLDEN, "d3q19_mrtmat.json")))["rows"]
    return np.array(rows, dtype=np.int64)  # [moment k][direction i]


# ------------------------------------------------------------- storage ---
def test_f16_matches_numpy_rounding(O):
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.standard_normal(4000).astype(np.float32) * np.float32(1e-2),
        rng.uniform(-70000, 70000, 2000).astype(np.float32),
        (rng.standard_normal(2000) * 1e-6).astype(np.float32),   # subnormal halves
        np.float32([0.0, -0.0, 65504.0, 65519.0, 65520.0, 6.1035156e-05, 5.9604645e-08, 2.9802322e-08,
                    np.inf, -np.inf]),
    ])
    # exact ties: halfway between two halves
    h = rng.integers(0, 0x7bff, 500).astype(np.uint16)
    a = h.view(np.float16).astype(np.float32)
    b = (h + 1).view(np.float16).astype(np.float32)
    vals = np.concatenate([vals, (a + (b - a) / 2).astype(np.float32)])
    L = O.lib()
    got = np.array([L.orc_f32_to_f16(float(v)) for v in vals], dtype=np.uint16)
    with np.errstate(over="ignore"):
        want = vals.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, want)


def test_f16_to_f32_every_pattern(O):
    L = O.lib()
    pats = np.arange(0, 0x10000, dtype=np.uint32).astype(np.uint16)
    f = pats.view(np.float16).astype(np.float32)
    finite = np.isfinite(f)
    got = np.array([L.orc_f16_to_f32(int(p)) for p in pats[finite]], dtype=np.float32)
    assert np.array_equal(got.view(np.uint32), f[finite].view(np.uint32))


def test_bf16_round_to_nearest_even(O):
    rng = np.random.default_rng(1)
    vals = np.concatenate([rng.standard_normal(3000) * 1e-3, rng.standard_normal(1000) * 10]).astype(np.float32)
    hi = (rng.integers(0, 0x7f00, 300).astype(np.uint32) << 16)
    ties = (hi | 0x8000).view(np.float32)
    vals = np.concatenate([vals, ties])
    L = O.lib()
    for v in vals:
        b = L.orc_f32_to_bf16(float(v))
        got = np.float64(np.uint32(b << 16).view(np.float32))
        u = np.float32(v).view(np.uint32)
        lo = np.float64((u & np.uint32(0xFFFF0000)).view(np.float32))
        up = np.float64(((u & np.uint32(0xFFFF0000)) + np.uint32(0x10000)).view(np.float32))
        x = np.float64(v)
        dlo, dup = abs(x - lo), abs(up - x)
        if dlo < dup:
            want = lo
        elif dup < dlo:
            want = up
        else:
            want = lo if ((u >> np.uint32(16)) & np.uint32(1)) == 0 else up
        assert got == want, (v, got, want)


# ------------------------------------------------------ d3q19 equilibrium ---
@pytest.fixture(scope="module")
def req_sympy():
    """Req(rho, J) by the algorithm of src/lib/d3q19.R:24-62."""
    sp = pytest.importorskip("sympy")
    Mk = mrtmat()
    M = sp.Matrix(19, 19, lambda i, k: int(Mk[k][i]))  # MRTMAT[i, k]
    U = M[:, [3, 5, 7]]
    p = [[2 if U[j, a] < 0 else int(U[j, a]) for a in range(3)] for j in range(19)]
    W = sp.Matrix(19, 19, lambda b, i: sp.prod([sp.Integer(U[b, a]) ** p[i][a] for a in range(3)]))
    rho, Jx, Jy, Jz, sig = sp.symbols("rho Jx Jy Jz sigma")
    J = [Jx, Jy, Jz]
    H = []
    for j in range(19):
        h = rho
        for a in range(3):
            if p[j][a] == 1:
                h = h * J[a] / rho
            if p[j][a] == 2:
                h = h * (J[a] ** 2 / rho ** 2 + sig)
        H.append(h)
    feq = sp.Matrix(1, 19, H) * W.inv()
    Req = feq.subs(sig, sp.Rational(1, 3)) * M
    out = []
    for k in range(19):
        e = sp.expand(Req[0, k])
        keep = [t for t in sp.Add.make_args(e) if e != 0 and sp.Poly(t * rho ** 10, Jx, Jy, Jz).total_degree() < 3]
        out.append(sp.Add(*keep))
    v = [int(sum(Mk[k][i] ** 2 for i in range(19))) for k in range(19)]
    Minv = sp.diag(*[sp.Rational(1, vk) for vk in v]) * M.T
    feq_trunc = sp.Matrix(1, 19, out) * Minv
    return {
        "req": sp.lambdify((rho, Jx, Jy, Jz), out, "numpy"),
        "feq": sp.lambdify((rho, Jx, Jy, Jz), list(feq_trunc), "numpy"),
        "v": np.array(v, dtype=float),
    }


def test_mrtmat_is_orthogonal():
    M = mrtmat().astype(float)
    G = M @ M.T
    assert np.array_equal(G, np.diag(np.diag(G)))
    assert list(np.diag(G)) == [19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8]


def test_d3q19_feq_matches_product_form_derivation(O, req_sympy):
    rng = np.random.default_rng(2)
    for _ in range(20):
        rho = 1 + 0.1 * rng.standard_normal()
        J = 0.05 * rng.standard_normal(3)
        got = O.d3q19_feq(rho, *J)
        want = np.array(req_sympy["feq"](rho, *J), dtype=float)
        assert np.allclose(got, want, rtol=0, atol=1e-15)


def test_d3q19_feq_weights_at_rest(O):
    f = O.d3q19_feq(1.0, 0, 0, 0)
    assert np.allclose(f, [1/3] + [1/18] * 6 + [1/36] * 12, rtol=0, atol=1e-16)


def _random_f(rng, Q=19, amp=0.01):
    f = np.array([1/3] + [1/18] * 6 + [1/36] * 12) if Q == 19 else np.array([4/9] + [1/9] * 4 + [1/36] * 4)
    return f * (1 + amp * rng.standard_normal(Q))


def test_d3q19_mrt_equal_rates_is_bgk(O):
    """With every rate equal to omega, MRT reduces to BGK with the same
    (product-form, truncated) equilibrium."""
    rng = np.random.default_rng(3)
    M = mrtmat()
    for w in (0.7, 1.3, 1.9):
        f = _random_f(rng)
        params = [w, 0, 1, 0, 0, 0, w, w, w, w, w]
        out = O.d3q19_collide(f, params)
        rho, J = f.sum(), M[[3, 5, 7]] @ f
        want = f + w * (O.d3q19_feq(rho, *J) - f)
        assert np.allclose(out, want, rtol=0, atol=2e-16 * 10)


def test_d3q19_moment_groups_relax_at_reference_rates(O, req_sympy):
    """R'_k - Req_k = gamma_k (R_k - Req_k) with gamma1 = 1-w on e, eps, nu,
    pi moments and gamma2 = 1 - 8(2-w)/(8-w) on q and m moments
    (models/flow/d3q19/Dynamics.c.Rt:246-252)."""
    rng = np.random.default_rng(4)
    M = mrtmat().astype(float)
    w = 1.37
    g1, g2 = 1 - w, 1 - 8 * (2 - w) / (8 - w)
    gam = np.array([0, g1, g1, 0, g2, 0, g2, 0, g2, g1, g1, g1, g1, g1, g1, g1, g2, g2, g2])
    f = _random_f(rng, amp=0.05)
    out = O.d3q19_collide(f, [w, 0, 1, 0, 0, 0] + [0] * 5)
    R, R2 = M @ f, M @ out
    rho, J = R[0], R[[3, 5, 7]]
    req = np.array(req_sympy["req"](rho, *J), dtype=float)
    for k in [1, 2, 4, 6, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18]:
        assert abs((R2[k] - req[k]) - gam[k] * (R[k] - req[k])) < 1e-14, k
    assert abs(R2[0] - R[0]) < 1e-15
    assert np.allclose(R2[[3, 5, 7]], J, atol=1e-16)


def test_d3q19_free_rates_map_to_groups(O, req_sympy):
    rng = np.random.default_rng(5)
    M = mrtmat().astype(float)
    w, se, seps, sq, spi, sm = 1.1, 1.19, 1.4, 1.2, 1.05, 1.6
    f = _random_f(rng, amp=0.05)
    out = O.d3q19_collide(f, [w, 0, 1, 0, 0, 0, se, seps, sq, spi, sm])
    R, R2 = M @ f, M @ out
    req = np.array(req_sympy["req"](R[0], *R[[3, 5, 7]]), dtype=float)
    rate = {1: se, 2: seps, 4: sq, 6: sq, 8: sq, 9: w, 11: w, 13: w, 14: w, 15: w, 10: spi, 12: spi,
            16: sm, 17: sm, 18: sm}
    for k, s in rate.items():
        assert abs((R2[k] - req[k]) - (1 - s) * (R[k] - req[k])) < 1e-14, k


def test_d3q19_force_adds_rho_F(O):
    rng = np.random.default_rng(6)
    M = mrtmat().astype(float)
    f = _random_f(rng)
    F = np.array([1e-4, -2e-4, 3e-5])
    out = O.d3q19_collide(f, [1.2, 0, 1, *F] + [0] * 5)
    assert abs(out.sum() - f.sum()) < 1e-15
    assert np.allclose(M[[3, 5, 7]] @ out, M[[3, 5, 7]] @ f + f.sum() * F, rtol=0, atol=1e-16)


# ------------------------------------------------------------ lattices ---
def test_one_hot_streaming_d2q9(O):
    ex = [0, 1, 0, -1, 0, 1, -1, -1, 1]
    ey = [0, 0, 1, 0, -1, 1, 1, -1, -1]
    nx, ny = 5, 4
    for (x, y) in [(0, 0), (4, 3), (2, 1)]:
        L = O.OracleLattice("d2q9_SRT", nx, ny, 1)
        L.set_params([1.0, 1.0])
        L.set_flags(np.zeros(nx * ny, np.uint16))
        f = np.zeros((9, 1, ny, nx))
        for q in range(9):
            f[q, 0, y, x] = q + 1
        L.set_f(f)
        L.iterate(1)
        g = L.get_f()
        for q in range(9):
            want = np.zeros((1, ny, nx))
            want[0, (y + ey[q]) % ny, (x + ex[q]) % nx] = q + 1
            assert np.array_equal(g[q], want), q


def test_one_hot_streaming_d3q19(O):
    M = mrtmat()
    e = M[[3, 5, 7]].T
    n = (4, 3, 5)  # nx, ny, nz
    for (x, y, z) in [(0, 0, 0), (3, 2, 4), (1, 1, 2)]:
        L = O.OracleLattice("d3q19", *n)
        L.set_params([1.0, 0, 1] + [0] * 8)
        L.set_flags(np.zeros(np.prod(n), np.uint16))
        f = np.zeros((19, n[2], n[1], n[0]))
        for q in range(19):
            f[q, z, y, x] = q + 1
        L.set_f(f)
        L.iterate(1)
        g = L.get_f()
        for q in range(19):
            want = np.zeros((n[2], n[1], n[0]))
            want[(z + e[q][2]) % n[2], (y + e[q][1]) % n[1], (x + e[q][0]) % n[0]] = q + 1
            assert np.array_equal(g[q], want), q


def _d2q9_periodic(O, f, omega=1.3, grav=(0.0, 0.0), steps=1, storage="f64"):
    ny, nx = f.shape[2], f.shape[3]
    L = O.OracleLattice("d2q9_SRT", nx, ny, 1, storage)
    L.set_params([omega, 1.0])
    L.set_zone_table(np.array([[0, 0, 0, grav[0], grav[1]]]))
    L.set_flags(np.full(nx * ny, 8, np.uint16))
    L.set_f(f)
    L.iterate(steps)
    return L


def test_d2q9_bgk_fixed_point(O):
    feq = O.d2q9_feq(1.1, 0.03, -0.02)
    f = np.broadcast_to(feq[:, None, None, None], (9, 1, 4, 6)).copy()
    L = _d2q9_periodic(O, f, steps=5)
    assert np.allclose(L.get_f(), f, rtol=0, atol=4e-16)


def test_d2q9_conservation_and_body_force(O):
    rng = np.random.default_rng(7)
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
    ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
    f = _random_f(rng, 9)[:, None, None, None] * (1 + 0.01 * rng.standard_normal((9, 1, 6, 7)))
    mass = f.sum()
    J = np.array([(f * ex[:, None, None, None]).sum(), (f * ey[:, None, None, None]).sum()])
    L = _d2q9_periodic(O, f, steps=3)
    g = L.get_f()
    assert abs(g.sum() - mass) < 1e-13
    assert np.allclose([(g * ex[:, None, None, None]).sum(), (g * ey[:, None, None, None]).sum()], J, atol=1e-14)
    G = (2e-4, -1e-4)
    L = _d2q9_periodic(O, f, grav=G, steps=1)
    g = L.get_f()
    Jn = np.array([(g * ex[:, None, None, None]).sum(), (g * ey[:, None, None, None]).sum()])
    assert np.allclose(Jn, J + mass * np.array(G), rtol=0, atol=1e-14)


D2_EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
D2_EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
D2_W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def test_d2q9_srt_feq_is_the_second_order_polynomial(O):
    """SetEquilibrium (models/flow/d2q9_SRT/Dynamics.c:120-130) is the
    standard second-order equilibrium w_i rho (1 + 3 e.u + 9/2 (e.u)^2 -
    3/2 u^2), quadratic terms included (they set the momentum flux; a BGK
    fixed point alone pins only moments 0 and 1)."""
    rng = np.random.default_rng(41)
    for _ in range(50):
        rho = 0.8 + 0.4 * rng.random()
        ux, uy = 0.2 * rng.standard_normal(2)
        eu = D2_EX * ux + D2_EY * uy
        want = D2_W * rho * (1 + 3 * eu + 4.5 * eu * eu - 1.5 * (ux * ux + uy * uy))
        got = O.d2q9_feq(rho, ux, uy)
        assert np.max(np.abs(got - want)) <= 4e-16 * rho, (rho, ux, uy)
        # the second moment: Pi = rho/3 I + rho u u
        Pxy = (got * D2_EX * D2_EY).sum()
        Pxx = (got * D2_EX * D2_EX).sum()
        assert abs(Pxy - rho * ux * uy) < 1e-15 and abs(Pxx - rho / 3 - rho * ux * ux) < 1e-15


def _srt_channel(O, tau, g, ny, steps):
    """x-invariant d2q9_SRT channel: Wall rows (bounce-back, no collision:
    mask ALL) at y = 0 and ny-1, BGK fluid between, zonal GravitationX = g."""
    L = O.OracleLattice("d2q9_SRT", 1, ny, 1)
    L.set_params([1 / tau, 1.0])
    L.set_zone_table(np.array([[0, 0, 0, g, 0]]))
    fl = np.full((1, ny, 1), 8, np.uint16)
    fl[0, 0, 0] = fl[0, ny - 1, 0] = 6
    L.set_flags(fl)
    L.init()
    L.iterate(steps)
    return L


@pytest.mark.parametrize("tau", [0.56, 0.8, 0.5 + np.sqrt(3) / 4, 1.2])
def test_d2q9_srt_poiseuille_with_bounce_back_slip(O, tau):
    """BGK force-driven Poiseuille flow between bounce-back walls.  The
    steady x-invariant solution of the lattice BGK equation is the parabola
    with walls half-way between nodes, u = g/(2 nu)(y - 1/2)(ny - 3/2 - y),
    nu = (tau - 1/2)/3, plus a uniform slip u_s = g (16 L - 3)/(24 nu) with
    L = (tau - 1/2)^2: zero at L = 3/16, where bounce-back is exact for
    Poiseuille flow (Ginzburg & d'Humieres, Phys. Rev. E 68, 066614, 2003;
    the tau-dependent bounce-back slip of He, Zou, Luo & Dembo, J. Stat.
    Phys. 87, 115, 1997).  U is J/rho + G/2 (Dynamics.c:107-116).  A wrong
    viscosity, a force not shifted by G/omega, or a missing G/2 in U fails
    this by far more than the 1e-8 bound."""
    ny, g = 12, 1e-6
    nu = (tau - 0.5) / 3
    lam = (tau - 0.5) ** 2
    L = _srt_channel(O, tau, g, ny, 80000)
    u = L.quantity("U")[0, 1:-1, 0, 0]
    y = np.arange(1, ny - 1)
    want = g / (2 * nu) * (y - 0.5) * (ny - 1.5 - y) + g * (16 * lam - 3) / (24 * nu)
    assert np.max(np.abs(u - want)) / np.max(np.abs(want)) < 1e-8, np.max(np.abs(u - want))
    # no transverse flow, density uniform
    assert np.max(np.abs(L.quantity("U")[0, 1:-1, 0, 1])) < 1e-14
    rho = L.quantity("Rho")[0, 1:-1, 0]
    assert np.ptp(rho) < 1e-12


@pytest.mark.parametrize("g", [(1e-5, 0.0), (4e-6, -3e-6)])
def test_d2q9_srt_drag_balances_the_body_force(O, g):
    """d2q9_SRT_drag's Drag/Lift (momentum exchange 2 sum_i f_i e_i of the
    populations arriving at Solid nodes) against the momentum balance of a
    periodic box: at steady state the body force puts G sum(rho) into the
    fluid each step (BGK nodes only), and all of it leaves through the
    body, so Drag = G_x M and Lift = G_y M with M the fluid mass, whatever
    the body's shape.  A factor or sign slip in the exchange, or a sum over
    the populations after the swap, fails this."""
    nx, ny = 24, 16
    yy, xx = np.mgrid[0:ny, 0:nx]
    solid = ((xx + 0.5 - 9) ** 2 + (yy + 0.5 - 7.3) ** 2 < 3.2 ** 2) | ((xx == 15) & (yy > 3) & (yy < 9))
    fl = np.where(solid, 5, 8).astype(np.uint16).reshape(1, ny, nx)
    L = O.OracleLattice("d2q9_SRT_drag", nx, ny, 1)
    L.set_params([1.0, 1.0])
    L.set_zone_table(np.array([[0, 0, 0, g[0], g[1], 0, 0]]))
    L.set_flags(fl)
    L.init()
    L.iterate(30000)
    glob = L.iterate(1)
    rho = L.quantity("Rho")[0]
    mass = rho[~solid].sum()
    assert abs(glob[0] - g[0] * mass) <= 1e-9 * abs(g[0] * mass), (glob[0], g[0] * mass)
    scale = abs(g[0] * mass)
    assert abs(glob[1] - g[1] * mass) <= 1e-9 * scale, (glob[1], g[1] * mass)


def _node_after(O, model, flag, f_node, params, zone_row=None):
    """Post-boundary state of one node with NodeType `flag` on a 3x3(x3)
    lattice whose other nodes have flag 0 (stream only)."""
    if model == "d3q19":
        L = O.OracleLattice("d3q19", 3, 3, 3)
        n = 27
        shape = (19, 3, 3, 3)
        idx = (slice(None), 1, 1, 1)
    else:
        L = O.OracleLattice(model, 3, 3, 1)
        n = 9
        shape = (9, 1, 3, 3)
        idx = (slice(None), 0, 1, 1)
    L.set_params(params)
    if zone_row is not None:
        L.set_zone_table(np.array([zone_row]))
    flags = np.zeros(n, np.uint16)
    flags[n // 2] = flag
    L.set_flags(flags)
    f = np.zeros(shape)
    # node (1,1,1) pulls f_i from (1,1,1) - e_i: fill every node with the
    # same populations so the pulled values are f_node
    f[:] = np.asarray(f_node).reshape((-1,) + (1,) * (len(shape) - 1))
    L.set_f(f)
    L.iterate(1)
    return L.get_f()[idx]


D2_EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
D2_EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])


@pytest.mark.parametrize("kind", ["WVelocity", "EVelocity"])
def test_d2q9_zou_he_velocity(O, kind):
    rng = np.random.default_rng(8)
    flag = {"WVelocity": 3, "EVelocity": 4}[kind]
    u0 = 0.037
    f = _random_f(rng, 9, 0.05)
    g = _node_after(O, "d2q9_SRT", flag, f, [1.0, 1.0], [0, u0, 0, 0, 0])
    rho = g.sum()
    assert abs((g * D2_EX).sum() - rho * u0) < 1e-15
    assert abs((g * D2_EY).sum()) < 1e-15


@pytest.mark.parametrize("kind", ["WPressure", "EPressure"])
def test_d2q9_zou_he_pressure(O, kind):
    rng = np.random.default_rng(9)
    flag = {"WPressure": 2, "EPressure": 1}[kind]
    f = _random_f(rng, 9, 0.05)
    g = _node_after(O, "d2q9_SRT", flag, f, [1.0, 1.03])
    assert abs(g.sum() - 1.03) < 1e-15
    assert abs((g * D2_EY).sum()) < 1e-15


def test_d2q9_bounce_back_reverses(O):
    rng = np.random.default_rng(10)
    f = _random_f(rng, 9, 0.05)
    g = _node_after(O, "d2q9_SRT", 6, f, [1.0, 1.0])
    opp = [0, 3, 4, 1, 2, 7, 8, 5, 6]
    assert np.array_equal(g, f[opp])


def _d3e():
    M = mrtmat()
    return M[3], M[5], M[7]


@pytest.mark.parametrize("kind,flag", [("WVelocity", 0x18), ("WPressure", 0x10), ("EPressure", 0x04),
                                       ("WPressureL", 0x14)])
def test_d3q19_zou_he(O, kind, flag):
    rng = np.random.default_rng(11)
    ex, ey, ez = _d3e()
    f = _random_f(rng, 19, 0.05)
    uin, rin = 0.031, 1.02
    g = _node_after(O, "d3q19", flag, f, [1.0, uin, rin] + [0] * 8)
    rho = g.sum()
    jx, jy, jz = (g * ex).sum(), (g * ey).sum(), (g * ez).sum()
    assert abs(jy) < 1e-15 and abs(jz) < 1e-15
    if kind == "WVelocity":
        assert abs(jx - rho * uin) < 1e-15
    elif kind == "WPressure":
        assert abs(rho - rin) < 1e-15
    elif kind == "EPressure":
        assert abs(rho - 1.0) < 1e-15
    else:  # limited: pressure inlet unless the velocity would exceed InletVelocity
        assert abs(rho - rin) < 1e-15 or abs(jx - rho * uin) < 1e-15


def test_d3q19_bounce_back_reverses(O):
    rng = np.random.default_rng(12)
    ex, ey, ez = _d3e()
    f = _random_f(rng, 19, 0.05)
    g = _node_after(O, "d3q19", 0x0C, f, [1.0, 0, 1] + [0] * 8)
    opp = [next(j for j in range(19) if ex[j] == -ex[i] and ey[j] == -ey[i] and ez[j] == -ez[i]) for i in range(19)]
    assert np.array_equal(g, f[opp])


def test_d3q19_globals_on_a_periodic_box(O):
    """All-MRT periodic box, one step with Globals: the reference evaluates
    them after `R /= Mw`, so VOLrho = sum(rho)/19, VOLpx = sum(Jx + rho Fx)/10
    and the velocities are (J/10)/(rho/19) (Dynamics.c.Rt:262-289)."""
    rng = np.random.default_rng(13)
    nx, ny, nz = 4, 5, 3
    ex, ey, ez = _d3e()
    base = np.array([1/3] + [1/18] * 6 + [1/36] * 12)
    f = base[:, None, None, None] * (1 + 0.02 * rng.standard_normal((19, nz, ny, nx)))
    F = (1e-4, 0, -5e-5)
    L = O.OracleLattice("d3q19", nx, ny, nz)
    L.set_params([1.3, 0, 1, *F] + [0] * 5)
    flags = np.full((nz, ny, nx), 0x20, np.uint16)
    flags[:, 2, :] |= 1  # XYslice on one row of nodes
    L.set_flags(flags)
    L.set_f(f)
    g = L.iterate(1)
    names = ["Flux"] + [p + s for p in ("XY", "XZ", "YZ") for s in ("vx", "vy", "vz", "rho", "area")] + \
            ["VOL" + s for s in ("vx", "vy", "vz", "px", "py", "pz", "rho", "volume")] + ["Objective", "MaxV"]
    G = dict(zip(names, g))
    mass = f.sum()
    # pulled rho/J per node: streaming permutes populations
    rho = np.zeros((nz, ny, nx))
    J = np.zeros((3, nz, ny, nx))
    for q in range(19):
        s = np.roll(f[q], shift=(ez[q], ey[q], ex[q]), axis=(0, 1, 2))
        rho += s
        J[0] += s * ex[q]; J[1] += s * ey[q]; J[2] += s * ez[q]
    Jn = J + rho[None] * np.array(F)[:, None, None, None]
    u = (Jn / 10) / (rho / 19)
    assert G["VOLvolume"] == nx * ny * nz
    assert abs(G["VOLrho"] - mass / 19) < 1e-13
    assert np.allclose([G["VOLpx"], G["VOLpy"], G["VOLpz"]], Jn.sum(axis=(1, 2, 3)) / 10, rtol=0, atol=1e-15)
    assert np.allclose([G["VOLvx"], G["VOLvy"], G["VOLvz"]], u.sum(axis=(1, 2, 3)), rtol=1e-12, atol=1e-15)
    assert abs(G["MaxV"] - np.sqrt((u ** 2).sum(axis=0)).max()) < 1e-15
    assert G["XYarea"] == nx * nz
    assert abs(G["XYrho"] - rho[:, 2, :].sum() / 19) < 1e-14
    assert abs(G["XYvx"] - u[0][:, 2, :].sum()) < 1e-14
    assert G["XZarea"] == 0 and G["Flux"] == 0 and G["Objective"] == 0


def test_d3q19_objective_adds_zone_weights(O):
    nx, ny, nz = 3, 3, 2
    L = O.OracleLattice("d3q19", nx, ny, nz)
    L.set_params([1.0, 0, 1] + [0] * 8)
    flags = np.full((nz, ny, nx), 0x20 | (1 << 6), np.uint16)  # zone 1
    L.set_flags(flags)
    table = np.zeros((2, 24))
    table[1, 23] = 0.5   # VOLvolumeInObj
    table[1, 16] = 2.0   # VOLvxInObj
    L.set_zone_table(table)
    L.init()
    g = L.iterate(1)
    assert abs(g[24] - (0.5 + 2.0) * nx * ny * nz) < 1e-14


def test_d3q19_poiseuille_halfway_bounce_back(O):
    """Force-driven channel between bounce-back walls on rows 0 and ny-1:
    with the reference's q rate 8(2-w)/(8-w) the steady profile is the exact
    parabola u(y) = g/(2 nu) (y - 1/2)(ny - 3/2 - y) (walls half-way between
    nodes); u taken as J/rho + F/2 (mean of pre- and post-force momentum)."""
    ny, g, tau = 12, 1e-6, 0.8
    nu = (tau - 0.5) / 3
    L = O.OracleLattice("d3q19", 1, ny, 1)
    L.set_params([1 / tau, 0, 1, g, 0, 0] + [0] * 5)
    flags = np.full((1, ny, 1), 0x20, np.uint16)
    flags[0, 0, 0] = flags[0, ny - 1, 0] = 0x0C
    L.set_flags(flags)
    L.init()
    L.iterate(6000)
    U = L.quantity("U")[0, :, 0, 0]
    rho = 1 + 3 * L.quantity("P")[0, :, 0]
    y = np.arange(1, ny - 1)
    u = U[1:-1] + g / 2 / rho[1:-1]
    want = g / (2 * nu) * (y - 0.5) * (ny - 1.5 - y)
    assert np.max(np.abs(u - want)) / want.max() < 1e-9


# ------------------------------------------------ d3q19_lid MovingWall ---
def _couette(O, omega, U, nx=3, ny=10, nz=3, steps=20000):
    import tclb_amd
    m = tclb_amd.Model("d3q19_lid")
    fl = np.full((nz, ny, nx), m.flag("MRT"), np.uint16)
    fl[:, 0, :] = m.flag("Wall")
    fl[:, ny - 1, :] = m.flag("MovingWall")
    L = O.OracleLattice("d3q19_lid", nx, ny, nz, "f64")
    L.set_params([omega, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, U, 0, 0])
    L.set_zone_table(np.zeros((1, 28)))
    L.set_flags(fl)
    L.init()
    L.iterate(steps - 1)
    return L, L.iterate(1)


@pytest.mark.parametrize("omega", [1.0, 1.6])
def test_lid_couette_profile_and_shear(O, omega):
    """Plane Couette flow between a Wall at y=0 and a MovingWall at y=ny-1
    (periodic in x and z): bounce-back puts the walls half-way to the first
    fluid node, so the steady profile is u_x = U (y - 1/2)/(ny - 2) exactly,
    the fluid drags the lid with the shear stress rho nu U/(ny - 2) per lid
    node (nu = (1/omega - 1/2)/3), presses on it with p = rho/3, and the z
    torque about the lattice centre is -(y_lid - y_c) Fx."""
    U, nx, ny, nz = 0.01, 3, 10, 3
    L, g = _couette(O, omega, U, nx, ny, nz)
    u = L.quantity("U")
    y = np.arange(ny)
    want = U * (y - 0.5) / (ny - 2)
    assert np.max(np.abs(u[:, 1:-1, :, 0] - want[None, 1:-1, None])) < 1e-11 * U
    assert np.max(np.abs(u[:, 1:-1, :, 1:])) < 1e-14
    assert np.all(u[:, ny - 1, :, 0] == U)           # MovingWall reports the wall velocity
    nu = (1 / omega - 0.5) / 3
    fx = -nu * U / (ny - 2) * nx * nz
    assert abs(g[24] - fx) < 1e-10 * abs(fx)
    assert abs(g[25] - nx * nz / 3.0) < 1e-10 * nx * nz / 3.0
    assert abs(g[26]) < 1e-14
    assert abs(g[27] - (-(ny - 1 - 0.5 * (ny - 1)) * g[24])) < 1e-10 * abs(g[27])
    assert g[28] == 0.0   # Objective: no InObj weights


def test_lid_at_rest_is_a_wall(O):
    """With zero wall velocity a MovingWall is the reference's Wall: the
    populations of a random case are bit-identical to the d3q19 run with the
    MovingWall nodes replaced by Wall nodes."""
    import tclb_amd
    rng = np.random.default_rng(3)
    nx, ny, nz = 7, 6, 5
    m = tclb_amd.Model("d3q19_lid")
    kinds = np.array([m.flag("MRT")] * 6 + [m.flag("Wall"), m.flag("MovingWall"), m.flag("MovingWall", "MRT")],
                     np.uint16)
    fl = kinds[rng.integers(0, len(kinds), nx * ny * nz)].reshape(nz, ny, nx)
    fw = fl.copy()
    fw[(fl & 0x1C) == 0x1C] = (fw[(fl & 0x1C) == 0x1C] & ~np.uint16(0x1C)) | np.uint16(0x0C)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    f0 = w[:, None, None, None] * (1 + 0.02 * rng.standard_normal((19, nz, ny, nx)))
    p = [1.3, 0, 1, 1e-5, 0, 0, 0, 0, 0, 0, 0]
    a = O.OracleLattice("d3q19_lid", nx, ny, nz, "f64")
    a.set_params(p + [0.0, 0.0, 0.0])
    a.set_zone_table(np.zeros((1, 28)))
    a.set_flags(fl)
    a.init()
    a.set_f(f0)
    b = O.OracleLattice("d3q19", nx, ny, nz, "f64")
    b.set_params(p)
    b.set_zone_table(np.zeros((1, 24)))
    b.set_flags(fw)
    b.init()
    b.set_f(f0)
    ga, gb = a.iterate(7), b.iterate(7)
    assert np.array_equal(a.get_f(), b.get_f())
    assert np.array_equal(ga[:24], gb[:24]) and ga[29] == gb[25]


def test_lid_conserves_mass(O):
    """The moving-wall term sums to zero over the directions, so a closed
    cavity keeps its total mass to round-off."""
    import tclb_amd
    from tclb_amd import cases
    c = cases.d3q19_cavity(n=12, ny=10, nz=8, u_lid=0.05, tau=0.52)
    L = O.OracleLattice("d3q19_lid", c.nx, c.ny, c.nz, "f64")
    L.set_params([1 / 0.52, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0.05, 0, 0])
    L.set_zone_table(np.zeros((1, 28)))
    L.set_flags(c.flags)
    L.init()
    m0 = L.get_f().sum()
    L.iterate(500)
    assert abs(L.get_f().sum() - m0) < 1e-11 * m0


# ------------------------------------------------------------ d2q9 MRT ---
def _d2q9m_basis_from_recipe():
    """The moment matrix of models/flow/d2q9/Dynamics.c.Rt:11-20 restated
    numerically: monomial moments ex^a ey^b ordered by a+b (src/lib/feq.R
    MRT_polyMatrix), weights from the rest equilibrium, Cholesky factor of
    M^-1 diag(1/w) M^-T taken on the reversed order, rows rescaled."""
    U = np.stack([D2_EX, D2_EY], 1)
    p = np.where(U < 0, 2, U)
    p = p[np.argsort(p.sum(1), kind="stable")]
    W = np.stack([np.prod(U.astype(float) ** pp, axis=1) for pp in p], 1)
    req0 = np.array([np.prod([1 / 3 if e == 2 else (0.0 if e == 1 else 1.0) for e in pp]) for pp in p])
    w = req0 @ np.linalg.inv(W)
    A = np.linalg.inv(W) @ np.diag(1 / w) @ np.linalg.inv(W.T)
    r = np.arange(9)[::-1]
    H = np.linalg.cholesky(A[np.ix_(r, r)]).T[np.ix_(r, r)]
    H = H * np.array([1, 3 ** .5 / 3, 3 ** .5 / 3, 2 ** .5, 2 ** .5, 1, 6 ** .5 / 3, 6 ** .5 / 3, 2])[:, None]
    B = W @ H.T
    order = [max(p.sum(1)[np.abs(np.linalg.inv(W) @ B[:, k]) > 1e-10]) for k in range(9)]
    return B, np.array(order), w


def _d2q9m_collide(O, f, params):
    L = O.OracleLattice("d2q9", 1, 1, 1)
    L.set_params(params)
    L.set_zone_table(np.zeros((1, 6)))
    L.set_flags(np.array([0x10], np.uint16))
    L.set_f(np.asarray(f).reshape(9, 1, 1, 1))
    g = L.iterate(1)
    return L.get_f().reshape(9), g


def _feq_std(rho, jx, jy):
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    eu = D2_EX * jx + D2_EY * jy
    return w * (rho + 3 * eu + 4.5 * eu ** 2 / rho - 1.5 * (jx * jx + jy * jy) / rho)


def test_d2q9_mrt_relaxes_each_moment_order_at_its_rate(O):
    """In the basis the recipe produces, the collision keeps rho, adds rho G
    to J, relaxes the non-equilibrium part of the 2nd, 3rd and 4th order
    moments by S2, S3, S4 and rebuilds around the equilibrium at the new J;
    the equilibrium moments are those of the standard second-order feq."""
    B, order, w = _d2q9m_basis_from_recipe()
    assert np.allclose(B, np.round(B), atol=1e-12)            # an integer basis
    assert list(order) == [0, 1, 1, 2, 2, 2, 3, 3, 4]
    assert np.allclose(w, [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, atol=1e-15)
    rng = np.random.default_rng(21)
    S = {2: 0.37, 3: -0.21, 4: 0.55}
    G = (2e-4, -1e-4)
    for _ in range(5):
        f = _random_f(rng, 9, 0.05)
        g, _ = _d2q9m_collide(O, f, [S[2], S[3], S[4], *G])
        m0, m1 = B.T @ f, B.T @ g
        rho, J = m0[0], m0[1:3]
        Jn = J + rho * np.array(G)
        eq0, eq1 = B.T @ _feq_std(rho, *J), B.T @ _feq_std(rho, *Jn)
        assert abs(m1[0] - rho) < 1e-15 and np.allclose(m1[1:3], Jn, atol=1e-15)
        for k in range(3, 9):
            assert abs(m1[k] - (S[order[k]] * (m0[k] - eq0[k]) + eq1[k])) < 1e-14, k


def test_d2q9_mrt_equal_rates_is_bgk(O):
    """S2 = S3 = S4 = 1 - omega: f' = feq(rho, J + rho G) + (1 - omega)(f - feq(rho, J))."""
    rng = np.random.default_rng(22)
    om, G = 1.37, (1e-4, 3e-4)
    for _ in range(5):
        f = _random_f(rng, 9, 0.05)
        g, _ = _d2q9m_collide(O, f, [1 - om] * 3 + list(G))
        rho, jx, jy = f.sum(), (f * D2_EX).sum(), (f * D2_EY).sum()
        want = _feq_std(rho, jx + rho * G[0], jy + rho * G[1]) + (1 - om) * (f - _feq_std(rho, jx, jy))
        assert np.max(np.abs(g - want)) < 1e-15


@pytest.mark.parametrize("side", ["W", "E", "N", "S"])
@pytest.mark.parametrize("kind", ["Velocity", "Pressure"])
def test_d2q9_mrt_zou_he_solves_the_reference_equations(O, side, kind):
    """ZouHeRewrite (src/lib/boundary.R:155-207): the populations entering
    through the side become feq(R) + f_opp - feq_opp(R); R is the solution of
    the velocity equations (sum fs e = V sum fs) or the pressure equations
    (sum fs = 1 + 3P, no tangential momentum) -- solved here with numpy."""
    flags = {"EPressure": 1, "WPressure": 2, "NVelocity": 3, "SVelocity": 4, "WVelocity": 5, "EVelocity": 6}
    name = side + kind
    if name not in flags:
        pytest.skip("the model has no " + name)
    n = {"W": (-1, 0), "E": (1, 0), "N": (0, 1), "S": (0, -1)}[side]
    rng = np.random.default_rng(23)
    f = _random_f(rng, 9, 0.05)
    V, P = (0.031, -0.012), 0.004
    g = _node_after(O, "d2q9", flags[name], f, [0, 0, 0, 0, 0], [V[0], V[1], P, 0, 0, 0])
    inc = (D2_EX * n[0] + D2_EY * n[1]) < 0
    opp = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6])

    def fs_of(R):
        fs = f.copy()
        fs[inc] = (_feq_std(1.0, *R) + f[opp] - _feq_std(1.0, *R)[opp])[inc]
        return fs

    def eqs(R):
        fs = fs_of(R)
        mom = np.array([(fs * D2_EX).sum(), (fs * D2_EY).sum()])
        if kind == "Velocity":
            return mom - np.array(V) * fs.sum()
        a = 0 if n[0] else 1
        e = mom.copy()
        e[a] = fs.sum() - (1 + 3 * P)
        return e

    b = eqs(np.zeros(2))
    A = np.stack([eqs(np.eye(2)[0]) - b, eqs(np.eye(2)[1]) - b], 1)
    R = np.linalg.solve(A, -b)
    assert np.max(np.abs(g - fs_of(R))) < 1e-15
    assert np.array_equal(g[~inc], f[~inc])


def test_d2q9_mrt_symmetry_and_bounce_back(O):
    rng = np.random.default_rng(24)
    f = _random_f(rng, 9, 0.05)
    g = _node_after(O, "d2q9", 7, f, [0] * 5, [0] * 6)          # NSymmetry
    assert np.array_equal(g[[4, 8, 7]], f[[2, 5, 6]]) and np.array_equal(g[[0, 1, 2, 3, 5, 6]], f[[0, 1, 2, 3, 5, 6]])
    g = _node_after(O, "d2q9", 8, f, [0] * 5, [0] * 6)          # SSymmetry
    assert np.array_equal(g[[2, 5, 6]], f[[4, 8, 7]])
    g = _node_after(O, "d2q9", 10, f, [0] * 5, [0] * 6)         # Wall
    assert np.array_equal(g, f[[0, 3, 4, 1, 2, 7, 8, 5, 6]])


def test_d2q9_mrt_inlet_outlet_globals(O):
    """On Outlet (Inlet) nodes CollisionMRT adds Jx/rho^2 to OutletFlux
    (InletFlux) and -(+) Jx/rho^2 ((rho-1)/3 + |J|^2/rho^3/2) to
    PressureLoss, before relaxation (Dynamics.c.Rt:221-236)."""
    rng = np.random.default_rng(25)
    f = _random_f(rng, 9, 0.05)
    rho, jx, jy = f.sum(), (f * D2_EX).sum(), (f * D2_EY).sum()
    loss = jx / rho / rho * ((rho - 1) / 3 + (jx * jx + jy * jy) / rho ** 2 / rho / 2)
    for flag, k, sign in ((0x10 | 0x40, 1, -1), (0x10 | 0x20, 2, 1)):
        L = O.OracleLattice("d2q9", 1, 1, 1)
        L.set_params([0.3, 0, 0, 0, 0])
        L.set_zone_table(np.array([[0, 0, 0, 0.5, 0.25, 0.125]]))
        L.set_flags(np.array([flag], np.uint16))
        L.set_f(f.reshape(9, 1, 1, 1))
        g = L.iterate(1)
        assert abs(g[k] - jx / rho / rho) < 1e-16 and abs(g[0] - sign * loss) < 1e-16
        assert g[3] == 0.5 + [0, 0.25, 0.125][k]     # Objective: the zone's InObj weights


def test_d2q9_mrt_init(O):
    """Init: the standard equilibrium at rho = 1 + 3 Pressure and velocity
    (VelocityX, VelocityY) of the node's zone; Wall nodes at rest, rho = 1."""
    L = O.OracleLattice("d2q9", 2, 1, 1)
    L.set_params([0] * 5)
    L.set_zone_table(np.array([[0.02, -0.01, 0.003, 0, 0, 0]]))
    L.set_flags(np.array([[0x10, 10]], np.uint16))
    L.init()
    f = L.get_f()[:, 0, 0]
    rho = 1 + 3 * 0.003
    assert np.max(np.abs(f[:, 0] - _feq_std(rho, 0.02 * rho, -0.01 * rho))) < 1e-16
    assert np.max(np.abs(f[:, 1] - _feq_std(1.0, 0, 0))) < 1e-16


@pytest.mark.parametrize("tau", [0.6, 0.8])
def test_d2q9_mrt_poiseuille_exact_with_magic_rates(O, tau):
    """A force-driven channel between bounce-back walls: when the 3rd-order
    rate s3 satisfies (1/s2 - 1/2)(1/s3 - 1/2) = 3/16 (s = 1 - S), the MRT
    steady state is the exact parabola with walls half-way between nodes,
    u = g/(2 nu) (y - 1/2)(ny - 3/2 - y) (a rate given to a wrong moment
    order or a misplaced force breaks this)."""
    ny, g = 12, 1e-6
    om = 1 / tau
    nu = (tau - 0.5) / 3
    s3 = 1 / ((3 / 16) / (1 / om - 0.5) + 0.5)
    L = O.OracleLattice("d2q9", 1, ny, 1)
    L.set_params([1 - om, 1 - s3, 0.0, g, 0])
    L.set_zone_table(np.zeros((1, 6)))
    fl = np.full((1, ny, 1), 0x10, np.uint16)
    fl[0, 0, 0] = fl[0, ny - 1, 0] = 10
    L.set_flags(fl)
    L.init()
    L.iterate(40000)
    u = L.quantity("U")[0, 1:-1, 0, 0]
    y = np.arange(1, ny - 1)
    want = g / (2 * nu) * (y - 0.5) * (ny - 1.5 - y)
    assert np.max(np.abs(u - want)) / want.max() < 1e-9


def test_d2q9_drag_is_d2q9_plus_momentum_exchange(O):
    """d2q9_drag evolves exactly as d2q9; on Solid nodes it adds
    2 sum_i f_i e_i of the incoming populations to Drag/Lift (and the zone's
    DragInObj/LiftInObj to Objective)."""
    rng = np.random.default_rng(26)
    nx, ny = 9, 7
    kinds = np.array([0x10] * 6 + [9, 10, 0x10 | 0x20], np.uint16)
    fl = kinds[rng.integers(0, len(kinds), nx * ny)].reshape(1, ny, nx)
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    f0 = w[:, None, None, None] * (1 + 0.03 * rng.standard_normal((9, 1, ny, nx)))
    runs = []
    for model, ncol in (("d2q9", 6), ("d2q9_drag", 8)):
        L = O.OracleLattice(model, nx, ny, 1)
        L.set_params([0.3, 0.1, -0.2, 1e-4, 0])
        L.set_zone_table(np.array([[0, 0, 0, 0.5, 0, 0, 0.25, 0.125][:ncol]]))
        L.set_flags(fl)
        L.init()
        L.set_f(f0)
        runs.append((L.iterate(1), L.get_f()))
    (g0, f_a), (g1, f_b) = runs
    assert np.array_equal(f_a, f_b)
    assert np.array_equal(g0[:3], g1[:3])
    # the pulled populations of the Solid nodes
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1]); ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
    fx = fy = 0.0
    for y, x in zip(*np.nonzero(fl[0] == 9)):
        fin = np.array([f0[q, 0, (y - ey[q]) % ny, (x - ex[q]) % nx] for q in range(9)])
        fx += 2 * (fin * ex).sum(); fy += 2 * (fin * ey).sum()
    assert abs(g1[3] - fx) < 1e-14 and abs(g1[4] - fy) < 1e-14
    nsolid = int(np.count_nonzero(fl == 9))
    assert abs(g1[5] - (g0[3] + nsolid * (0.25 + 0.125))) < 1e-12


# --------------------------------------------------------------- d3q27 ---
E27 = np.array([(i % 3 - 1, i // 3 % 3 - 1, i // 9 - 1) for i in range(27)], float)
W27 = np.array([[8 / 27, 2 / 27, 1 / 54, 1 / 216][int((e != 0).sum())] for e in E27])


def _feq27(rho, J):
    """the textbook second-order D3Q27 equilibrium"""
    u = np.asarray(J, float) / rho
    cu = E27 @ u
    return W27 * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u))


def _f27(rng, amp=0.02):
    rho, J = 1 + 0.01 * rng.standard_normal(), 0.03 * rng.standard_normal(3)
    return _feq27(rho, J) * (1 + amp * rng.standard_normal(27))


def test_d3q27_basis_and_equilibrium(O):
    """The moments are the 27 monomials ex^p ey^q ez^r, p,q,r in {0,1,2},
    by nondecreasing order (rho, then Jx, Jy, Jz); Req M^-1 is the textbook
    second-order equilibrium and the rest weights 8/27, 2/27, 1/54, 1/216."""
    M, order, pw = O.d3q27_basis()
    assert sorted(map(tuple, pw)) == sorted((p, q, r) for p in range(3) for q in range(3) for r in range(3))
    assert list(order) == sorted(order) and list(order) == list(pw.sum(1))
    assert [tuple(v) for v in pw[:4]] == [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    mono = np.array([[np.prod([1.0 if p == 0 else e[d] ** (1 if p == 1 else 2) for d, p in enumerate(c)])
                      for c in pw] for e in E27])
    assert np.array_equal(M, mono)
    assert np.allclose(O.d3q27_feq(1, 0, 0, 0), W27, rtol=0, atol=4e-16)
    rng = np.random.default_rng(270)
    for _ in range(5):
        rho, J = 1 + 0.05 * rng.standard_normal(), 0.05 * rng.standard_normal(3)
        assert np.allclose(O.d3q27_feq(rho, *J), _feq27(rho, J), rtol=0, atol=4e-16)


def test_d3q27_mrt_is_bgk_with_momentum_shift(O):
    """Every moment of order >= 2 relaxes by 1 - omega: f' = feq(rho, J + F)
    + (1 - omega)(f - feq(rho, J)); the flux Globals add (J + F + F/2)/rho
    and their zone weights; a Solid node ends with zero momentum."""
    rng = np.random.default_rng(271)
    omega, F = 1.3, np.array([1e-4, -2e-5, 3e-5])
    zone = [0, 0, 0, 0.5, 0.25, 0.125, 0, 0, 0]
    for nt, solid in ((0x10, False), (0x15, True)):
        f = _f27(rng)
        rho, J = f.sum(), E27.T @ f
        out, g = O.d3q27_node(f, nt, [omega, 0, *F], zone)
        want = _feq27(rho, np.zeros(3) if solid else J + F) + (1 - omega) * (f - _feq27(rho, J))
        assert np.allclose(out, want, rtol=0, atol=1e-15)
        assert np.allclose(g[:3], (J + 1.5 * F) / rho, rtol=1e-14, atol=0)
        assert abs(g[6] - 0.875) < 1e-15 and not g[3:6].any()


def test_d3q27_smagorinsky_rate(O):
    """Smagorinsky nodes relax at 1/tau with tau = (tau0 + sqrt(tau0^2 +
    18 Smag |Pi|))/2, |Pi| the Frobenius norm of the non-equilibrium
    stress sum_i e_a e_b (f_i - feq_i)."""
    rng = np.random.default_rng(272)
    omega, smag = 1.6, 0.17
    f = _f27(rng, 0.05)
    rho, J = f.sum(), E27.T @ f
    fneq = f - _feq27(rho, J)
    Pi = np.einsum("ia,ib,i->ab", E27, E27, fneq)
    tau0 = 1 / omega
    tau = (tau0 + np.sqrt(tau0 ** 2 + 18 * smag * np.sqrt((Pi ** 2).sum()))) / 2
    out, _ = O.d3q27_node(f, 0x50, [omega, smag, 0, 0, 0])
    assert np.allclose(out, _feq27(rho, J) + (1 - 1 / tau) * fneq, rtol=0, atol=1e-15)
    out0, _ = O.d3q27_node(f, 0x50, [omega, 0, 0, 0, 0])
    assert np.allclose(out0, _feq27(rho, J) + (1 - omega) * fneq, rtol=0, atol=1e-15)


def test_d3q27_entropic_rate(O):
    """Stab nodes relax the order > 2 part dh of the non-equilibrium at
    gamma2 = -gamma (ds.P.dh)/(dh.P.dh), ds its order-2 part, P =
    M^-1 diag(1/w) M^-T: numpy's inverse of the monomial matrix and the
    textbook equilibrium give the same update."""
    rng = np.random.default_rng(273)
    omega = 1.2
    M, order, _ = O.d3q27_basis()
    MI = np.linalg.inv(M)
    f = _f27(rng, 0.05)
    rho, J = f.sum(), E27.T @ f
    neq = (f - _feq27(rho, J)) @ M
    ds = np.where(order == 2, neq, 0)
    dh = np.where(order > 2, neq, 0)
    P = MI @ np.diag(1 / W27) @ MI.T
    g = 1 - omega
    g2 = -g * (ds @ P @ dh) / (dh @ P @ dh)
    want = _feq27(rho, J) + (g * ds + g2 * dh) @ MI
    out, _ = O.d3q27_node(f, 0x30, [omega, 0, 0, 0, 0])
    assert np.allclose(out, want, rtol=0, atol=1e-15)
    assert abs(g2 - g) > 0.05          # the stabiliser changed the rate


@pytest.mark.parametrize("nt,sgn,vel", [(8, 1, True), (7, 1, False), (4, -1, False)])
def test_d3q27_zou_he(O, nt, sgn, vel):
    """W/E Zou-He: the entering populations (sign*ex > 0) become
    f_opp + 6 w_i (e_i . J) for one vector J, and the node then has the
    prescribed density (pressure 3P+1) or velocity (rho V along x) and no
    tangential momentum; a velocity inlet's density is A/(1 - V)."""
    rng = np.random.default_rng(274 + nt)
    V, Pr = 0.04, 0.003
    f = _f27(rng)
    out, _ = O.d3q27_node(f, nt, [1.0, 0, 0, 0, 0], [V, Pr, 0])
    sel = sgn * E27[:, 0] > 0
    assert np.array_equal(out[~sel], f[~sel])
    d = out[sel] - f[::-1][sel]                     # opposite of i is 26 - i
    A = 6 * W27[sel, None] * E27[sel]
    Jfit, *_ = np.linalg.lstsq(A, d, rcond=None)
    assert np.allclose(A @ Jfit, d, rtol=0, atol=1e-15)
    rho, Jn = out.sum(), E27.T @ out
    if vel:
        known = f[E27[:, 0] == 0].sum() + 2 * f[sgn * E27[:, 0] < 0].sum()
        assert abs(rho - known / (1 - sgn * V)) < 1e-14
        assert abs(Jn[0] - rho * V) < 1e-15
    else:
        assert abs(rho - (3 * Pr + 1)) < 1e-14
    assert np.allclose(Jn[1:], 0, atol=1e-15)
    assert abs(Jfit[0] - Jn[0]) < 1e-14           # x: the node's own momentum


@pytest.mark.parametrize("nt,axis,sgn", [(1, 1, -1), (2, 1, 1), (3, 2, 1)])
def test_d3q27_symmetry(O, nt, axis, sgn):
    """N/S/I symmetry: the populations with sign*e_axis > 0 become their
    mirror images in the plane normal to the axis, the rest stay."""
    rng = np.random.default_rng(280 + nt)
    f = _f27(rng)
    out, _ = O.d3q27_node(f, nt, [1.0, 0, 0, 0, 0])
    for i in range(27):
        e = E27[i].copy()
        if sgn * e[axis] > 0:
            e[axis] = -e[axis]
            j = int(np.nonzero((E27 == e).all(1))[0][0])
            assert out[i] == f[j]
        else:
            assert out[i] == f[i]


def test_d3q27_wall_bounce_back_and_drag(O):
    """Wall nodes reverse every population; the drag Globals add
    2 sum_i f'_i e_i = -2 J of the incoming populations, and the zone's
    XDragForceInObj.. weights go to Objective."""
    rng = np.random.default_rng(285)
    f = _f27(rng)
    out, g = O.d3q27_node(f, 6, [1.0, 0, 0, 0, 0], [0, 0, 0, 0, 0, 0, 0.5, 0.25, 0.125])
    assert np.array_equal(out, f[::-1])
    assert np.allclose(g[3:6], -2 * (E27.T @ f), rtol=0, atol=1e-15)
    assert not g[:3].any() and abs(g[6] - 0.875) < 1e-15


# ------------------------------------------------------ d3q27_cumulant ---
_E1 = np.array([0, 1, -1])
EC = np.array([(_E1[i % 3], _E1[i // 3 % 3], _E1[i // 9]) for i in range(27)], float)
WC = np.array([[8 / 27, 2 / 27, 1 / 54, 1 / 216][int((e != 0).sum())] for e in EC])


def _feqc(rho, J):
    u = np.asarray(J, float) / rho
    cu = EC @ u
    return WC * rho * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u @ u))


def _cumulants(f, orders):
    """cumulants c_(p,q,r) of f/sum(f): derivatives of log(sum_i f_i exp(t.e_i)) at t = 0
    (sympy, exact rational arithmetic on the given values)"""
    sp = pytest.importorskip("sympy")
    t = sp.symbols("tx ty tz")
    fr = [sp.Rational(float(v)) for v in f]
    M = sum(fr[i] * sp.exp(sum(int(EC[i][a]) * t[a] for a in range(3))) for i in range(27))
    K = sp.log(M)
    out = {}
    for p in orders:
        d = K
        for a in range(3):
            if p[a]:
                d = sp.diff(d, t[a], p[a])
        out[p] = float(d.subs({t[0]: 0, t[1]: 0, t[2]: 0}))
    return out


def test_cumulant_collision_relaxes_the_cumulants(O):
    """CollisionMRT of d3q27_cumulant, read through cumulants computed from
    the moment generating function: mass kept, mean velocity + F, the
    deviatoric second cumulants relaxed by 1 - w0, the trace set to 1 (w1 =
    1), the off-diagonal ones by 1 - w0, the third order by 1 - Omega, and
    the fourth order and above set to zero (GalileanCorrection = 0)."""
    rng = np.random.default_rng(290)
    f = _feqc(1.02, [0.03, -0.02, 0.01]) * (1 + 0.05 * rng.standard_normal(27))
    nu, Omega, F = 0.03, 0.4, np.array([1e-4, -2e-4, 3e-4])
    w0 = 1 / (3 * nu + 0.5)
    out, _ = O.cum_node(f, 0x80, [nu, 0.01, 0.0, *F, Omega])
    orders = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1),
              (2, 1, 0), (1, 1, 1), (0, 1, 2), (2, 2, 0), (2, 1, 1), (1, 2, 2)]
    c0, c1 = _cumulants(f, orders), _cumulants(out, orders)
    assert abs(out.sum() - f.sum()) < 1e-15
    for k, p in enumerate([(1, 0, 0), (0, 1, 0), (0, 0, 1)]):
        assert abs(c1[p] - (c0[p] + F[k])) < 1e-14
    tr0 = c0[(2, 0, 0)] + c0[(0, 2, 0)] + c0[(0, 0, 2)]
    assert abs(c1[(2, 0, 0)] + c1[(0, 2, 0)] + c1[(0, 0, 2)] - 1) < 1e-14
    assert abs((c1[(2, 0, 0)] - c1[(0, 2, 0)]) - (1 - w0) * (c0[(2, 0, 0)] - c0[(0, 2, 0)])) < 1e-14
    assert abs((c1[(2, 0, 0)] - c1[(0, 0, 2)]) - (1 - w0) * (c0[(2, 0, 0)] - c0[(0, 0, 2)])) < 1e-14
    for p in [(1, 1, 0), (1, 0, 1), (0, 1, 1)]:
        assert abs(c1[p] - (1 - w0) * c0[p]) < 1e-14
    for p in [(2, 1, 0), (1, 1, 1), (0, 1, 2)]:
        assert abs(c1[p] - (1 - Omega) * c0[p]) < 1e-14, p
    for p in [(2, 2, 0), (2, 1, 1), (1, 2, 2)]:
        assert abs(c1[p]) < 1e-14, p
    assert abs(tr0 - 1) > 1e-3          # the test state was off equilibrium


def test_cumulant_buffer_nodes_use_nubuffer(O):
    """Any BOUNDARY type (here Buffer, which has no boundary condition) makes
    the collision relax at 1/(3 nubuffer + 0.5)."""
    rng = np.random.default_rng(291)
    f = _feqc(1.0, [0.02, 0.0, -0.01]) * (1 + 0.05 * rng.standard_normal(27))
    a, _ = O.cum_node(f, 0x80 | (1 << 3), [0.05, 0.2, 0.0, 0, 0, 0, 1.0])
    b, _ = O.cum_node(f, 0x80, [0.2, 0.05, 0.0, 0, 0, 0, 1.0])
    assert np.allclose(a, b, rtol=0, atol=1e-16)


def test_cumulant_galilean_correction_vanishes_at_rest(O):
    """The Galilean correction terms are proportional to u^2: at zero
    velocity and force they change nothing; with flow they change only the
    diagonal second-order cumulants (not the mass, the momentum or the
    off-diagonal ones)."""
    rng = np.random.default_rng(292)
    f = _feqc(1.0, [0.0, 0.0, 0.0]) * (1 + 0.05 * rng.standard_normal(27))
    J = EC.T @ f
    f0 = f - WC * 3 * (EC @ J)                     # remove the first moment
    a, _ = O.cum_node(f0, 0x80, [0.05, 0.01, 1.0, 0, 0, 0, 1.0])
    b, _ = O.cum_node(f0, 0x80, [0.05, 0.01, 0.0, 0, 0, 0, 1.0])
    assert np.abs(EC.T @ f0).max() < 1e-16
    assert np.allclose(a, b, rtol=0, atol=1e-16)
    f1 = _feqc(1.0, [0.08, -0.05, 0.03]) * (1 + 0.05 * rng.standard_normal(27))
    a, _ = O.cum_node(f1, 0x80, [0.05, 0.01, 1.0, 0, 0, 0, 1.0])
    b, _ = O.cum_node(f1, 0x80, [0.05, 0.01, 0.0, 0, 0, 0, 1.0])
    assert abs(a.sum() - b.sum()) < 1e-15 and np.allclose(EC.T @ a, EC.T @ b, rtol=0, atol=1e-15)
    ca, cb = (_cumulants(v, [(2, 0, 0), (1, 1, 0), (0, 1, 1)]) for v in (a, b))
    assert abs(ca[(1, 1, 0)] - cb[(1, 1, 0)]) < 1e-14 and abs(ca[(0, 1, 1)] - cb[(0, 1, 1)]) < 1e-14
    assert abs(ca[(2, 0, 0)] - cb[(2, 0, 0)]) > 1e-9


def _longitudinal_decay(O, U, gc, nu=0.01, nx=64, amp=1e-3, steps=400):
    """Decay rate / k^2 of a longitudinal (sound) wave u_x = U + amp sin(kx)
    on a periodic 64x1x1 d3q27_cumulant lattice: the slope of log(|rho_k|^2
    c_s^2 + |u_k|^2) over steps 55-395, u measured in the frame moving at U."""
    L = O.OracleLattice("d3q27_cumulant", nx, 1, 1)
    L.set_params([nu, 0.01, gc, 0, 0, 0, 1.0])
    L.set_zone_table(np.zeros((1, 8)))
    L.set_flags(np.full((1, 1, nx), 0x80, np.uint16))
    L.init()
    x = np.arange(nx)
    f = np.stack([_feqc(1.0, [U + amp * np.sin(2 * np.pi * xi / nx), 0, 0]) for xi in x], axis=1)
    L.set_f(f.reshape(27, 1, 1, nx))
    k = 2 * np.pi / nx
    ts, es = [], []
    for t in range(steps):
        L.iterate(1, globals_last=False)
        if t > 50 and t % 5 == 0:
            g = L.get_f()[:, 0, 0, :]
            rho = g.sum(0)
            ux = (g * EC[:, 0:1]).sum(0) / rho - U
            es.append(np.log(abs(np.fft.fft(rho - 1)[1]) ** 2 / 3 + abs(np.fft.fft(ux)[1]) ** 2))
            ts.append(t)
    return -np.polyfit(ts, es, 1)[0] / 2 / k ** 2


def test_cumulant_galilean_correction_restores_galilean_invariance(O):
    """What the correction terms of Dynamics.c.Rt:405-422 are for (Geier,
    Schoenherr, Pasquali & Krafczyk, Comput. Math. Appl. 70, 507, 2015,
    sec. 4.3): on D3Q27, e_x^3 = e_x leaves the equilibrium's third moment
    short of rho u_x^3, so the viscous normal stress, and with it the decay
    of a longitudinal wave, depends on the mean velocity U as U^2.  Measured
    on the oracle: without the correction the decay rate moves by about
    -4.5 U^2 (relative), a clean U^2 law (x4 from U = 0.05 to 0.1); with it
    the move is at least 20x smaller.  A sign slip, a wrong factor or a
    swapped w0/w1 in the terms leaves a U^2 defect of the uncorrected size."""
    base = {gc: _longitudinal_decay(O, 0.0, gc) for gc in (0.0, 1.0)}
    assert abs(base[0.0] - base[1.0]) <= 1e-6 * base[0.0]   # the terms vanish at U = 0
    d = {(gc, U): _longitudinal_decay(O, U, gc) / base[gc] - 1 for gc in (0.0, 1.0) for U in (0.05, 0.1)}
    assert -0.06 < d[(0.0, 0.1)] < -0.03
    assert 3.6 < d[(0.0, 0.1)] / d[(0.0, 0.05)] < 4.4
    for U in (0.05, 0.1):
        assert abs(d[(1.0, U)]) <= abs(d[(0.0, U)]) / 20, (U, d)


@pytest.mark.parametrize("btype,axis,sgn,vel", [(12, 0, 1, True), (2, 0, 1, True), (8, 0, -1, True),
                                               (11, 0, 1, False), (7, 0, -1, False), (3, 1, -1, True),
                                               (4, 1, 1, True), (5, 1, -1, False), (6, 1, 1, False)])
def test_cumulant_zou_he(O, btype, axis, sgn, vel):
    """Zou-He on the W/E and S/N faces: entering populations f_opp +
    6 w_i (e_i . J) for one J; density 3P+1 or velocity V along the normal,
    no tangential momentum."""
    rng = np.random.default_rng(300 + btype)
    V, Pr = 0.03, -0.002
    f = _feqc(1.0, 0.02 * rng.standard_normal(3)) * (1 + 0.02 * rng.standard_normal(27))
    out, _ = O.cum_node(f, btype << 3, [0.1, 0.1, 1.0, 0, 0, 0, 1.0], [V, Pr, 0])
    sel = sgn * EC[:, axis] > 0
    opp = [int(np.nonzero((EC == -EC[i]).all(1))[0][0]) for i in range(27)]
    assert np.array_equal(out[~sel], f[~sel])
    d = out[sel] - f[opp][sel]
    A = 6 * WC[sel, None] * EC[sel]
    Jfit, *_ = np.linalg.lstsq(A, d, rcond=None)
    assert np.allclose(A @ Jfit, d, rtol=0, atol=1e-15)
    rho, Jn = out.sum(), EC.T @ out
    if vel:
        assert abs(Jn[axis] - rho * V) < 1e-15
    else:
        assert abs(rho - (3 * Pr + 1)) < 1e-14
    assert np.allclose(np.delete(Jn, axis), 0, atol=1e-15)


def test_cumulant_symmetry_and_body_drag(O):
    """N/S symmetry (ADDITIONALS) mirror the populations entering through the
    plane; a Wall node reverses every population and, marked Body, adds
    -(p_after - p_before) = 2 J to Drag (x), Lateral (y), Lift (z)."""
    rng = np.random.default_rng(310)
    f = _feqc(1.0, 0.02 * rng.standard_normal(3)) * (1 + 0.02 * rng.standard_normal(27))
    for nt, sgn in ((1, -1), (2, 1)):
        out, _ = O.cum_node(f, nt, [0.1, 0.1, 1.0, 0, 0, 0, 1.0])
        for i in range(27):
            e = EC[i].copy()
            if sgn * e[1] > 0:
                e[1] = -e[1]
                assert out[i] == f[int(np.nonzero((EC == e).all(1))[0][0])]
            else:
                assert out[i] == f[i]
    zone = [0, 0, 0, 0, 0, 0.5, 0.25, 0.125, 0, 0, 0, 0]
    out, g = O.cum_node(f, (10 << 3) | 0x4, [0.1, 0.1, 1.0, 0, 0, 0, 1.0], zone)
    J = EC.T @ f
    opp = [int(np.nonzero((EC == -EC[i]).all(1))[0][0]) for i in range(27)]
    assert np.array_equal(out, f[opp])
    assert np.allclose(g[[2, 4, 3]], 2 * J, rtol=0, atol=1e-15)
    assert abs(g[9] - 0.875) < 1e-15
    _, g = O.cum_node(f, 10 << 3, [0.1, 0.1, 1.0, 0, 0, 0, 1.0], zone)
    assert not g.any()


# ------------------------------------------------------------- autosym ---
D2E = np.array([(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)])
D2W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def _mirror_map(axis):
    """q -> the density with the velocity component on `axis` (0 x, 1 y) negated"""
    out = []
    for e in D2E:
        m = e.copy()
        m[axis] = -m[axis]
        out.append(int(np.nonzero((D2E == m).all(1))[0][0]))
    return np.array(out)


def _extend(f, axis, kind):
    """the periodic mirror extension of f [9, 1, ny, nx] across its two ends
    along `axis` (0 x, 1 y): planes through the end nodes (kind 1, period
    2(n-1)) or half a node beyond them (kind 2, period 2n)"""
    ax = 3 - axis
    mp = _mirror_map(axis)
    n = f.shape[ax]
    idx = list(range(n)) + (list(range(n - 2, 0, -1)) if kind == 1 else list(range(n - 1, -1, -1)))
    g = np.take(f, idx, axis=ax).copy()
    back = np.take(g, range(n, len(idx)), axis=ax)[mp]
    sl = [slice(None)] * 4
    sl[ax] = slice(n, len(idx))
    g[tuple(sl)] = back
    return g


def _symmetrize_ends(f, axis):
    ax = 3 - axis
    mp = _mirror_map(axis)
    for end in (0, f.shape[ax] - 1):
        sl = [slice(None)] * 4
        sl[ax] = end
        v = f[tuple(sl)]
        f[tuple(sl)] = 0.5 * (v + v[mp])
    return f


@pytest.mark.parametrize("kind", [1, 2])
@pytest.mark.parametrize("axes", [(1,), (0,), (0, 1)])
def test_autosym_equals_the_mirror_extended_lattice(O, kind, axes):
    """d2q9_autosym (planes through the end nodes) and d2q9_autosym2 (planes
    half a node beyond them): a lattice with Symmetry*_minus on its first
    and *_plus on its last row/column evolves exactly like the part of the
    periodic mirror-extended lattice it covers, along y, x or both
    (corners: X then Y)."""
    model = "d2q9_autosym" if kind == 1 else "d2q9_autosym2"
    nx, ny = 7, 6
    rng = np.random.default_rng(320 + 10 * kind + len(axes) + axes[0])
    f = D2W[:, None, None, None] * (1 + 0.05 * rng.standard_normal((9, 1, ny, nx)))
    if kind == 1:
        for a in axes:
            f = _symmetrize_ends(f, a)
    flags = np.full((1, ny, nx), 0x10, np.uint16)            # MRT
    for a in axes:
        sh = 7 if a == 0 else 9                                 # SYMX / SYMY bits
        if a == 0:
            flags[0, :, 0] |= 2 << sh; flags[0, :, -1] |= 1 << sh
        else:
            flags[0, 0, :] |= 2 << sh; flags[0, -1, :] |= 1 << sh
    full = f
    for a in axes:
        full = _extend(full, a, kind)
    params = [1 - 1 / (3 * 0.05 + 0.5), 0.3, -0.2, 0, 0]
    runs = []
    for g, fl in ((f, flags), (full, np.full(full.shape[1:], 0x10, np.uint16))):
        L = O.OracleLattice(model, g.shape[3], g.shape[2], 1)
        L.set_params(params)
        L.set_zone_table(np.zeros((1, 8)))
        L.set_flags(fl)
        L.init()
        L.set_f(g)
        L.iterate(25, globals_last=False)
        runs.append(L.get_f())
    half, ext = runs
    got = ext[:, :, :ny, :nx]
    assert np.max(np.abs(half - got)) < 1e-14
    assert np.max(np.abs(half - f)) > 1e-4          # the state did evolve


def test_autosym_models_contract(O):
    """Both autosym builds add the groups SYMX and SYMY after d2q9's, with
    <name>X_plus = 1, X_minus = 2 in each (src/conf.R:433-441), and move
    the zone to bit 11."""
    import tclb_amd
    for model, name in (("d2q9_autosym", "Symmetry"), ("d2q9_autosym2", "SymmetryEdge")):
        m = tclb_amd.Model(model)
        assert m.nodetype(name + "X_plus") == (0x80, 0x180)
        assert m.nodetype(name + "X_minus") == (0x100, 0x180)
        assert m.nodetype(name + "Y_plus") == (0x200, 0x600)
        assert m.nodetype(name + "Y_minus") == (0x400, 0x600)
        assert m.group("SETTINGZONE") == (0xF800, 11)
        d = tclb_amd.Model("d2q9")
        assert m.settings == d.settings and m.global_names == d.global_names


def test_cumulant_smag_build(O):
    """d3q27_cumulant_SMAG: with Smag = 0 it is the plain model; otherwise
    the deviatoric second cumulants relax by 1 - w0 with w0 = 1/(tau0 +
    tau_t), tau_t = (sqrt(tau0^2 + 18 Smag^2 q) - tau0)/2 and q the norm of
    the second cumulants' deviation from 1/3 (off-diagonal twice), taken
    from cumulants computed independently of the oracle."""
    rng = np.random.default_rng(330)
    f = _feqc(1.01, [0.02, -0.01, 0.015]) * (1 + 0.08 * rng.standard_normal(27))
    nu, smag = 0.02, 0.3
    base, _ = O.cum_node(f, 0x80, [nu, 0.01, 0.0, 0, 0, 0, 1.0])
    zero, _ = O.cum_node(f, 0x80, [nu, 0.01, 0.0, 0, 0, 0, 1.0, 0.0], smag=True)
    assert np.array_equal(base, zero)
    out, _ = O.cum_node(f, 0x80, [nu, 0.01, 0.0, 0, 0, 0, 1.0, smag], smag=True)
    ords = [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
    c0, c1 = _cumulants(f, ords), _cumulants(out, ords)
    q = np.sqrt(sum((c0[p] - 1 / 3) ** 2 for p in ords[:3]) + 2 * sum(c0[p] ** 2 for p in ords[3:]))
    tau0 = 3 * nu + 0.5
    w0 = 1 / (tau0 + 0.5 * (np.sqrt(tau0 ** 2 + 18 * smag ** 2 * q) - tau0))
    assert abs(w0 - 1 / tau0) > 1e-3
    assert abs((c1[(2, 0, 0)] - c1[(0, 2, 0)]) - (1 - w0) * (c0[(2, 0, 0)] - c0[(0, 2, 0)])) < 1e-14
    assert abs(c1[(1, 1, 0)] - (1 - w0) * c0[(1, 1, 0)]) < 1e-14
    buf, _ = O.cum_node(f, 0x80 | (1 << 3), [nu, 0.01, 0.0, 0, 0, 0, 1.0, smag], smag=True)
    plain_buf, _ = O.cum_node(f, 0x80 | (1 << 3), [nu, 0.01, 0.0, 0, 0, 0, 1.0])
    assert np.array_equal(buf, plain_buf)       # buffer nodes keep nubuffer


# ------------------------------------------------ interpolated bounce-back ---
NO_CUT, CUT_MAX = 65535, 65000


def _opp(i):
    return int(np.nonzero((EC == -EC[i]).all(1))[0][0])


def test_ib_halfway_cut_is_plain_bounce_back(O):
    """A wall halfway along every link (cut = CUT_MAX/2, p = 1) gives the
    halfway bounce-back: each incoming population is the node's own
    post-collision population of the opposite direction, whatever streamed
    in, and the rest population stays."""
    rng = np.random.default_rng(340)
    f = _feqc(1.0, [0.02, -0.01, 0.01]) * (1 + 0.05 * rng.standard_normal(27))
    old = _feqc(1.0, [0.01, 0.02, -0.01]) * (1 + 0.05 * rng.standard_normal(27))
    out, g = O.cum_ib(f, 0x100 | 0x4, np.full(26, CUT_MAX // 2), old)
    assert out[0] == f[0]
    for i in range(1, 27):
        assert out[_opp(i)] == old[i]
    assert not g.any()          # (f_b - fa) vanishes link by link


@pytest.mark.parametrize("cut", [0, 1, 9750, 32499, 32500, 32501, 48100, 64999, 65000])
def test_ib_single_link_is_bouzidi(O, cut):
    """One cut link: the population coming back along it follows Bouzidi,
    Firdaouss & Lallemand (Phys. Fluids 13, 3452 (2001)), linear
    interpolation with the wall at the fraction q of the link:
      q < 1/2:  f_opp(x) = 2q f*_i(x) + (1-2q) f*_i(x - e_i)
      q >= 1/2: f_opp(x) = f*_i(x)/(2q) + (2q-1)/(2q) f*_opp(x)
    (f* post-collision; f*_i(x - e_i) is what streamed in), every other
    population is untouched, and on a Body node (f_opp - f*_i) e_i goes to
    Drag, Lateral and Lift."""
    rng = np.random.default_rng(341 + cut)
    q = cut / CUT_MAX
    zone = [0, 0, 0, 0, 0, 0.5, 0.25, 0.125, 0, 0, 0, 0]
    for i in range(1, 27):
        f = _feqc(1.0, [0.02, -0.01, 0.01]) * (1 + 0.05 * rng.standard_normal(27))
        old = _feqc(1.0, [0.01, 0.02, -0.01]) * (1 + 0.05 * rng.standard_normal(27))
        c = np.full(26, NO_CUT)
        c[i - 1] = cut
        b = _opp(i)
        want = 2 * q * old[i] + (1 - 2 * q) * f[i] if q < 0.5 else old[i] / (2 * q) + (2 * q - 1) / (2 * q) * old[b]
        out, g = O.cum_ib(f, 0x100 | 0x4, c, old, zone)
        assert abs(out[b] - want) <= 4e-16, (i, out[b], want)
        assert np.array_equal(np.delete(out, b), np.delete(f, b))
        F = (out[b] - old[i]) * EC[i]
        assert np.allclose(g[[2, 4, 3]], F, rtol=0, atol=1e-17)
        assert abs(g[9] - 0.875) < 1e-15
        _, g = O.cum_ib(f, 0x100, c, old, zone)     # not a Body node: no force
        assert not g.any()


def test_ib_links_are_taken_in_order(O):
    """With both directions of a line cut, the lower-numbered direction is
    taken first (Dynamics.c.Rt:190-214 loops over the densities in order),
    so the later one interpolates on what the first wrote when its wall is
    nearer than half a link."""
    rng = np.random.default_rng(350)
    f = _feqc(1.0, [0.02, -0.01, 0.01]) * (1 + 0.05 * rng.standard_normal(27))
    old = _feqc(1.0, [0.01, 0.02, -0.01]) * (1 + 0.05 * rng.standard_normal(27))
    cuts = rng.integers(0, CUT_MAX + 1, 26)
    cuts[rng.random(26) < 0.3] = NO_CUT
    out, _ = O.cum_ib(f, 0x100, cuts, old)
    want = f.copy()
    for i in range(1, 27):            # the reference's loop, written out
        if cuts[i - 1] == NO_CUT:
            continue
        p = cuts[i - 1] * (2.0 / CUT_MAX)
        b = _opp(i)
        want[b] = old[i] / p + old[b] * (p - 1) / p if p > 1 else old[i] * p + want[i] * (1 - p)
    assert np.allclose(out, want, rtol=0, atol=1e-16)


def test_ib_poiseuille_walls_off_the_grid(O):
    """Force-driven flow between two planes at y = 2.3 and 19.6 (between
    nodes), their cut distances taken from the geometry of the planes: the
    steady profile is the parabola g/(2 nu) (y - ya)(yb - y) with the walls
    where they are, to 0.2 % of its peak; the same run with each cut
    measured from the wrong end of its link (q -> 1 - q) misses by several
    per cent."""
    import tclb_amd
    m = tclb_amd.Model("d3q27_cumulant_IB")
    ny, ya, yb, g, nu = 24, 2.3, 19.6, 1e-6, 1 / 6
    cuts = np.full((26, 1, ny, 1), NO_CUT, np.uint16)
    for y in range(ny):
        for d in range(1, 27):
            ey = EC[d][1]
            r = [(w - y) / ey for w in (ya, yb, ya + ny, yb - ny)] if ey else []
            r = [v for v in r if 0 <= v <= 1]
            if r:
                cuts[d - 1, 0, y, 0] = int(min(r) * CUT_MAX)
    flags = np.full((1, ny, 1), m.flag("MRT"), np.uint16)
    flags[0, :, 0] |= np.where((cuts[:, 0, :, 0] != NO_CUT).any(0), m.flag("IB"), 0).astype(np.uint16)
    assert (flags & m.flag("IB")).astype(bool).sum() == 4     # the nodes either side of each wall

    def peak_error(c):
        L = O.OracleLattice("d3q27_cumulant_IB", 1, ny, 1)
        L.set_params([nu, 0.01, 1.0, g, 0, 0, 1.0])
        L.set_zone_table(np.zeros((1, 32)))
        L.set_flags(flags)
        L.set_cuts(c)
        L.init()
        L.iterate(10000, False)
        ux = L.quantity("U").reshape(-1, 3)[:, 0]
        y = np.arange(ny)
        want = g / (2 * nu) * (y - ya) * (yb - y)
        inside = (y > ya) & (y < yb)
        return np.abs(ux - want)[inside].max() / want.max()

    assert peak_error(cuts) < 2e-3
    wrong = np.where(cuts != NO_CUT, CUT_MAX - cuts, cuts).astype(np.uint16)
    assert peak_error(wrong) > 3e-2


def test_ib_contract():
    """d3q27_cumulant_IB and _IB_SMAG (src/models.R:54-66 names option
    builds model_<options>): NodeType IB in its own group HO_BOUNDARY, after
    COLLISION in the alphabetical packing, so the zone moves up one bit; the
    rest of the model is the plain (or SMAG) one."""
    import tclb_amd
    for name, base in (("d3q27_cumulant_IB", "d3q27_cumulant"), ("d3q27_cumulant_IB_SMAG", "d3q27_cumulant_SMAG")):
        m, b = tclb_amd.Model(name), tclb_amd.Model(base)
        assert m.nodetype("IB") == (0x100, 0x100)
        assert m.group("HO_BOUNDARY") == (0x100, 8)
        assert m.group("SETTINGZONE")[1] == 9 and b.group("SETTINGZONE")[1] == 8
        for t in ("Wall", "MRT", "Body", "NSymmetry", "WVelocity", "Buffer"):
            assert m.nodetype(t) == b.nodetype(t)
        assert m.settings == b.settings and m.global_names == b.global_names


# ------------------------------------------------------------ averages ---
def test_avg_build_accumulates_the_post_collision_state(O):
    """d3q27_cumulant_AVG on a periodic box of collision, buffer and wall
    nodes, checked against sums kept in numpy from the states alone: each
    step adds, on collision nodes, P and u (0 on Wall nodes) of the
    post-collision populations (the stored snapshot) and u_i u_j, and
    dxu^2, dyv^2, dzw^2 from the second cumulants of the pulled populations
    (after the wall's bounce-back), taken here as central moments; the
    quantities divide by the iterations since the last reset, and a reset
    starts the sums again."""
    import tclb_amd
    m = tclb_amd.Model("d3q27_cumulant_AVG")
    assert m.averages == ["avgP", "varUX", "varUY", "varUZ", "varUXUY", "varUXUZ", "varUYUZ", "avgdxu2",
                          "avgdyv2", "avgdzw2", "avgUX", "avgUY", "avgUZ"]
    assert m.quantities == ["P", "U", "Solid", "KinE", "ReStr", "Dissipation", "avgU", "varU", "averageP"]
    nx, ny, nz = 6, 5, 4
    rng = np.random.default_rng(360)
    kinds = np.array([m.flag("MRT")] * 6 + [m.flag("Buffer", "MRT"), m.flag("Wall", "MRT"), m.flag("Wall")],
                     np.uint16)
    flags = kinds[rng.integers(0, len(kinds), nx * ny * nz)].reshape(nz, ny, nx)
    nu, nub, F = 0.04, 0.1, np.array([2e-5, -1e-5, 3e-6])
    L = O.OracleLattice("d3q27_cumulant_AVG", nx, ny, nz)
    L.set_params([nu, nub, 1.0, *F, 0.8])
    L.set_zone_table(np.zeros((1, 32)))
    L.set_flags(flags)
    L.init()
    f0 = np.stack([_feqc(1.0, 0.02 * rng.standard_normal(3)) for _ in range(nx * ny * nz)], 1)
    f0 = (f0 * (1 + 0.03 * rng.standard_normal(f0.shape))).reshape(27, nz, ny, nx)
    L.set_f(f0)
    coll = (flags & m.flag("MRT")) != 0
    wall = (flags & 0x78) == m.flag("Wall")
    bnd = (flags & 0x78) != 0
    opp = [_opp(i) for i in range(27)]

    def pulled(f):
        g = np.empty_like(f)
        for i in range(27):
            g[i] = np.roll(f[i], shift=(int(EC[i][2]), int(EC[i][1]), int(EC[i][0])), axis=(0, 1, 2))
        return np.where(wall[None], g[opp], g)          # Wall: bounce-back before the collision

    def run(steps):
        S = {k: np.zeros((nz, ny, nx)) for k in O.AVERAGES}
        for _ in range(steps):
            p = pulled(L.get_f())
            m0 = p.sum(0)
            mu = np.einsum("i...,ia->a...", p, EC) / m0
            c2 = np.einsum("i...,ia->a...", p, EC ** 2) / m0 - mu ** 2      # c200, c020, c002
            w0 = np.where(bnd, 1 / (3 * nub + 0.5), 1 / (3 * nu + 0.5))
            dxu = -w0 / 2 * (2 * c2[0] - c2[1] - c2[2]) - 0.5 * (c2.sum(0) - 1)
            dyv = dxu + 1.5 * w0 * (c2[0] - c2[1])
            dzw = dxu + 1.5 * w0 * (c2[0] - c2[2])
            L.iterate(1, False)
            f = L.get_f()
            rho = f.sum(0)
            u = (np.einsum("i...,ia->a...", f, EC) + F[:, None, None, None] / 2) / rho
            u = np.where(wall[None], 0.0, u)
            add = {"avgP": (rho - 1) / 3, "avgUX": u[0], "avgUY": u[1], "avgUZ": u[2], "varUX": u[0] ** 2,
                   "varUY": u[1] ** 2, "varUZ": u[2] ** 2, "varUXUY": u[0] * u[1], "varUXUZ": u[0] * u[2],
                   "varUYUZ": u[1] * u[2], "avgdxu2": dxu ** 2, "avgdyv2": dyv ** 2, "avgdzw2": dzw ** 2}
            for k, v in add.items():
                S[k] += np.where(coll, v, 0.0)
        return S

    for steps in (7, 4):
        if steps == 4:
            L.reset_average()
        S = run(steps)
        got = L.averages()
        for k in O.AVERAGES:
            assert np.allclose(got[k], S[k], rtol=1e-11, atol=1e-16), k
        n = float(steps)
        mu = np.stack([S["avgUX"], S["avgUY"], S["avgUZ"]], -1) / n
        var = np.stack([S["varUX"], S["varUY"], S["varUZ"]], -1) / n - mu ** 2
        rs = np.stack([S["varUYUZ"] / n - mu[..., 1] * mu[..., 2], S["varUXUZ"] / n - mu[..., 0] * mu[..., 2],
                       S["varUXUY"] / n - mu[..., 0] * mu[..., 1]], -1)
        diss = 0
        for k, (a, ax) in enumerate((("avgUX", 2), ("avgUY", 1), ("avgUZ", 0))):
            g = np.roll(S[a], -1, ax) - np.roll(S[a], 1, ax)
            diss = diss + nu * (S[("avgdxu2", "avgdyv2", "avgdzw2")[k]] / n - g * g / (4 * n * n))
        want = {"avgU": mu, "varU": var, "ReStr": rs, "KinE": 0.5 * var.sum(-1), "averageP": S["avgP"] / n,
                "Dissipation": diss}
        for q, w in want.items():
            assert np.allclose(L.quantity(q), w, rtol=1e-10, atol=1e-18), q


# ------------------------------------------------------- d2q9 bc builds ---
def _d2q9_lattice(O, model, flags, settings=(), zonal=None, f=None):
    import tclb_amd
    from tclb_amd import cases
    ny, nx = flags.shape
    c = cases.Case("t", model, nx, ny, 1, flags.reshape(1, ny, nx), list(settings), zonal or [],
                   init_f=None if f is None else f.reshape(9, 1, ny, nx))
    return O.lattice_for_case(c)


def _d2q9_random_f(rng, ny, nx):
    u = 0.02 * rng.standard_normal((2, ny, nx))
    cu = np.einsum("ia,a...->i...", D2E, u)
    return D2W[:, None, None] * (1 + 3 * cu + 4.5 * cu ** 2 - 1.5 * (u ** 2).sum(0)) * \
        (1 + 0.01 * rng.standard_normal((9, ny, nx)))


def test_bc_build_forces_like_gravitation(O):
    """d2q9_bc: BC on a node without a BOUNDARY type is a body force added to
    the gravitation (J += (G + BC) rho) and adds BC/2 to U, so uniform BC on a
    periodic box of MRT nodes is d2q9 with that much more gravitation."""
    import tclb_amd
    m = tclb_amd.Model("d2q9_bc")
    assert m.params == ["BC[0]", "BC[1]"] and m.extras == ["BC[0]", "BC[1]"]
    rng = np.random.default_rng(370)
    ny, nx = 9, 13
    flags = np.full((ny, nx), m.flag("MRT"), np.uint16)
    f = _d2q9_random_f(rng, ny, nx)
    G, B = (2e-5, -1e-5), (3e-5, 4e-5)
    st = [("Viscosity", 0.05), ("S3", 0.1), ("GravitationX", G[0]), ("GravitationY", G[1])]
    a = _d2q9_lattice(O, "d2q9_bc", flags, st, f=f)
    a.set_extra(np.stack([np.full((1, ny, nx), B[0]), np.full((1, ny, nx), B[1])]))
    b = _d2q9_lattice(O, "d2q9", flags, [("Viscosity", 0.05), ("S3", 0.1), ("GravitationX", G[0] + B[0]),
                                          ("GravitationY", G[1] + B[1])], f=f)
    a.iterate(7, False)
    b.iterate(7, False)
    assert np.allclose(a.get_f(), b.get_f(), rtol=0, atol=2e-17)
    assert np.allclose(a.quantity("U"), b.quantity("U"), rtol=0, atol=1e-17)
    assert np.array_equal(a.extra()[:, 0], np.stack([np.full((ny, nx), B[0]), np.full((ny, nx), B[1])]))


def test_bc_build_walls_inlets_and_init(O):
    """d2q9_bc at Init: BC = J/rho = the zonal velocity on BOUNDARY nodes, 0
    elsewhere.  A Wall node then records 2 sum f e of the populations it
    reverses (the pulled ones, before the swap), or 0 when they sum to
    exactly 1 + 3 Pressure; a WVelocity node imposes BC as its velocity, so
    setting BC there is d2q9 with that zonal velocity; boundary nodes take
    no gravitation."""
    import tclb_amd
    m = tclb_amd.Model("d2q9_bc")
    rng = np.random.default_rng(371)
    ny, nx = 8, 12
    MRT, Wall = m.flag("MRT"), m.flag("Wall")
    flags = np.full((ny, nx), MRT, np.uint16)
    flags[0, :] = flags[-1, :] = Wall
    flags[1:-1, 0] = m.flag("WVelocity", "MRT", zone=1)
    zon = [("VelocityX", 1, 0.02), ("VelocityY", 1, -0.004)]
    L = _d2q9_lattice(O, "d2q9_bc", flags, [("Viscosity", 0.05)], zon)
    e = L.extra()[:, 0]
    inlet = flags == m.flag("WVelocity", "MRT", zone=1)
    assert np.all(e[0][inlet] == 0.02) and np.all(e[1][inlet] == -0.004)   # (u rho)/rho, exact here
    assert not e[:, ~inlet].any()
    # walls: the momentum they reverse
    f = _d2q9_random_f(rng, ny, nx)
    L.set_f(f.reshape(9, 1, ny, nx))
    pulled = np.stack([np.roll(f[i], shift=(D2E[i][1], D2E[i][0]), axis=(0, 1)) for i in range(9)])
    L.iterate(1, False)
    e = L.extra()[:, 0]
    for y in (0, ny - 1):
        for x in range(nx):
            p = pulled[:, y, x]
            want = 2 * (D2E.T @ p) if p.sum() != 1.0 else np.zeros(2)
            assert np.allclose(e[:, y, x], want, rtol=1e-14, atol=1e-18), (y, x)
    # a wall node whose pulled populations sum to exactly 1 + 3 Pressure (= 1) records 0
    f2 = f.copy()
    for i in range(9):   # the sources of node (0, 4)
        sy, sx = (0 - D2E[i][1]) % ny, (4 - D2E[i][0]) % nx
        f2[i, sy, sx] = [0.5, 0.0625, 0.0625, 0.0625, 0.0625, 0.0625, 0.0625, 0.0625, 0.0625][i]
    L.set_f(f2.reshape(9, 1, ny, nx))
    L.iterate(1, False)
    assert not L.extra()[:, 0, 0, 4].any()
    # inlets read BC as their velocity
    bx, by = 0.03, 0.007
    a = _d2q9_lattice(O, "d2q9_bc", flags, [("Viscosity", 0.05), ("GravitationX", 1e-5)], zon, f=f)
    e = a.extra()
    e[0][:, inlet] = bx
    e[1][:, inlet] = by
    a.set_extra(e)
    b = _d2q9_lattice(O, "d2q9", flags, [("Viscosity", 0.05), ("GravitationX", 1e-5)],
                      [("VelocityX", 1, bx), ("VelocityY", 1, by)], f=f)
    a.iterate(1, False)
    b.iterate(1, False)
    fa, fb = a.get_f()[:, 0], b.get_f()[:, 0]
    inner = flags == MRT
    assert np.allclose(fa[:, inner], fb[:, inner], rtol=0, atol=2e-17)   # fluid: same force, same inlet
    # on the inlet nodes the plain model adds G rho to J, the bc build does not
    rho = fb[:, inlet].sum(0)
    assert np.allclose((D2E[:, 0] @ fb[:, inlet]) - (D2E[:, 0] @ fa[:, inlet]), 1e-5 * rho, rtol=1e-9, atol=0)


# ------------------------------------------- stochastic rounding (16-bit) ---
def _sr_neighbours_f16(x):
    """the binary16 values either side of float32 x (numpy's float16)"""
    h = np.float16(x)
    if float(h) == float(x):
        return float(h), float(h)
    a = np.nextafter(h, np.float16(np.inf if float(h) < float(x) else -np.inf))
    lo, hi = sorted([float(h), float(a)], key=abs)
    return lo, hi


@pytest.mark.parametrize("x", [1.0e-3 + 3.1e-7, -2.7e-3 - 1.3e-7, 3.0e-6 + 1.1e-8, -4.5e-7, 0.15 + 1e-5, 1.0 / 3.0,
                               2.3e-8, -4.1e-8, 6.09e-5, 6.13e-5, -1234.567])
def test_stochastic_rounding_f16_is_unbiased(O, x):
    """The oracle's stochastic binary16 rounding (the engine's option for the
    16-bit formats): every result is one of the two neighbours lo, hi of x
    (normal and subnormal binary16 alike, either side of the smallest normal
    2^-14, and below the smallest subnormal 2^-24, where lo = 0), the mean
    over the hash's bits is x
    within 4.5 standard errors, and the variance is (|hi| - |x|)(|x| - |lo|),
    the two-point distribution's.  Round-to-nearest would fail the mean."""
    L = O.lib()
    xf = float(np.float32(x))
    lo, hi = _sr_neighbours_f16(xf)
    n = 40000
    vals = np.array([L.orc_f16_to_f32(L.orc_f32_to_f16_sr(xf, L.orc_sr_bits(k, 3, 17, 5))) for k in range(n)])
    assert set(np.unique(vals)) <= {lo, hi}
    var = (abs(hi) - abs(xf)) * (abs(xf) - abs(lo))
    assert abs(vals.mean() - xf) <= 4.5 * np.sqrt(var / n) + 1e-12 * abs(xf)
    assert abs(vals.var() - var) <= 0.05 * var
    nearest = float(np.float16(xf))
    assert abs(nearest - xf) > 10 * np.sqrt(var / n)   # the bias stochastic rounding removes
    # over all 2^13 values of the 13 bits the rounding takes, hi comes out
    # floor(2^13 (|x| - |lo|)/(|hi| - |lo|)) times (exact where the float's
    # bits below binary16's spacing are 13, within 2^-13 of it elsewhere)
    ups = sum(abs(L.orc_f16_to_f32(L.orc_f32_to_f16_sr(xf, t << 19))) == abs(hi) for t in range(8192))
    assert ups == int(np.floor(8192 * (abs(xf) - abs(lo)) / (abs(hi) - abs(lo))))


@pytest.mark.parametrize("x", [1.0e-3 + 3.1e-7, -2.7e-3 - 1.3e-7, 0.15 + 1e-5])
def test_stochastic_rounding_bf16_is_unbiased(O, x):
    """The same for bfloat16: lo and hi are x's bits truncated and truncated
    plus one unit in the last kept place."""
    L = O.lib()
    xf = np.float32(x)
    u = int(xf.view(np.uint32))
    lo = float(np.uint32(u & 0xFFFF0000).view(np.float32))
    hi = float(np.uint32((u & 0xFFFF0000) + 0x10000).view(np.float32))
    n = 40000
    vals = np.array([L.orc_bf16_to_f32(L.orc_f32_to_bf16_sr(float(xf), L.orc_sr_bits(k, 11, 3, 9))) for k in range(n)])
    assert set(np.unique(vals)) <= {lo, hi}
    xv = float(xf)
    var = (abs(hi) - abs(xv)) * (abs(xv) - abs(lo))
    assert abs(vals.mean() - xv) <= 4.5 * np.sqrt(var / n)
    assert abs(vals.var() - var) <= 0.05 * var
    # over all 2^16 values of the 16 bits the rounding takes, hi comes out
    # (|x| - |lo|)/(|hi| - |lo|) 2^16 times: the mean is x exactly
    ups = sum(L.orc_f32_to_bf16_sr(float(xf), t << 16) != (u >> 16) for t in range(65536))
    assert ups == round(65536 * (abs(xv) - abs(lo)) / (abs(hi) - abs(lo)))


def test_stochastic_rounding_matches_the_gfx950_conversions(O):
    """The oracle's two stochastic roundings against the vectors gfx950's
    v_cvt_sr_f16_f32 and v_cvt_sr_bf16_f32 wrote
    (tests/golden/sr_conversions.npz, made by tests/golden/make_sr_fixture.py):
    random (x, r) pairs at every float exponent from -45 to 17, structured
    words (0, ~0, single bits, low and high masks), zeros, binary16's largest
    values and overflow, its subnormal limits, inf and NaN.  Bit-exact, NaN
    compared as NaN."""
    L = O.lib()
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "sr_conversions.npz"))
    x, r, h, b = d["x"], d["r"], d["f16"], d["bf16"]
    assert len(x) > 10000
    xf = x.view(np.float32)
    nan = (x & 0x7FFFFFFF) > 0x7F800000
    got_h = np.array([L.orc_f32_to_f16_sr(float(v), int(w)) for v, w in zip(xf, r)], dtype=np.uint16)
    got_b = np.array([L.orc_f32_to_bf16_sr(float(v), int(w)) for v, w in zip(xf, r)], dtype=np.uint16)
    assert np.array_equal(got_h[~nan], h[~nan])
    assert np.array_equal(got_b[~nan], b[~nan])
    assert np.all((got_h[nan] & 0x7FFF) > 0x7C00) and np.all((h[nan] & 0x7FFF) > 0x7C00)
    assert np.all((got_b[nan] & 0x7FFF) > 0x7F80) and np.all((b[nan] & 0x7FFF) > 0x7F80)
    # the cases the fixture holds: subnormal results, flushes to zero below
    # 2^-37, overflow to inf, and carries into the next binade
    e = ((x >> 23) & 0xFF).astype(int) - 127
    assert np.any((e < -14) & (e > -25) & ((h & 0x7FFF) != 0)) and np.any((e < -37) & ((h & 0x7FFF) == 0))
    assert np.any((h & 0x7FFF) == 0x7C00) and np.any(((h >> 10) & 31) != np.clip(e + 15, 0, 31))


def test_stochastic_rounding_bits_are_uniform_and_independent(O):
    """The random words of (node, direction, step, seed), in the bits the
    roundings take (binary16 the top 13, bfloat16 the top 16): uniform in
    mean and variance, and uncorrelated between the directions of a node,
    between neighbouring nodes and between consecutive steps (|Pearson r|
    below 4.5/sqrt(n) for every pair); direction 2j + 1 takes the low half
    of direction 2j's word, so the two use disjoint bits."""
    L = O.lib()
    n, Q = 4000, 19
    bound = 4.5 / np.sqrt(n)
    for shift in (19, 16):
        w = {st: np.array([[L.orc_sr_bits(k, q, st, 3) for q in range(Q)] for k in range(n + 1)], dtype=np.uint64)
             for st in (7, 8)}
        u = {st: (w[st] >> np.uint64(shift)).astype(np.float64) / 2.0 ** (32 - shift) for st in w}
        x = u[7][:n]
        assert np.all(np.abs(x.mean(0) - 0.5) < 4.5 * np.sqrt(1 / 12 / n))
        assert np.all(np.abs(x.var(0) - 1 / 12) < 0.006)
        c = np.corrcoef(x.T)
        assert np.max(np.abs(c - np.eye(Q))) < bound                     # directions of a node
        for q in range(Q):
            assert abs(np.corrcoef(u[7][:n, q], u[7][1:, q])[0, 1]) < bound   # neighbouring nodes
            assert abs(np.corrcoef(u[7][:n, q], u[8][:n, q])[0, 1]) < bound   # consecutive steps
        for j in range(Q // 2):
            assert np.all(w[7][:, 2 * j + 1] == (w[7][:, 2 * j] << np.uint64(16)) & np.uint64(0xFFFFFFFF))
