"""CPU restatement of TCLB's case-file unit system and NodeType rasterizer.

TEST INFRASTRUCTURE ONLY: the checker of the engine's C++ rasterizer
(tclb_amd/csrc/geometry.cpp, units.cpp).  Only tests/ import this module.

It restates, independently of the engine's code (Python's own XML parser,
numpy over each primitive's box of nodes), what the reference computes:
  * UnitEnv (src/unit.cpp:64-279, src/unit.h:152-195): readText, readUnit,
    readUnitAlpha, alt with +/- terms, makeGauge with GaussSolve;
  * Geometry::load / Draw (src/Geometry.cpp.Rt:141-330, 758-1184): flags,
    masks, modes, zones, nested regions, the primitives Box, Sphere,
    HalfSphere, OffgridSphere, OffgridPipe, XPipe, YPipe, ZPipe, XAnnulus,
    Pipe, PipeY, PipeZ, Cylinder, Wedge, Text and named zones, with the
    defaults of src/def.cpp.Rt:10-33.
"""
import math
import xml.etree.ElementTree as ET

import numpy as np

BASE = ["m", "s", "kg", "K", "x", "y", "z", "A", "t"]


class UV:
    """value * prod(base_k ** p_k)"""

    def __init__(self, val=0.0, p=None):
        self.val = float(val)
        self.p = list(p) if p is not None else [0.0] * 9

    def __mul__(self, o):
        return UV(self.val * o.val, [a + b for a, b in zip(self.p, o.p)])

    def __truediv__(self, o):
        return UV(self.val / o.val, [a - b for a, b in zip(self.p, o.p)])

    def pw(self, e):
        return UV(math.pow(self.val, e) if not (self.val == 0 and e < 0) else float("inf"),
                  [a * e for a in self.p])


class Units:
    def __init__(self):
        self.scale = [1.0] * 9
        self.units = {}
        for k, n in enumerate(BASE):
            p = [0.0] * 9
            p[k] = 1.0
            self.units[n] = UV(1.0, p)
        for n, t in [("N", "1kgm/s2"), ("Pa", "1N/m2"), ("J", "1Nm"), ("W", "1J/s"), ("V", "1kgm2/t3/A"),
                     ("C", "1tA"), ("nm", "1e-9m"), ("um", "1e-6m"), ("mm", "1e-3m"), ("cm", "1e-2m"),
                     ("km", "1e+3m"), ("h", "3600s"), ("ns", "1e-9s"), ("us", "1e-6s"), ("ms", "1e-3s"),
                     ("g", "1e-3kg"), ("mg", "1e-6kg")]:
            self.units[n] = self.text(t)
        self.units["d"] = UV(math.atan(1.0) * 4.0 / 180.)
        self.units["%"] = UV(1. / 100.)
        self.units["An"] = UV(6.022 * 100000000000000000000000.)
        self.gauges = {}

    def one(self, s):
        return self.units.get(s, UV(0.0))

    def alpha(self, s, e):
        r1 = self.one(s[:1])
        if len(s) < 2:
            return r1.pw(e)
        r1 = r1 * self.alpha(s[1:], e)
        r2 = self.one(s[:2])
        r2 = r2 * self.alpha(s[2:], e) if len(s) > 2 else r2.pw(e)
        if r1.val == 0:
            return UV(0.0) if r2.val == 0 else r2
        if r2.val == 0:
            return r1
        if s[0] == "m":
            return r2
        raise ValueError("ambiguous unit " + s)

    def unit(self, s):
        ret, sign, i, n = UV(1.0), 1, 0, len(s)
        while i < n:
            j = i
            while i < n and s[i].isascii() and s[i].isalpha():
                i += 1
            k = i
            while i < n and (s[i].isdigit() or s[i] == "."):
                i += 1
            e = float(s[k:i]) if i > k else 1.0
            last = self.alpha(s[j:k], e) if k > j else UV(1.0)
            ret = ret * last if sign > 0 else ret / last
            j = i
            while i < n and not (s[i].isascii() and s[i].isalnum()):
                i += 1
            if i - j > 1:
                raise ValueError("too many separators in " + s)
            if i - j == 1:
                if s[j] != "/":
                    raise ValueError("only / in units")
                sign = -1
        return ret

    def text(self, s):
        i = 0
        while i < len(s) and s[i] in "+-0123456789.eE":
            i += 1
        r = self.unit(s[i:])
        if i > 0:
            r = r * UV(_atof(s[:i]))
        return r

    def alt_uv(self, v):
        r = v.val
        for k in range(9):
            r *= math.pow(self.scale[k], v.p[k])
        return r

    def alt(self, s, default=None):
        if s == "" and default is not None:
            return float(default)
        terms, i, j = [], 0, 0
        n = len(s)
        while i < n:
            c = s[j] if j < n else ""
            if c in ("+", "-", ""):
                if j > i:
                    terms.append(s[i:j])
                i = j
            elif c in ("e", "E") and j + 1 < n and s[j + 1] in "+-":
                j += 1
            j += 1
        ret = 0.0
        for t in terms:
            ret += self.alt_uv(self.text(t))
        return ret

    def gauge(self, name, value, gauge):
        self.gauges[name] = self.text(value) / self.text(gauge) / UV(1.0)

    def make_gauge(self):
        n = 9
        A = [0.0] * (n * n)   # A[col*n + row]
        b = [0.0] * n
        x = [0.0] * n
        rows = 0
        for name in sorted(self.gauges):
            g = self.gauges[name]
            for j in range(n):
                A[j * n + rows] = g.p[j]
            b[rows] = math.log(g.val)
            rows += 1
        for j in range(n):
            if not any(A[j * n + l] != 0 for l in range(rows)):
                A[j * n + rows] = 1.0
                b[rows] = 0.0
                rows += 1
        assert rows == n, "gauge system is not square"
        z = lambda v: abs(v) < 1e-10
        for k in range(n - 1):
            if z(A[k * n + k]):
                for i in range(k + 1, n):
                    if not z(A[k * n + i]):
                        for j in range(n):
                            A[j * n + k], A[j * n + i] = A[j * n + i], A[j * n + k]
                        b[k], b[i] = b[i], b[k]
                        break
            for i in range(k + 1, n):
                m = A[k * n + i] / A[k * n + k]
                for j in range(n):
                    A[j * n + i] -= m * A[j * n + k]
                b[i] -= m * b[k]
        for i in range(n - 1, -1, -1):
            s = 0.0
            for j in range(n - 1, i, -1):
                s += A[j * n + i] * x[j]
            x[i] = (b[i] - s) / A[i * n + i]
        self.scale = [math.exp(-v) for v in x]


def _atof(s):
    """C atof: the longest prefix that parses as a number (0 if none)."""
    for k in range(len(s), 0, -1):
        try:
            return float(s[:k])
        except ValueError:
            pass
    return 0.0


def _round(v):
    return int(v + 0.5) if v > 0 else int(v - 0.5)


def _as_int(s):
    """pugixml as_int: sign, decimal or 0x digits, stop at the first other."""
    s = s.lstrip(" \t\r\n")
    neg = s[:1] == "-"
    if s[:1] in "+-" and s:
        s = s[1:]
    if s[:2].lower() == "0x":
        d = ""
        for ch in s[2:]:
            if ch in "0123456789abcdefABCDEF":
                d += ch
            else:
                break
        v = int(d, 16) if d else 0
    else:
        d = ""
        for ch in s:
            if ch.isdigit():
                d += ch
            else:
                break
        v = int(d) if d else 0
    return -v if neg else v


# ----------------------------------------------------- NodeType packing ---
# AddNodeType calls of the models (models/flow/d2q9_SRT/Dynamics.R:31-37,
# models/flow/d3q19/Dynamics.R:83-95); d3q19_lid appends MovingWall to
# BOUNDARY (an engine model, no reference counterpart)
NODETYPES = {
    "d2q9_SRT": [("EPressure", "BOUNDARY"), ("WPressure", "BOUNDARY"), ("WVelocity", "BOUNDARY"),
                 ("EVelocity", "BOUNDARY"), ("Solid", "BOUNDARY"), ("Wall", "BOUNDARY"), ("BGK", "COLLISION")],
    "d3q19": [("XYslice", "ADDITIONALS"), ("XZslice", "ADDITIONALS"), ("YZslice", "ADDITIONALS"),
              ("EPressure", "BOUNDARY"), ("Solid", "BOUNDARY"), ("Wall", "BOUNDARY"), ("WPressure", "BOUNDARY"),
              ("WPressureL", "BOUNDARY"), ("WVelocity", "BOUNDARY"), ("MRT", "COLLISION")],
}
# models/flow/d2q9/Dynamics.R:62-79
NODETYPES["d2q9"] = [("EPressure", "BOUNDARY"), ("WPressure", "BOUNDARY"), ("NVelocity", "BOUNDARY"),
                     ("SVelocity", "BOUNDARY"), ("WVelocity", "BOUNDARY"), ("EVelocity", "BOUNDARY"),
                     ("NSymmetry", "BOUNDARY"), ("SSymmetry", "BOUNDARY"), ("Inlet", "OBJECTIVE"),
                     ("Outlet", "OBJECTIVE"), ("Solid", "BOUNDARY"), ("Wall", "BOUNDARY"), ("MRT", "COLLISION")]
NODETYPES["d2q9_SRT_drag"] = NODETYPES["d2q9_SRT"]
NODETYPES["d2q9_drag"] = NODETYPES["d2q9"]
NODETYPES["d3q19_lid"] = NODETYPES["d3q19"][:9] + [("MovingWall", "BOUNDARY"), ("MRT", "COLLISION")]


class Packing:
    """src/conf.R:600-699: groups in alphabetical order, ceil(log2(n+1))
    bits each, values 1..n; the zone takes the rest of the 16-bit flag;
    then None (NONE, 0/0) and Clear (ALL, 0/0xFFFF)."""

    def __init__(self, model):
        types = NODETYPES[model]
        groups = sorted(set(g for _, g in types))
        self.types, self.masks = [], []   # (name, value, group), (group, mask)
        shift, bits = 1, 0
        for g in groups:
            names = [t for t, gg in types if gg == g]
            l = math.ceil(math.log2(len(names) + 1))
            self.masks.append((g, shift * (2 ** l - 1)))
            self.types += [(t, shift * (k + 1), g) for k, t in enumerate(names)]
            shift *= 2 ** l
            bits += l
        self.zone_shift = bits
        self.zone_mask = (2 ** (16 - bits) - 1) * shift
        self.masks.append(("SETTINGZONE", self.zone_mask))
        self.types.append(("DefaultZone", 0, "SETTINGZONE"))
        self.masks += [("NONE", 0), ("ALL", 0xFFFF)]
        self.types += [("None", 0, "NONE"), ("Clear", 0, "ALL")]

    def value(self, name):
        return next(v for t, v, _ in self.types if t == name)


# ------------------------------------------------------------- geometry ---
def _defaults(pk):
    """<Geometry> defaults of src/def.cpp.Rt:10-33 for a Packing."""
    box = lambda *a: ("Box", dict(zip(["dx", "dy", "dz", "fx", "fy", "fz"], a)), [])
    zones = [
        ("Inlet", [box("0", "0", "0", "0", "-1", "-1")]),
        ("Outlet", [box("-1", "0", "0", "-1", "-1", "-1")]),
        ("Channel", [box("0", "0", "0", "-1", "0", "-1"), box("0", "-1", "0", "-1", "-1", "-1")]),
        ("Tunnel", [box("0", "0", "0", "-1", "0", "-1"), box("0", "-1", "0", "-1", "-1", "-1"),
                    box("0", "0", "0", "-1", "-1", "0"), box("0", "0", "-1", "-1", "-1", "-1")]),
        ("Tunnel", [box("0", "0", "0", "0", "-1", "-1")]),
        ("Inlet", [box("0", "0", "0", "0", "-1", "-1")]),
    ]
    out = [("Zone", {"name": n}, kids) for n, kids in zones]
    out.append(("Mask", {"name": "NONE", "value": "0"}, []))
    out.append(("Type", {"name": "None", "value": "0", "mask": "NONE"}, []))
    # the reference prints masks with %f ("7.000000"), which as_int reads as 7
    for g, mask in pk.masks:
        out.append(("Mask", {"name": g, "value": f"{float(mask):f}"}, []))
    for t, v, g in pk.types:
        out.append(("Type", {"name": t, "value": str(v), "mask": g}, []))
    return out


class Node:
    """element, attributes in order, children, parent"""

    def __init__(self, tag, attrib, parent=None):
        self.tag, self.attrib, self.children, self.parent = tag, dict(attrib), [], parent

    @classmethod
    def of(cls, el, parent=None):
        n = cls(el.tag, el.attrib, parent)
        n.children = [cls.of(c, n) for c in el if isinstance(c.tag, str)]
        return n

    def find(self, tag, name):
        for c in self.children:
            if c.tag == tag and c.attrib.get("name") == name:
                return c
        return None


class Rasterizer:
    def __init__(self, model, nx, ny, nz, units):
        self.pk, self.u = Packing(model), units
        self.shape = (nx, ny, nz)
        self.geom = np.zeros((nz, ny, nx), dtype=np.uint16)
        self.zones = {"DefaultZone": 0}
        self.fg = self.mask = 0
        self.mode = "overwrite"

    # --- flag state
    def set_flag(self, name):
        t = self.root.find("Type", name)
        if t is None:
            raise KeyError("Unknown flag (in xml): " + name)
        self.fg = _as_int(t.attrib.get("value", "")) & 0xFFFF
        self.set_mask(t.attrib.get("mask", ""))
        self.mode = "overwrite"

    def set_mask(self, name):
        t = self.root.find("Mask", name)
        if t is None:
            raise KeyError("Unknown mask (in xml): " + name)
        self.mask = (_as_int(t.attrib["value"]) if "value" in t.attrib else -1) & 0xFFFF

    def set_zone(self, name):
        k = self.zones.setdefault(name, len(self.zones))
        zmask, shift = self.pk.zone_mask, self.pk.zone_shift
        self.fg = (self.fg & ~zmask & 0xFFFF) | ((k << shift) & 0xFFFF)
        self.mask = self.mask | zmask

    # --- regions
    def region(self, n):
        if n is None:
            return [0, 0, 0, 1, 1, 1]
        if n.tag == "Geometry":
            return [0, 0, 0] + list(self.shape)
        r = self.region(n.parent)
        for ax, d in enumerate(["dx", "dy", "dz"]):
            if d in n.attrib:
                v = n.attrib[d]
                side = v[:1] if v[:1] in ("<", ">") else "+"
                w = _round(self.u.alt(v[1:] if side != "+" else v))
                if side == "<" or (side == "+" and w < 0):
                    w = r[3 + ax] + w
                r[ax] += w
                r[3 + ax] -= w
        for ax, f in enumerate(["fx", "fy", "fz"]):
            if f in n.attrib:
                w = _round(self.u.alt(n.attrib[f]))
                if w < 0:
                    w = r[3 + ax] + w + r[ax]
                r[3 + ax] = w - r[ax] + 1
        for ax, s in enumerate(["nx", "ny", "nz"]):
            if s in n.attrib:
                r[3 + ax] = _round(self.u.alt(n.attrib[s]))
        return r

    # --- painting: every node of `sel` (a boolean array over the box
    # starting at (x0,y0,z0)) that lies inside the lattice
    def paint(self, x0, y0, z0, sel):
        nx, ny, nz = self.shape
        zz, yy, xx = np.nonzero(sel)
        xx, yy, zz = xx + x0, yy + y0, zz + z0
        ok = (xx >= 0) & (xx < nx) & (yy >= 0) & (yy < ny) & (zz >= 0) & (zz < nz)
        xx, yy, zz = xx[ok], yy[ok], zz[ok]
        # a box visits each node once, so the order of the nodes is free
        g = self.geom[zz, yy, xx]
        hit = (g & self.mask) != 0
        keep = ~hit if self.mode == "fill" else (hit if self.mode == "change" else np.ones_like(hit))
        self.geom[zz[keep], yy[keep], xx[keep]] = (g[keep] & (~self.mask & 0xFFFF)) | self.fg

    @staticmethod
    def grid(x0, y0, z0, nx, ny, nz):
        X = (np.arange(max(nx, 0)) + x0).astype(np.float64)[None, None, :]
        Y = (np.arange(max(ny, 0)) + y0).astype(np.float64)[None, :, None]
        Z = (np.arange(max(nz, 0)) + z0).astype(np.float64)[:, None, None]
        return X, Y, Z

    @staticmethod
    def sphere(a, b, c):
        a = 2 * a - 1
        b = 2 * b - 1
        c = 2 * c - 1
        return (a * a + b * b + c * c) < 1

    def draw(self, node):
        for n in node.children:
            r = self.region(n)
            dx, dy, dz, nx, ny, nz = r
            t = n.tag
            with np.errstate(divide="ignore", invalid="ignore"):
                if t == "Box":
                    lo = [max(dx, 0), max(dy, 0), max(dz, 0)]
                    hi = [min(dx + nx, self.shape[0]), min(dy + ny, self.shape[1]), min(dz + nz, self.shape[2])]
                    if all(h > l for l, h in zip(lo, hi)):
                        self.paint(lo[0], lo[1], lo[2], np.ones((hi[2] - lo[2], hi[1] - lo[1], hi[0] - lo[0]), bool))
                elif t in ("Sphere", "HalfSphere"):
                    X, Y, Z = self.grid(dx, dy, dz, nx, ny, nz)
                    a = (.5 + X - dx) / nx
                    b = (.5 + Y - dy) / ny
                    c = (.5 + Z - dz) / nz
                    if t == "HalfSphere":
                        b = 0.5 - (.5 + Y - dy) / ny / 2.
                    self.paint(dx, dy, dz, np.broadcast_to(self.sphere(a, b, c), (max(nz, 0), max(ny, 0), max(nx, 0))))
                elif t == "OffgridSphere":
                    x0, y0, z0 = (self.u.alt(n.attrib.get(k, "")) for k in "xyz")
                    if "R" in n.attrib:
                        Rx = Ry = Rz = self.u.alt(n.attrib["R"])
                    else:
                        Rx, Ry, Rz = (self.u.alt(n.attrib.get(k, "")) for k in ("Rx", "Ry", "Rz"))
                    bx, by, bz = int(x0 - Rx - 5), int(y0 - Ry - 5), int(z0 - Rz - 5)
                    sx, sy, sz = int(2 * Rx + 10), int(2 * Ry + 10), int(2 * Rz + 10)
                    X, Y, Z = self.grid(bx, by, bz, sx, sy, sz)
                    a, b, c = .5 + X - x0, .5 + Y - y0, .5 + Z - z0
                    sel = a * a / (Rx * Rx) + b * b / (Ry * Ry) + c * c / (Rz * Rz) < 1.
                    self.paint(bx, by, bz, sel)
                elif t in ("OffgridPipe", "XPipe", "YPipe", "ZPipe"):
                    x0, y0, z0 = (self.u.alt(n.attrib.get(k, "")) for k in "xyz")
                    ra, rb = {"XPipe": ("Ry", "Rz"), "YPipe": ("Rx", "Rz")}.get(t, ("Rx", "Ry"))
                    if "R" in n.attrib:
                        Ra = Rb = self.u.alt(n.attrib["R"])
                    else:
                        Ra, Rb = self.u.alt(n.attrib.get(ra, "")), self.u.alt(n.attrib.get(rb, ""))
                    if t == "XPipe":
                        dy, dz, ny, nz = int(y0 - Ra - 5), int(z0 - Rb - 5), int(2 * Ra + 10), int(2 * Rb + 10)
                    elif t == "YPipe":
                        dx, dz, nx, nz = int(x0 - Ra - 5), int(z0 - Rb - 5), int(2 * Ra + 10), int(2 * Rb + 10)
                    else:
                        dx, dy, nx, ny = int(x0 - Ra - 5), int(y0 - Rb - 5), int(2 * Ra + 10), int(2 * Rb + 10)
                    X, Y, Z = self.grid(dx, dy, dz, nx, ny, nz)
                    if t == "OffgridPipe":
                        p, q = .5 + X - x0, .5 + Y - y0
                        sel = p * p / (Ra * Ra) + q * q / (Rb * Rb) < 1.
                    elif t == "XPipe":
                        p, q = .5 + Y - y0, .5 + Z - z0
                        sel = q * q / (Rb * Rb) + p * p / (Ra * Ra) < 1.
                    elif t == "YPipe":   # x measured from the y attribute, as the reference does
                        p, q = .5 + X - y0, .5 + Z - z0
                        sel = q * q / (Rb * Rb) + p * p / (Ra * Ra) < 1.
                    else:                # ZPipe: likewise
                        p, q = .5 + X - y0, .5 + Y - y0
                        sel = q * q / (Rb * Rb) + p * p / (Ra * Ra) < 1.
                    self.paint(dx, dy, dz, np.broadcast_to(sel, (max(nz, 0), max(ny, 0), max(nx, 0))))
                elif t == "XAnnulus":
                    y0, z0 = self.u.alt(n.attrib.get("y", "")), self.u.alt(n.attrib.get("z", ""))
                    if "Ro" in n.attrib:
                        Ryo = Rzo = self.u.alt(n.attrib["Ro"])
                        Ryi = Rzi = self.u.alt(n.attrib.get("Ri", ""))
                    else:
                        Ryi, Rzi, Ryo, Rzo = (self.u.alt(n.attrib.get(k, "")) for k in ("Ry_i", "Rz_i", "Ry_o", "Rz_o"))
                    dy, dz, ny, nz = int(y0 - Ryo - 5), int(z0 - Rzo - 5), int(2 * Ryo + 10), int(2 * Rzo + 10)
                    X, Y, Z = self.grid(dx, dy, dz, nx, ny, nz)
                    p, q = .5 + Y - y0, .5 + Z - z0
                    sel = (q * q / (Rzo * Rzo) + p * p / (Ryo * Ryo) < 1.) & (q * q / (Rzi * Rzi) + p * p / (Ryi * Ryi) > 1.)
                    self.paint(dx, dy, dz, np.broadcast_to(sel, (max(nz, 0), max(ny, 0), max(nx, 0))))
                elif t in ("Pipe", "PipeY", "PipeZ", "Cylinder"):
                    # one axis over the region, the other two one node wider on each side
                    ax = {"Pipe": 0, "Cylinder": 0, "PipeY": 1, "PipeZ": 2}[t]
                    lo = [dx - 1, dy - 1, dz - 1]
                    sz = [nx + 2, ny + 2, nz + 2]
                    lo[ax] += 1
                    sz[ax] -= 2
                    X, Y, Z = self.grid(lo[0], lo[1], lo[2], sz[0], sz[1], sz[2])
                    if t == "Pipe":
                        sel = ~self.sphere(0.5, (.5 + Y - dy) / ny, (.5 + Z - dz) / nz)
                    elif t == "PipeY":
                        sel = ~self.sphere(0.5, (.5 + X - dx) / nx, (.5 + Z - dz) / nz)
                    elif t == "PipeZ":
                        sel = ~self.sphere(0.5, (.5 + Y - dy) / ny, (.5 + X - dx) / nx)
                    else:
                        sel = self.sphere((.5 + X - dx) / nx, (.5 + Y - dy) / ny, 0.5)
                    self.paint(lo[0], lo[1], lo[2], np.broadcast_to(sel, (max(sz[2], 0), max(sz[1], 0), max(sz[0], 0))))
                elif t == "Wedge":
                    X, Y, Z = self.grid(dx, dy, dz, nx, ny, nz)
                    a, b, c = (X - dx) / (nx - 1.), (Y - dy) / (ny - 1.), (Z - dz) / (nz - 1.)
                    d = n.attrib.get("direction", "") or "UpperLeft"
                    if d in ("UpperRight", "LowerRight", "UpperRightXZ", "LowerRightXZ"):
                        a = 1. - a
                    if d in ("LowerLeft", "LowerRight"):
                        b = 1. - b
                    if d in ("LowerLeftXZ", "LowerRightXZ"):
                        c = 1. - c
                    if d in ("UpperLeft", "UpperRight", "LowerLeft", "LowerRight"):
                        delta = a - b
                    elif d in ("UpperLeftXZ", "UpperRightXZ", "LowerLeftXZ", "LowerRightXZ"):
                        delta = a - c
                    else:
                        delta = np.zeros(1)
                    sel = np.broadcast_to(delta < 1e-10, (max(nz, 0), max(ny, 0), max(nx, 0)))
                    self.paint(dx, dy, dz, sel)
                elif t == "Text":
                    crop = self.region(n.parent)
                    vals = [int(v) for v in open(n.attrib["file"]).read().split()]
                    sel = np.array(vals[: nx * ny * nz]).reshape(nx, ny, nz).transpose(2, 1, 0) != 0
                    X, Y, Z = self.grid(dx, dy, dz, nx, ny, nz)
                    inc = ((X >= max(crop[0], 0)) & (X < min(crop[0] + crop[3], self.shape[0])) &
                           (Y >= max(crop[1], 0)) & (Y < min(crop[1] + crop[4], self.shape[1])) &
                           (Z >= max(crop[2], 0)) & (Z < min(crop[2] + crop[5], self.shape[2])))
                    self.paint(dx, dy, dz, sel & inc)
                else:
                    z = self.root.find("Zone", t)
                    if z is None:
                        raise KeyError("Unknown geometry element: " + t)
                    self.draw(z)

    def load(self, gnode):
        self.root = gnode
        for tag, attrib, kids in _defaults(self.pk):
            if gnode.find(tag, attrib["name"]) is not None:
                continue
            z = Node(tag, attrib, gnode)
            z.children = [_mk(k, z) for k in kids]
            gnode.children.insert(0, z)
        for n in list(gnode.children):
            if n.tag in ("Zone", "Type", "Mask"):
                continue
            self.set_flag(n.tag)
            for k, v in n.attrib.items():
                if k == "name":
                    self.set_zone(v)
                elif k == "mask":
                    self.set_mask(v)
                elif k == "mode":
                    if v not in ("overwrite", "fill", "change"):
                        raise KeyError("Unknown mode (in xml): " + v)
                    self.mode = v
            if "zone" in n.attrib:
                z = gnode.find("Zone", n.attrib["zone"])
                if z is not None:
                    self.draw(z)
            self.draw(n)


def _mk(spec, parent):
    tag, attrib, kids = spec
    n = Node(tag, attrib, parent)
    n.children = [_mk(k, n) for k in kids]
    return n


def read_case(path_or_text, model, is_text=False):
    """Units and the NodeType map of every <Geometry> of a case:
    returns (flags [nz,ny,nx] uint16, zones dict, Units)."""
    root = ET.fromstring(path_or_text) if is_text else ET.parse(path_or_text).getroot()
    config = Node.of(root)
    u = Units()
    units = next((c for c in config.children if c.tag == "Units"), None)
    if units is not None:
        for i, p in enumerate([c for c in units.children if c.tag == "Param"]):
            u.gauge(p.attrib.get("name", f"unnamed{i + 1}"), p.attrib["value"], p.attrib["gauge"])
        u.make_gauge()
    geoms = [c for c in config.children if c.tag == "Geometry"]
    g0 = geoms[0]
    nx, ny, nz = (_round(u.alt(g0.attrib.get(k, ""), 1)) for k in ("nx", "ny", "nz"))
    R = Rasterizer(model, nx, ny, nz, u)
    for g in geoms:
        R.load(g)
    return R.geom, R.zones, u
