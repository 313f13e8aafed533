"""The reference's binary state layout (<file>_<rank>.pri), host side.

The reference keeps a snapshot as 27 margin blocks (src/conf.R:771-1012):
block sx + 3 sy + 9 sz + 1 with each side 0 below the domain, 1 the domain,
2 above it.  saveSolution writes the non-empty ones in order
(src/Lattice.cu.Rt:642-661,708-737).  These tests pin the engine's packing
(tclbx_pri_*) three ways:

- the file sizes and block lists of d2q9, d3q19 and d3q27 against closed
  forms counted by hand from the velocity sets: a margin block holds the
  densities whose velocity points out of it;
- the reference's GET rule, restated here from get_offsets (src/conf.R:964-
  978) and the pull of LatticeAccess::pop (src/LatticeAccess.inc.cpp.Rt:311-
  325): reading density i at node x from the packed file, from x - e_i with
  the margin blocks standing in for the wrapped neighbours, gives the pulled
  value plane_i[x - e_i (periodic)] at every node, for every model family
  (plain, cumulant with its f fields also read at the node, autosym builds
  with their mirror reaches, the averages that read one node either side);
- unpack inverts pack.
Velocities come from the reference's Dynamics.R files (d3q19 from the
MRTMAT fixture), not from the engine.
"""
import json
import os

import numpy as np
import pytest

import tclb_amd

HERE = os.path.dirname(os.path.abspath(__file__))

D2Q9 = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0), (1, -1, 0)]
D3Q27 = [(x, y, z) for z in (-1, 0, 1) for y in (-1, 0, 1) for x in (-1, 0, 1)]          # expand.grid(-1:1,...)
_c = (0, 1, -1)
CUMULANT = [(_c[x], _c[y], _c[z]) for z in range(3) for y in range(3) for x in range(3)]  # f<x><y><z>, digits 0,1,2


def d3q19_velocities():
    rows = json.load(open(os.path.join(HERE, "golden", "d3q19_mrtmat.json")))["rows"]
    # rows are the moments; moments 3, 5, 7 are ex, ey, ez (selU = c(4,6,8))
    return [(rows[3][i], rows[5][i], rows[7][i]) for i in range(19)]


def velocities(model):
    if model.startswith("d2q9"):
        return D2Q9
    if model.startswith("d3q19"):
        return d3q19_velocities()
    if model.startswith("d3q27_cumulant"):
        return CUMULANT
    return D3Q27


# ---------------------------------------------------------- closed forms ---

@pytest.mark.parametrize("n", [(4, 5, 6), (7, 3, 2), (1, 1, 1)])
def test_d3q19_blocks_in_closed_form(n):
    """d3q19 has no velocity with three nonzero components: the 8 corner
    blocks are empty, 19 remain.  A face block holds the 5 densities
    leaving through it, an edge block the 1 diagonal leaving through both
    faces, the domain block all 19."""
    nx, ny, nz = n
    N = nx * ny * nz
    blocks = dict(tclb_amd.pri_blocks("d3q19", nx, ny, nz))
    corners = {1, 3, 7, 9, 19, 21, 25, 27}
    assert set(blocks) == set(range(1, 28)) - corners
    assert [b for b, _ in tclb_amd.pri_blocks("d3q19", nx, ny, nz)] == sorted(blocks)
    assert blocks[14] == 19 * N
    for b, size in blocks.items():
        s = [(b - 1) % 3, (b - 1) // 3 % 3, (b - 1) // 9]
        margins = [a for a in range(3) if s[a] != 1]
        span = int(np.prod([n[a] for a in range(3) if s[a] == 1]))
        if len(margins) == 1:
            assert size == 5 * span, b
        elif len(margins) == 2:
            assert size == 1 * span, b
    total = 19 * N + 10 * (ny * nz + nx * nz + nx * ny) + 4 * (nx + ny + nz)
    assert tclb_amd.pri_bytes("d3q19", "f64", nx, ny, nz) == 8 * total
    assert tclb_amd.pri_bytes("d3q19", "f32", nx, ny, nz) == 4 * total


def test_d2q9_and_d3q27_sizes_in_closed_form():
    nx, ny = 11, 6
    # d2q9: 9 in the domain, 3 per side, 1 per corner (z stays in the domain)
    assert tclb_amd.pri_bytes("d2q9_SRT", "f64", nx, ny, 1) == 8 * (9 * nx * ny + 3 * 2 * (nx + ny) + 4)
    assert [b for b, _ in tclb_amd.pri_blocks("d2q9_SRT", nx, ny, 1)] == [10, 11, 12, 13, 14, 15, 16, 17, 18]
    # d3q27: 27 in the domain, 9 per face, 3 per edge, 1 per corner
    nx, ny, nz = 5, 4, 3
    total = 27 * nx * ny * nz + 9 * 2 * (ny * nz + nx * nz + nx * ny) + 3 * 4 * (nx + ny + nz) + 8
    assert tclb_amd.pri_bytes("d3q27", "f64", nx, ny, nz) == 8 * total
    assert len(tclb_amd.pri_blocks("d3q27", nx, ny, nz)) == 27
    # the cumulant build adds its 13 averages, in the domain only except
    # avgUX/Y/Z, which read one node either side along their axis
    base = tclb_amd.pri_bytes("d3q27_cumulant", "f64", nx, ny, nz)
    assert base == 8 * total
    extra = 13 * nx * ny * nz + 2 * (ny * nz + nx * nz + nx * ny)
    assert tclb_amd.pri_bytes("d3q27_cumulant_AVG", "f64", nx, ny, nz) == base + 8 * extra
    # the bc builds: BC[0], BC[1] in the domain
    assert tclb_amd.pri_bytes("d2q9_bc", "f64", 7, 4, 1) == tclb_amd.pri_bytes("d2q9", "f64", 7, 4, 1) + 8 * 2 * 28


def test_formats_without_a_reference_counterpart_are_refused():
    for model, storage in [("d3q19", "f16s"), ("d3q19", "f32s"), ("d2q9", "bf16s"), ("d3q27_cumulant", "f16s")]:
        with pytest.raises(tclb_amd.TclbxError, match="reference"):
            tclb_amd.pri_bytes(model, storage, 4, 4, 4 if model[1] == "3" else 1)
    for model in ("d2q9", "d2q9_bc_autosym2", "d2q9_drag"):
        assert tclb_amd.pri_bytes(model, "f16s", 5, 4, 1) == tclb_amd.pri_bytes(model, "f64", 5, 4, 1) // 4


# ----------------------------------------------- the reference's get rule ---

def reach(model):
    """Each field's reach (min, max per axis) as the reference declares it:
    -e for a density (AddDensity -> AddField(dx = -dx)), the cumulant's f
    fields also at the node, the averages avgUX/Y/Z one node either side,
    and the autosym mirror closure (src/conf.R:400-445)."""
    E = velocities(model)
    fields = [[(-e[a], -e[a]) for a in range(3)] for e in E]
    if model.startswith("d3q27_cumulant"):
        fields = [[(min(lo, 0), max(hi, 0)) for lo, hi in f] for f in fields]
    spec = tclb_amd.Model(model)
    for name in spec.extras:
        r = [(0, 0)] * 3
        for a, ax in enumerate("XYZ"):
            if name == "avgU" + ax:
                r = list(r)
                r[a] = (-1, 1)
        fields.append(r)
    sym = 2 if model.endswith("autosym2") else (1 if model.endswith("autosym") else 0)
    if sym:
        shift = 1 if sym == 2 else 0
        Q = len(E)
        dirs = [set() for _ in fields]
        for k, f in enumerate(fields):
            for x in range(f[0][0], f[0][1] + 1):
                for y in range(f[1][0], f[1][1] + 1):
                    for z in range(f[2][0], f[2][1] + 1):
                        dirs[k].add((x, y, z))

        def mirror(k, a):
            if k >= Q or E[k] == (0, 0, 0):
                return k
            m = list(E[k])
            m[a] = -m[a]
            return E.index(tuple(m))

        changed = True
        while changed:
            changed = False
            for k in range(len(fields)):
                for a in range(3):
                    j = mirror(k, a)
                    for d in list(dirs[k]):
                        if d[a] == 0:
                            continue
                        m = list(d)
                        m[a] = -d[a] - shift if d[a] < 0 else -d[a] + shift
                        if tuple(m) not in dirs[j]:
                            dirs[j].add(tuple(m))
                            changed = True
        fields = [[(min(d[a] for d in dirs[k]), max(d[a] for d in dirs[k])) for a in range(3)]
                  for k in range(len(fields))]
    return fields


def block_table(fields, n):
    """{block: (offset, [(field, nsize)] in order, sizes)}: MarginNSize and
    Size of src/conf.R:887-893,1000-1012."""
    table, off = {}, 0
    for b in range(27):
        s = (b % 3, b // 3 % 3, b // 9)
        nodes = int(np.prod([n[a] if s[a] == 1 else 1 for a in range(3)]))
        layers = []
        for k, f in enumerate(fields):
            th = [1 if s[a] == 1 else (max(-f[a][0], 0) if s[a] == 0 else max(f[a][1], 0)) for a in range(3)]
            if np.prod(th) > 0:
                layers.append((k, th))
        if layers:
            table[b + 1] = (off, layers, nodes, s)
            off += nodes * sum(int(np.prod(th)) for _, th in layers)
    return table, off


def get_value(buf, table, fields, n, k, w, dw):
    """Field k at node w shifted by dw, the reference's way: per axis, a
    position beyond the domain is read from the margin block on that side,
    at its offset into the margin (get_offsets, conditions w + dw < 0 and
    w + dw >= n)."""
    side, pos = [], []
    for a in range(3):
        p = w[a] + dw[a]
        if p < 0:
            side.append(0)
            pos.append(p + max(-fields[k][a][0], 0))
        elif p >= n[a]:
            side.append(2)
            pos.append(p - n[a])
        else:
            side.append(1)
            pos.append(p)
    b = side[0] + 3 * side[1] + 9 * side[2] + 1
    off, layers, nodes, s = table[b]
    for kk, th in layers:
        ext = [n[a] if s[a] == 1 else th[a] for a in range(3)]
        if kk == k:
            return buf[off + pos[0] + ext[0] * (pos[1] + ext[1] * pos[2])]
        off += nodes * int(np.prod(th))
    raise AssertionError(f"field {k} not in block {b}")


MODELS = [("d2q9_SRT", (6, 5, 1)), ("d2q9", (5, 4, 1)), ("d2q9_autosym", (5, 4, 1)), ("d2q9_bc_autosym2", (6, 3, 1)),
          ("d3q19", (4, 3, 5)), ("d3q27", (3, 4, 2)), ("d3q27_cumulant_AVG", (3, 3, 4)), ("d3q19", (1, 2, 1))]


@pytest.mark.parametrize("model,n", MODELS)
def test_packed_file_streams_like_the_reference(model, n):
    nx, ny, nz = n
    spec = tclb_amd.Model(model)
    E = velocities(model)
    assert len(E) == spec.Q
    rng = np.random.default_rng(7)
    planes = rng.standard_normal((spec.Q, nz, ny, nx))
    extra = rng.standard_normal((len(spec.extras), nz, ny, nx))
    raw = tclb_amd.pri_pack(model, "f64", planes, extra if spec.extras else None)
    buf = raw.view(np.float64)
    fields = reach(model)
    table, total = block_table(fields, n)
    assert total == buf.size
    assert [(b, table[b][2] * sum(int(np.prod(th)) for _, th in table[b][1])) for b in sorted(table)] == \
        tclb_amd.pri_blocks(model, nx, ny, nz)
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                w = (x, y, z)
                for i, e in enumerate(E):
                    # pop: density i from x - e_i
                    src = ((x - e[0]) % nx, (y - e[1]) % ny, (z - e[2]) % nz)
                    got = get_value(buf, table, fields, n, i, w, (-e[0], -e[1], -e[2]))
                    assert got == planes[i, src[2], src[1], src[0]], (model, i, w)
                    # every offset inside the field's reach reads the wrapped node
                    for dx in range(fields[i][0][0], fields[i][0][1] + 1):
                        for dy in range(fields[i][1][0], fields[i][1][1] + 1):
                            for dz in range(fields[i][2][0], fields[i][2][1] + 1):
                                s = ((x + dx) % nx, (y + dy) % ny, (z + dz) % nz)
                                got = get_value(buf, table, fields, n, i, w, (dx, dy, dz))
                                assert got == planes[i, s[2], s[1], s[0]], (model, i, w, (dx, dy, dz))
                for k, name in enumerate(spec.extras):
                    f = fields[spec.Q + k]
                    for a in range(3):
                        for d in (f[a][0], f[a][1]):
                            dw = [0, 0, 0]
                            dw[a] = d
                            s = [(w[b] + dw[b]) % n[b] for b in range(3)]
                            got = get_value(buf, table, fields, n, spec.Q + k, w, dw)
                            assert got == extra[k, s[2], s[1], s[0]], (name, w, dw)


@pytest.mark.parametrize("model,storage,n", [("d3q19", "f32", (5, 3, 4)), ("d2q9_bc", "f16s", (7, 5, 1)),
                                             ("d3q27_cumulant_AVG_IB", "f64", (3, 4, 3)),
                                             ("d2q9_autosym2", "f32s", (4, 6, 1))])
def test_unpack_inverts_pack(model, storage, n):
    nx, ny, nz = n
    spec = tclb_amd.Model(model)
    rng = np.random.default_rng(11)
    planes = rng.standard_normal((spec.Q, nz, ny, nx)).astype(tclb_amd.STORAGE_DTYPE[storage])
    # extras travel in the stored type: values that type holds exactly
    extra = rng.standard_normal((len(spec.extras), nz, ny, nx)).astype(tclb_amd.STORAGE_DTYPE[storage])
    extra = extra.astype(np.float64)
    raw = tclb_amd.pri_pack(model, storage, planes, extra if spec.extras else None)
    assert raw.size == tclb_amd.pri_bytes(model, storage, nx, ny, nz)
    back, ex = tclb_amd.pri_unpack(model, storage, nx, ny, nz, raw, spec.Q, len(spec.extras))
    assert np.array_equal(back.view(np.uint8), planes.view(np.uint8))
    assert np.array_equal(ex, extra)
