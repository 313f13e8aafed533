// pri.cpp -- the reference's margin-block layout of a snapshot, packed and
// unpacked on the host.  The rules restated here are src/conf.R:137-170
// (AddField reach), :373-445 (autosym mirrors), :771-1012 (margins, sizes,
// put and get offsets in the default layout, X_MOD = 0, CPU_LAYOUT = FALSE)
// and src/Lattice.cu.Rt:642-661 (listTabs: the non-empty blocks in order).
#include "pri.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <set>
#include <stdexcept>

namespace tclbx {

namespace {

bool is_cumulant(const ModelDef* m) { return m->name.compare(0, 14, "d3q27_cumulant") == 0; }

bool is_d2q9_mrt(const ModelDef* m) {
    switch (m->id) {
    case MODEL_D2Q9: case MODEL_D2Q9_DRAG: case MODEL_D2Q9_AUTOSYM: case MODEL_D2Q9_AUTOSYM2:
    case MODEL_D2Q9_BC: case MODEL_D2Q9_BC_AUTOSYM: case MODEL_D2Q9_BC_AUTOSYM2:
        return true;
    default:
        return false;
    }
}

int autosym_of(const ModelDef* m) {
    switch (m->id) {
    case MODEL_D2Q9_AUTOSYM: case MODEL_D2Q9_BC_AUTOSYM: return 1;
    case MODEL_D2Q9_AUTOSYM2: case MODEL_D2Q9_BC_AUTOSYM2: return 2;
    default: return 0;
    }
}

using Dir = std::array<int, 3>;

}  // namespace

std::vector<PriField> pri_fields(const ModelDef* m) {
    std::vector<PriField> f;
    // streamed densities: AddDensity(dx = e) declares the field at -e
    for (int q = 0; q < m->Q; q++) {
        PriField p;
        p.plane = q;
        const int e[3] = {m->ex[q], m->ey[q], m->ez[q]};
        for (int a = 0; a < 3; a++) p.mn[a] = p.mx[a] = -e[a];
        // d3q27_cumulant adds every f at (0,0,0) too (its Dynamics.R:51)
        if (is_cumulant(m))
            for (int a = 0; a < 3; a++) { p.mn[a] = std::min(p.mn[a], 0); p.mx[a] = std::max(p.mx[a], 0); }
        f.push_back(p);
    }
    // the extra densities follow in the engine's order, which is their
    // declaration order: the averages (avgUX/Y/Z read one node either side
    // along their axis, Dynamics.R:88-90) or the BC parameters
    const int nav = (int)m->averages.size();
    for (int k = 0; k < m->nextra(); k++) {
        PriField p;
        p.plane = m->Q + k;
        for (int a = 0; a < 3; a++) p.mn[a] = p.mx[a] = 0;
        if (k < nav) {
            const std::string& n = m->averages[k];
            int axis = n == "avgUX" ? 0 : n == "avgUY" ? 1 : n == "avgUZ" ? 2 : -1;
            if (axis >= 0) { p.mn[axis] = -1; p.mx[axis] = 1; }
        }
        f.push_back(p);
    }
    // autosym: each field's box of directions gains its mirror partner's
    // directions reflected (through the node for autosym 1, half a node out
    // for autosym 2), to a fixed point; the partner of density i along an
    // axis is the density of the same group whose velocity is e_i with that
    // component negated
    const int as = autosym_of(m);
    if (as) {
        const int shift = as == 2 ? 1 : 0;
        const int nf = (int)f.size();
        std::vector<std::set<Dir>> dirs(nf);
        for (int i = 0; i < nf; i++)
            for (int x = f[i].mn[0]; x <= f[i].mx[0]; x++)
                for (int y = f[i].mn[1]; y <= f[i].mx[1]; y++)
                    for (int z = f[i].mn[2]; z <= f[i].mx[2]; z++) dirs[i].insert(Dir{x, y, z});
        auto partner = [&](int i, int a) {
            if (f[i].plane >= m->Q) return i;   // a parameter density mirrors onto itself
            const int q = f[i].plane;
            int e[3] = {m->ex[q], m->ey[q], m->ez[q]};
            if (e[0] == 0 && e[1] == 0 && e[2] == 0) return i;
            e[a] = -e[a];
            for (int j = 0; j < m->Q; j++)
                if (m->ex[j] == e[0] && m->ey[j] == e[1] && m->ez[j] == e[2]) return j;
            throw std::logic_error("autosym: no mirror density");
        };
        for (bool changed = true; changed;) {
            changed = false;
            for (int i = 0; i < nf; i++)
                for (int a = 0; a < 3; a++) {
                    const int j = partner(i, a);
                    const std::set<Dir> src = dirs[i];
                    for (Dir d : src) {
                        if (d[a] < 0) d[a] = -d[a] - shift;
                        else if (d[a] > 0) d[a] = -d[a] + shift;
                        else continue;
                        if (dirs[j].insert(d).second) changed = true;
                    }
                }
        }
        for (int i = 0; i < nf; i++)
            for (int a = 0; a < 3; a++) {
                f[i].mn[a] = 1 << 30;
                f[i].mx[a] = -(1 << 30);
                for (const Dir& d : dirs[i]) {
                    f[i].mn[a] = std::min(f[i].mn[a], d[a]);
                    f[i].mx[a] = std::max(f[i].mx[a], d[a]);
                }
            }
    }
    return f;
}

namespace {

// thickness of a field's layer on one side of an axis
int thick(const PriField& f, int a, int side) {
    if (side == 1) return 1;
    return side == 0 ? std::max(-f.mn[a], 0) : std::max(f.mx[a], 0);
}

}  // namespace

PriLayout pri_layout(const ModelDef* m, int nx, int ny, int nz) {
    PriLayout L;
    L.fields = pri_fields(m);
    const int n[3] = {nx, ny, nz};
    size_t off = 0;
    for (int b = 0; b < 27; b++) {
        PriBlock B;
        B.index = b + 1;
        B.side[0] = b % 3;
        B.side[1] = (b / 3) % 3;
        B.side[2] = b / 9;
        B.nodes = 1;
        for (int a = 0; a < 3; a++) if (B.side[a] == 1) B.nodes *= (size_t)n[a];
        size_t layers = 0;
        for (int i = 0; i < (int)L.fields.size(); i++) {
            int ns = 1;
            for (int a = 0; a < 3; a++) ns *= thick(L.fields[i], a, B.side[a]);
            if (ns > 0) {
                B.fields.push_back(i);
                B.layers.push_back(ns);
                layers += (size_t)ns;
            }
        }
        if (layers == 0) continue;
        B.elems = B.nodes * layers;
        B.offset = off;
        off += B.elems;
        L.blocks.push_back(B);
    }
    L.elems = off;
    return L;
}

bool pri_exact(const ModelDef* m, int storage, std::string* why) {
    switch (storage) {
    case ST_F64: case ST_F32:
        return true;
    case ST_F32S: case ST_F16S:
        // float-shift / half-shift store f - shift, with the shift only of
        // the fields that declare one: the d2q9 (MRT) densities, shift = w_i
        if (is_d2q9_mrt(m)) return true;
        if (why) *why = std::string("the reference's ") + (storage == ST_F32S ? "float" : "half") +
                        "-shift build stores " + m->name + "'s densities unshifted, the engine's " +
                        storage_name(storage) + " stores f - w_i";
        return false;
    default:
        if (why) *why = std::string("the reference has no ") + storage_name(storage) + " storage";
        return false;
    }
}

size_t pri_bytes(const ModelDef* m, int storage, int nx, int ny, int nz) {
    return pri_layout(m, nx, ny, nz).elems * (size_t)storage_bytes(storage);
}

namespace {

// an extra density (FP64 in the engine) as a stored value of the format
void extra_to_storage(int storage, double v, void* dst) {
    switch (storage) {
    case ST_F64: std::memcpy(dst, &v, 8); break;
    case ST_F32: case ST_F32S: { float f = (float)v; std::memcpy(dst, &f, 4); break; }
    case ST_F16S: { _Float16 h = (_Float16)(float)v; std::memcpy(dst, &h, 2); break; }
    default: throw std::invalid_argument("storage without a reference layout");
    }
}

double storage_to_extra(int storage, const void* src) {
    switch (storage) {
    case ST_F64: { double v; std::memcpy(&v, src, 8); return v; }
    case ST_F32: case ST_F32S: { float f; std::memcpy(&f, src, 4); return f; }
    case ST_F16S: { _Float16 h; std::memcpy(&h, src, 2); return (double)(float)h; }
    default: throw std::invalid_argument("storage without a reference layout");
    }
}

// Where the put of node coordinate w (along one axis of extent n) sends a
// field, per side of the axis: src/conf.R:907-918 (put_tab, put_sel) with
// the conditions of :979-983.  A node can write several sides (the domain
// block and a margin copy) when the field reaches both ways.
bool put_writes(const PriField& f, int a, int side, int w, int n) {
    const int mn = std::min(f.mn[a], 0), mx = std::max(f.mx[a], 0);
    switch (side) {
    case 0:   // below: only fields reaching below, from the last -mn nodes
        return f.mn[a] < 0 && w >= n + mn;
    case 2:   // above: only fields reaching above, from the first mx nodes
        return f.mx[a] > 0 && w < mx;
    default:  // the domain: all nodes, except the edge a one-way field leaves through
        if (f.mn[a] == f.mx[a] && f.mx[a] > 0 && w < mx) return false;
        if (f.mn[a] == f.mx[a] && f.mn[a] < 0 && w >= n + mn) return false;
        return true;
    }
}

// The node coordinate a position of a block's layer holds: below, the
// position counts the last -mn nodes; in the domain it is the node; above,
// the first mx nodes.
int node_of(const PriField& f, int a, int side, int p, int n) {
    if (side == 0) return n + std::min(f.mn[a], 0) + p;
    return p;
}

template <class Fn>
void for_each_slot(const PriLayout& L, const int n[3], Fn fn) {
    for (const PriBlock& B : L.blocks) {
        size_t base = B.offset;
        for (size_t k = 0; k < B.fields.size(); k++) {
            const PriField& f = L.fields[B.fields[k]];
            int ext[3];
            for (int a = 0; a < 3; a++) ext[a] = B.side[a] == 1 ? n[a] : thick(f, a, B.side[a]);
            size_t idx = base;
            for (int pz = 0; pz < ext[2]; pz++)
                for (int py = 0; py < ext[1]; py++)
                    for (int px = 0; px < ext[0]; px++, idx++) {
                        const int p[3] = {px, py, pz};
                        int w[3];
                        bool written = true;
                        for (int a = 0; a < 3; a++) {
                            w[a] = node_of(f, a, B.side[a], p[a], n[a]);
                            written = written && put_writes(f, a, B.side[a], w[a], n[a]);
                        }
                        const size_t node = (size_t)w[0] + (size_t)n[0] * ((size_t)w[1] + (size_t)n[1] * w[2]);
                        fn(f, idx, node, written);
                    }
            base += (size_t)B.layers[k] * B.nodes;
        }
    }
}

void check_format(const ModelDef* m, int storage) {
    std::string why;
    if (!pri_exact(m, storage, &why))
        throw std::invalid_argument("no reference layout for " + m->name + " in " + storage_name(storage) + ": " + why);
}

}  // namespace

void pri_pack(const ModelDef* m, int storage, int nx, int ny, int nz, const void* planes, size_t pitch,
              const double* extra, void* out) {
    check_format(m, storage);
    const PriLayout L = pri_layout(m, nx, ny, nz);
    const int n[3] = {nx, ny, nz};
    const size_t S = (size_t)storage_bytes(storage);
    const size_t N = (size_t)nx * ny * nz;
    const char* src = (const char*)planes;
    char* dst = (char*)out;
    std::memset(dst, 0, L.elems * S);
    for_each_slot(L, n, [&](const PriField& f, size_t idx, size_t node, bool written) {
        if (!written) return;
        if (f.plane < m->Q)
            std::memcpy(dst + idx * S, src + ((size_t)f.plane * pitch + node) * S, S);
        else
            extra_to_storage(storage, extra[(size_t)(f.plane - m->Q) * N + node], dst + idx * S);
    });
}

void pri_unpack(const ModelDef* m, int storage, int nx, int ny, int nz, const void* in, void* planes,
                size_t pitch, double* extra) {
    check_format(m, storage);
    const PriLayout L = pri_layout(m, nx, ny, nz);
    const int n[3] = {nx, ny, nz};
    const size_t S = (size_t)storage_bytes(storage);
    const size_t N = (size_t)nx * ny * nz;
    const char* src = (const char*)in;
    char* dst = (char*)planes;
    // A node's value is read from the slot its put wrote that is not a
    // margin copy: per axis, the domain side when the put writes there,
    // else the one margin side its edge sends it to.
    std::vector<size_t> block_of(28, (size_t)-1);
    for (size_t b = 0; b < L.blocks.size(); b++) block_of[L.blocks[b].index] = b;
    for (size_t fi = 0; fi < L.fields.size(); fi++) {
        const PriField& f = L.fields[fi];
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    const int w[3] = {x, y, z};
                    int side[3], p[3];
                    for (int a = 0; a < 3; a++) {
                        if (put_writes(f, a, 1, w[a], n[a])) { side[a] = 1; p[a] = w[a]; }
                        else if (put_writes(f, a, 0, w[a], n[a])) { side[a] = 0; p[a] = w[a] - (n[a] + std::min(f.mn[a], 0)); }
                        else { side[a] = 2; p[a] = w[a]; }
                    }
                    const int bi = side[0] + 3 * side[1] + 9 * side[2] + 1;
                    const PriBlock& B = L.blocks.at(block_of[bi]);
                    size_t base = B.offset;
                    size_t k = 0;
                    for (; k < B.fields.size() && B.fields[k] != (int)fi; k++) base += (size_t)B.layers[k] * B.nodes;
                    if (k == B.fields.size()) throw std::logic_error("pri: field missing from its block");
                    int ext[3];
                    for (int a = 0; a < 3; a++) ext[a] = B.side[a] == 1 ? n[a] : thick(f, a, B.side[a]);
                    const size_t idx = base + (size_t)p[0] + (size_t)ext[0] * ((size_t)p[1] + (size_t)ext[1] * p[2]);
                    const size_t node = (size_t)x + (size_t)nx * ((size_t)y + (size_t)ny * z);
                    if (f.plane < m->Q)
                        std::memcpy(dst + ((size_t)f.plane * pitch + node) * S, src + idx * S, S);
                    else
                        extra[(size_t)(f.plane - m->Q) * N + node] = storage_to_extra(storage, src + idx * S);
                }
    }
}

}  // namespace tclbx
