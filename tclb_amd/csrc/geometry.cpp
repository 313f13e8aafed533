// geometry.cpp -- see geometry.h.  The primitives and their arithmetic
// follow src/Geometry.cpp.Rt:322-378 (inSphere, inWedge) and :758-1124
// (Draw); every test is evaluated in double with the same operations so a
// case file gives the reference's NodeType map bit for bit.
#include "geometry.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>

namespace tclbx {

GeomRegion GeomRegion::intersect(const GeomRegion& w) const {
    GeomRegion r;
    r.dx = std::max(dx, w.dx);
    r.dy = std::max(dy, w.dy);
    r.dz = std::max(dz, w.dz);
    r.nx = std::min(dx + nx, w.dx + w.nx) - r.dx;
    r.ny = std::min(dy + ny, w.dy + w.ny) - r.dy;
    r.nz = std::min(dz + nz, w.dz + w.nz) - r.dz;
    if (r.nx <= 0 || r.ny <= 0 || r.nz <= 0) r.nx = r.ny = r.nz = 0;
    return r;
}

namespace {

[[noreturn]] void fail(const std::string& m) { throw std::runtime_error(m); }

// myround of src/utils.h:34-37: halves away from zero, then truncation
int myround(double v) { return v > 0 ? (int)(v + 0.5) : (int)(v - 0.5); }

// (2x-1)^2 + (2y-1)^2 + (2z-1)^2 < 1
bool in_sphere(double x, double y, double z) {
    x = 2 * x - 1;
    y = 2 * y - 1;
    z = 2 * z - 1;
    return (x * x + y * y + z * z) < 1;
}

bool in_wedge(double x, double y, double z, const std::string& type_in) {
    const std::string type = type_in.empty() ? "UpperLeft" : type_in;
    double delta = 0;
    if (type == "UpperLeft") {
        delta = x - y;
    } else if (type == "UpperRight") {
        x = 1. - x; delta = x - y;
    } else if (type == "LowerLeft") {
        y = 1. - y; delta = x - y;
    } else if (type == "LowerRight") {
        x = 1. - x; y = 1. - y; delta = x - y;
    } else if (type == "UpperLeftXZ") {
        delta = x - z;
    } else if (type == "UpperRightXZ") {
        x = 1. - x; delta = x - z;
    } else if (type == "LowerLeftXZ") {
        z = 1. - z; delta = x - z;
    } else if (type == "LowerRightXZ") {
        x = 1. - x; z = 1. - z; delta = x - z;
    }
    // any other direction leaves delta = 0 and fills the whole box, as the
    // reference does
    return delta < 1e-10;
}

XmlNode* add(XmlNode* p, const char* name, std::initializer_list<std::pair<const char*, std::string>> at) {
    XmlNode* n = p->append_child(name);
    for (auto& a : at) n->attrs.push_back({a.first, a.second});
    return n;
}

void box(XmlNode* z, const char* dx, const char* dy, const char* dz, const char* fx, const char* fy, const char* fz) {
    add(z, "Box", {{"dx", dx}, {"dy", dy}, {"dz", dz}, {"fx", fx}, {"fy", fy}, {"fz", fz}});
}

}  // namespace

std::unique_ptr<XmlNode> geometry_defaults(const ModelDef& m) {
    auto g = std::make_unique<XmlNode>();
    g->name = "Geometry";
    XmlNode* z;
    z = add(g.get(), "Zone", {{"name", "Inlet"}});
    box(z, "0", "0", "0", "0", "-1", "-1");
    z = add(g.get(), "Zone", {{"name", "Outlet"}});
    box(z, "-1", "0", "0", "-1", "-1", "-1");
    z = add(g.get(), "Zone", {{"name", "Channel"}});
    box(z, "0", "0", "0", "-1", "0", "-1");
    box(z, "0", "-1", "0", "-1", "-1", "-1");
    z = add(g.get(), "Zone", {{"name", "Tunnel"}});
    box(z, "0", "0", "0", "-1", "0", "-1");
    box(z, "0", "-1", "0", "-1", "-1", "-1");
    box(z, "0", "0", "0", "-1", "-1", "0");
    box(z, "0", "0", "-1", "-1", "-1", "-1");
    // the reference lists a second Tunnel and a second Inlet; the first
    // definition of a name is the one used, so these never take effect
    z = add(g.get(), "Zone", {{"name", "Tunnel"}});
    box(z, "0", "0", "0", "0", "-1", "-1");
    z = add(g.get(), "Zone", {{"name", "Inlet"}});
    box(z, "0", "0", "0", "0", "-1", "-1");
    add(g.get(), "Mask", {{"name", "NONE"}, {"value", "0"}});
    add(g.get(), "Type", {{"name", "None"}, {"value", "0"}, {"mask", "NONE"}});
    for (auto& gr : m.groups) add(g.get(), "Mask", {{"name", gr.name}, {"value", std::to_string(gr.mask)}});
    for (auto& t : m.nodetypes)
        add(g.get(), "Type", {{"name", t.name}, {"value", std::to_string(t.value)}, {"mask", t.group}});
    return g;
}

Geometry::Geometry(const ModelDef& m, int nx, int ny, int nz, const Units& units)
    : model_(m), units_(units) {
    if (nx <= 0 || ny <= 0 || nz <= 0) fail("geometry size must be positive");
    region_.nx = nx; region_.ny = ny; region_.nz = nz;
    total_ = region_;
    geom_.assign((size_t)nx * ny * nz, 0);
    zones_["DefaultZone"] = 0;
}

int Geometry::val(const XmlAttr* a) const {
    if (!a) fail("Attribute without value and default");
    return myround(units_.alt(a->value));
}

int Geometry::val_p(const XmlAttr* a, char* side) const {
    std::string t = a->value;
    *side = t.empty() ? '\0' : t[0];
    if (*side == '>' || *side == '<') t = t.substr(1);
    else *side = '+';
    return myround(units_.alt(t));
}

double Geometry::val_d(const XmlNode& n, const char* name) const {
    return units_.alt(n.get(name));
}

void Geometry::set_flag(const std::string& name) {
    const XmlNode* t = fg_xml_->find_child_by_attribute("Type", "name", name);
    if (!t) fail("Unknown flag (in xml): " + name);
    fg_ = (uint16_t)xml_as_int(t->get("value"));
    set_mask(t->get("mask"));
    mode_ = OVERWRITE;
}

void Geometry::set_mask(const std::string& name) {
    const XmlNode* t = fg_xml_->find_child_by_attribute("Mask", "name", name);
    if (!t) fail("Unknown mask (in xml): " + name);
    fg_mask_ = (uint16_t)(t->has("value") ? xml_as_int(t->get("value")) : -1);
}

void Geometry::set_mode(const std::string& mode) {
    if (mode == "overwrite") mode_ = OVERWRITE;
    else if (mode == "fill") mode_ = FILL;
    else if (mode == "change") mode_ = CHANGE;
    else fail("Unknown mode (in xml): " + mode);
}

void Geometry::set_zone(const std::string& name) {
    int k;
    auto it = zones_.find(name);
    if (it != zones_.end()) {
        k = it->second;
    } else {
        k = (int)zones_.size();
        zones_[name] = k;
    }
    if (k >= model_.zone_max) fail("too many setting zones (" + std::to_string(k + 1) + ")");
    const GroupDef* zg = model_.group("SETTINGZONE");
    fg_ = (uint16_t)((fg_ & ~zg->mask) | ((uint32_t)k << model_.zone_shift));
    fg_mask_ = (uint16_t)(fg_mask_ | zg->mask);
}

GeomRegion Geometry::get_region(const XmlNode* n) const {
    if (!n) return GeomRegion();
    if (n->name == "Geometry") return total_;
    GeomRegion r = get_region(n->parent);
    char side;
    if (const XmlAttr* a = n->attr("dx")) {
        int w = val_p(a, &side);
        if (side == '<') w = r.nx + w;
        else if (side == '+' && w < 0) w = r.nx + w;
        r.dx += w;
        r.nx -= w;
    }
    if (const XmlAttr* a = n->attr("dy")) {
        int w = val_p(a, &side);
        if (side == '<') w = r.ny + w;
        else if (side == '+' && w < 0) w = r.ny + w;
        r.dy += w;
        r.ny -= w;
    }
    if (const XmlAttr* a = n->attr("dz")) {
        int w = val_p(a, &side);
        if (side == '<') w = r.nz + w;
        else if (side == '+' && w < 0) w = r.nz + w;
        r.dz += w;
        r.nz -= w;
    }
    if (const XmlAttr* a = n->attr("fx")) {
        int w = val(a);
        if (w < 0) w = r.nx + w + r.dx;
        r.nx = w - r.dx + 1;
    }
    if (const XmlAttr* a = n->attr("fy")) {
        int w = val(a);
        if (w < 0) w = r.ny + w + r.dy;
        r.ny = w - r.dy + 1;
    }
    if (const XmlAttr* a = n->attr("fz")) {
        int w = val(a);
        if (w < 0) w = r.nz + w + r.dz;
        r.nz = w - r.dz + 1;
    }
    if (const XmlAttr* a = n->attr("nx")) r.nx = val(a);
    if (const XmlAttr* a = n->attr("ny")) r.ny = val(a);
    if (const XmlAttr* a = n->attr("nz")) r.nz = val(a);
    return r;
}

void Geometry::dot(int x, int y, int z) {
    if (!region_.in(x, y, z)) return;
    uint16_t& g = geom_[(size_t)(x - region_.dx) +
                        (size_t)region_.nx * ((size_t)(y - region_.dy) + (size_t)region_.ny * (z - region_.dz))];
    if (g & fg_mask_) {
        if (mode_ == FILL) return;
    } else {
        if (mode_ == CHANGE) return;
    }
    g = (uint16_t)((g & ~fg_mask_) | fg_);
}

void Geometry::draw(const XmlNode& node) {
    for (auto& cp : node.children) {
        const XmlNode& n = *cp;
        GeomRegion reg = get_region(&n);
        const std::string& e = n.name;
        if (e == "Box") {
            reg = region_.intersect(reg);
            for (int z = reg.dz; z < reg.dz + reg.nz; z++)
                for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                    for (int x = reg.dx; x < reg.dx + reg.nx; x++) dot(x, y, z);
        } else if (e == "HalfSphere") {
            for (int z = reg.dz; z < reg.dz + reg.nz; z++)
                for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                    for (int x = reg.dx; x < reg.dx + reg.nx; x++)
                        if (in_sphere((.5 + x - reg.dx) / reg.nx, 0.5 - (.5 + y - reg.dy) / reg.ny / 2.,
                                      (.5 + z - reg.dz) / reg.nz))
                            dot(x, y, z);
        } else if (e == "Sphere") {
            for (int z = reg.dz; z < reg.dz + reg.nz; z++)
                for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                    for (int x = reg.dx; x < reg.dx + reg.nx; x++)
                        if (in_sphere((.5 + x - reg.dx) / reg.nx, (.5 + y - reg.dy) / reg.ny, (.5 + z - reg.dz) / reg.nz))
                            dot(x, y, z);
        } else if (e == "OffgridSphere") {
            const double x0 = val_d(n, "x"), y0 = val_d(n, "y"), z0 = val_d(n, "z");
            double Rx, Ry, Rz;
            if (!n.has("R")) {
                Rx = val_d(n, "Rx"); Ry = val_d(n, "Ry"); Rz = val_d(n, "Rz");
            } else {
                Rx = Ry = Rz = val_d(n, "R");
            }
            const double Rx2 = Rx * Rx, Ry2 = Ry * Ry, Rz2 = Rz * Rz;
            reg.dx = (int)(x0 - Rx - 5); reg.dy = (int)(y0 - Ry - 5); reg.dz = (int)(z0 - Rz - 5);
            reg.nx = (int)(2 * Rx + 10); reg.ny = (int)(2 * Ry + 10); reg.nz = (int)(2 * Rz + 10);
            for (int z = reg.dz; z < reg.dz + reg.nz; z++)
                for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                    for (int x = reg.dx; x < reg.dx + reg.nx; x++) {
                        const double xx = (.5 + x - x0), yy = (.5 + y - y0), zz = (.5 + z - z0);
                        if (xx * xx / Rx2 + yy * yy / Ry2 + zz * zz / Rz2 < 1.) dot(x, y, z);
                    }
        } else if (e == "OffgridPipe" || e == "ZPipe" || e == "YPipe" || e == "XPipe") {
            const double x0 = val_d(n, "x"), y0 = val_d(n, "y"), z0 = val_d(n, "z");
            // the two radii across the pipe's axis
            const bool xp = e == "XPipe", yp = e == "YPipe";
            const char* ra = xp ? "Ry" : "Rx";
            const char* rb = (xp || yp) ? "Rz" : "Ry";
            double Ra, Rb;
            if (!n.has("R")) {
                Ra = val_d(n, ra); Rb = val_d(n, rb);
            } else {
                Ra = Rb = val_d(n, "R");
            }
            const double Ra2 = Ra * Ra, Rb2 = Rb * Rb;
            if (xp) {
                reg.dy = (int)(y0 - Ra - 5); reg.dz = (int)(z0 - Rb - 5);
                reg.ny = (int)(2 * Ra + 10); reg.nz = (int)(2 * Rb + 10);
            } else if (yp) {
                reg.dx = (int)(x0 - Ra - 5); reg.dz = (int)(z0 - Rb - 5);
                reg.nx = (int)(2 * Ra + 10); reg.nz = (int)(2 * Rb + 10);
            } else {
                reg.dx = (int)(x0 - Ra - 5); reg.dy = (int)(y0 - Rb - 5);
                reg.nx = (int)(2 * Ra + 10); reg.ny = (int)(2 * Rb + 10);
            }
            for (int z = reg.dz; z < reg.dz + reg.nz; z++)
                for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                    for (int x = reg.dx; x < reg.dx + reg.nx; x++) {
                        bool in;
                        if (e == "OffgridPipe") {
                            const double xx = (.5 + x - x0), yy = (.5 + y - y0);
                            in = xx * xx / Ra2 + yy * yy / Rb2 < 1.;
                        } else if (xp) {
                            const double yy = (.5 + y - y0), zz = (.5 + z - z0);
                            in = zz * zz / Rb2 + yy * yy / Ra2 < 1.;
                        } else if (yp) {
                            // the reference centres x on the y attribute here
                            const double xx = (.5 + x - y0), zz = (.5 + z - z0);
                            in = zz * zz / Rb2 + xx * xx / Ra2 < 1.;
                        } else {
                            // ZPipe: likewise x measured from y0
                            const double xx = (.5 + x - y0), yy = (.5 + y - y0);
                            in = yy * yy / Rb2 + xx * xx / Ra2 < 1.;
                        }
                        if (in) dot(x, y, z);
                    }
        } else if (e == "XAnnulus") {
            const double y0 = val_d(n, "y"), z0 = val_d(n, "z");
            (void)val_d(n, "x");
            double Ry_o, Rz_o, Ry_i, Rz_i;
            if (!n.has("Ro")) {
                Ry_i = val_d(n, "Ry_i"); Rz_i = val_d(n, "Rz_i");
                Ry_o = val_d(n, "Ry_o"); Rz_o = val_d(n, "Rz_o");
            } else {
                const double Ro = val_d(n, "Ro"), Ri = val_d(n, "Ri");
                Ry_o = Rz_o = Ro;
                Ry_i = Rz_i = Ri;
            }
            const double Ry2_o = Ry_o * Ry_o, Rz2_o = Rz_o * Rz_o, Ry2_i = Ry_i * Ry_i, Rz2_i = Rz_i * Rz_i;
            reg.dy = (int)(y0 - Ry_o - 5); reg.dz = (int)(z0 - Rz_o - 5);
            reg.ny = (int)(2 * Ry_o + 10); reg.nz = (int)(2 * Rz_o + 10);
            for (int z = reg.dz; z < reg.dz + reg.nz; z++)
                for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                    for (int x = reg.dx; x < reg.dx + reg.nx; x++) {
                        const double yy = (.5 + y - y0), zz = (.5 + z - z0);
                        if ((zz * zz / Rz2_o + yy * yy / Ry2_o < 1.) && (zz * zz / Rz2_i + yy * yy / Ry2_i > 1.))
                            dot(x, y, z);
                    }
        } else if (e == "Pipe") {
            for (int z = reg.dz - 1; z < reg.dz + reg.nz + 1; z++)
                for (int y = reg.dy - 1; y < reg.dy + reg.ny + 1; y++)
                    for (int x = reg.dx; x < reg.dx + reg.nx; x++)
                        if (!in_sphere(0.5, (.5 + y - reg.dy) / reg.ny, (.5 + z - reg.dz) / reg.nz)) dot(x, y, z);
        } else if (e == "PipeY") {
            for (int z = reg.dz - 1; z < reg.dz + reg.nz + 1; z++)
                for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                    for (int x = reg.dx - 1; x < reg.dx + reg.nx + 1; x++)
                        if (!in_sphere(0.5, (.5 + x - reg.dx) / reg.nx, (.5 + z - reg.dz) / reg.nz)) dot(x, y, z);
        } else if (e == "PipeZ") {
            for (int z = reg.dz; z < reg.dz + reg.nz; z++)
                for (int y = reg.dy - 1; y < reg.dy + reg.ny + 1; y++)
                    for (int x = reg.dx - 1; x < reg.dx + reg.nx + 1; x++)
                        if (!in_sphere(0.5, (.5 + y - reg.dy) / reg.ny, (.5 + x - reg.dx) / reg.nx)) dot(x, y, z);
        } else if (e == "Cylinder") {
            for (int z = reg.dz - 1; z < reg.dz + reg.nz + 1; z++)
                for (int y = reg.dy - 1; y < reg.dy + reg.ny + 1; y++)
                    for (int x = reg.dx; x < reg.dx + reg.nx; x++)
                        if (in_sphere((.5 + x - reg.dx) / reg.nx, (.5 + y - reg.dy) / reg.ny, 0.5)) dot(x, y, z);
        } else if (e == "Wedge") {
            const std::string type = n.get("direction");
            for (int z = reg.dz; z < reg.dz + reg.nz; z++)
                for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                    for (int x = reg.dx; x < reg.dx + reg.nx; x++)
                        if (in_wedge((x - reg.dx) / (reg.nx - 1.), (y - reg.dy) / (reg.ny - 1.),
                                     (z - reg.dz) / (reg.nz - 1.), type))
                            dot(x, y, z);
        } else if (e == "Text") {
            // integers, z fastest then y then x, one per node of the region;
            // nonzero ones inside the parent's region are drawn
            const GeomRegion crop = region_.intersect(get_region(n.parent));
            if (!n.has("file")) fail("No 'file' attribute in 'Text' element in xml conf");
            FILE* f = std::fopen(n.get("file").c_str(), "rt");
            if (!f) fail("Could not open file: " + n.get("file"));
            for (int x = reg.dx; x < reg.dx + reg.nx; x++)
                for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                    for (int z = reg.dz; z < reg.dz + reg.nz; z++) {
                        int v;
                        if (std::fscanf(f, "%d", &v) == EOF) {
                            std::fclose(f);
                            fail("File (" + n.get("file") + ") ended while reading");
                        }
                        if (v != 0 && crop.in(x, y, z)) dot(x, y, z);
                    }
            std::fclose(f);
        } else if (e == "STL") {
            load_stl(reg, n);
        } else if (e == "Sweep") {
            load_sweep(reg, n);
        } else {
            const XmlNode* zn = fg_xml_->find_child_by_attribute("Zone", "name", e);
            if (!zn) fail("Unknown geometry element: " + e);
            draw(*zn);
        }
    }
}

// ------------------------------------------------------------------ STL ---
namespace {

// a binary STL record (50 bytes): normal, three vertices, attribute
#pragma pack(push, 1)
struct StlTri {
    float norm[3], p1[3], p2[3], p3[3];
    short v;
};
#pragma pack(pop)
static_assert(sizeof(StlTri) == 50, "STL record");

// GaussSolve of src/unit.cpp:13-49 on a column-major n x n matrix
// (A[col*n + row]), row swap only on a zero pivot
void gauss_solve(double* A, double* b, double* x, int n) {
    auto z = [](double v) { return std::fabs(v) < 1e-10; };
    auto I = [n](int r, int c) { return c * n + r; };
    for (int k = 0; k < n - 1; k++) {
        if (z(A[I(k, k)]))
            for (int i = k + 1; i < n; i++)
                if (!z(A[I(i, k)])) {
                    for (int j = 0; j < n; j++) std::swap(A[I(k, j)], A[I(i, j)]);
                    std::swap(b[k], b[i]);
                    break;
                }
        for (int i = k + 1; i < n; i++) {
            const double m = A[I(i, k)] / A[I(k, k)];
            for (int j = 0; j < n; j++) A[I(i, j)] -= m * A[I(k, j)];
            b[i] -= m * b[k];
        }
    }
    for (int i = n - 1; i >= 0; i--) {
        double sum = 0.0;
        int j;
        for (j = n - 1; j >= i + 1; j--) sum += A[I(i, j)] * x[j];
        x[i] = (b[i] - sum) / A[I(i, j)];
    }
}

// calcCut (src/Geometry.cpp.Rt:479-504): where does the link from (x,y,z)
// along d cross the triangle?  Solves x + r3 d = r0 p1 + r1 p2 + r2 p3 with
// r0 + r1 + r2 = 1; a cut when every r >= 0 and r3 <= 1, at
// r3 * CUT_MAX (truncated to cut_t), else NO_CUT.  A singular system (a
// link in the triangle's plane) gives NaNs, which pass every test as they do
// there; the reference's conversion of NaN * CUT_MAX to cut_t (a 32-bit
// truncating conversion on x86-64, then the low 16 bits) gives 0, so such a
// link is cut at 0.
uint16_t calc_cut(const StlTri& t, double x, double y, double z, int dx, int dy, int dz) {
    double m[16], b[4] = {0, 0, 0, 1}, r[4] = {0, 0, 0, 0};
    m[0] = x - t.p1[0]; m[1] = y - t.p1[1]; m[2] = z - t.p1[2]; m[3] = 1;
    m[4] = x - t.p2[0]; m[5] = y - t.p2[1]; m[6] = z - t.p2[2]; m[7] = 1;
    m[8] = x - t.p3[0]; m[9] = y - t.p3[1]; m[10] = z - t.p3[2]; m[11] = 1;
    m[12] = dx; m[13] = dy; m[14] = dz; m[15] = 0;
    gauss_solve(m, b, r, 4);
    if (r[0] < 0 || r[1] < 0 || r[2] < 0 || r[3] < 0 || r[3] > 1) return NO_CUT;
    const double q = r[3] * CUT_MAX;
    if (std::isnan(q)) return 0;
    return (uint16_t)(int)q;
}

}  // namespace

// Geometry::loadSTL with transformSTL (src/Geometry.cpp.Rt:381-472,507-686):
// the mesh is rotated (Xrot, Yrot, Zrot), scaled, moved (x, y, z), rounded
// to 1e-5 and shifted by -1/2 plus a small fixed offset, all on float
// vertices.  side="in"/"out": a node is drawn when a ray along ray_axis from
// the region's start crosses the mesh an odd (in) or even (out) number of
// times before it; side="surface": when any of its 26 links crosses a
// triangle, and each link keeps the nearest crossing as its cut distance.
void Geometry::load_stl(const GeomRegion& reg, const XmlNode& n) {
    int inside_out = 0, axis = 1;
    if (!n.has("file")) fail("No 'file' attribute in 'STL' element in xml conf");
    if (n.has("side")) {
        const std::string side = n.get("side");
        if (side == "in") inside_out = 0;
        else if (side == "out") inside_out = 1;
        else if (side == "surface") inside_out = 2;
        else fail("'side' in 'STL' element have to be 'in', 'out' or 'surface'");
    }
    if (n.has("ray_axis")) {
        if (inside_out > 1) fail("the 'ray_axis' setting in 'STL' makes sens only with 'in' or 'out' side");
        const std::string a = n.get("ray_axis");
        if (a == "x") axis = 0;
        else if (a == "y") axis = 1;
        else if (a == "z") axis = 2;
        else fail("'ray_axis' in 'STL' element have to be 'x', 'y' or 'z'");
    }
    FILE* f = std::fopen(n.get("file").c_str(), "rb");
    if (!f) fail("'STL' element: " + n.get("file") + " doesn't exists or cannot be opened");
    char header[80];
    int ntri = 0;
    const bool ok = std::fread(header, 1, 80, f) == 80 && std::fread(&ntri, sizeof(int), 1, f) == 1;
    if (!ok || std::strncmp(header, "solid", 5) == 0) {
        std::fclose(f);
        fail("'STL' element " + n.get("file") + " is not in binary format!");
    }
    if (ntri < 0) { std::fclose(f); fail("bad STL triangle count"); }
    std::vector<StlTri> tri((size_t)ntri);
    const size_t got = std::fread(tri.data(), sizeof(StlTri), (size_t)ntri, f);
    std::fclose(f);
    if (got != (size_t)ntri) fail("STL file " + n.get("file") + " ended early");

    // transformSTL
    const char* axes[3] = {"X", "Y", "Z"};
    const int id1s[3] = {1, 2, 0}, id2s[3] = {2, 0, 1};
    for (int d = 0; d < 3; d++) {
        const XmlAttr* a = n.attr(std::string(axes[d]) + "rot");
        if (!a) continue;
        const double v = units_.alt(a->value);
        const int i1 = id1s[d], i2 = id2s[d];
        for (auto& t : tri)
            for (float* p : {t.p1, t.p2, t.p3}) {
                const double yy = p[i1], zz = p[i2];
                p[i1] = (float)(std::cos(v) * yy - std::sin(v) * zz);
                p[i2] = (float)(std::sin(v) * yy + std::cos(v) * zz);
            }
    }
    if (n.has("scale")) {
        const double v = units_.alt(n.get("scale"));
        for (auto& t : tri)
            for (float* p : {t.p1, t.p2, t.p3})
                for (int j = 0; j < 3; j++) p[j] = (float)(v * p[j]);
    }
    for (int j = 0; j < 3; j++) {
        const char* nm[3] = {"x", "y", "z"};
        if (!n.has(nm[j])) continue;
        const double v = units_.alt(n.get(nm[j]));
        for (auto& t : tri)
            for (float* p : {t.p1, t.p2, t.p3}) p[j] = (float)(p[j] + v);
    }
    const double sm_diff[3] = {0.1403e-4, 0.1687e-4, 0.1987e-4};
    for (auto& t : tri)
        for (int j = 0; j < 3; j++)
            for (float* p : {t.p1, t.p2, t.p3}) {
                p[j] = (float)(myround(p[j] * 1e5) * 1e-5);
                p[j] = (float)(p[j] + (sm_diff[j] - 0.5));
            }

    std::vector<signed char> lev((size_t)std::max(0LL, reg.size()), (signed char)inside_out);
    // rays through a triangle; edge hits; resolved negatively; unresolved
    size_t hits[4] = {0, 0, 0, 0};
    auto roff = [&reg](int x, int y, int z) {
        return (size_t)(x - reg.dx) + (size_t)reg.nx * ((size_t)(y - reg.dy) + (size_t)reg.ny * (z - reg.dz));
    };
    static const int d27[27][3] = {
        {0, 0, 0},   {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},   {1, 1, 0},   {-1, 1, 0},  {0, -1, 0},
        {1, -1, 0},  {-1, -1, 0}, {0, 0, 1},  {1, 0, 1},   {-1, 0, 1},  {0, 1, 1},   {1, 1, 1},
        {-1, 1, 1},  {0, -1, 1},  {1, -1, 1}, {-1, -1, 1}, {0, 0, -1},  {1, 0, -1},  {-1, 0, -1},
        {0, 1, -1},  {1, 1, -1},  {-1, 1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, -1, -1}};
    for (const StlTri& t : tri) {
        int mn[3], mx[3];
        for (int j = 0; j < 3; j++) {
            mn[j] = (int)std::ceil(t.p1[j]);
            if (t.p2[j] < mn[j]) mn[j] = (int)std::ceil(t.p2[j]);
            if (t.p3[j] < mn[j]) mn[j] = (int)std::ceil(t.p3[j]);
            mx[j] = (int)std::floor(t.p1[j]);
            if (t.p2[j] > mx[j]) mx[j] = (int)std::floor(t.p2[j]);
            if (t.p3[j] > mx[j]) mx[j] = (int)std::floor(t.p3[j]);
            mn[j] -= 1;
            mx[j] += 1;
        }
        if (inside_out == 2) {
            for (int x = mn[0]; x <= mx[0]; x++)
                for (int z = mn[2]; z <= mx[2]; z++)
                    for (int y = mn[1]; y <= mx[1]; y++) {
                        if (!region_.in(x, y, z)) continue;
                        // ActivateCuts: the first surface allocates the cuts, all NO_CUT
                        if (cuts_.empty()) cuts_.assign(26 * (size_t)region_.size(), NO_CUT);
                        const size_t k = (size_t)(x - region_.dx) +
                                         (size_t)region_.nx * ((size_t)(y - region_.dy) + (size_t)region_.ny * (z - region_.dz));
                        for (int d = 1; d < 27; d++) {
                            const uint16_t q = calc_cut(t, x, y, z, d27[d][0], d27[d][1], d27[d][2]);
                            uint16_t& c = cuts_[(size_t)(d - 1) * (size_t)region_.size() + k];
                            if (q < c) c = q;  // the nearest crossing along the link
                            if (q != NO_CUT) dot(x, y, z);
                        }
                    }
            continue;
        }
        const int a1 = axis, a2 = (axis + 1) % 3, a3 = (axis + 2) % 3;
        const int l1 = a1 == 0 ? reg.dx : (a1 == 1 ? reg.dy : reg.dz);
        // the differences are taken in float, as the vertices are floats
        const double v1[2] = {t.p2[a2] - t.p1[a2], t.p2[a3] - t.p1[a3]};
        const double v2[2] = {t.p3[a2] - t.p1[a2], t.p3[a3] - t.p1[a3]};
        const double c0 = v1[0] * v2[1] - v1[1] * v2[0];
        const double dv[2] = {-0.5694552, 0.8220224};  // tie-break direction for rays through edges
        for (int x2 = mn[a2]; x2 <= mx[a2]; x2++)
            for (int x3 = mn[a3]; x3 <= mx[a3]; x3++) {
                const double v[2] = {x2 - t.p1[a2], x3 - t.p1[a3]};  // float
                double c1 = v1[0] * v[1] - v1[1] * v[0];
                double c2 = v[0] * v2[1] - v[1] * v2[0];
                c1 /= c0;
                c2 /= c0;
                const double c3 = 1. - c1 - c2;
                const double dc1 = (v1[0] * dv[1] - v1[1] * dv[0]) / c0;
                const double dc2 = (dv[0] * v2[1] - dv[1] * v2[0]) / c0;
                const double dc3 = -dc1 - dc2;
                const double cs[3] = {c1, c2, c3}, dcs[3] = {dc1, dc2, dc3};
                int topo = 0;
                for (int k = 0; k < 3; k++) {
                    if (cs[k] == 0) {
                        // the ray runs through an edge: counted as the ray
                        // moved a little along dv would be
                        hits[1]++;
                        if (dcs[k] > 0) topo++;
                        else if (dcs[k] == 0) hits[3]++;
                        else hits[2]++;
                    } else if (cs[k] > 0) {
                        topo++;
                    }
                }
                if (topo != 3) continue;
                hits[0]++;
                const double h = t.p1[a1] * c3 + t.p2[a1] * c2 + t.p3[a1] * c1;
                for (int x1 = l1; x1 <= h; x1++) {
                    int x, y, z;
                    if (axis == 0) { x = x1; y = x2; z = x3; }
                    else if (axis == 1) { x = x3; y = x1; z = x2; }
                    else { x = x2; y = x3; z = x1; }
                    if (reg.in(x, y, z)) lev[roff(x, y, z)]++;
                }
            }
    }
    if (inside_out != 2) {
        for (int z = reg.dz; z < reg.dz + reg.nz; z++)
            for (int y = reg.dy; y < reg.dy + reg.ny; y++)
                for (int x = reg.dx; x < reg.dx + reg.nx; x++)
                    if (lev[roff(x, y, z)] % 2 == 1) dot(x, y, z);
        // the reference's report of the ray hits (loadSTL's output/notice)
        note("STL: triangle hits: " + std::to_string(hits[0]));
        if (hits[1] > 0) {
            note("STL: \\_ edge hits: " + std::to_string(hits[1]));
            note("STL:    \\_ resolved negatively: " + std::to_string(hits[2]));
            note("STL:    \\_ resolved positively: " + std::to_string(hits[1] - hits[3] - hits[2]));
            if (hits[3] > 0)
                note("STL:    \\_ could not be resolved: " + std::to_string(hits[3]) + " (this can cause problems!)");
        }
    }
}

// ---------------------------------------------------------------- Sweep ---
namespace {

// knots of a clamped uniform B-spline: k+1 zeros, (i-k)/(n-k), ones
// (knot_bs of src/spline.h:9-14)
double knot(int i, int n, int k) {
    if (i < k + 1) return 0;
    if (i < n) return 1.0 * (i - k) / (n - k);
    return 1;
}

// de Boor's recursion for the order-k clamped B-spline through the control
// values p at x in [0,1) (bspline_mod of src/spline.h:16-35)
double bspline(double x, std::vector<double> p, int k) {
    const int n = (int)p.size();
    int i = (int)std::floor(x * (n - k)) + k;
    if (k > n - 1) k = n - 1;
    if (i > n - 1) i = n - 1;
    if (i < k) i = k;
    for (int j = k; j > 0; j--)
        for (int l = 0; l < j; l++) {
            const double a = (x - knot(i - l, n, k)) / (knot(i - l + j, n, k) - knot(i - l, n, k));
            p[i - l] = a * p[i - l] + (1 - a) * p[i - l - 1];
        }
    return p[i];
}

}  // namespace

// Geometry::loadSweep (src/Geometry.cpp.Rt:688-755): balls of radius r
// centred on a B-spline through the <Point x y z r> children, sampled every
// `step` of the curve parameter (default 1e-4, or 1/steps); a sample is
// drawn only when centre or radius moved by 1/4 node or more in some
// coordinate since the last drawn one.
void Geometry::load_sweep(const GeomRegion& reg, const XmlNode& node) {
    int order = 1;
    if (node.has("order")) order = xml_as_int(node.get("order"));
    double dl = 1e-4;
    if (node.has("step")) dl = xml_as_double(node.get("step"));
    if (node.has("steps")) dl = 1 / units_.alt(node.get("steps"));
    double def_r = 1;
    if (node.has("r")) def_r = units_.alt(node.get("r"));
    std::vector<double> px, py, pz, pr;
    for (const auto& c : node.children) {
        if (c->name == "Point") {
            px.push_back(units_.alt(c->get("x")));
            py.push_back(units_.alt(c->get("y")));
            pz.push_back(units_.alt(c->get("z")));
            pr.push_back(c->has("r") ? units_.alt(c->get("r"), def_r) : def_r);
        } else {
            note("Unknown element " + c->name + " in Sweep");
        }
    }
    const int n = (int)px.size();
    if (n == 0) {
        note("No points in Sweep");
        return;
    }
    if (order > n - 1) {
        note("For order " + std::to_string(order) + " b-spline you need at least " + std::to_string(order + 1) +
             " points");
        order = n - 1;
    }
    if (!(dl > 0)) fail("Sweep: the step has to be positive");
    double x0_ = -999.0, y0_ = -999.0, z0_ = -999.0, r0_ = -999.0;
    for (double l = 0; l < 1; l += dl) {
        const double x0 = bspline(l, px, order), y0 = bspline(l, py, order);
        const double z0 = bspline(l, pz, order), r = bspline(l, pr, order);
        if (std::fabs(x0 - x0_) < 0.25 && std::fabs(y0 - y0_) < 0.25 && std::fabs(z0 - z0_) < 0.25 &&
            std::fabs(r - r0_) < 0.25)
            continue;
        x0_ = x0; y0_ = y0; z0_ = z0; r0_ = r;
        GeomRegion ball;
        ball.dx = (int)(x0 - r - 1); ball.dy = (int)(y0 - r - 1); ball.dz = (int)(z0 - r - 1);
        ball.nx = ball.ny = ball.nz = (int)(2 * r + 2);
        const GeomRegion t = reg.intersect(ball);
        for (int x = t.dx; x < t.dx + t.nx; x++)
            for (int y = t.dy; y < t.dy + t.ny; y++)
                for (int z = t.dz; z < t.dz + t.nz; z++)
                    if ((.5 + x - x0) * (.5 + x - x0) + (.5 + y - y0) * (.5 + y - y0) + (.5 + z - z0) * (.5 + z - z0) <
                        r * r)
                        dot(x, y, z);
    }
}

void Geometry::load_zone(const std::string& name) {
    const XmlNode* zn = fg_xml_->find_child_by_attribute("Zone", "name", name);
    // the reference prints an error for an unknown zone= and goes on
    if (zn) draw(*zn);
}

void Geometry::load(XmlNode& node) {
    fg_xml_ = &node;
    auto defs = geometry_defaults(model_);
    for (auto& z : defs->children) {
        const XmlAttr* a = z->attr("name");
        if (!a) continue;
        if (node.find_child_by_attribute(z->name, "name", a->value)) continue;
        node.prepend_copy(*z);
    }
    for (auto& cp : node.children) {
        XmlNode& n = *cp;
        if (n.name == "Zone" || n.name == "Type" || n.name == "Mask") continue;
        set_flag(n.name);
        for (auto& a : n.attrs) {
            if (a.name == "name") set_zone(a.value);
            else if (a.name == "mask") set_mask(a.value);
            else if (a.name == "mode") set_mode(a.value);
        }
        if (n.has("zone")) load_zone(n.get("zone"));
        draw(n);
    }
}

}  // namespace tclbx
