// geometry.cpp -- see geometry.h.  The primitives and their arithmetic
// follow src/Geometry.cpp.Rt:322-378 (inSphere, inWedge) and :758-1124
// (Draw); every test is evaluated in double with the same operations so a
// case file gives the reference's NodeType map bit for bit.
#include "geometry.h"

#include <algorithm>
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
        } else if (e == "STL" || e == "Sweep") {
            fail("Geometry element " + e + " is not supported by this engine");
        } else {
            const XmlNode* zn = fg_xml_->find_child_by_attribute("Zone", "name", e);
            if (!zn) fail("Unknown geometry element: " + e);
            draw(*zn);
        }
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
