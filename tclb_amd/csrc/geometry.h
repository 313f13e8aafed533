// geometry.h -- the NodeType rasterizer of TCLB case files.
//
// Geometry::load (src/Geometry.cpp.Rt:1140-1184) walks the children of the
// case's <Geometry> element: each child names a NodeType (<Wall>, <MRT>, ...)
// whose value and group mask become the foreground flag; its attributes
// name= (a setting zone), mask= (the bits it may overwrite) and mode=
// (overwrite / fill / change) adjust it, in attribute order; zone= draws a
// named <Zone>; then its children are drawn: primitives (Box, Sphere,
// Wedge, ...) and references to named zones.  Every drawn node gets
// geom = (geom & ~mask) | flag, subject to the mode (src/Geometry.cpp.Rt:
// 307-319).  Regions nest: dx/dy/dz, fx/fy/fz and nx/ny/nz of an element are
// relative to its parent's region (src/Geometry.cpp.Rt:216-304).
//
// The flags are computed on the host into a uint16 array, x fastest, the
// layout FlagOverwrite takes.
#pragma once
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "model.h"
#include "units.h"
#include "xml.h"

namespace tclbx {

// lbRegion (src/Region.h:4-42): origin and size
// cut distances (src/types.h:25-28): cut_t = uint16_t, a crossing at the
// fraction r of a link is r * CUT_MAX, NO_CUT marks an uncut link
constexpr uint16_t NO_CUT = 65535;
constexpr int CUT_MAX = 65000;

struct GeomRegion {
    int dx = 0, dy = 0, dz = 0, nx = 1, ny = 1, nz = 1;
    bool in(int x, int y, int z) const {
        return x >= dx && y >= dy && z >= dz && x - dx < nx && y - dy < ny && z - dz < nz;
    }
    GeomRegion intersect(const GeomRegion& w) const;
    long long size() const { return (long long)nx * ny * nz; }
};

// The <Geometry> defaults of src/def.cpp.Rt:10-33 for model m: the zones
// Inlet, Outlet, Channel and Tunnel, the Mask of every NodeType group and
// the Type of every NodeType.
std::unique_ptr<XmlNode> geometry_defaults(const ModelDef& m);

class Geometry {
public:
    Geometry(const ModelDef& m, int nx, int ny, int nz, const Units& units);
    // rasterizes the children of `geom` (a <Geometry> element, which gets the
    // defaults prepended as the reference does); throws std::runtime_error
    // with the reference's message on an unknown flag, mask, mode, zone or
    // element
    void load(XmlNode& geom);
    const std::vector<uint16_t>& flags() const { return geom_; }
    // cut distances (Geometry::Q, src/Geometry.h:29): empty until a surface
    // STL is drawn, then cuts[d][node] for the 26 links d = direction - 1 in
    // the d3q27 order, NO_CUT where no triangle crosses the link
    const std::vector<uint16_t>& cuts() const { return cuts_; }
    // setting zones by name -> number, DefaultZone = 0, in order of first use
    const std::map<std::string, int>& zones() const { return zones_; }
    const GeomRegion& region() const { return region_; }
    // notices the reference prints while drawing (e.g. about a Sweep's
    // points), in order; cleared by nothing
    const std::vector<std::string>& notes() const { return notes_; }

private:
    enum Mode { OVERWRITE, FILL, CHANGE };
    int val(const XmlAttr* a) const;             // rounded unit value
    int val_p(const XmlAttr* a, char* side) const;
    double val_d(const XmlNode& n, const char* name) const;
    void set_flag(const std::string& name);
    void set_mask(const std::string& name);
    void set_mode(const std::string& mode);
    void set_zone(const std::string& name);
    GeomRegion get_region(const XmlNode* n) const;
    void dot(int x, int y, int z);
    void note(const std::string& m) { notes_.push_back(m); }
    void draw(const XmlNode& n);
    void load_zone(const std::string& name);
    void load_stl(const GeomRegion& reg, const XmlNode& n);
    void load_sweep(const GeomRegion& reg, const XmlNode& n);

    const ModelDef& model_;
    const Units& units_;
    GeomRegion region_, total_;
    std::vector<uint16_t> geom_;
    std::vector<uint16_t> cuts_;
    std::map<std::string, int> zones_;
    XmlNode* fg_xml_ = nullptr;
    uint32_t fg_ = 0, fg_mask_ = 0;
    Mode mode_ = OVERWRITE;
    std::vector<std::string> notes_;
};

}  // namespace tclbx
