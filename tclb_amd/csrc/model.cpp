// model.cpp -- model tables (see model.h for the contract they follow).
#include "model.h"

#include <algorithm>
#include <cmath>
#include <map>

namespace tclbx {

int ModelDef::setting_id(const std::string& n) const {
    for (size_t i = 0; i < settings.size(); i++) if (settings[i].name == n) return (int)i;
    return -1;
}
int ModelDef::zonal_id(const std::string& n) const {
    for (size_t i = 0; i < zonal.size(); i++) if (zonal[i].name == n) return (int)i;
    return -1;
}
int ModelDef::global_id(const std::string& n) const {
    for (size_t i = 0; i < globals.size(); i++) if (globals[i].name == n) return (int)i;
    return -1;
}
int ModelDef::quantity_id(const std::string& n) const {
    for (size_t i = 0; i < quantities.size(); i++) if (quantities[i].name == n) return (int)i;
    return -1;
}
const NodeTypeDef* ModelDef::nodetype(const std::string& n) const {
    for (auto& t : nodetypes) if (t.name == n) return &t;
    return nullptr;
}
const GroupDef* ModelDef::group(const std::string& n) const {
    for (auto& g : groups) if (g.name == n) return &g;
    return nullptr;
}

namespace {

// NodeType packing of src/conf.R:600-699: groups sorted by name, each gets
// ceil(log2(n+1)) bits and values 1..n shifted; the zone takes the remaining
// bits of the 16-bit flag; None (0/0) and Clear (0/ALL) are appended.
void pack_nodetypes(ModelDef& m, const std::vector<std::pair<std::string, std::string>>& types) {
    std::map<std::string, std::vector<std::string>> by_group;  // sorted by group name
    for (auto& t : types) by_group[t.second].push_back(t.first);
    uint32_t shift_val = 1;
    int shift_num = 0;
    for (auto& g : by_group) {
        int n = (int)g.second.size();
        int l = (int)std::ceil(std::log2((double)n + 1.0));
        GroupDef gd;
        gd.name = g.first;
        gd.mask = shift_val * ((1u << l) - 1);
        gd.shift = shift_num;
        gd.bits = l;
        gd.max = n;
        m.groups.push_back(gd);
        for (int i = 0; i < n; i++) {
            NodeTypeDef t;
            t.name = g.second[i];
            t.group = g.first;
            t.value = shift_val * (uint32_t)(i + 1);
            t.mask = gd.mask;
            m.nodetypes.push_back(t);
        }
        shift_val <<= l;
        shift_num += l;
    }
    m.flag_bits = 16;
    m.zone_shift = shift_num;
    m.zone_max = (1 << (16 - shift_num)) - 1;
    GroupDef zg;
    zg.name = "SETTINGZONE";
    zg.mask = (uint32_t)m.zone_max << shift_num;
    zg.shift = shift_num;
    zg.bits = 16 - shift_num;
    zg.max = m.zone_max;
    m.groups.push_back(zg);
    NodeTypeDef dz;
    dz.name = "DefaultZone"; dz.group = "SETTINGZONE"; dz.value = 0; dz.mask = zg.mask;
    m.nodetypes.push_back(dz);
    GroupDef none; none.name = "NONE";
    m.groups.push_back(none);
    NodeTypeDef nn; nn.name = "None"; nn.group = "NONE";
    m.nodetypes.push_back(nn);
    GroupDef all; all.name = "ALL"; all.mask = 0xFFFF; all.bits = 16; all.max = 0xFFFF;
    m.groups.push_back(all);
    NodeTypeDef cl; cl.name = "Clear"; cl.group = "ALL"; cl.value = 0; cl.mask = 0xFFFF;
    m.nodetypes.push_back(cl);
}

// Globals of src/conf.R:740-761: SUM globals, Objective, MAX globals; one
// zonal <G>InObj weight per SUM global other than Objective.
void finish_globals(ModelDef& m, const std::vector<GlobalDef>& declared) {
    for (auto& g : declared) if (!g.is_max) m.globals.push_back(g);
    m.objective = (int)m.globals.size();
    m.globals.push_back({"Objective", false});
    m.sum_globals = (int)m.globals.size();
    for (auto& g : declared) if (g.is_max) m.globals.push_back(g);
    for (auto& g : declared)
        if (!g.is_max) m.zonal.push_back({g.name + "InObj", 0.0, -1, 0});
}

ModelDef make_d2q9(bool drag) {
    ModelDef m;
    m.id = drag ? MODEL_D2Q9_SRT_DRAG : MODEL_D2Q9_SRT;
    m.name = drag ? "d2q9_SRT_drag" : "d2q9_SRT";
    m.Q = 9; m.dim = 2;
    // models/flow/d2q9_SRT/Dynamics.R:6-14
    m.ex = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    m.ey = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    m.ez = std::vector<int>(9, 0);
    m.w = {4. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 36, 1. / 36, 1. / 36, 1. / 36};
    // Dynamics.R:22-29; Threshold from src/conf.R:769
    m.settings = {{"omega", 0.0, -1, 0}, {"nu", 0.16666666, 0, 1}, {"Density", 1.0, -1, 0},
                  {"Threshold", 0.5, -1, 0}};
    m.zonal = {{"Velocity", 0, -1, 0}, {"Velocity_x", 0, -1, 0}, {"Velocity_y", 0, -1, 0},
               {"GravitationX", 0, -1, 0}, {"GravitationY", 0, -1, 0}};
    std::vector<GlobalDef> g;
    if (drag) g = {{"Drag", false}, {"Lift", false}};
    finish_globals(m, g);
    m.quantities = {{"U", true, "m/s"}, {"Rho", false, "kg/m3"}};  // Dynamics.R:18-19
    if (drag) m.globals[0].unit = m.globals[1].unit = "N";
    pack_nodetypes(m, {{"EPressure", "BOUNDARY"}, {"WPressure", "BOUNDARY"}, {"WVelocity", "BOUNDARY"},
                       {"EVelocity", "BOUNDARY"}, {"Solid", "BOUNDARY"}, {"Wall", "BOUNDARY"},
                       {"BGK", "COLLISION"}});
    return m;
}

// d3q19, and with `lid` the model d3q19_lid: d3q19 plus a MovingWall node
// type (bounce-back with Ladd's wall-velocity correction), the wall velocity
// settings MovingWallVelocityX/Y/Z and the Globals MovingWallFx/Fy/Fz (force
// of the fluid on the MovingWall nodes, momentum exchange) and
// MovingWallTorque (z torque of that force about the axis through the
// lattice centre).  The reference has no moving wall; d3q19 itself keeps the
// reference's settings, Globals and NodeType bits unchanged.
ModelDef make_d3q19(bool lid) {
    ModelDef m;
    m.id = lid ? MODEL_D3Q19_LID : MODEL_D3Q19;
    m.name = lid ? "d3q19_lid" : "d3q19";
    m.Q = 19; m.dim = 3;
    // columns 4/6/8 of MRTMAT, models/flow/d3q19/Dynamics.R:1-35
    m.ex = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    m.ey = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    m.ez = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    m.w.assign(19, 1. / 36);
    m.w[0] = 1. / 3;
    for (int i = 1; i <= 6; i++) m.w[i] = 1. / 18;
    // Dynamics.R:40-48, then the free MRT rates of this engine (0 = the
    // reference's fixed rate of that moment group), then Threshold
    m.settings = {{"omega", 0.0, -1, 0},
                  {"nu", 0.16666666, 0, 1, "1m2/s"},
                  {"InletVelocity", 0.0, -1, 0, "1m/s"},
                  {"InletPressure", 0.0, 4, 2, "1Pa"},
                  {"InletDensity", 1.0, -1, 0, "1kg/m3"},
                  {"ForceX", 0.0, -1, 0},
                  {"ForceY", 0.0, -1, 0},
                  {"ForceZ", 0.0, -1, 0}};
    for (const char* r : {"MRT_Se", "MRT_Seps", "MRT_Sq", "MRT_Spi", "MRT_Sm"})
        m.settings.push_back({r, 0.0, -1, 0, "1", !lid});
    if (lid)
        for (const char* c : {"X", "Y", "Z"})
            m.settings.push_back({std::string("MovingWallVelocity") + c, 0.0, -1, 0, "1m/s"});
    m.settings.push_back({"Threshold", 0.5, -1, 0});
    // Dynamics.R:50-81 (units as declared there)
    std::vector<GlobalDef> g = {{"Flux", false, "m3/s"}};
    for (const char* p : {"XY", "XZ", "YZ"}) {
        for (const char* s : {"vx", "vy", "vz"}) g.push_back({std::string(p) + s, false, "m3/s"});
        g.push_back({std::string(p) + "rho", false, "kg/m"});
        g.push_back({std::string(p) + "area", false, "m2"});
    }
    for (const char* s : {"vx", "vy", "vz"}) g.push_back({std::string("VOL") + s, false, "m4/s"});
    for (const char* s : {"px", "py", "pz"}) g.push_back({std::string("VOL") + s, false, "mkg/s"});
    g.push_back({"VOLrho", false, "kg"});
    g.push_back({"VOLvolume", false, "m3"});
    if (lid) {
        for (const char* s : {"Fx", "Fy", "Fz"}) g.push_back({std::string("MovingWall") + s, false, "N"});
        g.push_back({"MovingWallTorque", false, "Nm"});
    }
    g.push_back({"MaxV", true, "m3"});
    finish_globals(m, g);
    m.quantities = {{"P", false, "Pa"}, {"U", true, "m/s"}};  // Dynamics.R:37-38
    std::vector<std::pair<std::string, std::string>> types = {
        {"XYslice", "ADDITIONALS"}, {"XZslice", "ADDITIONALS"}, {"YZslice", "ADDITIONALS"},
        {"EPressure", "BOUNDARY"},  {"Solid", "BOUNDARY"},      {"Wall", "BOUNDARY"},
        {"WPressure", "BOUNDARY"},  {"WPressureL", "BOUNDARY"}, {"WVelocity", "BOUNDARY"}};
    // the seventh BOUNDARY type still fits the group's 3 bits (value 0x1C)
    if (lid) types.push_back({"MovingWall", "BOUNDARY"});
    types.push_back({"MRT", "COLLISION"});
    pack_nodetypes(m, types);
    return m;
}

// d2q9: MRT in the weighted-orthogonal (Hermite) moment basis the reference
// generates from src/lib/feq.R (models/flow/d2q9/Dynamics.c.Rt:1-40), stored
// shifted by w_i, Zou-He conditions on all four sides, symmetry planes,
// pressure-loss and flux Globals on Inlet/Outlet nodes
// (models/flow/d2q9/Dynamics.R).  With `drag` the model d2q9_drag: the
// same plus Drag/Lift Globals by momentum exchange on Solid nodes, as
// d2q9_SRT_drag adds them to d2q9_SRT.
// autosym = 1, 2: the reference's d2q9_autosym / d2q9_autosym2 builds
// (models/flow/d2q9/conf.mk OPT="bc*autosym", src/conf.R:373-445): NodeType
// groups SYMX and SYMY with <name>X_plus/_minus, <name>Y_plus/_minus, name
// Symmetry (plane through the node) or SymmetryEdge (plane half a node out),
// acting in the streaming step.  bc: the bc builds (d2q9_bc,
// d2q9_bc_autosym, d2q9_bc_autosym2): parameter densities BC[0], BC[1]
// (Dynamics.R:23-26) that the velocity conditions read, Wall/Solid nodes
// fill with the momentum they reverse, and that force the fluid nodes
// (Dynamics.c.Rt:28-34,49-53,80-87,146-159,254-261).
ModelDef make_d2q9_mrt(bool drag, int autosym = 0, bool bc = false) {
    ModelDef m;
    static const ModelId plain[3] = {MODEL_D2Q9, MODEL_D2Q9_AUTOSYM, MODEL_D2Q9_AUTOSYM2};
    static const ModelId withbc[3] = {MODEL_D2Q9_BC, MODEL_D2Q9_BC_AUTOSYM, MODEL_D2Q9_BC_AUTOSYM2};
    m.id = drag ? MODEL_D2Q9_DRAG : (bc ? withbc[autosym] : plain[autosym]);
    // option builds are named model_<options in conf.mk's order, bc*autosym>
    m.name = drag ? "d2q9_drag"
                  : std::string("d2q9") + (bc ? "_bc" : "") +
                        (autosym == 1 ? "_autosym" : autosym == 2 ? "_autosym2" : "");
    m.Q = 9; m.dim = 2;
    m.ex = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    m.ey = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    m.ez = std::vector<int>(9, 0);
    m.w = {4. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 36, 1. / 36, 1. / 36, 1. / 36};
    // Dynamics.R:38-60: RelaxationRate sets S2 = 1 - RelaxationRate,
    // Viscosity sets RelaxationRate = 1/(3 Viscosity + 0.5)
    m.settings = {{"RelaxationRate", 0.0, 4, 3}, {"Viscosity", 0.16666666, 0, 1},
                  {"GravitationX", 0.0, -1, 0},  {"GravitationY", 0.0, -1, 0},
                  {"S2", 0.0, -1, 0},            {"S3", 0.0, -1, 0},
                  {"S4", 0.0, -1, 0},            {"Threshold", 0.5, -1, 0}};
    m.zonal = {{"VelocityX", 0, -1, 0}, {"VelocityY", 0, -1, 0}, {"Pressure", 0, -1, 0}};
    std::vector<GlobalDef> g = {{"PressureLoss", false, "1mPa"}, {"OutletFlux", false, "1m2/s"},
                                {"InletFlux", false, "1m2/s"}};
    if (drag) {
        g.push_back({"Drag", false, "N"});
        g.push_back({"Lift", false, "N"});
    }
    finish_globals(m, g);
    m.quantities = {{"Rho", false, "kg/m3"}, {"U", true, "m/s"}};
    if (bc) m.params = {"BC[0]", "BC[1]"};
    std::vector<std::pair<std::string, std::string>> types = {
        {"EPressure", "BOUNDARY"}, {"WPressure", "BOUNDARY"}, {"NVelocity", "BOUNDARY"},
        {"SVelocity", "BOUNDARY"}, {"WVelocity", "BOUNDARY"}, {"EVelocity", "BOUNDARY"},
        {"NSymmetry", "BOUNDARY"}, {"SSymmetry", "BOUNDARY"}, {"Inlet", "OBJECTIVE"},
        {"Outlet", "OBJECTIVE"}, {"Solid", "BOUNDARY"}, {"Wall", "BOUNDARY"},
        {"MRT", "COLLISION"}};
    if (autosym) {
        const std::string s = autosym == 1 ? "Symmetry" : "SymmetryEdge";
        for (const char* t : {"X_plus", "X_minus"}) types.push_back({s + t, "SYMX"});
        for (const char* t : {"Y_plus", "Y_minus"}) types.push_back({s + t, "SYMY"});
    }
    pack_nodetypes(m, types);
    return m;
}

// d3q27 (models/flow/d3q27/Dynamics.R): all 27 velocities of {-1,0,1}^3,
// density i at ((i%3)-1, (i/3)%3-1, i/9-1) as expand.grid orders them; MRT
// in raw moments relaxing every order >= 2 (BGK unless Stab or Smagorinsky
// nodes change the rates), Zou-He on W/E, symmetry planes, drag by momentum
// exchange on Wall nodes, flux Globals.
ModelDef make_d3q27() {
    ModelDef m;
    m.id = MODEL_D3Q27;
    m.name = "d3q27";
    m.Q = 27; m.dim = 3;
    for (int i = 0; i < 27; i++) {
        m.ex.push_back(i % 3 - 1);
        m.ey.push_back(i / 3 % 3 - 1);
        m.ez.push_back(i / 9 - 1);
        const int nz = (m.ex[i] != 0) + (m.ey[i] != 0) + (m.ez[i] != 0);
        m.w.push_back(nz == 0 ? 8. / 27 : nz == 1 ? 2. / 27 : nz == 2 ? 1. / 54 : 1. / 216);
    }
    // Dynamics.R:16-26 (no units declared), Threshold from src/conf.R:769
    m.settings = {{"omega", 0.0, -1, 0}, {"nu", 0.16666666, 0, 1}, {"Smag", 0.0, -1, 0},
                  {"ForceX", 0.0, -1, 0}, {"ForceY", 0.0, -1, 0}, {"ForceZ", 0.0, -1, 0},
                  {"Threshold", 0.5, -1, 0}};
    m.zonal = {{"Velocity", 0, -1, 0}, {"Pressure", 0, -1, 0}, {"Turbulence", 0, -1, 0}};
    std::vector<GlobalDef> g;
    for (const char* a : {"X", "Y", "Z"}) g.push_back({std::string(a) + "Flux", false, "m3/s"});
    for (const char* a : {"X", "Y", "Z"}) g.push_back({std::string(a) + "DragForce", false, "N"});
    finish_globals(m, g);
    m.quantities = {{"P", false, "Pa"}, {"U", true, "m/s"}, {"Fd", true, "N"}};
    pack_nodetypes(m, {{"Smagorinsky", "LES"}, {"Stab", "ENTROPIC"}, {"NSymmetry", "BOUNDARY"},
                       {"SSymmetry", "BOUNDARY"}, {"ISymmetry", "BOUNDARY"}, {"EPressure", "BOUNDARY"},
                       {"Solid", "BOUNDARY"}, {"Wall", "BOUNDARY"}, {"WPressure", "BOUNDARY"},
                       {"WVelocity", "BOUNDARY"}, {"MRT", "COLLISION"}});
    return m;
}

// d3q27_cumulant (models/flow/d3q27_cumulant/Dynamics.R, no build options):
// density i = px + 3 py + 9 pz at e = ({0,1,-1}[px], {0,1,-1}[py],
// {0,1,-1}[pz]) (named f<px><py><pz> there); the cumulant collision with a
// Galilean correction, Zou-He on the W/E and S/N faces, symmetry planes,
// momentum-exchange drag on Wall nodes marked Body, buffer-layer viscosity.
// smag: the SMAG build (d3q27_cumulant_SMAG): a Smag setting and the
// Smagorinsky viscosity of Dynamics.c.Rt:362-374.  ib: the IB build
// (d3q27_cumulant_IB): NodeType IB, interpolated bounce-back on the cut
// distances of the lattice (Dynamics.c.Rt:180-222).  avg: the AVG build
// (d3q27_cumulant_AVG): 13 averaged densities and the quantities built on
// them (Dynamics.R:66-90, Dynamics.c.Rt:62-99,410-414,507-518).  Option
// builds are named <model>_<options in the order of conf.mk's OPT>
// (src/models.R:54-66).
ModelDef make_d3q27_cumulant(bool smag = false, bool ib = false, bool avg = false) {
    ModelDef m;
    static const ModelId ids[8] = {MODEL_D3Q27_CUMULANT,     MODEL_D3Q27_CUMULANT_SMAG,
                                   MODEL_D3Q27_CUMULANT_IB,  MODEL_D3Q27_CUMULANT_IB_SMAG,
                                   MODEL_D3Q27_CUMULANT_AVG, MODEL_D3Q27_CUMULANT_AVG_SMAG,
                                   MODEL_D3Q27_CUMULANT_AVG_IB, MODEL_D3Q27_CUMULANT_AVG_IB_SMAG};
    m.id = ids[(avg ? 4 : 0) + (ib ? 2 : 0) + (smag ? 1 : 0)];
    m.name = std::string("d3q27_cumulant") + (avg ? "_AVG" : "") + (ib ? "_IB" : "") + (smag ? "_SMAG" : "");
    m.Q = 27; m.dim = 3;
    const int e1[3] = {0, 1, -1};
    for (int i = 0; i < 27; i++) {
        m.ex.push_back(e1[i % 3]);
        m.ey.push_back(e1[i / 3 % 3]);
        m.ez.push_back(e1[i / 9]);
        const int nz = (i % 3 != 0) + (i / 3 % 3 != 0) + (i / 9 != 0);
        m.w.push_back(nz == 0 ? 8. / 27 : nz == 1 ? 2. / 27 : nz == 2 ? 1. / 54 : 1. / 216);
    }
    // Dynamics.R:18-28 (nu has no derived setting here), Threshold from src/conf.R:769
    m.settings = {{"nu", 0.16666666, -1, 0},         {"nubuffer", 0.01, -1, 0}, {"GalileanCorrection", 1.0, -1, 0},
                  {"ForceX", 0.0, -1, 0},            {"ForceY", 0.0, -1, 0},    {"ForceZ", 0.0, -1, 0},
                  {"Omega", 1.0, -1, 0}};
    if (smag) m.settings.push_back({"Smag", 0.0, -1, 0});
    m.settings.push_back({"Threshold", 0.5, -1, 0});
    m.zonal = {{"Velocity", 0, -1, 0}, {"Pressure", 0, -1, 0}, {"Turbulence", 0, -1, 0}};
    std::vector<GlobalDef> g = {{"Density", false, "kg/m3"}, {"Flux", false, "m3/s"},     {"Drag", false, "N"},
                                {"Lift", false, "N"},        {"Lateral", false, "N"},     {"Mass", false, "kg"},
                                {"XMomentum", false, "kgm/s"}, {"YMomentum", false, "kgm/s"},
                                {"ZMomentum", false, "kgm/s"}};
    finish_globals(m, g);
    m.quantities = {{"P", false, "Pa"}, {"U", true, "m/s"}, {"Solid", false, "1"}};
    if (avg) {
        m.quantities.insert(m.quantities.end(), {{"KinE", false, "1"}, {"ReStr", true, "1"}, {"Dissipation", false, "1"},
                                                 {"avgU", true, "m/s"}, {"varU", true, "1"}, {"averageP", false, "Pa"}});
        m.averages = {"avgP",    "varUX",   "varUY",   "varUZ",   "varUXUY", "varUXUZ", "varUYUZ",
                      "avgdxu2", "avgdyv2", "avgdzw2", "avgUX",   "avgUY",   "avgUZ"};
    }
    std::vector<std::pair<std::string, std::string>> types = {
        {"Buffer", "BOUNDARY"},       {"WVelocityTurbulent", "BOUNDARY"}, {"NVelocity", "BOUNDARY"},
        {"SVelocity", "BOUNDARY"},    {"NPressure", "BOUNDARY"},          {"SPressure", "BOUNDARY"},
        {"NSymmetry", "ADDITIONALS"}, {"SSymmetry", "ADDITIONALS"},       {"Body", "BODY"}};
    // the IB build adds NodeType IB in its own group HO_BOUNDARY (Dynamics.R:61-63)
    if (ib) types.push_back({"IB", "HO_BOUNDARY"});
    for (const char* t : {"EPressure", "EVelocity", "Solid", "Wall", "WPressure", "WVelocity"})
        types.push_back({t, "BOUNDARY"});
    types.push_back({"MRT", "COLLISION"});
    pack_nodetypes(m, types);
    return m;
}

const std::vector<ModelDef>& registry() {
    static std::vector<ModelDef> r = {make_d2q9(false), make_d2q9(true), make_d3q19(false), make_d3q19(true),
                                      make_d2q9_mrt(false), make_d2q9_mrt(true), make_d3q27(),
                                      make_d3q27_cumulant(), make_d2q9_mrt(false, 1), make_d2q9_mrt(false, 2),
                                      make_d3q27_cumulant(true), make_d3q27_cumulant(false, true),
                                      make_d3q27_cumulant(true, true), make_d3q27_cumulant(false, false, true),
                                      make_d3q27_cumulant(false, true, true), make_d3q27_cumulant(true, false, true),
                                      make_d3q27_cumulant(true, true, true), make_d2q9_mrt(false, 0, true),
                                      make_d2q9_mrt(false, 1, true), make_d2q9_mrt(false, 2, true)};
    return r;
}

}  // namespace

double derived_value(const SettingDef& s, double v) {
    switch (s.derive_kind) {
    case 1: return 1.0 / (3 * v + 0.5);  // omega / RelaxationRate = 1/(3 nu + 0.5)
    case 2: return 1.0 + v * 3;          // InletDensity = 1 + 3 InletPressure
    case 3: return 1 - v;                // S2 = 1 - RelaxationRate
    default: return v;
    }
}

const ModelDef* model_by_name(const std::string& name) {
    for (auto& m : registry()) if (m.name == name) return &m;
    return nullptr;
}

const ModelDef* model_by_id(int id) {
    if (id < 0 || id >= (int)registry().size()) return nullptr;
    return &registry()[id];
}

int storage_bytes(int s) {
    switch (s) {
    case ST_F64: return 8;
    case ST_F32: case ST_F32S: return 4;
    case ST_F16S: case ST_BF16S: return 2;
    default: return 0;
    }
}

const char* storage_name(int s) {
    static const char* n[] = {"f64", "f32", "f32s", "f16s", "bf16s"};
    return (s >= 0 && s < ST_COUNT) ? n[s] : "?";
}

int storage_by_name(const std::string& name) {
    for (int s = 0; s < ST_COUNT; s++) if (name == storage_name(s)) return s;
    return -1;
}

}  // namespace tclbx
