// model.h -- the model contract of the engine: Densities, Settings (with
// derived settings), zonal Settings, Globals, Quantities and NodeType flags
// for each model the node kernel implements.
//
// It mirrors what TCLB generates per model from Dynamics.R through
// src/conf.R (AddDensity/AddSetting/AddGlobal/AddQuantity/AddNodeType,
// src/conf.R:65-345) and exposes through Model_m (src/Lists.h:13-165):
//   * NodeType bits packed per group, groups in alphabetical order, each group
//     ceil(log2(n+1)) bits, values 1..n (src/conf.R:600-699);
//   * Globals ordered SUM globals, then Objective, then MAX globals, with one
//     zonal "<Global>InObj" weight per SUM global (src/conf.R:740-761);
//   * derived settings: setting nu also sets omega = 1/(3 nu + 0.5)
//     (models/flow/*/Dynamics.R, src/Lattice.cu.Rt:1350-1358).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace tclbx {

enum ModelId {
    MODEL_D2Q9_SRT = 0,
    MODEL_D2Q9_SRT_DRAG = 1,
    MODEL_D3Q19 = 2,
    MODEL_D3Q19_LID = 3,
    MODEL_D2Q9 = 4,
    MODEL_D2Q9_DRAG = 5,
    MODEL_D3Q27 = 6,
    MODEL_D3Q27_CUMULANT = 7,
    MODEL_D2Q9_AUTOSYM = 8,
    MODEL_D2Q9_AUTOSYM2 = 9,
    MODEL_D3Q27_CUMULANT_SMAG = 10,
    MODEL_D3Q27_CUMULANT_IB = 11,
    MODEL_D3Q27_CUMULANT_IB_SMAG = 12,
    MODEL_D3Q27_CUMULANT_AVG = 13,
    MODEL_D3Q27_CUMULANT_AVG_IB = 14,
    MODEL_D3Q27_CUMULANT_AVG_SMAG = 15,
    MODEL_D3Q27_CUMULANT_AVG_IB_SMAG = 16,
    MODEL_D2Q9_BC = 17,
    MODEL_D2Q9_BC_AUTOSYM = 18,
    MODEL_D2Q9_BC_AUTOSYM2 = 19,
    MODEL_COUNT = 20
};

// storage format of the distributions in HBM
enum Storage { ST_F64 = 0, ST_F32 = 1, ST_F32S = 2, ST_F16S = 3, ST_BF16S = 4, ST_COUNT = 5 };
// kernel instantiations of the 16-bit formats whose step stores round
// stochastically (a lattice option), kept apart so that the plain step
// carries none of that code; never a lattice's storage
enum { ST_F16S_SR = 5, ST_BF16S_SR = 6 };

struct SettingDef {
    std::string name;
    double def;            // default value
    int derived;           // index of the setting this one sets, or -1
    int derive_kind;       // 0 none, 1 1/(3x+0.5), 2 1+3x, 3 1-x
    std::string unit = "1";  // unit of the value (Log's _si column divides by it)
    bool extension = false;  // added by this engine to a reference model (not in its Log)
};

struct NodeTypeDef {
    std::string name, group;
    uint32_t value = 0, mask = 0;
};

struct GroupDef {
    std::string name;
    uint32_t mask = 0;
    int shift = 0, bits = 0, max = 0;
};

struct GlobalDef {
    std::string name;
    bool is_max;
    std::string unit = "1";
};

struct QuantityDef {
    std::string name;
    bool vector;
    std::string unit = "1";
};

struct ModelDef {
    ModelId id;
    std::string name;
    int Q, dim;
    std::vector<int> ex, ey, ez;      // direction of density i
    std::vector<double> w;            // lattice weight = storage shift of density i
    std::vector<SettingDef> settings; // global settings
    std::vector<SettingDef> zonal;    // zonal settings (incl. <Global>InObj)
    std::vector<GlobalDef> globals;   // SUM..., Objective, MAX...
    std::vector<QuantityDef> quantities;
    // densities that are not streamed, kept in FP64 planes beside the
    // distributions and updated in place (densities Q, Q+1, ...): first
    // those added with average=TRUE (zeroed by resetAverage), then the
    // parameter densities (parameter=TRUE, set from outside)
    std::vector<std::string> averages;
    std::vector<std::string> params;
    int nextra() const { return (int)(averages.size() + params.size()); }
    std::vector<NodeTypeDef> nodetypes;
    std::vector<GroupDef> groups;
    int zone_shift = 0, zone_max = 0, flag_bits = 16;
    int objective = -1, sum_globals = 0;

    int setting_id(const std::string& n) const;
    int zonal_id(const std::string& n) const;
    int global_id(const std::string& n) const;
    int quantity_id(const std::string& n) const;
    const NodeTypeDef* nodetype(const std::string& n) const;
    const GroupDef* group(const std::string& n) const;
};

// value a setting gives the setting it derives (Model::Setting::derivedValue,
// the equations of AddSetting in Dynamics.R)
double derived_value(const SettingDef& s, double v);

const ModelDef* model_by_name(const std::string& name);
const ModelDef* model_by_id(int id);
int storage_bytes(int storage);
const char* storage_name(int storage);
int storage_by_name(const std::string& name);

}  // namespace tclbx
