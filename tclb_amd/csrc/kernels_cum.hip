// kernels_cum.hip -- instantiates and dispatches the node kernels of one
// family of models (see kernels_impl.h); compiled apart so the families
// build in parallel.
#include "kernels_impl.h"

namespace tclbx {

hipError_t TCLBX_FAM(step_cum)(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s) {
    switch (model) {
    case MODEL_D3Q27_CUMULANT: return step_m<MODEL_D3Q27_CUMULANT>(storage, globals, a, block, s);
    case MODEL_D3Q27_CUMULANT_SMAG: return step_m<MODEL_D3Q27_CUMULANT_SMAG>(storage, globals, a, block, s);
    case MODEL_D3Q27_CUMULANT_IB: return step_m<MODEL_D3Q27_CUMULANT_IB>(storage, globals, a, block, s);
    case MODEL_D3Q27_CUMULANT_IB_SMAG: return step_m<MODEL_D3Q27_CUMULANT_IB_SMAG>(storage, globals, a, block, s);
    }
    return hipErrorNotFound;
}

hipError_t TCLBX_FAM(init_cum)(int model, int storage, const StepArgs& a, int block, hipStream_t s) {
    switch (model) {
    case MODEL_D3Q27_CUMULANT: return init_m<MODEL_D3Q27_CUMULANT>(storage, a, block, s);
    case MODEL_D3Q27_CUMULANT_SMAG: return init_m<MODEL_D3Q27_CUMULANT_SMAG>(storage, a, block, s);
    case MODEL_D3Q27_CUMULANT_IB: return init_m<MODEL_D3Q27_CUMULANT_IB>(storage, a, block, s);
    case MODEL_D3Q27_CUMULANT_IB_SMAG: return init_m<MODEL_D3Q27_CUMULANT_IB_SMAG>(storage, a, block, s);
    }
    return hipErrorNotFound;
}

hipError_t TCLBX_FAM(batch_cum)(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s) {
    switch (model) {
    case MODEL_D3Q27_CUMULANT: return batch_m<MODEL_D3Q27_CUMULANT>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D3Q27_CUMULANT_SMAG: return batch_m<MODEL_D3Q27_CUMULANT_SMAG>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D3Q27_CUMULANT_IB: return batch_m<MODEL_D3Q27_CUMULANT_IB>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D3Q27_CUMULANT_IB_SMAG: return batch_m<MODEL_D3Q27_CUMULANT_IB_SMAG>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    }
    return hipErrorNotFound;
}

hipError_t TCLBX_FAM(quantity_cum)(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s) {
    switch (model) {
    case MODEL_D3Q27_CUMULANT: return quantity_m<MODEL_D3Q27_CUMULANT>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D3Q27_CUMULANT_SMAG: return quantity_m<MODEL_D3Q27_CUMULANT_SMAG>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D3Q27_CUMULANT_IB: return quantity_m<MODEL_D3Q27_CUMULANT_IB>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D3Q27_CUMULANT_IB_SMAG: return quantity_m<MODEL_D3Q27_CUMULANT_IB_SMAG>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    }
    return hipErrorNotFound;
}

}  // namespace tclbx
