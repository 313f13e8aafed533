// kernels_d3q19.hip -- instantiates and dispatches the node kernels of one
// family of models (see kernels_impl.h); compiled apart so the families
// build in parallel.
#include "kernels_impl.h"

namespace tclbx {

hipError_t TCLBX_FAM(step_d3q19)(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s) {
    switch (model) {
    case MODEL_D3Q19: return step_m<MODEL_D3Q19>(storage, globals, a, block, s);
    case MODEL_D3Q19_LID: return step_m<MODEL_D3Q19_LID>(storage, globals, a, block, s);
    }
    return hipErrorNotFound;
}

hipError_t TCLBX_FAM(init_d3q19)(int model, int storage, const StepArgs& a, int block, hipStream_t s) {
    switch (model) {
    case MODEL_D3Q19: return init_m<MODEL_D3Q19>(storage, a, block, s);
    case MODEL_D3Q19_LID: return init_m<MODEL_D3Q19_LID>(storage, a, block, s);
    }
    return hipErrorNotFound;
}

hipError_t TCLBX_FAM(batch_d3q19)(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s) {
    switch (model) {
    case MODEL_D3Q19: return batch_m<MODEL_D3Q19>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D3Q19_LID: return batch_m<MODEL_D3Q19_LID>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    }
    return hipErrorNotFound;
}

hipError_t TCLBX_FAM(quantity_d3q19)(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s) {
    switch (model) {
    case MODEL_D3Q19: return quantity_m<MODEL_D3Q19>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D3Q19_LID: return quantity_m<MODEL_D3Q19_LID>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    }
    return hipErrorNotFound;
}

}  // namespace tclbx
