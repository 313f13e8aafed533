// kernels_2d.hip -- instantiates and dispatches the node kernels of one
// family of models (see kernels_impl.h); compiled apart so the families
// build in parallel.
#include "kernels_impl.h"

namespace tclbx {

hipError_t TCLBX_FAM(step_2d)(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s) {
    switch (model) {
    case MODEL_D2Q9_SRT: return step_m<MODEL_D2Q9_SRT>(storage, globals, a, block, s);
    case MODEL_D2Q9_SRT_DRAG: return step_m<MODEL_D2Q9_SRT_DRAG>(storage, globals, a, block, s);
    case MODEL_D2Q9: return step_m<MODEL_D2Q9>(storage, globals, a, block, s);
    case MODEL_D2Q9_DRAG: return step_m<MODEL_D2Q9_DRAG>(storage, globals, a, block, s);
    case MODEL_D2Q9_AUTOSYM: return step_m<MODEL_D2Q9_AUTOSYM>(storage, globals, a, block, s);
    case MODEL_D2Q9_AUTOSYM2: return step_m<MODEL_D2Q9_AUTOSYM2>(storage, globals, a, block, s);
    case MODEL_D2Q9_BC: return step_m<MODEL_D2Q9_BC>(storage, globals, a, block, s);
    case MODEL_D2Q9_BC_AUTOSYM: return step_m<MODEL_D2Q9_BC_AUTOSYM>(storage, globals, a, block, s);
    case MODEL_D2Q9_BC_AUTOSYM2: return step_m<MODEL_D2Q9_BC_AUTOSYM2>(storage, globals, a, block, s);
    }
    return hipErrorNotFound;
}

hipError_t TCLBX_FAM(init_2d)(int model, int storage, const StepArgs& a, int block, hipStream_t s) {
    switch (model) {
    case MODEL_D2Q9_SRT: return init_m<MODEL_D2Q9_SRT>(storage, a, block, s);
    case MODEL_D2Q9_SRT_DRAG: return init_m<MODEL_D2Q9_SRT_DRAG>(storage, a, block, s);
    case MODEL_D2Q9: return init_m<MODEL_D2Q9>(storage, a, block, s);
    case MODEL_D2Q9_DRAG: return init_m<MODEL_D2Q9_DRAG>(storage, a, block, s);
    case MODEL_D2Q9_AUTOSYM: return init_m<MODEL_D2Q9_AUTOSYM>(storage, a, block, s);
    case MODEL_D2Q9_AUTOSYM2: return init_m<MODEL_D2Q9_AUTOSYM2>(storage, a, block, s);
    case MODEL_D2Q9_BC: return init_m<MODEL_D2Q9_BC>(storage, a, block, s);
    case MODEL_D2Q9_BC_AUTOSYM: return init_m<MODEL_D2Q9_BC_AUTOSYM>(storage, a, block, s);
    case MODEL_D2Q9_BC_AUTOSYM2: return init_m<MODEL_D2Q9_BC_AUTOSYM2>(storage, a, block, s);
    }
    return hipErrorNotFound;
}

hipError_t TCLBX_FAM(batch_2d)(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s) {
    switch (model) {
    case MODEL_D2Q9_SRT: return batch_m<MODEL_D2Q9_SRT>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D2Q9_SRT_DRAG: return batch_m<MODEL_D2Q9_SRT_DRAG>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D2Q9: return batch_m<MODEL_D2Q9>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D2Q9_DRAG: return batch_m<MODEL_D2Q9_DRAG>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D2Q9_AUTOSYM: return batch_m<MODEL_D2Q9_AUTOSYM>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D2Q9_AUTOSYM2: return batch_m<MODEL_D2Q9_AUTOSYM2>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D2Q9_BC: return batch_m<MODEL_D2Q9_BC>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D2Q9_BC_AUTOSYM: return batch_m<MODEL_D2Q9_BC_AUTOSYM>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case MODEL_D2Q9_BC_AUTOSYM2: return batch_m<MODEL_D2Q9_BC_AUTOSYM2>(storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
    }
    return hipErrorNotFound;
}

hipError_t TCLBX_FAM(quantity_2d)(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s) {
    switch (model) {
    case MODEL_D2Q9_SRT: return quantity_m<MODEL_D2Q9_SRT>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D2Q9_SRT_DRAG: return quantity_m<MODEL_D2Q9_SRT_DRAG>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D2Q9: return quantity_m<MODEL_D2Q9>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D2Q9_DRAG: return quantity_m<MODEL_D2Q9_DRAG>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D2Q9_AUTOSYM: return quantity_m<MODEL_D2Q9_AUTOSYM>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D2Q9_AUTOSYM2: return quantity_m<MODEL_D2Q9_AUTOSYM2>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D2Q9_BC: return quantity_m<MODEL_D2Q9_BC>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D2Q9_BC_AUTOSYM: return quantity_m<MODEL_D2Q9_BC_AUTOSYM>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case MODEL_D2Q9_BC_AUTOSYM2: return quantity_m<MODEL_D2Q9_BC_AUTOSYM2>(storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    }
    return hipErrorNotFound;
}

}  // namespace tclbx
