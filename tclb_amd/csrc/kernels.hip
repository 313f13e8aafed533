// kernels.hip -- host side of the node kernels (kernels_impl.h): the
// non-template kernels (density access, the Globals reduction, the copy),
// and the launch entry points of kernels.h, which hand each model to the
// compile unit of its family (kernels_<family>.hip).
#include "kernels_impl.h"

namespace tclbx {

// the compile unit of a family that holds the kernels of this storage format
#define TCLBX_PICK(name, storage) ((storage) == ST_F64 ? name##_f64 : name##_f32)

int kernel_nglobals(int model) {
    switch (model) {
    case MODEL_D2Q9_SRT: return Lat<MODEL_D2Q9_SRT>::G;
    case MODEL_D2Q9_SRT_DRAG: return Lat<MODEL_D2Q9_SRT_DRAG>::G;
    case MODEL_D3Q19: return Lat<MODEL_D3Q19>::G;
    case MODEL_D3Q19_LID: return Lat<MODEL_D3Q19_LID>::G;
    case MODEL_D2Q9: return Lat<MODEL_D2Q9>::G;
    case MODEL_D2Q9_DRAG: return Lat<MODEL_D2Q9_DRAG>::G;
    case MODEL_D3Q27: return Lat<MODEL_D3Q27>::G;
    case MODEL_D3Q27_CUMULANT: return Lat<MODEL_D3Q27_CUMULANT>::G;
    case MODEL_D3Q27_CUMULANT_SMAG: return Lat<MODEL_D3Q27_CUMULANT_SMAG>::G;
    case MODEL_D3Q27_CUMULANT_IB: return Lat<MODEL_D3Q27_CUMULANT_IB>::G;
    case MODEL_D3Q27_CUMULANT_IB_SMAG: return Lat<MODEL_D3Q27_CUMULANT_IB_SMAG>::G;
    case MODEL_D3Q27_CUMULANT_AVG: return Lat<MODEL_D3Q27_CUMULANT_AVG>::G;
    case MODEL_D3Q27_CUMULANT_AVG_IB: return Lat<MODEL_D3Q27_CUMULANT_AVG_IB>::G;
    case MODEL_D3Q27_CUMULANT_AVG_SMAG: return Lat<MODEL_D3Q27_CUMULANT_AVG_SMAG>::G;
    case MODEL_D3Q27_CUMULANT_AVG_IB_SMAG: return Lat<MODEL_D3Q27_CUMULANT_AVG_IB_SMAG>::G;
    case MODEL_D2Q9_AUTOSYM: return Lat<MODEL_D2Q9_AUTOSYM>::G;
    case MODEL_D2Q9_AUTOSYM2: return Lat<MODEL_D2Q9_AUTOSYM2>::G;
    case MODEL_D2Q9_BC: return Lat<MODEL_D2Q9_BC>::G;
    case MODEL_D2Q9_BC_AUTOSYM: return Lat<MODEL_D2Q9_BC_AUTOSYM>::G;
    case MODEL_D2Q9_BC_AUTOSYM2: return Lat<MODEL_D2Q9_BC_AUTOSYM2>::G;
    }
    return 0;
}

template <int ST>
__global__ void k_get_density(const void* buf, long long pitch, long long n, int q, double w, double* out) {
    using S = Store<ST>;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (double)S::get(((const typename S::T*)buf)[q * pitch + i], (typename S::R)w);
}

template <int ST>
__global__ void k_set_density(void* buf, long long pitch, long long n, int q, double w, const double* in) {
    using S = Store<ST>;
    using R = typename S::R;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ((typename S::T*)buf)[q * pitch + i] = S::put((R)in[i], (R)w);
}

// deterministic second pass: block (Global gi, lattice c) sums slots
// [first_c, first_c + count_c) of that Global's partials contiguously, TwoSum
// pairs throughout, and adds the result to out[c*nglob + gi] (cleared before
// a Globals step, accumulated over steps with ITER_GLOBS).  The order depends
// only on the lattice's own tiles, so a lattice gets the same Globals alone
// and in a batch.
__global__ void __launch_bounds__(256) k_reduce_globals(const double* gpart, int nslots, int nglob,
                                                        int sum_globals, const int* first, const int* count,
                                                        double* out) {
    const int gi = blockIdx.x;
    const int c = blockIdx.y;
    const bool is_sum = gi < sum_globals;
    const int n0 = first ? first[c] : 0;
    const int nb = count ? count[c] : nslots;
    const double* p = gpart + ((size_t)gi * nslots + n0) * 2;
    double v = 0.0, cc = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) {
        if (is_sum) dd_add(v, cc, p[2 * b], p[2 * b + 1]);
        else v = fmax(v, p[2 * b]);
    }
    __shared__ double lds[4][2];
    if (is_sum) wave_sum_dd(v, cc);
    else v = wave_max(v);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) { lds[wave][0] = v; lds[wave][1] = cc; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = lds[0][0], rc = lds[0][1];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
            if (is_sum) dd_add(r, rc, lds[w][0], lds[w][1]);
            else r = fmax(r, lds[w][0]);
        }
        double* o = out + (size_t)c * nglob + gi;
        *o = is_sum ? *o + (r + rc) : fmax(*o, r);
    }
}

// plain device copy, one 16-byte vector per thread (the fastest of the copy
// shapes measured on MI355X): the bandwidth ceiling the node kernel is
// compared with, and a kernel of known traffic for calibrating the HBM byte
// counters
__global__ void __launch_bounds__(256) k_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
    const size_t n = bytes / sizeof(uint4);
    const size_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, s, (const uint4*)src, (uint4*)dst, n);
    return hipGetLastError();
}

namespace {
double weight(int model, int q) {
    if (model == MODEL_D3Q27) return Lat<MODEL_D3Q27>::w(q);
    if ((model >= MODEL_D3Q27_CUMULANT_IB && model <= MODEL_D3Q27_CUMULANT_AVG_IB_SMAG) ||
        model == MODEL_D3Q27_CUMULANT || model == MODEL_D3Q27_CUMULANT_SMAG)
        return Lat<MODEL_D3Q27_CUMULANT>::w(q);
    return (model == MODEL_D3Q19 || model == MODEL_D3Q19_LID) ? Lat<MODEL_D3Q19>::w(q) : Lat<MODEL_D2Q9_SRT>::w(q);
}
}  // namespace

int step_blocks(const StepArgs& a, bool globals, int block) {
    const int rpb = globals ? a.rpb : a.rpt;
    if (a.flat) {
        const long long chunks = ((long long)a.nx * a.ny * a.nz + block - 1) / block;
        return (int)((chunks + rpb - 1) / rpb);
    }
    const int rows = a.ny * a.nz;
    return a.bpr * ((rows + rpb - 1) / rpb);
}

hipError_t launch_step(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s) {
    if (hipError_t e = TCLBX_PICK(step_2d, storage)(model, storage, globals, a, block, s); e != hipErrorNotFound) return e;
    if (hipError_t e = TCLBX_PICK(step_d3q19, storage)(model, storage, globals, a, block, s); e != hipErrorNotFound) return e;
    if (hipError_t e = TCLBX_PICK(step_d3q27, storage)(model, storage, globals, a, block, s); e != hipErrorNotFound) return e;
    if (hipError_t e = TCLBX_PICK(step_cum, storage)(model, storage, globals, a, block, s); e != hipErrorNotFound) return e;
    if (hipError_t e = TCLBX_PICK(step_cumavg, storage)(model, storage, globals, a, block, s); e != hipErrorNotFound) return e;
    return hipErrorInvalidValue;
}

hipError_t launch_init(int model, int storage, const StepArgs& a, int block, hipStream_t s) {
    if (hipError_t e = TCLBX_PICK(init_2d, storage)(model, storage, a, block, s); e != hipErrorNotFound) return e;
    if (hipError_t e = TCLBX_PICK(init_d3q19, storage)(model, storage, a, block, s); e != hipErrorNotFound) return e;
    if (hipError_t e = TCLBX_PICK(init_d3q27, storage)(model, storage, a, block, s); e != hipErrorNotFound) return e;
    if (hipError_t e = TCLBX_PICK(init_cum, storage)(model, storage, a, block, s); e != hipErrorNotFound) return e;
    if (hipError_t e = TCLBX_PICK(init_cumavg, storage)(model, storage, a, block, s); e != hipErrorNotFound) return e;
    return hipErrorInvalidValue;
}

hipError_t launch_reduce_globals(const double* gpart, int nblocks, int nglob, int sum_globals, double* out,
                                 hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_globals, dim3(nglob, 1), dim3(256), 0, s, gpart, nblocks, nglob, sum_globals,
                       (const int*)nullptr, (const int*)nullptr, out);
    return hipGetLastError();
}

hipError_t launch_reduce_globals_cases(const double* gpart, int nslots, int nglob, int sum_globals, int ncases,
                                       const int* first, const int* count, double* out, hipStream_t s) {
    if (ncases <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reduce_globals, dim3(nglob, ncases), dim3(256), 0, s, gpart, nslots, nglob, sum_globals,
                       first, count, out);
    return hipGetLastError();
}

hipError_t launch_batch_step(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles,
                             int ntiles, int phase, double* gpart, int grid, int block, hipStream_t s) {
    if (hipError_t e = TCLBX_PICK(batch_2d, storage)(model, storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
        e != hipErrorNotFound)
        return e;
    if (hipError_t e = TCLBX_PICK(batch_d3q19, storage)(model, storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
        e != hipErrorNotFound)
        return e;
    if (hipError_t e = TCLBX_PICK(batch_d3q27, storage)(model, storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
        e != hipErrorNotFound)
        return e;
    if (hipError_t e = TCLBX_PICK(batch_cum, storage)(model, storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
        e != hipErrorNotFound)
        return e;
    if (hipError_t e = TCLBX_PICK(batch_cumavg, storage)(model, storage, globals, cases, tiles, ntiles, phase, gpart, grid, block, s);
        e != hipErrorNotFound)
        return e;
    return hipErrorInvalidValue;
}

hipError_t launch_quantity(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx,
                           int ry, int rz, double scale, double* out, hipStream_t s) {
    if (hipError_t e = TCLBX_PICK(quantity_2d, storage)(model, storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
        e != hipErrorNotFound)
        return e;
    if (hipError_t e = TCLBX_PICK(quantity_d3q19, storage)(model, storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
        e != hipErrorNotFound)
        return e;
    if (hipError_t e = TCLBX_PICK(quantity_d3q27, storage)(model, storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
        e != hipErrorNotFound)
        return e;
    if (hipError_t e = TCLBX_PICK(quantity_cum, storage)(model, storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
        e != hipErrorNotFound)
        return e;
    if (hipError_t e = TCLBX_PICK(quantity_cumavg, storage)(model, storage, q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
        e != hipErrorNotFound)
        return e;
    return hipErrorInvalidValue;
}

hipError_t launch_get_density(int model, int storage, const void* buf, long long pitch, long long n, int q,
                              double* out, hipStream_t s) {
    const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
    const double w = weight(model, q);
    switch (storage) {
    case ST_F64: hipLaunchKernelGGL(k_get_density<ST_F64>, grid, blk, 0, s, buf, pitch, n, q, w, out); break;
    case ST_F32: hipLaunchKernelGGL(k_get_density<ST_F32>, grid, blk, 0, s, buf, pitch, n, q, w, out); break;
    case ST_F32S: hipLaunchKernelGGL(k_get_density<ST_F32S>, grid, blk, 0, s, buf, pitch, n, q, w, out); break;
    case ST_F16S: hipLaunchKernelGGL(k_get_density<ST_F16S>, grid, blk, 0, s, buf, pitch, n, q, w, out); break;
    case ST_BF16S: hipLaunchKernelGGL(k_get_density<ST_BF16S>, grid, blk, 0, s, buf, pitch, n, q, w, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_set_density(int model, int storage, void* buf, long long pitch, long long n, int q,
                              const double* in, hipStream_t s) {
    const dim3 grid((unsigned)((n + 255) / 256)), blk(256);
    const double w = weight(model, q);
    switch (storage) {
    case ST_F64: hipLaunchKernelGGL(k_set_density<ST_F64>, grid, blk, 0, s, buf, pitch, n, q, w, in); break;
    case ST_F32: hipLaunchKernelGGL(k_set_density<ST_F32>, grid, blk, 0, s, buf, pitch, n, q, w, in); break;
    case ST_F32S: hipLaunchKernelGGL(k_set_density<ST_F32S>, grid, blk, 0, s, buf, pitch, n, q, w, in); break;
    case ST_F16S: hipLaunchKernelGGL(k_set_density<ST_F16S>, grid, blk, 0, s, buf, pitch, n, q, w, in); break;
    case ST_BF16S: hipLaunchKernelGGL(k_set_density<ST_BF16S>, grid, blk, 0, s, buf, pitch, n, q, w, in); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace tclbx
