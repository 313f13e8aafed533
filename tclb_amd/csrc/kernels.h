// kernels.h -- host-side interface of the node kernels (kernels_impl.h,
// launched through kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace tclbx {

// Model parameters passed by value in every launch (no per-step constant
// upload).  d2q9: [omega, Density].  d3q19: [omega, InletVelocity,
// InletDensity, ForceX, ForceY, ForceZ, then the relaxation factors of the
// e, eps, q, pi and m moment groups, then gamma1 = 1 - omega (nu group),
// then (d3q19_lid) MovingWallVelocityX/Y/Z].  d2q9 (MRT): [S2, S3, S4,
// GravitationX, GravitationY].  d3q27: [omega, Smag, ForceX, ForceY, ForceZ].
// d3q27_cumulant and its builds: [nu, nubuffer, GalileanCorrection, ForceX,
// ForceY, ForceZ, Omega, Smag].
constexpr int MAX_PAR = 16;
// side of the tiles of the temporally blocked step (nodes, one per thread);
// it takes at most TB_TILE / 2 - 1 steps per launch
constexpr int TB_TILE = 32;
constexpr int MAX_GLOBALS = 32;

struct StepArgs {
    const void* fin;       // current snapshot: Q planes of `pitch` elements
    void* fout;            // next snapshot
    const uint16_t* flags; // NodeType per node, x fastest
    const uint16_t* cuts;  // cut distances [26][node] (IB builds), or null: no cuts
    double* extra;         // densities that are not streamed, [k][node], FP64, updated in place
                           // (averages of the AVG builds, BC[0..1] of the bc builds), or null
    double avg_n;          // average_iter = iterations since the last average reset (quantities)
    const double* zone;    // zonal settings [zone][zcols]
    double* gpart;         // per-block partial Globals [nglob][nblocks][2] (sum, compensation)
    long long pitch;       // elements per density plane
    int nx, ny, nz;
    int bpr;               // blocks per (y,z) row
    int rpb;               // rows walked by one block of a Globals step
    int rpt;               // rows walked by one block of a plain step
    int flat;              // 1: tiles are runs of the linear node index, not row segments
    double inv_nx;         // 1/nx (flat tiles: node index -> row)
    double inv_ny;         // 1/ny (flat tiles: row -> y, z)
    int zcols, nzones;
    unsigned ib_mask;      // NodeType bit of IB (the cumulant IB builds), 0 for every other model
    // NodeType-sorted dispatch of a plain step: mode 0 updates every node of
    // the tiles; mode 1 (bulk) only the nodes whose class (NodeType bits
    // below the zone) is bulk_val, the plain collision node, so that every
    // wavefront takes the same branch; mode 2 the list[0..list_n) of the
    // other nodes, sorted by class, one thread each.  Mode 3: tb_k plain
    // steps in one launch of the temporally blocked step (2D models), tiles
    // of TB_TILE x TB_TILE nodes, tb_bx of them along x (set by the launch)
    int mode;
    unsigned class_mask, bulk_val;
    const unsigned* list;
    unsigned list_n;
    int tb_k, tb_bx;
    // stochastic rounding of the stores of a 16-bit format (0: nearest even);
    // sr_key = sr_key(iteration the step computes, the lattice's seed)
    int sr;
    unsigned sr_key;
    // a launch of several steps (the temporally blocked step) keys step s
    // of the launch itself: sr_key(sr_step + s, sr_seed)
    unsigned sr_step, sr_seed;
    double par[MAX_PAR];
    // float forms the float collisions take, derived from par on the host
    // (model_params)
    float fpar[MAX_PAR];
};

// One lattice of a batch: its step arguments (fin/fout = snapshots 0/1, rpb
// of its Globals tiles) and which snapshot is current when the batch phase
// is 0.
struct BatchCase {
    StepArgs args;
    int parity;
    // stochastic rounding (the SR instantiations of the batch step): the
    // case's iteration when the current sequence of launches began, and its
    // seed; launch k of the sequence computes iteration sr_iter + k, keyed
    // as a lone lattice keys it
    unsigned sr_iter, sr_seed;
    int pad_;
};

// Blocks (tiles) of one step launch of the lattice `a` with `block`
// threads per block (a Globals step walks a.rpb rows per block).
int step_blocks(const StepArgs& a, bool globals, int block);
// One time step: pull streaming + NodeType dispatch + collision (+ Globals).
hipError_t launch_step(int model, int storage, bool globals, const StepArgs& a, int block,
                       hipStream_t s);
// One step of a batch: a persistent grid of `grid` blocks walks `ntiles`
// tiles {case, tile of the case}; Globals partials go to slot = tile index.
hipError_t launch_batch_step(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles,
                             int ntiles, int phase, double* gpart, int grid, int block, hipStream_t s);
// Per-case Globals: case c sums partial slots [first[c], first[c]+count[c])
// into out[c*nglob ...].
hipError_t launch_reduce_globals_cases(const double* gpart, int nslots, int nglob, int sum_globals, int ncases,
                                       const int* first, const int* count, double* out, hipStream_t s);
// Init(): equilibrium from the settings, written to a.fout.
hipError_t launch_init(int model, int storage, const StepArgs& a, int block, hipStream_t s);
// Deterministic reduction of the per-block partials into out[nglob]
// (fixed order: independent of timing, identical run to run).
hipError_t launch_reduce_globals(const double* gpart, int nblocks, int nglob, int sum_globals,
                                 double* out, hipStream_t s);
// Quantity `q` of the model over the box region [x0,x0+rx)x[y0..)x[z0..),
// pulled populations of a.fin, times scale; vector quantities write 3
// doubles per node.
hipError_t launch_quantity(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0,
                           int rx, int ry, int rz, double scale, double* out, hipStream_t s);
// Density plane q of snapshot `buf` <-> double array over the whole lattice.
hipError_t launch_get_density(int model, int storage, const void* buf, long long pitch, long long n,
                              int q, double* out, hipStream_t s);
hipError_t launch_set_density(int model, int storage, void* buf, long long pitch, long long n,
                              int q, const double* in, hipStream_t s);

// plain copy of `bytes` (a multiple of 16) on the device
hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);

int kernel_nglobals(int model);

}  // namespace tclbx
