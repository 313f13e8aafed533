// kernels_impl.h -- the per-time-step node kernel for gfx950: the device
// code and the launch templates, included by kernels.hip and by the
// compile units kernels_<family>.hip that instantiate each family of
// models (so the families compile in parallel).
//
// One thread = one node.  A launch processes the whole lattice for one step:
// pull streaming from snapshot A (f_i(x) = A_i(x - e_i), periodic wrap on
// every axis, as the reference's self-aliased margins give on one rank:
// src/LatticeAccess.inc.cpp.Rt:319-335, src/Lattice.cu.Rt:445-453), the
// NodeType boundary switch, the collision, and a store to snapshot B at x
// (push, src/LatticeAccess.inc.cpp.Rt:337-347).  Node order inside Run()
// follows src/cuda.cu.Rt:322-327 (Load -> Run -> Save -> Glob).
//
// Layout in HBM: SoA, density plane q at base + q*pitch, node n = x + nx*(y +
// ny*z) inside a plane (the reference's block14[f*N + x + nx*(y+ny*z)],
// src/Lattice.cu.Rt:1138); pitch is N rounded up so every plane starts on a
// 256-byte boundary.  Blocks cover x-segments of one (y,z) row, so a
// wavefront's 64 stores fill whole 128-byte lines and its pulls are 64
// consecutive elements shifted by at most one.
//
// Globals: per-thread FP64 accumulators, reduced by wavefront shuffles, then
// across the block's wavefronts through LDS, one partial per block; a second
// launch sums the partials in a fixed order (deterministic, unlike the
// reference's per-warp atomicAdd, src/cross.hpp:202-218).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "kernels.h"
#include "model.h"
#include "storage.h"

// Which storage formats this compile unit instantiates: TCLBX_FP_GROUP 64 =
// FP64 only, 32 = the float-collision formats only, unset = all.
#ifndef TCLBX_FP_GROUP
#define TCLBX_FP_GROUP 0
#endif
#define TCLBX_HAS_F64 (TCLBX_FP_GROUP != 32)
#define TCLBX_HAS_FLOAT (TCLBX_FP_GROUP != 64)

namespace tclbx {

// ------------------------------------------------------------ lattices ---
template <int MODEL> struct Lat;
// Q densities, G Globals of which the first SUMG are sums (the rest MAX),
// zone shift of the NodeType flag
template <> struct Lat<MODEL_D2Q9_SRT> {
    static constexpr int Q = 9, G = 1, SUMG = 1, ZSHIFT = 4;
    __host__ __device__ static constexpr int ex(int i) { return i == 1 || i == 5 || i == 8 ? 1 : (i == 3 || i == 6 || i == 7 ? -1 : 0); }
    __host__ __device__ static constexpr int ey(int i) { return i == 2 || i == 5 || i == 6 ? 1 : (i == 4 || i == 7 || i == 8 ? -1 : 0); }
    __host__ __device__ static constexpr int ez(int) { return 0; }
    __host__ __device__ static constexpr double w(int i) { return i == 0 ? 4. / 9 : (i < 5 ? 1. / 9 : 1. / 36); }
};
template <> struct Lat<MODEL_D2Q9_SRT_DRAG> : Lat<MODEL_D2Q9_SRT> {
    static constexpr int G = 3, SUMG = 3;
};
// d2q9 MRT: same directions; NodeType groups BOUNDARY (4 bits), COLLISION,
// OBJECTIVE (2 bits), zone from bit 7
template <> struct Lat<MODEL_D2Q9> : Lat<MODEL_D2Q9_SRT> {
    static constexpr int G = 4, SUMG = 4, ZSHIFT = 7;
};
// d2q9 + Drag, Lift (Globals 3, 4; Objective 5)
template <> struct Lat<MODEL_D2Q9_DRAG> : Lat<MODEL_D2Q9> {
    static constexpr int G = 6, SUMG = 6;
};
// d2q9_autosym / d2q9_autosym2: d2q9 plus the groups SYMX (bits 7-8) and
// SYMY (bits 9-10) of the reference's autosym builds; zone from bit 11
template <> struct Lat<MODEL_D2Q9_AUTOSYM> : Lat<MODEL_D2Q9> {
    static constexpr int ZSHIFT = 11;
};
template <> struct Lat<MODEL_D2Q9_AUTOSYM2> : Lat<MODEL_D2Q9> {
    static constexpr int ZSHIFT = 11;
};
// the bc builds: the same NodeType bits, plus the parameter densities BC[0], BC[1]
template <> struct Lat<MODEL_D2Q9_BC> : Lat<MODEL_D2Q9> {};
template <> struct Lat<MODEL_D2Q9_BC_AUTOSYM> : Lat<MODEL_D2Q9_AUTOSYM> {};
template <> struct Lat<MODEL_D2Q9_BC_AUTOSYM2> : Lat<MODEL_D2Q9_AUTOSYM2> {};
template <int MODEL> constexpr bool is_d2q9_bc =
    MODEL == MODEL_D2Q9_BC || MODEL == MODEL_D2Q9_BC_AUTOSYM || MODEL == MODEL_D2Q9_BC_AUTOSYM2;
template <int MODEL> constexpr bool is_d2q9m = MODEL == MODEL_D2Q9 || MODEL == MODEL_D2Q9_DRAG ||
                                              MODEL == MODEL_D2Q9_AUTOSYM || MODEL == MODEL_D2Q9_AUTOSYM2 ||
                                              is_d2q9_bc<MODEL>;
// 1: autosym (planes through the node), 2: autosym2 (planes half a node out)
template <int MODEL> constexpr int autosym_kind =
    (MODEL == MODEL_D2Q9_AUTOSYM || MODEL == MODEL_D2Q9_BC_AUTOSYM) ? 1
    : (MODEL == MODEL_D2Q9_AUTOSYM2 || MODEL == MODEL_D2Q9_BC_AUTOSYM2) ? 2 : 0;
template <> struct Lat<MODEL_D3Q19> {
    static constexpr int Q = 19, G = 26, SUMG = 25, ZSHIFT = 6;
    __host__ __device__ static constexpr int ex(int i) {
        return (i == 1 || i == 7 || i == 9 || i == 11 || i == 13) ? 1
             : (i == 2 || i == 8 || i == 10 || i == 12 || i == 14) ? -1 : 0;
    }
    __host__ __device__ static constexpr int ey(int i) {
        return (i == 3 || i == 7 || i == 8 || i == 15 || i == 17) ? 1
             : (i == 4 || i == 9 || i == 10 || i == 16 || i == 18) ? -1 : 0;
    }
    __host__ __device__ static constexpr int ez(int i) {
        return (i == 5 || i == 11 || i == 12 || i == 15 || i == 16) ? 1
             : (i == 6 || i == 13 || i == 14 || i == 17 || i == 18) ? -1 : 0;
    }
    __host__ __device__ static constexpr double w(int i) { return i == 0 ? 1. / 3 : (i < 7 ? 1. / 18 : 1. / 36); }
};
// d3q19 + MovingWall: Globals MovingWallFx/Fy/Fz/Torque (24..27) before
// Objective (28) and MaxV (29)
template <> struct Lat<MODEL_D3Q19_LID> : Lat<MODEL_D3Q19> {
    static constexpr int G = 30, SUMG = 29;
};
template <int MODEL> constexpr bool is_d3q19 = MODEL == MODEL_D3Q19 || MODEL == MODEL_D3Q19_LID;
// d3q27: density i at e = (i%3 - 1, i/3%3 - 1, i/9 - 1) (expand.grid order,
// models/flow/d3q27/Dynamics.R:1-10), opposite 26 - i; NodeType groups
// BOUNDARY (4 bits), COLLISION, ENTROPIC, LES, zone from bit 7; Globals
// XFlux, YFlux, ZFlux, XDragForce, YDragForce, ZDragForce, Objective
template <> struct Lat<MODEL_D3Q27> {
    static constexpr int Q = 27, G = 7, SUMG = 7, ZSHIFT = 7;
    __host__ __device__ static constexpr int ex(int i) { return i % 3 - 1; }
    __host__ __device__ static constexpr int ey(int i) { return i / 3 % 3 - 1; }
    __host__ __device__ static constexpr int ez(int i) { return i / 9 - 1; }
    // direction with the component on axis ax (3: every axis) negated
    __host__ __device__ static constexpr int mirror(int i, int ax) {
        return ax == 3 ? 26 - i : i + (2 - 2 * (i / (ax == 0 ? 1 : ax == 1 ? 3 : 9) % 3)) * (ax == 0 ? 1 : ax == 1 ? 3 : 9);
    }
    __host__ __device__ static constexpr double w(int i) {
        return ((i % 3 != 1) + (i / 3 % 3 != 1) + (i / 9 != 1)) == 0 ? 8. / 27
             : ((i % 3 != 1) + (i / 3 % 3 != 1) + (i / 9 != 1)) == 1 ? 2. / 27
             : ((i % 3 != 1) + (i / 3 % 3 != 1) + (i / 9 != 1)) == 2 ? 1. / 54 : 1. / 216;
    }
};

// d3q27_cumulant: density i = px + 3py + 9pz at e = {0,1,-1}[p] per axis
// (models/flow/d3q27_cumulant/Dynamics.R:1-13); NodeType groups ADDITIONALS
// (2 bits), BODY, BOUNDARY (4 bits from bit 3), COLLISION (bit 7), zone
// from bit 8; Globals Density, Flux, Drag, Lift, Lateral, Mass, X/Y/Z
// Momentum, Objective
template <int MODEL> constexpr bool is_cum = MODEL == MODEL_D3Q27_CUMULANT || MODEL == MODEL_D3Q27_CUMULANT_SMAG ||
                                            MODEL == MODEL_D3Q27_CUMULANT_IB || MODEL == MODEL_D3Q27_CUMULANT_IB_SMAG ||
                                            MODEL == MODEL_D3Q27_CUMULANT_AVG || MODEL == MODEL_D3Q27_CUMULANT_AVG_IB ||
                                            MODEL == MODEL_D3Q27_CUMULANT_AVG_SMAG ||
                                            MODEL == MODEL_D3Q27_CUMULANT_AVG_IB_SMAG;
template <> struct Lat<MODEL_D3Q27_CUMULANT> {
    static constexpr int Q = 27, G = 10, SUMG = 10, ZSHIFT = 8;
    __host__ __device__ static constexpr int e1(int p) { return p == 0 ? 0 : (p == 1 ? 1 : -1); }
    __host__ __device__ static constexpr int ex(int i) { return e1(i % 3); }
    __host__ __device__ static constexpr int ey(int i) { return e1(i / 3 % 3); }
    __host__ __device__ static constexpr int ez(int i) { return e1(i / 9); }
    // digit p -> (3 - p) % 3 negates a component
    __host__ __device__ static constexpr int neg(int i, int st) {
        return i + ((3 - i / st % 3) % 3 - i / st % 3) * st;
    }
    __host__ __device__ static constexpr int mirror(int i, int ax) {
        return ax == 3 ? neg(neg(neg(i, 1), 3), 9) : neg(i, ax == 0 ? 1 : ax == 1 ? 3 : 9);
    }
    __host__ __device__ static constexpr double w(int i) {
        return ((i % 3 != 0) + (i / 3 % 3 != 0) + (i / 9 != 0)) == 0 ? 8. / 27
             : ((i % 3 != 0) + (i / 3 % 3 != 0) + (i / 9 != 0)) == 1 ? 2. / 27
             : ((i % 3 != 0) + (i / 3 % 3 != 0) + (i / 9 != 0)) == 2 ? 1. / 54 : 1. / 216;
    }
};
template <> struct Lat<MODEL_D3Q27_CUMULANT_SMAG> : Lat<MODEL_D3Q27_CUMULANT> {};
// the IB builds: NodeType IB is group HO_BOUNDARY at bit 8, zone from bit 9
template <> struct Lat<MODEL_D3Q27_CUMULANT_IB> : Lat<MODEL_D3Q27_CUMULANT> {
    static constexpr int ZSHIFT = 9;
};
template <> struct Lat<MODEL_D3Q27_CUMULANT_IB_SMAG> : Lat<MODEL_D3Q27_CUMULANT_IB> {};
// the AVG builds keep the NodeType bits of the build without AVG
template <> struct Lat<MODEL_D3Q27_CUMULANT_AVG> : Lat<MODEL_D3Q27_CUMULANT> {};
template <> struct Lat<MODEL_D3Q27_CUMULANT_AVG_SMAG> : Lat<MODEL_D3Q27_CUMULANT> {};
template <> struct Lat<MODEL_D3Q27_CUMULANT_AVG_IB> : Lat<MODEL_D3Q27_CUMULANT_IB> {};
template <> struct Lat<MODEL_D3Q27_CUMULANT_AVG_IB_SMAG> : Lat<MODEL_D3Q27_CUMULANT_IB> {};
template <int MODEL> constexpr bool is_cum_ib = MODEL == MODEL_D3Q27_CUMULANT_IB ||
                                               MODEL == MODEL_D3Q27_CUMULANT_IB_SMAG ||
                                               MODEL == MODEL_D3Q27_CUMULANT_AVG_IB ||
                                               MODEL == MODEL_D3Q27_CUMULANT_AVG_IB_SMAG;
template <int MODEL> constexpr bool is_cum_smag = MODEL == MODEL_D3Q27_CUMULANT_SMAG ||
                                                 MODEL == MODEL_D3Q27_CUMULANT_IB_SMAG ||
                                                 MODEL == MODEL_D3Q27_CUMULANT_AVG_SMAG ||
                                                 MODEL == MODEL_D3Q27_CUMULANT_AVG_IB_SMAG;
template <int MODEL> constexpr bool is_cum_avg = MODEL == MODEL_D3Q27_CUMULANT_AVG ||
                                                MODEL == MODEL_D3Q27_CUMULANT_AVG_IB ||
                                                MODEL == MODEL_D3Q27_CUMULANT_AVG_SMAG ||
                                                MODEL == MODEL_D3Q27_CUMULANT_AVG_IB_SMAG;


// A pointer into HBM as a global (address space 1) pointer.  The lattices
// of a batch come from a case table in memory, so the compiler sees generic
// pointers there and would use flat loads, which count against the scalar
// memory counter as well: each wait for a case field read by a scalar load
// would then drain the density loads in flight, one by one.
template <class T> using gptr = __attribute__((address_space(1))) T*;

// ---------------------------------------------------------------- d2q9 ---
// models/flow/d2q9_SRT/Dynamics.c, same operations in the same order.
template <class R>
__device__ __forceinline__ void d2q9_feq(R* f_eq, R d, R ux, R uy) {
    f_eq[0] = ( R(2.) + ( -uy*uy - ux*ux )*R(3.) )*d*R(2.)/R(9.);
    f_eq[1] = ( R(2.) + ( -uy*uy + ( 1 + ux )*ux*R(2.) )*R(3.) )*d/R(18.);
    f_eq[2] = ( R(2.) + ( -ux*ux + ( 1 + uy )*uy*R(2.) )*R(3.) )*d/R(18.);
    f_eq[3] = ( R(2.) + ( -uy*uy + ( -1 + ux )*ux*R(2.) )*R(3.) )*d/R(18.);
    f_eq[4] = ( R(2.) + ( -ux*ux + ( -1 + uy )*uy*R(2.) )*R(3.) )*d/R(18.);
    f_eq[5] = ( R(1.) + ( ( 1 + uy )*uy + ( 1 + ux + uy*R(3.) )*ux )*R(3.) )*d/R(36.);
    f_eq[6] = ( R(1.) + ( ( 1 + uy )*uy + ( -1 + ux - uy*R(3.) )*ux )*R(3.) )*d/R(36.);
    f_eq[7] = ( R(1.) + ( ( -1 + uy )*uy + ( -1 + ux + uy*R(3.) )*ux )*R(3.) )*d/R(36.);
    f_eq[8] = ( R(1.) + ( ( -1 + uy )*uy + ( 1 + ux - uy*R(3.) )*ux )*R(3.) )*d/R(36.);
}

// Run(), Dynamics.c:40-67, plus (DRAG) Ladd momentum exchange on Solid nodes.
// zone columns: Velocity, Velocity_x, Velocity_y, GravitationX,
// GravitationY, [DragInObj, LiftInObj].
// The float BGK collision, f += omega (feq - f) (Dynamics.c:57-66), on the
// deviations d = f - w, written so that the update Delta conserves mass and
// gives momentum rho G exactly up to unbiased roundings: Delta is evaluated
// for directions 1, 2, 5, 6, 7, 8 from feq - w = w_i (rho - 1 + 3 e.U +
// (9/2 (e.U)^2 - 3/2 U^2)/rho), U = J + rho G/omega (the reference's
// rho u with u = J/rho + G/omega), direction 3 and 4 from the momentum
// balance and the rest population from the mass balance.  Evaluated
// directly, the rounded weights 1/9, 1/36 and an approximate 1/rho create
// mass and momentum each step.
__device__ __forceinline__ void d2q9_bgk_dev(float* d, float omega, const double* zr) {
    const float s = ((d[0] + d[1]) + (d[2] + d[3])) + ((d[4] + d[5]) + (d[6] + d[7])) + d[8];
    const float rho = 1.f + s;
    const float d57 = d[5] - d[7], d68 = d[6] - d[8];
    const float Jx = (d[1] - d[3]) + d57 - d68, Jy = (d[2] - d[4]) + d57 + d68;
    const float Gx = (float)zr[3], Gy = (float)zr[4];
    const float rw = rho * __builtin_amdgcn_rcpf(omega);
    const float Ux = __builtin_fmaf(Gx, rw, Jx), Uy = __builtin_fmaf(Gy, rw, Jy);
    const float ir = __builtin_amdgcn_rcpf(rho), h = 4.5f * ir;
    const float A = s - 1.5f * (Ux * Ux + Uy * Uy) * ir;   // rho - 1 - 3/2 U^2/rho
    const float p = Ux + Uy, m = Uy - Ux;
    const float D1 = omega * (__builtin_fmaf(Ux, __builtin_fmaf(h, Ux, 3.f), A) * (1.f / 9.f) - d[1]);
    const float D2 = omega * (__builtin_fmaf(Uy, __builtin_fmaf(h, Uy, 3.f), A) * (1.f / 9.f) - d[2]);
    const float D5 = omega * (__builtin_fmaf(p, __builtin_fmaf(h, p, 3.f), A) * (1.f / 36.f) - d[5]);
    const float D7 = omega * (__builtin_fmaf(p, __builtin_fmaf(h, p, -3.f), A) * (1.f / 36.f) - d[7]);
    const float D6 = omega * (__builtin_fmaf(m, __builtin_fmaf(h, m, 3.f), A) * (1.f / 36.f) - d[6]);
    const float D8 = omega * (__builtin_fmaf(m, __builtin_fmaf(h, m, -3.f), A) * (1.f / 36.f) - d[8]);
    const float E57 = D5 - D7, E68 = D6 - D8;
    const float D3 = (D1 + E57) - (E68 + rho * Gx);
    const float D4 = (D2 + E57) + (E68 - rho * Gy);
    const float D0 = -(((D1 + D3) + (D2 + D4)) + ((D5 + D7) + (D6 + D8)));
    d[0] += D0; d[1] += D1; d[2] += D2; d[3] += D3; d[4] += D4;
    d[5] += D5; d[6] += D6; d[7] += D7; d[8] += D8;
}

template <class R, bool DRAG, bool GLOB, bool COLLIDE = true>
__device__ __forceinline__ void d2q9_node(R* f, unsigned nt, R omega, R Density, const double* zr,
                                          double* g) {
    R uf;
    switch (nt & 0x7u) {
    case 5:
    case 6:
        if (DRAG && GLOB && (nt & 0x7u) == 5) {
            R fx = R(2.0) * (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]);
            R fy = R(2.0) * (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]);
            g[0] += (double)fx; g[2] += zr[5];
            g[1] += (double)fy; g[2] += zr[6];
        }
        uf = f[3]; f[3] = f[1]; f[1] = uf;
        uf = f[4]; f[4] = f[2]; f[2] = uf;
        uf = f[7]; f[7] = f[5]; f[5] = uf;
        uf = f[8]; f[8] = f[6]; f[6] = uf;
        break;
    case 4: {  // EVelocity
        R ux0 = (R)zr[1];
        R rho = ( f[0] + f[2] + f[4] + R(2.)*(f[1] + f[5] + f[8]) ) / (R(1.) + ux0);
        R ru = rho * ux0;
        f[3] = f[1] - (R(2.)/R(3.)) * ru;
        f[7] = f[5] - (R(1.)/R(6.)) * ru + (R(1.)/R(2.))*(f[2] - f[4]);
        f[6] = f[8] - (R(1.)/R(6.)) * ru + (R(1.)/R(2.))*(f[4] - f[2]);
        break; }
    case 2: {  // WPressure
        R rho = Density;
        R ux0 = R(-1.) + ( f[0] + f[2] + f[4] + R(2.)*(f[3] + f[7] + f[6]) ) / rho;
        R ru = rho * ux0;
        f[1] = f[3] - (R(2.)/R(3.)) * ru;
        f[5] = f[7] - (R(1.)/R(6.)) * ru + (R(1.)/R(2.))*(f[4] - f[2]);
        f[8] = f[6] - (R(1.)/R(6.)) * ru + (R(1.)/R(2.))*(f[2] - f[4]);
        break; }
    case 3: {  // WVelocity
        R u0 = (R)zr[1];
        R rho = ( f[0] + f[2] + f[4] + R(2.)*(f[3] + f[7] + f[6]) ) / (R(1.) - u0);
        R ru = rho * u0;
        f[1] = f[3] + (R(2.)/R(3.)) * ru;
        f[5] = f[7] + (R(1.)/R(6.)) * ru + (R(1.)/R(2.))*(f[4] - f[2]);
        f[8] = f[6] + (R(1.)/R(6.)) * ru + (R(1.)/R(2.))*(f[2] - f[4]);
        break; }
    case 1: {  // EPressure
        R rho = Density;
        R ux0 = R(-1.) + ( f[0] + f[2] + f[4] + R(2.)*(f[1] + f[5] + f[8]) ) / rho;
        R ru = rho * ux0;
        f[3] = f[1] - (R(2.)/R(3.)) * ru;
        f[7] = f[5] - (R(1.)/R(6.)) * ru + (R(1.)/R(2.))*(f[2] - f[4]);
        f[6] = f[8] - (R(1.)/R(6.)) * ru + (R(1.)/R(2.))*(f[4] - f[2]);
        break; }
    default: break;
    }
    if (COLLIDE && (nt & 0x8u)) {  // BGK
        if constexpr (sizeof(R) != 8) {
            // float: on the deviations (f32 storage holds f itself)
            using L = Lat<MODEL_D2Q9_SRT>;
#pragma unroll
            for (int i = 0; i < 9; i++) f[i] -= (R)L::w(i);
            d2q9_bgk_dev(f, omega, zr);
#pragma unroll
            for (int i = 0; i < 9; i++) f[i] += (R)L::w(i);
            return;
        }
        R d = f[8]+f[7]+f[6]+f[5]+f[4]+f[3]+f[2]+f[1]+f[0];
        R f_eq[9];
        R ux = (( f[8]-f[7]-f[6]+f[5]-f[3]+f[1] )/d + (R)zr[3]/omega );
        R uy = ((-f[8]-f[7]+f[6]+f[5]-f[4]+f[2] )/d + (R)zr[4]/omega );
        d2q9_feq(f_eq, d, ux, uy);
#pragma unroll
        for (int i = 0; i < 9; i++) f[i] = f[i] + omega*(f_eq[i]-f[i]);
    }
}

// d2q9_SRT node of a shifted storage format, on the deviations d = f - w:
// boundary nodes add w back for their boundary condition, BGK runs on d.
template <bool DRAG, bool GLOB>
__device__ __forceinline__ void d2q9_node_dev(float* d, unsigned nt, float omega, float Density, const double* zr,
                                              double* g) {
    using L = Lat<MODEL_D2Q9_SRT>;
    if ((nt & 0x7u) != 0u) {
#pragma unroll
        for (int i = 0; i < 9; i++) d[i] += (float)L::w(i);
        d2q9_node<float, DRAG, GLOB, false>(d, nt, omega, Density, zr, g);
#pragma unroll
        for (int i = 0; i < 9; i++) d[i] -= (float)L::w(i);
    }
    if (nt & 0x8u) d2q9_bgk_dev(d, omega, zr);
}

// ------------------------------------------------------------ d2q9 MRT ---
// models/flow/d2q9/Dynamics.c.Rt.  The moment basis the reference generates
// (src/lib/feq.R MRT_eq with the Cholesky weighting of Dynamics.c.Rt:11-20)
// is the integer Hermite basis: moment k of direction i is kB9[i][k], for
// e = (ex, ey): 1, ex, ey, 3ex^2-1, 3ey^2-1, 3ex ey, ey(3ex^2-1),
// ex(3ey^2-1), (3ex^2-1)(3ey^2-1), of orders 0,1,1,2,2,2,3,3,4.  Its
// equilibrium moments are rho, Jx, Jy, 3Jx^2/rho, 3Jy^2/rho, 3JxJy/rho, 0,
// 0, 0, and the inverse is kB9i[k][i] = w_i kB9[i][k] / |moment k|^2
// (exact fractions below).
__device__ constexpr int kB9[9][9] = {
    {1, 0, 0, -1, -1, 0, 0, 0, 1},   {1, 1, 0, 2, -1, 0, 0, -1, -2},  {1, 0, 1, -1, 2, 0, -1, 0, -2},
    {1, -1, 0, 2, -1, 0, 0, 1, -2},  {1, 0, -1, -1, 2, 0, 1, 0, -2},  {1, 1, 1, 2, 2, 3, 2, 2, 4},
    {1, -1, 1, 2, 2, -3, 2, -2, 4},  {1, -1, -1, 2, 2, 3, -2, -2, 4}, {1, 1, -1, 2, 2, -3, -2, 2, 4}};
__device__ constexpr double kB9i[9][9] = {
    {4. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 9, 1. / 36, 1. / 36, 1. / 36, 1. / 36},
    {0, 1. / 3, 0, -1. / 3, 0, 1. / 12, -1. / 12, -1. / 12, 1. / 12},
    {0, 0, 1. / 3, 0, -1. / 3, 1. / 12, 1. / 12, -1. / 12, -1. / 12},
    {-2. / 9, 1. / 9, -1. / 18, 1. / 9, -1. / 18, 1. / 36, 1. / 36, 1. / 36, 1. / 36},
    {-2. / 9, -1. / 18, 1. / 9, -1. / 18, 1. / 9, 1. / 36, 1. / 36, 1. / 36, 1. / 36},
    {0, 0, 0, 0, 0, 1. / 12, -1. / 12, 1. / 12, -1. / 12},
    {0, 0, -1. / 6, 0, 1. / 6, 1. / 12, 1. / 12, -1. / 12, -1. / 12},
    {0, -1. / 6, 0, 1. / 6, 0, 1. / 12, -1. / 12, -1. / 12, 1. / 12},
    {1. / 9, -1. / 18, -1. / 18, -1. / 18, -1. / 18, 1. / 36, 1. / 36, 1. / 36, 1. / 36}};

// The basis is a tensor product: per axis (1, e, 3e^2 - 1) over e = -1, 0,
// 1, and moment (a, b) of that product is k = {0,1,3}[a] + {0,2,4}[b] for
// a + b <= 1, k = 5 (3 x the (1,1) product), 6 = (2,1), 7 = (1,2), 8 =
// (2,2).  One axis maps (f-, f0, f+) to (f- + f0 + f+, f+ - f-, 2(f+ + f-)
// - f0) and back by f0 = (2 M0 - M2)/3, f+- = (M0 + M2)/6 +- M1/2.  The
// float path transforms by these two passes (about half the operations
// of the dense products); FP64 keeps the dense products of the reference's
// generated code, and with them bit-for-bit parity with the oracle.
// Rows by e_y (-1, 0, +1) of the directions by e_x (-1, 0, +1):
__device__ constexpr int kD9rows[3][3] = {{7, 4, 8}, {3, 0, 1}, {6, 2, 5}};

template <class R>
__device__ __forceinline__ void d2q9m_moments_fact(const R* f, R* m) {
    R A[3][3];  // A[a][row]
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const R fm = f[kD9rows[r][0]], f0 = f[kD9rows[r][1]], fp = f[kD9rows[r][2]];
        const R s = fp + fm;
        A[0][r] = s + f0;
        A[1][r] = fp - fm;
        A[2][r] = R(2) * s - f0;
    }
    R T[3][3];  // T[a][b]
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const R s = A[a][2] + A[a][0];
        T[a][0] = s + A[a][1];
        T[a][1] = A[a][2] - A[a][0];
        T[a][2] = R(2) * s - A[a][1];
    }
    m[0] = T[0][0]; m[1] = T[1][0]; m[2] = T[0][1]; m[3] = T[2][0]; m[4] = T[0][2];
    m[5] = R(3) * T[1][1]; m[6] = T[2][1]; m[7] = T[1][2]; m[8] = T[2][2];
}

template <class R>
__device__ __forceinline__ void d2q9m_from_moments_fact(R* f, const R* m) {
    const R third = R(1) / R(3), sixth = R(1) / R(6);
    const R T[3][3] = {{m[0], m[2], m[4]}, {m[1], m[5] * third, m[7]}, {m[3], m[6], m[8]}};
    R A[3][3];  // A[a][row]
#pragma unroll
    for (int a = 0; a < 3; a++) {
        const R c = (T[a][0] + T[a][2]) * sixth, d = T[a][1] * R(0.5);
        A[a][1] = (R(2) * T[a][0] - T[a][2]) * third;
        A[a][2] = c + d;
        A[a][0] = c - d;
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        const R c = (A[0][r] + A[2][r]) * sixth, d = A[1][r] * R(0.5);
        f[kD9rows[r][1]] = (R(2) * A[0][r] - A[2][r]) * third;
        f[kD9rows[r][2]] = c + d;
        f[kD9rows[r][0]] = c - d;
    }
}

// f = R B^-1 over the nonzero coefficients in ascending k (FP64), or by the
// two passes above (float)
template <class R>
__device__ __forceinline__ void d2q9m_from_moments(R* f, const R* m) {
    if constexpr (!std::is_same<R, double>::value) {
        d2q9m_from_moments_fact(f, m);
        return;
    }
#pragma unroll
    for (int i = 0; i < 9; i++) {
        R acc = 0;
#pragma unroll
        for (int k = 0; k < 9; k++)
            if (kB9i[k][i] != 0) acc += m[k] * R(kB9i[k][i]);
        f[i] = acc;
    }
}

// SetEquilibrum(rho, Jx, Jy), Dynamics.c.Rt:73-79
template <class R>
__device__ __forceinline__ void d2q9m_feq(R* f, R rho, R Jx, R Jy) {
    R m[9] = {rho, Jx, Jy, R(3.) * Jx * Jx / rho, R(3.) * Jy * Jy / rho, R(3.) * Jx * Jy / rho, R(0), R(0), R(0)};
    d2q9m_from_moments(f, m);
}

// ZouHeRewrite (src/lib/boundary.R:155-207) for the side whose outward
// normal is s along axis A (0 = x, 1 = y): the populations entering the
// domain (e.n < 0) become f_opp + 6 w (e.R), R solved so that the node gets
// velocity (VelocityX, VelocityY) or density 1 + 3 Pressure with no
// tangential momentum.  With S0 = sum of the tangential f + 2 sum of the
// outgoing f and T their tangential momentum, the equations give
//   velocity: R_n = V_n S0 / (1 + s V_n),  R_t = 3 (V_t (S0 - s R_n) - T)
//   pressure: R_n = s (S0 - (3 P + 1)),    R_t = -3 T.
// d2q9 density with the x (y) component of its velocity negated
__host__ __device__ constexpr int d2q9_symx(int i) {
    return i == 1 ? 3 : i == 3 ? 1 : i == 5 ? 6 : i == 6 ? 5 : i == 7 ? 8 : i == 8 ? 7 : i;
}
__host__ __device__ constexpr int d2q9_symy(int i) {
    return i == 2 ? 4 : i == 4 ? 2 : i == 5 ? 8 : i == 8 ? 5 : i == 6 ? 7 : i == 7 ? 6 : i;
}
__host__ __device__ constexpr int d2q9_opp(int i) { return i == 0 ? 0 : (i < 5 ? (i + 1) % 4 + 1 : (i - 3) % 4 + 5); }

template <class R, int A, int SGN, bool VEL>
__device__ __forceinline__ void d2q9m_zouhe(R* f, R Vx, R Vy, R P) {
    using L = Lat<MODEL_D2Q9_SRT>;
    R S0 = 0, T = 0;
#pragma unroll
    for (int i = 0; i < 9; i++) {
        const int en = SGN * (A == 0 ? L::ex(i) : L::ey(i));
        const int et = A == 0 ? L::ey(i) : L::ex(i);
        if (en == 0) {
            S0 += f[i];
            T += f[i] * R(et);
        } else if (en > 0) {
            S0 += R(2.) * f[i];
        }
    }
    R Rn, Rt;
    if (VEL) {
        const R Vn = A == 0 ? Vx : Vy, Vt = A == 0 ? Vy : Vx;
        Rn = Vn * S0 / (R(1.) + R(SGN) * Vn);
        Rt = R(3.) * (Vt * (S0 - R(SGN) * Rn) - T);
    } else {
        Rn = R(SGN) * (S0 - (P * R(3.) + R(1.)));
        Rt = R(-3.) * T;
    }
    const R Rx = A == 0 ? Rn : Rt, Ry = A == 0 ? Rt : Rn;
#pragma unroll
    for (int i = 1; i < 9; i++) {
        const int en = SGN * (A == 0 ? L::ex(i) : L::ey(i));
        if (en < 0) {
            const R cu = R(L::ex(i)) * Rx + R(L::ey(i)) * Ry;
            f[i] = f[d2q9_opp(i)] + (i < 5 ? cu * R(2.) / R(3.) : cu / R(6.));
        }
    }
}

// Symmetry(2, -1 / +1) (src/lib/boundary.R:33-50): the populations entering
// through a north (south) symmetry plane copy their mirror images
template <class R, bool NORTH>
__device__ __forceinline__ void d2q9m_symmetry(R* f) {
    if (NORTH) { f[4] = f[2]; f[8] = f[5]; f[7] = f[6]; }
    else { f[2] = f[4]; f[5] = f[8]; f[6] = f[7]; }
}

// The float MRT collision (CollisionMRT, Dynamics.c.Rt:196-268) as an
// increment, f += B^-1 (m' - m), with m = f B the moments and m' the relaxed
// and forced ones.  The inverse transform's float constants 1/3 and 1/6 are
// both rounded up, so the product form f = B^-1 m' scales every population
// by about 1 + 6e-8 and creates that much mass each step (a drift that
// dominated the float formats' error on open channels); the increment has
// no mass and the force as its momentum, so those constants only scale the
// non-equilibrium change.  DEV: f holds the deviations f - w, whose moments
// are those of f except m0, which is rho - 1 (the rest populations w have
// only the zeroth moment).  Zone columns and Globals as d2q9m_node.
template <bool GLOB, bool DRAG, bool BCOPT, bool DEV>
__device__ __forceinline__ void d2q9m_collide_f(float* f, unsigned nt, const double* par, const double* zr,
                                                double* g, double* bc, size_t NN) {
    constexpr int OBJ = DRAG ? 5 : 3;
    float m[9];
    d2q9m_moments_fact(f, m);
    const float rho = DEV ? 1.f + m[0] : m[0];
    const float Jx = m[1], Jy = m[2];
    if (GLOB) {
        const unsigned obj = nt & 0x60u;
        if (obj == 0x40u || obj == 0x20u) {  // Outlet / Inlet
            const float Usq = (Jx * Jx + Jy * Jy) / (rho * rho);
            const float flux = Jx / rho / rho;
            const float loss = (Jx / rho) / rho * ((rho - 1.f) / 3.f + Usq / rho / 2.f);
            const int k = obj == 0x40u ? 1 : 2;
            g[k] += (double)flux; g[OBJ] += zr[3 + k];
            g[0] += (double)(obj == 0x40u ? -loss : loss); g[OBJ] += zr[3];
        }
    }
    // 3 a b / rho with one approximate reciprocal (second-order terms only)
    const float ir3 = 3.f * __builtin_amdgcn_rcpf(rho);
    const float S2 = (float)par[0], S3 = (float)par[1], S4 = (float)par[2];
    const float n3 = m[3] - Jx * Jx * ir3, n4 = m[4] - Jy * Jy * ir3, n5 = m[5] - Jx * Jy * ir3;
    float Gx = (float)par[3], Gy = (float)par[4];
    if constexpr (BCOPT) {
        // only the nodes with no BOUNDARY type take the force, G + BC
        // (Dynamics.c.Rt:254-258)
        if ((nt & 0xFu) == 0u) {
            Gx += bc ? (float)bc[0] : 0.f;
            Gy += bc ? (float)bc[NN] : 0.f;
        } else {
            Gx = 0.f;
            Gy = 0.f;
        }
    }
    const float Kx = __builtin_fmaf(Gx, rho, Jx), Ky = __builtin_fmaf(Gy, rho, Jy);
    float dm[9];
    dm[0] = 0.f;
    dm[1] = Kx - Jx;
    dm[2] = Ky - Jy;
    dm[3] = __builtin_fmaf(S2, n3, Kx * Kx * ir3) - m[3];
    dm[4] = __builtin_fmaf(S2, n4, Ky * Ky * ir3) - m[4];
    dm[5] = __builtin_fmaf(S2, n5, Kx * Ky * ir3) - m[5];
    dm[6] = __builtin_fmaf(S3, m[6], -m[6]);
    dm[7] = __builtin_fmaf(S3, m[7], -m[7]);
    dm[8] = __builtin_fmaf(S4, m[8], -m[8]);
    float df[9];
    d2q9m_from_moments_fact(df, dm);
#pragma unroll
    for (int i = 0; i < 9; i++) f[i] += df[i];
}

// Run(), Dynamics.c.Rt:104-138 and CollisionMRT, :196-268.  Zone columns:
// [VelocityX, VelocityY, Pressure, PressureLossInObj, OutletFluxInObj,
// InletFluxInObj, (d2q9_drag) DragInObj, LiftInObj]; Globals
// [PressureLoss, OutletFlux, InletFlux, (DRAG) Drag, Lift, Objective].
// DRAG adds the momentum exchange of Solid nodes, 2 sum_i f_i e_i of the
// incoming populations (as d2q9_SRT_drag).  BCOPT: the bc builds, with the
// node's BC[0] at bc[0] and BC[1] at bc[NN] (null: BC = 0, nothing kept).
template <class R, bool GLOB, bool DRAG, bool BCOPT = false, bool COLLIDE = true>
__device__ __forceinline__ void d2q9m_node(R* f, unsigned nt, const double* par, const double* zr, double* g,
                                           double* bc = nullptr, size_t NN = 0) {
    constexpr int OBJ = DRAG ? 5 : 3;
    R Vx = (R)zr[0], Vy = (R)zr[1];
    const R P = (R)zr[2];
    if constexpr (BCOPT) {  // BC_Velocity = (BC[0], BC[1]) (Dynamics.c.Rt:28-34)
        Vx = bc ? (R)bc[0] : R(0);
        Vy = bc ? (R)bc[NN] : R(0);
    }
    switch (nt & 0xFu) {
    case 9:   // Solid
    case 10: {  // Wall: FullBounceBack
        if constexpr (BCOPT) {
            // BC = 2 sum f e of what is reversed, 0 when the node holds
            // exactly 1 + 3 Pressure (Dynamics.c.Rt:146-159)
            const R temp = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
            const bool keep = temp != R(1) + P * R(3);
            if (bc) {
                bc[0] = keep ? (double)(R(2.) * (f[1] - f[3] + f[5] - f[6] - f[7] + f[8])) : 0.0;
                bc[NN] = keep ? (double)(R(2.) * (f[2] - f[4] + f[5] + f[6] - f[7] - f[8])) : 0.0;
            }
        }
        if (DRAG && GLOB && (nt & 0xFu) == 9u) {
            const R fx = R(2.0) * (f[1] - f[3] + f[5] - f[6] - f[7] + f[8]);
            const R fy = R(2.0) * (f[2] - f[4] + f[5] + f[6] - f[7] - f[8]);
            g[3] += (double)fx; g[OBJ] += zr[6];
            g[4] += (double)fy; g[OBJ] += zr[7];
        }
        R t;
        t = f[1]; f[1] = f[3]; f[3] = t;
        t = f[2]; f[2] = f[4]; f[4] = t;
        t = f[5]; f[5] = f[7]; f[7] = t;
        t = f[6]; f[6] = f[8]; f[8] = t;
        break; }
    case 6: d2q9m_zouhe<R, 0, 1, true>(f, Vx, Vy, P); break;    // EVelocity
    case 2: d2q9m_zouhe<R, 0, -1, false>(f, Vx, Vy, P); break;  // WPressure
    case 5: d2q9m_zouhe<R, 0, -1, true>(f, Vx, Vy, P); break;   // WVelocity
    case 1: d2q9m_zouhe<R, 0, 1, false>(f, Vx, Vy, P); break;   // EPressure
    case 3: d2q9m_zouhe<R, 1, 1, true>(f, Vx, Vy, P); break;    // NVelocity
    case 4: d2q9m_zouhe<R, 1, -1, true>(f, Vx, Vy, P); break;   // SVelocity
    case 7: d2q9m_symmetry<R, true>(f); break;                  // NSymmetry
    case 8: d2q9m_symmetry<R, false>(f); break;                 // SSymmetry
    default: break;
    }
    if (!COLLIDE || (nt & 0x10u) == 0) return;  // MRT
    if constexpr (!std::is_same<R, double>::value) {
        d2q9m_collide_f<GLOB, DRAG, BCOPT, false>(f, nt, par, zr, g, bc, NN);
        return;
    }
    R m[9];
    {
#pragma unroll
        for (int k = 0; k < 9; k++) {
            R acc = 0;
#pragma unroll
            for (int i = 0; i < 9; i++)
                if (kB9[i][k] != 0) acc += f[i] * R(kB9[i][k]);
            m[k] = acc;
        }
    }
    const R rho = m[0];
    R Jx = m[1], Jy = m[2];
    if (GLOB) {
        const unsigned obj = nt & 0x60u;
        if (obj == 0x40u || obj == 0x20u) {  // Outlet / Inlet
            const R Usq = (Jx * Jx + Jy * Jy) / (rho * rho);
            const R flux = Jx / rho / rho;
            const R loss = (Jx / rho) / rho * ((rho - R(1.)) / R(3.) + Usq / rho / R(2.));
            const int k = obj == 0x40u ? 1 : 2;
            g[k] += (double)flux; g[OBJ] += zr[3 + k];
            g[0] += (double)(obj == 0x40u ? -loss : loss); g[OBJ] += zr[3];
        }
    }
    // FP64 from here on (the float formats take d2q9m_collide_f): the
    // reference's operations, IEEE divisions included
    const auto q3 = [&](R u, R v) { return R(3.) * u * v / rho; };
    const R S2 = (R)par[0], S3 = (R)par[1], S4 = (R)par[2];
    m[3] = S2 * (m[3] - q3(Jx, Jx));
    m[4] = S2 * (m[4] - q3(Jy, Jy));
    m[5] = S2 * (m[5] - q3(Jx, Jy));
    m[6] = S3 * m[6];
    m[7] = S3 * m[7];
    m[8] = S4 * m[8];
    if constexpr (BCOPT) {
        // only the nodes with no BOUNDARY type take the force, G + BC
        // (Dynamics.c.Rt:254-258)
        if ((nt & 0xFu) == 0u) {
            Jx = Jx + ((R)par[3] + (bc ? (R)bc[0] : R(0))) * rho;
            Jy = Jy + ((R)par[4] + (bc ? (R)bc[NN] : R(0))) * rho;
        }
    } else {
        Jx = Jx + (R)par[3] * rho;
        Jy = Jy + (R)par[4] * rho;
    }
    m[1] = Jx;
    m[2] = Jy;
    m[3] = m[3] + q3(Jx, Jx);
    m[4] = m[4] + q3(Jy, Jy);
    m[5] = m[5] + q3(Jx, Jy);
    d2q9m_from_moments(f, m);
}

// d2q9 MRT node of a shifted storage format, on the deviations d = f - w:
// boundary nodes add w back for their boundary condition, the collision
// runs on d.
template <bool GLOB, bool DRAG, bool BCOPT>
__device__ __forceinline__ void d2q9m_node_dev(float* d, unsigned nt, const double* par, const double* zr,
                                               double* g, double* bc, size_t NN) {
    using L = Lat<MODEL_D2Q9_SRT>;
    if ((nt & 0xFu) != 0u) {
#pragma unroll
        for (int i = 0; i < 9; i++) d[i] += (float)L::w(i);
        d2q9m_node<float, GLOB, DRAG, BCOPT, false>(d, nt, par, zr, g, bc, NN);
#pragma unroll
        for (int i = 0; i < 9; i++) d[i] -= (float)L::w(i);
    }
    if ((nt & 0x10u) != 0u) d2q9m_collide_f<GLOB, DRAG, BCOPT, true>(d, nt, par, zr, g, bc, NN);
}

// --------------------------------------------------------------- d3q19 ---
// Equilibrium moments (src/lib/d3q19.R:24-62, product form truncated to
// O(J^2)); entries 0,3,5,7 are the conserved rho, J.
template <class R>
__device__ __forceinline__ void d3q19_req(R* q, R rho, R Jx, R Jy, R Jz) {
    R xx = Jx*Jx/rho, yy = Jy*Jy/rho, zz = Jz*Jz/rho;
    R j2 = xx + yy + zz;
    q[1] = R(-11.)*rho + R(19.)*j2;
    q[2] = R(3.)*rho - R(2.)*j2;
    q[4] = R(-2.)*Jx/R(3.);
    q[6] = R(-2.)*Jy/R(3.);
    q[8] = R(-2.)*Jz/R(3.);
    q[9] = R(2.)*xx - yy - zz;
    q[10] = -q[9];
    q[11] = yy - zz;
    q[12] = -q[11];
    q[13] = Jx*Jy/rho;
    q[14] = Jy*Jz/rho;
    q[15] = Jx*Jz/rho;
}

// The same moments for the float path: one reciprocal of rho and constant
// products instead of the nine IEEE divisions (the float formats are held to
// a tolerance of a few float ulps, not to identical roundings).
// `lin` is the density of the rho-linear terms (rho itself, or rho - 1 for
// deviations from the rest state), `rho` the one that divides J^2.
__device__ __forceinline__ void d3q19_req_fast(float* q, float rho, float Jx, float Jy, float Jz, float lin) {
    const float ir = __builtin_amdgcn_rcpf(rho);
    const float xx = Jx*Jx*ir, yy = Jy*Jy*ir, zz = Jz*Jz*ir;
    const float j2 = xx + yy + zz;
    q[1] = -11.f*lin + 19.f*j2;
    q[2] = 3.f*lin - 2.f*j2;
    q[4] = (-2.f/3.f)*Jx;
    q[6] = (-2.f/3.f)*Jy;
    q[8] = (-2.f/3.f)*Jz;
    q[9] = 2.f*xx - yy - zz;
    q[10] = -q[9];
    q[11] = yy - zz;
    q[12] = -q[11];
    q[13] = Jx*Jy*ir;
    q[14] = Jy*Jz*ir;
    q[15] = Jx*Jz*ir;
}

// f = sum_k r_k M[k][i] for the d3q19 moment basis of
// models/flow/d3q19/Dynamics.R:1-21, factored by symmetry groups.
template <class R>
__device__ __forceinline__ void d3q19_from_moments(R* f, const R* r) {
    R bf = r[0] - R(11.)*r[1] - R(4.)*r[2];
    R be = r[0] + R(8.)*r[1] + r[2];
    R d1x = r[3] - R(4.)*r[4], d1y = r[5] - R(4.)*r[6], d1z = r[7] - R(4.)*r[8];
    R t3 = r[3] + r[4], t5 = r[5] + r[6], t7 = r[7] + r[8];
    R p9 = R(2.)*r[9] - R(4.)*r[10];
    R q9 = R(2.)*r[10] - r[9];
    R e9 = r[9] + r[10];
    R w11 = r[11] - R(2.)*r[12];
    R e11 = r[11] + r[12];
    f[0] = r[0] - R(30.)*r[1] + R(12.)*r[2];
    f[1] = bf + d1x + p9;
    f[2] = bf - d1x + p9;
    f[3] = bf + d1y + q9 + w11;
    f[4] = bf - d1y + q9 + w11;
    f[5] = bf + d1z + q9 - w11;
    f[6] = bf - d1z + q9 - w11;
    R E0 = be + e9 + e11;
    f[7]  = E0 + t3 + t5 + r[13] + r[16] - r[17];
    f[8]  = E0 - t3 + t5 - r[13] - r[16] - r[17];
    f[9]  = E0 + t3 - t5 - r[13] + r[16] + r[17];
    f[10] = E0 - t3 - t5 + r[13] - r[16] + r[17];
    R F0 = be + e9 - e11;
    f[11] = F0 + t3 + t7 + r[15] - r[16] + r[18];
    f[12] = F0 - t3 + t7 - r[15] + r[16] + r[18];
    f[13] = F0 + t3 - t7 - r[15] - r[16] - r[18];
    f[14] = F0 - t3 - t7 + r[15] + r[16] - r[18];
    R G0 = be - R(2.)*e9;
    f[15] = G0 + t5 + t7 + r[14] + r[17] - r[18];
    f[16] = G0 - t5 + t7 - r[14] - r[17] - r[18];
    f[17] = G0 + t5 - t7 - r[14] + r[17] + r[18];
    f[18] = G0 - t5 - t7 + r[14] - r[17] + r[18];
}

// f = sum_k r_k M[k][i] in float: the groups of d3q19_from_moments with
// shared partial sums and FMAs for the integer coefficients.
__device__ __forceinline__ void d3q19_from_moments_f(float* f, const float* r) {
    const float bf = __builtin_fmaf(-11.f, r[1], __builtin_fmaf(-4.f, r[2], r[0]));
    const float be = __builtin_fmaf(8.f, r[1], r[0] + r[2]);
    const float d1x = __builtin_fmaf(-4.f, r[4], r[3]), d1y = __builtin_fmaf(-4.f, r[6], r[5]),
                d1z = __builtin_fmaf(-4.f, r[8], r[7]);
    const float t3 = r[3] + r[4], t5 = r[5] + r[6], t7 = r[7] + r[8];
    const float p9 = __builtin_fmaf(2.f, r[9], -4.f * r[10]);
    const float q9 = __builtin_fmaf(2.f, r[10], -r[9]);
    const float e9 = r[9] + r[10];
    const float w11 = __builtin_fmaf(-2.f, r[12], r[11]);
    const float e11 = r[11] + r[12];
    f[0] = __builtin_fmaf(-30.f, r[1], __builtin_fmaf(12.f, r[2], r[0]));
    const float bp = bf + p9, bq = bf + q9, bqp = bq + w11, bqm = bq - w11;
    f[1] = bp + d1x;
    f[2] = bp - d1x;
    f[3] = bqp + d1y;
    f[4] = bqp - d1y;
    f[5] = bqm + d1z;
    f[6] = bqm - d1z;
    const float be9 = be + e9;
    const float E0 = be9 + e11, F0 = be9 - e11, G0 = __builtin_fmaf(-2.f, e9, be);
    {   // 7..10: E0 +-t3 +-t5 with r13, r16, r17
        const float a = t3 + t5, b = t5 - t3, u = r[13] + r[16], v = r[16] - r[13];
        const float P = E0 - r[17], Qn = E0 + r[17];
        f[7] = (P + u) + a;
        f[8] = (P - u) + b;
        f[9] = (Qn + v) - b;
        f[10] = (Qn - v) - a;
    }
    {   // 11..14: F0 +-t3 +-t7 with r15, r16, r18
        const float a = t3 + t7, b = t7 - t3, x = r[15] - r[16], w = r[15] + r[16];
        const float P = F0 + r[18], Qn = F0 - r[18];
        f[11] = (P + x) + a;
        f[12] = (P - x) + b;
        f[13] = (Qn - w) - b;
        f[14] = (Qn + w) - a;
    }
    {   // 15..18: G0 +-t5 +-t7 with r14, r17, r18
        const float a = t5 + t7, b = t7 - t5, u = r[14] + r[17], v = r[17] - r[14];
        const float P = G0 - r[18], Qn = G0 + r[18];
        f[15] = (P + u) + a;
        f[16] = (P + b) - u;
        f[17] = (Qn + v) - b;
        f[18] = (Qn - v) - a;
    }
}

// SetEquilibrum, Dynamics.c.Rt:212-217: f = Req %*% diag(1/Mw) %*% t(M)
template <class R>
__device__ __forceinline__ void d3q19_feq(R* f, R rho, R Jx, R Jy, R Jz) {
    R q[19];
    d3q19_req(q, rho, Jx, Jy, Jz);
    q[0] = rho; q[3] = Jx; q[5] = Jy; q[7] = Jz;
    q[16] = 0; q[17] = 0; q[18] = 0;
    constexpr double iv[19] = {1./19, 1./2394, 1./252, 1./10, 1./40, 1./10, 1./40, 1./10, 1./40,
                               1./36, 1./72, 1./12, 1./24, 1./4, 1./4, 1./4, 1./8, 1./8, 1./8};
#pragma unroll
    for (int k = 0; k < 19; k++) q[k] = q[k] * R(iv[k]);
    d3q19_from_moments(f, q);
}

// Global i += x; Objective (index OBJ) += the zone's <Global>InObj weight
// (src/cuda.cu.Rt:114-119)
template <bool GLOB, int OBJ>
__device__ __forceinline__ void gadd(double* g, int i, double x, const double* zr) {
    if (GLOB) { g[i] += x; g[OBJ] += zr[i]; }
}

// XY/XZ/YZ slice Globals vx, vy, vz, rho, area starting at index BASE
template <bool GLOB, int OBJ, int BASE>
__device__ __forceinline__ void slice_add(double* g, double ux, double uy, double uz, double rs,
                                          const double* zr) {
    gadd<GLOB, OBJ>(g, BASE + 0, ux, zr);
    gadd<GLOB, OBJ>(g, BASE + 1, uy, zr);
    gadd<GLOB, OBJ>(g, BASE + 2, uz, zr);
    gadd<GLOB, OBJ>(g, BASE + 3, rs, zr);
    gadd<GLOB, OBJ>(g, BASE + 4, 1.0, zr);
}

// the volume and slice Globals of one collided node (Dynamics.c.Rt:266-289)
template <bool GLOB, bool LID, class R>
__device__ __forceinline__ void d3q19_globals(double* g, unsigned nt, R rs, R jx, R jy, R jz, const double* zr) {
    constexpr int OBJ = LID ? 28 : 24, MAXV = LID ? 29 : 25;
    const R ux = jx / rs, uy = jy / rs, uz = jz / rs;
    switch (nt & 0x3u) {  // slice Globals, register indices fixed at compile time
    case 1: slice_add<GLOB, OBJ, 1>(g, ux, uy, uz, rs, zr); break;
    case 2: slice_add<GLOB, OBJ, 6>(g, ux, uy, uz, rs, zr); break;
    case 3: slice_add<GLOB, OBJ, 11>(g, ux, uy, uz, rs, zr); break;
    }
    gadd<GLOB, OBJ>(g, 16, ux, zr);
    gadd<GLOB, OBJ>(g, 17, uy, zr);
    gadd<GLOB, OBJ>(g, 18, uz, zr);
    gadd<GLOB, OBJ>(g, 19, jx, zr);
    gadd<GLOB, OBJ>(g, 20, jy, zr);
    gadd<GLOB, OBJ>(g, 21, jz, zr);
    gadd<GLOB, OBJ>(g, 22, rs, zr);
    gadd<GLOB, OBJ>(g, 23, 1.0, zr);
    const double v = (double)sqrt(ux*ux + uy*uy + uz*uz);
    g[MAXV] = fmax(g[MAXV], v);
}

// CollisionMRT, Dynamics.c.Rt:223-290.  Rates by moment group (0-based):
// e {1}, eps {2}, q {4,6,8}, nu {9,11,13,14,15}, pi {10,12}, m {16,17,18};
// the reference uses gamma1 = 1-omega for e/eps/nu/pi and
// The float collision of d3q19's MRT (CollisionMRT, Dynamics.c.Rt:196-262):
// the moments R = f M, relaxation of the non-conserved ones toward Req,
// the body force added to J after relaxation, Req of the forced J added
// back, R /= Mw (Mw = colSums(M^2)) and f = R M^T, as d3q19_mrt_dense
// computes them in FP64, here with fewer float operations: the moments and
// the inverse transform from shared pair sums and differences, the factor
// 1/Mw folded into the relaxation factors (computed on the host, `fp`) and
// into the equilibrium constants, and explicit FMAs (the float formats are
// held to a tolerance of a few float ulps, not to identical roundings).
// fp = [ge/2394, geps/252, gq/40, gamma1/36, gpi/72, gamma1/12, gpi/24,
// gamma1/4, gm/8, ForceX, ForceY, ForceZ]: gamma1 = 1 - omega, gamma2 =
// 1-8(2-omega)/(8-omega), ge, geps, gq, gpi, gm the group factors (a
// positive MRT_S* setting gives its group 1 - S).
// DEV: f holds the deviations f - w of a shifted storage format (the rest
// populations w have moments rho = 1, m1 = -11, m2 = 3 and no others), and
// the collision is evaluated on them: the linear terms of Req take rho - 1,
// J^2/rho and rho F take the full rho, and the result is again f - w.  The
// large constant parts cancel exactly instead of being rounded in float.
template <class R, bool GLOB, bool LID, bool DEV = false>
__device__ __forceinline__ void d3q19_mrt(R* f, unsigned nt, const float* fp, const double* zr, double* g) {
    static_assert(std::is_same<R, float>::value, "the FP64 collision is d3q19_mrt_dense");
    (void)nt;
    // pair sums and differences of opposite-in-one-axis directions
    const float s12 = f[1] + f[2], d12 = f[1] - f[2];
    const float s34 = f[3] + f[4], d34 = f[3] - f[4];
    const float s56 = f[5] + f[6], d56 = f[5] - f[6];
    const float s78 = f[7] + f[8], d78 = f[7] - f[8];
    const float s910 = f[9] + f[10], d910 = f[9] - f[10];
    const float s1112 = f[11] + f[12], d1112 = f[11] - f[12];
    const float s1314 = f[13] + f[14], d1314 = f[13] - f[14];
    const float s1516 = f[15] + f[16], d1516 = f[15] - f[16];
    const float s1718 = f[17] + f[18], d1718 = f[17] - f[18];
    const float E = s78 + s910, F = s1112 + s1314, G = s1516 + s1718;
    const float A = s12 + (s34 + s56), B = E + (F + G);
    const float Px = d78 + d910, Qx = d1112 + d1314;     // bx = Px + Qx, m16 = Px - Qx
    const float Ry = s78 - s910, Ty = d1516 + d1718;     // by = Ry + Ty, m17 = Ty - Ry
    const float Sz = s1112 - s1314, Uz = s1516 - s1718;  // bz = Sz + Uz, m18 = Sz - Uz
    const float bx = Px + Qx, by = Ry + Ty, bz = Sz + Uz;
    const float m0 = f[0] + (A + B);
    const float m1 = __builtin_fmaf(-30.f, f[0], __builtin_fmaf(-11.f, A, 8.f * B));
    const float m2 = __builtin_fmaf(12.f, f[0], __builtin_fmaf(-4.f, A, B));
    const float Jx = d12 + bx, Jy = d34 + by, Jz = d56 + bz;
    const float m4 = __builtin_fmaf(-4.f, d12, bx), m6 = __builtin_fmaf(-4.f, d34, by),
                m8 = __builtin_fmaf(-4.f, d56, bz);
    const float EF = E + F, c23 = s34 + s56, c2m3 = s34 - s56, EmF = E - F;
    const float m9 = __builtin_fmaf(2.f, s12 - G, EF - c23);
    const float m10 = __builtin_fmaf(-4.f, s12, __builtin_fmaf(2.f, c23 - G, EF));
    const float m11 = c2m3 + EmF;
    const float m12 = __builtin_fmaf(-2.f, c2m3, EmF);
    const float m13 = d78 - d910, m14 = d1516 - d1718, m15 = d1112 - d1314;
    const float m16 = Px - Qx, m17 = Ty - Ry, m18 = Sz - Uz;

    // Req of the unforced J (d3q19_req: rho-linear terms on m0, which is
    // rho - 1 for deviations)
    const float rho = DEV ? 1.f + m0 : m0;
    const float ir = __builtin_amdgcn_rcpf(rho);
    float xx = Jx * Jx * ir, yy = Jy * Jy * ir, zz = Jz * Jz * ir;
    float j2 = xx + yy + zz, yz = yy + zz;
    const float n1 = m1 - __builtin_fmaf(19.f, j2, -11.f * m0);
    const float n2 = m2 - __builtin_fmaf(-2.f, j2, 3.f * m0);
    const float n4 = __builtin_fmaf(2.f / 3.f, Jx, m4), n6 = __builtin_fmaf(2.f / 3.f, Jy, m6),
                n8 = __builtin_fmaf(2.f / 3.f, Jz, m8);
    const float q9 = __builtin_fmaf(2.f, xx, -yz);
    const float n9 = m9 - q9, n10 = m10 + q9;
    const float n11 = m11 - (yy - zz), n12 = m12 + (yy - zz);
    const float n13 = m13 - Jx * Jy * ir, n14 = m14 - Jy * Jz * ir, n15 = m15 - Jx * Jz * ir;

    // forcing after relaxation (Dynamics.c.Rt:256-258), then Req of the
    // forced J, everything scaled by 1/Mw
    const float Fx = __builtin_fmaf(rho, fp[9], Jx), Fy = __builtin_fmaf(rho, fp[10], Jy),
                Fz = __builtin_fmaf(rho, fp[11], Jz);
    xx = Fx * Fx * ir; yy = Fy * Fy * ir; zz = Fz * Fz * ir;
    j2 = xx + yy + zz; yz = yy + zz;
    const float ir4 = ir * 0.25f, yzd = yy - zz;
    float M[19];
    M[0] = m0 * (1.f / 19.f);
    M[1] = __builtin_fmaf(fp[0], n1, __builtin_fmaf(19.f / 2394.f, j2, (-11.f / 2394.f) * m0));
    M[2] = __builtin_fmaf(fp[1], n2, __builtin_fmaf(-2.f / 252.f, j2, (3.f / 252.f) * m0));
    // J/10 corrected to the rounded quotient: the rounded constant 0.1 alone
    // would add 1.5e-8 of the momentum each step
    const auto tenth = [](float x) {
        const float q = x * 0.1f;
        return __builtin_fmaf(__builtin_fmaf(-10.f, q, x), 0.1f, q);
    };
    M[3] = tenth(Fx);
    M[5] = tenth(Fy);
    M[7] = tenth(Fz);
    M[4] = __builtin_fmaf(fp[2], n4, (-1.f / 60.f) * Fx);
    M[6] = __builtin_fmaf(fp[2], n6, (-1.f / 60.f) * Fy);
    M[8] = __builtin_fmaf(fp[2], n8, (-1.f / 60.f) * Fz);
    M[9] = __builtin_fmaf(fp[3], n9, __builtin_fmaf(1.f / 18.f, xx, (-1.f / 36.f) * yz));
    M[10] = __builtin_fmaf(fp[4], n10, __builtin_fmaf(-1.f / 36.f, xx, (1.f / 72.f) * yz));
    M[11] = __builtin_fmaf(fp[5], n11, (1.f / 12.f) * yzd);
    M[12] = __builtin_fmaf(fp[6], n12, (-1.f / 24.f) * yzd);
    M[13] = __builtin_fmaf(fp[7], n13, Fx * Fy * ir4);
    M[14] = __builtin_fmaf(fp[7], n14, Fy * Fz * ir4);
    M[15] = __builtin_fmaf(fp[7], n15, Fx * Fz * ir4);
    M[16] = fp[8] * m16;
    M[17] = fp[8] * m17;
    M[18] = fp[8] * m18;
    d3q19_from_moments_f(f, M);

    // Dynamics.c.Rt:266-289 evaluates the Globals after "R /= Mw", i.e. on
    // rho/19 and J/10 -- kept as the reference computes them
    if (GLOB) d3q19_globals<GLOB, LID>(g, nt, DEV ? rho * (1.f / 19.f) : M[0], M[3], M[5], M[7], zr);
}

// The d3q19 moment matrix M[k][i] (moment k of direction i), the listed
// rows of MRTMAT in models/flow/d3q19/Dynamics.R:1-21, and colSums(M^2).
__device__ constexpr int kM[19][19] = {
 {  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1},
 {-30,-11,-11,-11,-11,-11,-11,  8,  8,  8,  8,  8,  8,  8,  8,  8,  8,  8,  8},
 { 12, -4, -4, -4, -4, -4, -4,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1,  1},
 {  0,  1, -1,  0,  0,  0,  0,  1, -1,  1, -1,  1, -1,  1, -1,  0,  0,  0,  0},
 {  0, -4,  4,  0,  0,  0,  0,  1, -1,  1, -1,  1, -1,  1, -1,  0,  0,  0,  0},
 {  0,  0,  0,  1, -1,  0,  0,  1,  1, -1, -1,  0,  0,  0,  0,  1, -1,  1, -1},
 {  0,  0,  0, -4,  4,  0,  0,  1,  1, -1, -1,  0,  0,  0,  0,  1, -1,  1, -1},
 {  0,  0,  0,  0,  0,  1, -1,  0,  0,  0,  0,  1,  1, -1, -1,  1,  1, -1, -1},
 {  0,  0,  0,  0,  0, -4,  4,  0,  0,  0,  0,  1,  1, -1, -1,  1,  1, -1, -1},
 {  0,  2,  2, -1, -1, -1, -1,  1,  1,  1,  1,  1,  1,  1,  1, -2, -2, -2, -2},
 {  0, -4, -4,  2,  2,  2,  2,  1,  1,  1,  1,  1,  1,  1,  1, -2, -2, -2, -2},
 {  0,  0,  0,  1,  1, -1, -1,  1,  1,  1,  1, -1, -1, -1, -1,  0,  0,  0,  0},
 {  0,  0,  0, -2, -2,  2,  2,  1,  1,  1,  1, -1, -1, -1, -1,  0,  0,  0,  0},
 {  0,  0,  0,  0,  0,  0,  0,  1, -1, -1,  1,  0,  0,  0,  0,  0,  0,  0,  0},
 {  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  1, -1, -1,  1},
 {  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  1, -1, -1,  1,  0,  0,  0,  0},
 {  0,  0,  0,  0,  0,  0,  0,  1, -1,  1, -1, -1,  1, -1,  1,  0,  0,  0,  0},
 {  0,  0,  0,  0,  0,  0,  0, -1, -1,  1,  1,  0,  0,  0,  0,  1, -1,  1, -1},
 {  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  0,  1,  1, -1, -1, -1, -1,  1,  1}};
__device__ constexpr int kMw[19] = {19, 2394, 252, 10, 40, 10, 40, 10, 40, 36, 72, 12, 24, 4, 4, 4, 8, 8, 8};

// CollisionMRT in the reference's dense form (Dynamics.c.Rt:253-264):
// R = f M over the nonzero entries of M in ascending order, relaxation,
// forcing, R /= Mw by division, f = R M^T.  Used for FP64 storage, where it
// gives the same roundings as the CPU restatement of the reference.
template <class R, bool GLOB, bool LID>
__device__ __forceinline__ void d3q19_mrt_dense(R* f, unsigned nt, const double* par, const double* zr,
                                                double* g) {
    // relaxation factors computed on the host in double with the same
    // operations (gamma1 = 1 - omega, gamma2 = 1 - 8(2-omega)/(8-omega))
    const R ge = (R)par[6], geps = (R)par[7], gq = (R)par[8], gpi = (R)par[9], gm = (R)par[10];
    const R gamma1 = (R)par[11];
    R m[19];
#pragma unroll
    for (int k = 0; k < 19; k++) {
        R acc = 0;
#pragma unroll
        for (int i = 0; i < 19; i++)
            if (kM[k][i] != 0) acc += f[i] * R(kM[k][i]);
        m[k] = acc;
    }
    const R rho = m[0];
    R q[19];
    d3q19_req(q, rho, m[3], m[5], m[7]);
    m[1] = ge * (m[1] - q[1]);
    m[2] = geps * (m[2] - q[2]);
    m[4] = gq * (m[4] - q[4]);
    m[6] = gq * (m[6] - q[6]);
    m[8] = gq * (m[8] - q[8]);
    m[9] = gamma1 * (m[9] - q[9]);
    m[10] = gpi * (m[10] - q[10]);
    m[11] = gamma1 * (m[11] - q[11]);
    m[12] = gpi * (m[12] - q[12]);
    m[13] = gamma1 * (m[13] - q[13]);
    m[14] = gamma1 * (m[14] - q[14]);
    m[15] = gamma1 * (m[15] - q[15]);
    m[16] = gm * (m[16] - 0);
    m[17] = gm * (m[17] - 0);
    m[18] = gm * (m[18] - 0);
    m[3] += rho * (R)par[3];
    m[5] += rho * (R)par[4];
    m[7] += rho * (R)par[5];
    d3q19_req(q, rho, m[3], m[5], m[7]);
    m[1] += q[1]; m[2] += q[2]; m[4] += q[4]; m[6] += q[6]; m[8] += q[8];
#pragma unroll
    for (int k = 9; k < 16; k++) m[k] += q[k];
    m[16] += R(0); m[17] += R(0); m[18] += R(0);
#pragma unroll
    for (int k = 0; k < 19; k++) m[k] /= R(kMw[k]);
#pragma unroll
    for (int i = 0; i < 19; i++) {
        R acc = 0;
#pragma unroll
        for (int k = 0; k < 19; k++)
            if (kM[k][i] != 0) acc += m[k] * R(kM[k][i]);
        f[i] = acc;
    }
    if (GLOB) d3q19_globals<GLOB, LID>(g, nt, m[0], m[3], m[5], m[7], zr);
}

// Run(), Dynamics.c.Rt:181-210, with the Zou-He conditions of :98-179.
// Direction sets: X<0 {2,8,10,12,14}, X==0 {0,3,4,5,6,15,16,17,18},
// X>0 {1,7,9,11,13}; summed in ascending index order.
template <class R>
__device__ __forceinline__ void d3q19_inlet(R* f, R rho, R ux) {
    R Nxy = ((f[3] + f[15] + f[17]) - (f[4] + f[16] + f[18]))/R(2.);
    R Nxz = ((f[5] + f[15] + f[16]) - (f[6] + f[17] + f[18]))/R(2.);
    f[1]  = f[2]  + rho * ux / R(3.);
    f[9]  = f[8]  + rho * ux / R(6.) + Nxy;
    f[7]  = f[10] + rho * ux / R(6.) - Nxy;
    f[11] = f[14] + rho * ux / R(6.) - Nxz;
    f[13] = f[12] + rho * ux / R(6.) + Nxz;
}

template <class R>
__device__ __forceinline__ R sum_xm(const R* f) { return f[2] + f[8] + f[10] + f[12] + f[14]; }
template <class R>
__device__ __forceinline__ R sum_xp(const R* f) { return f[1] + f[7] + f[9] + f[11] + f[13]; }
template <class R>
__device__ __forceinline__ R sum_x0(const R* f) {
    return f[0] + f[3] + f[4] + f[5] + f[6] + f[15] + f[16] + f[17] + f[18];
}

// Bounce-back swap of every opposite pair (Dynamics.c.Rt:54-75)
template <class R>
__device__ __forceinline__ void d3q19_bounce(R* f) {
    R t;
    t = f[1]; f[1] = f[2]; f[2] = t;
    t = f[3]; f[3] = f[4]; f[4] = t;
    t = f[5]; f[5] = f[6]; f[6] = t;
    t = f[7]; f[7] = f[10]; f[10] = t;
    t = f[8]; f[8] = f[9]; f[9] = t;
    t = f[11]; f[11] = f[14]; f[14] = t;
    t = f[12]; f[12] = f[13]; f[13] = t;
    t = f[15]; f[15] = f[18]; f[18] = t;
    t = f[16]; f[16] = f[17]; f[17] = t;
}

// MovingWall (d3q19_lid only; the reference has no moving wall): full-way
// bounce-back with Ladd's correction for a wall moving at U (wall density
// 1): the population leaving in direction j gains 6 w_j (e_j . U), i.e.
// (e_j . U)/3 on the axes and (e_j . U)/6 on the diagonals.  With GLOB the
// force of the fluid on the node, momentum exchange over the links whose
// source node is fluid (not Solid, Wall or MovingWall):
// F = sum_i e_i (f_in,i + f_out,opp(i)), goes to MovingWallFx/Fy/Fz and its z
// torque about the axis through ((nx-1)/2, (ny-1)/2) to MovingWallTorque.
__host__ __device__ constexpr int d3q19_opp(int i) {
    return i == 0 ? 0 : (i < 7 ? ((i & 1) ? i + 1 : i - 1) : (i < 11 ? 17 - i : (i < 15 ? 25 - i : 33 - i)));
}

template <class R, bool GLOB>
__device__ __forceinline__ void d3q19_moving_wall(R* f, const double* par, const double* zr, double* g,
                                                  const uint16_t* flags, int x, int y, int z, int nx, int ny,
                                                  int nz) {
    using L = Lat<MODEL_D3Q19>;
    R fin[19];
    if (GLOB) {
#pragma unroll
        for (int i = 0; i < 19; i++) fin[i] = f[i];
    }
    d3q19_bounce(f);
    const R Ux = (R)par[12], Uy = (R)par[13], Uz = (R)par[14];
#pragma unroll
    for (int j = 1; j < 19; j++) {
        const R cu = (R)L::ex(j) * Ux + (R)L::ey(j) * Uy + (R)L::ez(j) * Uz;
        f[j] = f[j] + (j < 7 ? cu / R(3.) : cu / R(6.));
    }
    if (GLOB) {
        double Fx = 0, Fy = 0, Fz = 0;
#pragma unroll
        for (int i = 1; i < 19; i++) {
            int sx = x - L::ex(i), sy = y - L::ey(i), sz = z - L::ez(i);
            sx = sx < 0 ? sx + nx : (sx >= nx ? sx - nx : sx);
            sy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
            sz = sz < 0 ? sz + nz : (sz >= nz ? sz - nz : sz);
            const unsigned b = flags[(unsigned)sx + (unsigned)nx * ((unsigned)sy + (unsigned)ny * (unsigned)sz)] & 0x1Cu;
            if (b != 0x08u && b != 0x0Cu && b != 0x1Cu) {
                const double mom = (double)fin[i] + (double)f[d3q19_opp(i)];
                Fx += mom * L::ex(i); Fy += mom * L::ey(i); Fz += mom * L::ez(i);
            }
        }
        const double tq = (x - 0.5 * (nx - 1)) * Fy - (y - 0.5 * (ny - 1)) * Fx;
        g[24] += Fx; g[28] += zr[24];
        g[25] += Fy; g[28] += zr[25];
        g[26] += Fz; g[28] += zr[26];
        g[27] += tq; g[28] += zr[27];
    }
}

// x, y, z and the flags locate the node for the MovingWall of d3q19_lid;
// DEV: the boundary condition only (the collision follows on deviations)
template <class R, bool GLOB, bool LID, bool DEV = false>
__device__ __forceinline__ void d3q19_node(R* f, unsigned nt, const double* par, const double* zr,
                                           double* g, const uint16_t* flags, int x, int y, int z, int nx, int ny,
                                           int nz, const float* fp = nullptr) {
    const R InletVelocity = (R)par[1], InletDensity = (R)par[2];
    switch (nt & 0x1Cu) {
    case 0x14: {  // WPressureL
        R SF = sum_xm(f);
        SF = sum_x0(f) + 2 * SF;
        R rho = InletDensity;
        R ux = R(1.) - SF / rho;
        if (ux > InletVelocity) {
            ux = InletVelocity;
            rho = SF / (R(1.) - ux);
        }
        d3q19_inlet(f, rho, ux);
        break; }
    case 0x10: {  // WPressure
        R rho = InletDensity;
        R ux = sum_xm(f);
        ux = sum_x0(f) + 2 * ux;
        ux = R(1.) - ux / rho;
        d3q19_inlet(f, rho, ux);
        break; }
    case 0x18: {  // WVelocity
        R ux = InletVelocity;
        R rho = sum_xm(f);
        rho = sum_x0(f) + 2 * rho;
        rho = rho / (R(1.) - ux);
        d3q19_inlet(f, rho, ux);
        break; }
    case 0x04: {  // EPressure
        R rho = R(1.0);
        R ux = sum_xp(f);
        ux = sum_x0(f) + 2 * ux;
        ux = R(-1.) + ux / rho;
        R Nxy = ((f[3] + f[15] + f[17]) - (f[4] + f[16] + f[18]))/2;
        R Nxz = ((f[5] + f[15] + f[16]) - (f[6] + f[17] + f[18]))/2;
        f[2]  = f[1]  - rho * ux / R(3.0);
        f[8]  = f[9]  - rho * ux / R(6.0) - Nxy;
        f[10] = f[7]  - rho * ux / R(6.0) + Nxy;
        f[14] = f[11] - rho * ux / R(6.0) + Nxz;
        f[12] = f[13] - rho * ux / R(6.0) - Nxz;
        break; }
    case 0x0C:    // Wall
    case 0x08:    // Solid
        d3q19_bounce(f);
        break;
    case 0x1C:    // MovingWall
        if constexpr (LID) d3q19_moving_wall<R, GLOB>(f, par, zr, g, flags, x, y, z, nx, ny, nz);
        break;
    default: break;
    }
    if constexpr (!DEV) {
        if ((nt & 0x20u) == 0x20u) {
            if constexpr (sizeof(R) == 8) {
                d3q19_mrt_dense<R, GLOB, LID>(f, nt, par, zr, g);
            } else {
                // float on the deviations f - w: the rounded 1/19 then scales
                // rho - 1 only, and adds no mass each step
                using L = Lat<MODEL_D3Q19>;
#pragma unroll
                for (int i = 0; i < 19; i++) f[i] -= R(L::w(i));
                d3q19_mrt<R, GLOB, LID, true>(f, nt, fp, zr, g);
#pragma unroll
                for (int i = 0; i < 19; i++) f[i] += R(L::w(i));
            }
        }
    }
}

// Node of a shifted storage format, on the deviations d = f - w: boundary
// nodes add w back for their boundary condition, the collision runs on d.
template <class R, bool GLOB, bool LID>
__device__ __forceinline__ void d3q19_node_dev(R* d, unsigned nt, const double* par, const double* zr,
                                               double* g, const uint16_t* flags, int x, int y, int z, int nx,
                                               int ny, int nz, const float* fp) {
    using L = Lat<MODEL_D3Q19>;
    if ((nt & 0x1Cu) != 0u) {
#pragma unroll
        for (int i = 0; i < 19; i++) d[i] += R(L::w(i));
        d3q19_node<R, GLOB, LID, true>(d, nt, par, zr, g, flags, x, y, z, nx, ny, nz);
#pragma unroll
        for (int i = 0; i < 19; i++) d[i] -= R(L::w(i));
    }
    if ((nt & 0x20u) == 0x20u) d3q19_mrt<R, GLOB, LID, true>(d, nt, fp, zr, g);
}

// --------------------------------------------------------------- d3q27 ---
// The moment basis of MRT_eq(U, ortogonal=FALSE) (src/lib/feq.R:15-22,
// 37-84) is every monomial ex^p ey^q ez^r with p, q, r in {0, 1, 2}: the
// tensor product of the 1-D basis {1, e, e^2} on e in {-1, 0, 1}.  Moment
// k = p + 3q + 9r is computed by three passes of the 1-D transform
//   (f-, f0, f+) -> (f- + f0 + f+, f+ - f-, f+ + f-)
// along x, y and z in registers; its inverse is
//   (m0, m1, m2) -> ((m2 - m1)/2, m0 - m2, (m2 + m1)/2).
// The order of moment k is p + q + r.
__host__ __device__ constexpr int d3q27_order(int k) { return k % 3 + k / 3 % 3 + k / 9; }

template <class R>
__device__ __forceinline__ void d3q27_fwd(R* m) {
#pragma unroll
    for (int s = 1; s <= 9; s *= 3)
#pragma unroll
        for (int b = 0; b < 27; b++) {
            if ((b / s) % 3 != 0) continue;
            const R a = m[b], c = m[b + s], d = m[b + 2 * s];
            m[b] = a + c + d;
            m[b + s] = d - a;
            m[b + 2 * s] = d + a;
        }
}

template <class R>
__device__ __forceinline__ void d3q27_inv(R* m) {
#pragma unroll
    for (int s = 1; s <= 9; s *= 3)
#pragma unroll
        for (int b = 0; b < 27; b++) {
            if ((b / s) % 3 != 0) continue;
            const R m0 = m[b], m1 = m[b + s], m2 = m[b + 2 * s];
            m[b] = (m2 - m1) * R(0.5);
            m[b + s] = m0 - m2;
            m[b + 2 * s] = (m2 + m1) * R(0.5);
        }
}

// Equilibrium moment k: rho times the product over the axes of J_a/rho
// (power 1) or J_a^2/rho^2 + 1/3 (power 2), truncated to second order in J
// (MRT_eq, feq.R:40-57).  With n2 axes at power 2:
//   no power-1 axis:  rho/3^n2 + sum over power-2 axes of J_b^2/rho / 3^(n2-1)
//   one (a):          J_a / 3^n2
//   two (a, a'):      J_a J_a' / rho / 3^n2
//   three:            0
template <class R>
__device__ __forceinline__ R d3q27_req(int k, R rho, R Jx, R Jy, R Jz, R ir) {
    const int p[3] = {k % 3, k / 3 % 3, k / 9};
    const R J[3] = {Jx, Jy, Jz};
    int n1 = 0, n2 = 0, a1 = -1, a2 = -1;
    for (int d = 0; d < 3; d++) {
        if (p[d] == 1) { if (a1 < 0) a1 = d; else a2 = d; n1++; }
        if (p[d] == 2) n2++;
    }
    const R c = n2 == 0 ? R(1) : n2 == 1 ? R(1) / R(3) : n2 == 2 ? R(1) / R(9) : R(1) / R(27);
    if (n1 == 0) {
        R s = rho * c;
        for (int d = 0; d < 3; d++)
            if (p[d] == 2) s += J[d] * J[d] * ir * (c * R(3));
        return s;
    }
    if (n1 == 1) return J[a1] * c;
    if (n1 == 2) return J[a1] * J[a2] * ir * c;
    return R(0);
}

template <class L> __host__ __device__ constexpr int q27_e(int i, int ax) {
    return ax == 0 ? L::ex(i) : (ax == 1 ? L::ey(i) : L::ez(i));
}

// Zou-He on a face normal to axis AX (ZouHe(EQ, AX+1, sign, type),
// src/lib/boundary.R:66-109, with the equilibrium of MRT_eq(ortogonal=
// FALSE)): the entering populations (SGN*e_AX > 0) become f_opp + 6 w_i
// (e_i . J), the non-equilibrium bounce-back of the second-order
// equilibrium, with J solved so that the node has density rho and no
// tangential momentum.  With A = sum of the e_AX = 0 populations + 2 * the
// outgoing ones and K_t the tangential momentum of the e_AX = 0 populations:
//   velocity V: rho = A / (1 - SGN V), J_AX = rho V;  pressure P: rho =
//   3P + 1, J_AX = SGN (rho - A);  and J_t = -3 K_t.
template <class R, class L, int AX, int SGN, bool VEL>
__device__ __forceinline__ void q27_zouhe(R* f, R V, R P) {
    R A = 0, K[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 27; i++) {
        const int en = SGN * q27_e<L>(i, AX);
        if (en == 0) {
            A += f[i];
            K[0] += f[i] * R(L::ex(i));
            K[1] += f[i] * R(L::ey(i));
            K[2] += f[i] * R(L::ez(i));
        } else if (en < 0) {
            A += R(2) * f[i];
        }
    }
    R rho, Jn;
    if (VEL) {
        rho = A / (R(1) - R(SGN) * V);
        Jn = rho * V;
    } else {
        rho = P * R(3) + R(1);
        Jn = R(SGN) * (rho - A);
    }
    R J[3];
#pragma unroll
    for (int a = 0; a < 3; a++) J[a] = a == AX ? Jn : R(-3) * K[a];
#pragma unroll
    for (int i = 0; i < 27; i++)
        if (SGN * q27_e<L>(i, AX) > 0)
            f[i] = f[L::mirror(i, 3)] +
                   R(6 * L::w(i)) * (R(L::ex(i)) * J[0] + R(L::ey(i)) * J[1] + R(L::ez(i)) * J[2]);
}

// Symmetry(dir, sign) (src/lib/boundary.R:33-50): the populations with
// sign*e_dir > 0 copy their mirror images in the plane normal to dir (the
// indices are template constants, so f stays in registers)
template <class R, class L, int AX, int SGN, int... I>
__device__ __forceinline__ void q27_symmetry_seq(R* f, std::integer_sequence<int, I...>) {
    ((SGN * q27_e<L>(I, AX) > 0 ? (void)(f[I] = f[L::mirror(I, AX)]) : (void)0), ...);
}
template <class R, class L, int AX, int SGN>
__device__ __forceinline__ void q27_symmetry(R* f) {
    q27_symmetry_seq<R, L, AX, SGN>(f, std::make_integer_sequence<int, 27>{});
}

// Both symmetry planes normal to AX as selects: `minus` copies the e_AX > 0
// populations onto their e_AX < 0 mirrors (Symmetry(AX+1, -1)), `plus` the
// reverse (Symmetry(AX+1, 1)).  Written without branches so that the
// populations stay in registers behind a second NodeType switch.
template <class R, class L, int AX, int... I>
__device__ __forceinline__ void q27_symmetry_sel_seq(R* f, bool minus, bool plus, std::integer_sequence<int, I...>) {
    R g[27];
    ((g[I] = f[I]), ...);
    ((q27_e<L>(I, AX) < 0 ? (void)(f[I] = minus ? g[L::mirror(I, AX)] : g[I])
                          : (q27_e<L>(I, AX) > 0 ? (void)(f[I] = plus ? g[L::mirror(I, AX)] : g[I]) : (void)0)),
     ...);
}
template <class R, class L, int AX>
__device__ __forceinline__ void q27_symmetry_sel(R* f, bool minus, bool plus) {
    q27_symmetry_sel_seq<R, L, AX>(f, minus, plus, std::make_integer_sequence<int, 27>{});
}

// Run() and CollisionMRT of models/flow/d3q27/Dynamics.c.Rt:147-245.  par =
// [omega, Smag, ForceX, ForceY, ForceZ]; zone columns [Velocity, Pressure,
// Turbulence, XFluxInObj, YFluxInObj, ZFluxInObj, XDragForceInObj,
// YDragForceInObj, ZDragForceInObj].
template <class R, bool GLOB>
__device__ __forceinline__ void d3q27_node(R* f, unsigned nt, const double* par, const double* zr, double* g) {
    using L = Lat<MODEL_D3Q27>;
    switch (nt & 0xFu) {
    case 7: q27_zouhe<R, L, 0, 1, false>(f, (R)zr[0], (R)zr[1]); break;   // WPressure
    case 8: q27_zouhe<R, L, 0, 1, true>(f, (R)zr[0], (R)zr[1]); break;    // WVelocity
    case 4: q27_zouhe<R, L, 0, -1, false>(f, (R)zr[0], (R)zr[1]); break;  // EPressure
    case 1: q27_symmetry<R, L, 1, -1>(f); break;                          // NSymmetry
    case 2: q27_symmetry<R, L, 1, 1>(f); break;                           // SSymmetry
    case 3: q27_symmetry<R, L, 2, 1>(f); break;                           // ISymmetry
    case 6: {                                                         // Wall
#pragma unroll
        for (int i = 0; i < 13; i++) { const R t = f[i]; f[i] = f[26 - i]; f[26 - i] = t; }
        if (GLOB) {
            // Ladd's momentum exchange, 2 sum_i f_i e_i after the swap
            R fx = 0, fy = 0, fz = 0;
#pragma unroll
            for (int i = 0; i < 27; i++) {
                fx += f[i] * R(Lat<MODEL_D3Q27>::ex(i));
                fy += f[i] * R(Lat<MODEL_D3Q27>::ey(i));
                fz += f[i] * R(Lat<MODEL_D3Q27>::ez(i));
            }
            g[3] += (double)(R(2) * fx); g[6] += zr[6];
            g[4] += (double)(R(2) * fy); g[6] += zr[7];
            g[5] += (double)(R(2) * fz); g[6] += zr[8];
        }
        break; }
    default: break;
    }
    if ((nt & 0x10u) == 0) return;  // MRT
    R m[27];
#pragma unroll
    for (int i = 0; i < 27; i++) m[i] = f[i];
    d3q27_fwd(m);
    const R rho = m[0], ir = R(1) / rho;
    R Jx = m[1], Jy = m[3], Jz = m[9];
#pragma unroll
    for (int k = 0; k < 27; k++)
        if (d3q27_order(k) >= 2) m[k] -= d3q27_req<R>(k, rho, Jx, Jy, Jz, ir);
    R gamma = R(1) - (R)par[0];
    if (nt & 0x40u) {  // Smagorinsky: Q = sum of the squared non-equilibrium stress
        const R Q = m[2] * m[2] + m[6] * m[6] + m[18] * m[18] +
                    R(2) * (m[4] * m[4] + m[10] * m[10] + m[12] * m[12]);
        const R q = R(18) * sqrt(Q) * (R)par[1];
        const R tau0 = R(1) / (R(1) - gamma);
        R tau = sqrt(tau0 * tau0 + q);
        tau = (tau + tau0) / R(2);
        gamma = R(1) - R(1) / tau;
    }
    R gamma2 = gamma;
    if (nt & 0x20u) {
        // entropic stabiliser: a = ds.P.dh, b = dh.P.dh with ds, dh the order-2
        // and order>2 non-equilibrium moments and P = M^-1 diag(1/w) M^-T,
        // the tensor cube of the 1-D Gram matrix [[3/2,0,-3/2],[0,3,0],[-3/2,0,9/2]]
        R y[27];
#pragma unroll
        for (int k = 0; k < 27; k++) y[k] = d3q27_order(k) > 2 ? m[k] : R(0);
#pragma unroll
        for (int s = 1; s <= 9; s *= 3)
#pragma unroll
            for (int b = 0; b < 27; b++) {
                if ((b / s) % 3 != 0) continue;
                const R u0 = y[b], u1 = y[b + s], u2 = y[b + 2 * s];
                y[b] = R(1.5) * (u0 - u2);
                y[b + s] = R(3) * u1;
                y[b + 2 * s] = R(4.5) * u2 - R(1.5) * u0;
            }
        R a = 0, bb = 0;
#pragma unroll
        for (int k = 0; k < 27; k++) {
            if (d3q27_order(k) == 2) a += m[k] * y[k];
            if (d3q27_order(k) > 2) bb += m[k] * y[k];
        }
        gamma2 = -gamma2 * a / bb;
    }
#pragma unroll
    for (int k = 0; k < 27; k++) {
        if (d3q27_order(k) == 2) m[k] *= gamma;
        if (d3q27_order(k) > 2) m[k] *= gamma2;
    }
    const R Fx = (R)par[2], Fy = (R)par[3], Fz = (R)par[4];
    Jx += Fx; Jy += Fy; Jz += Fz;
    if (GLOB) {
        g[0] += (double)((Jx + Fx / R(2)) / rho); g[6] += zr[3];
        g[1] += (double)((Jy + Fy / R(2)) / rho); g[6] += zr[4];
        g[2] += (double)((Jz + Fz / R(2)) / rho); g[6] += zr[5];
    }
    if ((nt & 0xFu) == 5u) { Jx = 0; Jy = 0; Jz = 0; }  // Solid
#pragma unroll
    for (int k = 0; k < 27; k++)
        if (d3q27_order(k) >= 2) m[k] += d3q27_req<R>(k, rho, Jx, Jy, Jz, ir);
    m[0] = rho; m[1] = Jx; m[3] = Jy; m[9] = Jz;
    d3q27_inv(m);
#pragma unroll
    for (int i = 0; i < 27; i++) f[i] = m[i];
}

// SetEquilibrum(rho, Jx, Jy, Jz) = Req M^-1
template <class R>
__device__ __forceinline__ void d3q27_feq(R* f, R rho, R Jx, R Jy, R Jz) {
    const R ir = R(1) / rho;
#pragma unroll
    for (int k = 0; k < 27; k++) f[k] = d3q27_req<R>(k, rho, Jx, Jy, Jz, ir);
    d3q27_inv(f);
}

// ------------------------------------------------------ d3q27_cumulant ---
// Raw moments in place for the digit order (0, +1, -1) of this lattice:
// (f0, f+, f-) -> (f0 + f+ + f-, f+ - f-, f+ + f-) along x, y, z; moment
// i = px + 3py + 9pz has powers (px, py, pz).  Inverse per line:
// (m0, m1, m2) -> (m0 - m2, (m2 + m1)/2, (m2 - m1)/2).
template <class R>
__device__ __forceinline__ void cum_fwd(R* m) {
#pragma unroll
    for (int s = 1; s <= 9; s *= 3)
#pragma unroll
        for (int b = 0; b < 27; b++) {
            if ((b / s) % 3 != 0) continue;
            const R a = m[b], c = m[b + s], d = m[b + 2 * s];
            m[b] = a + c + d;
            m[b + s] = c - d;
            m[b + 2 * s] = c + d;
        }
}

template <class R>
__device__ __forceinline__ void cum_inv(R* m) {
#pragma unroll
    for (int s = 1; s <= 9; s *= 3)
#pragma unroll
        for (int b = 0; b < 27; b++) {
            if ((b / s) % 3 != 0) continue;
            const R m0 = m[b], m1 = m[b + s], m2 = m[b + 2 * s];
            m[b] = m0 - m2;
            m[b + s] = (m2 + m1) * R(0.5);
            m[b + 2 * s] = (m2 - m1) * R(0.5);
        }
}

// Moments <-> cumulants of the normalised distribution f/m0: with j the
// first axis on which n has a nonzero power and n' = n - e_j,
//   m_n = sum over k <= n' of C(n', k) c_{k+e_j} m_{n'-k},
// the Leibniz rule on the moment generating function M = m0 exp(K); the
// term k = n' is c_n m0.  The cumulants of models/flow/d3q27_cumulant/
// Dynamics.c.Rt:291-318 are this recursion written out.
__host__ __device__ constexpr int cum_first_axis(int i) { return i % 3 ? 0 : (i / 3 % 3 ? 1 : 2); }

// term k of the sum for moment i (all indices compile-time constants)
template <int I, int K>
struct CumTerm {
    static constexpr int j = cum_first_axis(I);
    static constexpr int st = j == 0 ? 1 : (j == 1 ? 3 : 9);
    static constexpr int nx = I % 3 - (j == 0), ny = I / 3 % 3 - (j == 1), nz = I / 9 - (j == 2);
    static constexpr int kx = K % 3, ky = K / 3 % 3, kz = K / 9;
    static constexpr bool on = kx <= nx && ky <= ny && kz <= nz && !(kx == nx && ky == ny && kz == nz);
    static constexpr int bin = (nx == 2 && kx == 1 ? 2 : 1) * (ny == 2 && ky == 1 ? 2 : 1) * (nz == 2 && kz == 1 ? 2 : 1);
    static constexpr int ck = on ? K + st : 0;
    static constexpr int mk = on ? (nx - kx) + 3 * (ny - ky) + 9 * (nz - kz) : 0;
};

template <class R, int I, int... K>
__device__ __forceinline__ R cum_sum(const R* c, const R* m, std::integer_sequence<int, K...>) {
    // fused multiply-adds: the float path of this model is arithmetic-bound
    R s = 0;
    ((CumTerm<I, K>::on
          ? (void)(s = fma(CumTerm<I, K>::bin == 1 ? c[CumTerm<I, K>::ck] : R(CumTerm<I, K>::bin) * c[CumTerm<I, K>::ck],
                           m[CumTerm<I, K>::mk], s))
          : (void)0), ...);
    return s;
}

template <class R, bool TO_CUM, int I>
__device__ __forceinline__ void cum_one(R* m, R* c, R m0, R im0) {
    const R s = cum_sum<R, I>(c, m, std::make_integer_sequence<int, 27>{});
    if constexpr (TO_CUM && !std::is_same<R, double>::value && (I == 1 || I == 3 || I == 9)) {
        // the velocity J/rho in float, corrected to the rounded quotient:
        // J (1/rho) rounds 1/rho once more, and with rho near 1 from below
        // that rounding goes up more often than down, so J/rho rho would
        // gain about 1e-8 of the momentum each step
        const R q = (m[I] - s) * im0;
        c[I] = fma(fma(-q, m0, m[I] - s), im0, q);
    } else if (TO_CUM) {
        c[I] = (m[I] - s) * im0;
    } else {
        m[I] = fma(c[I], m0, s);
    }
}

// one reciprocal of m0 instead of 26 divisions (the float path is
// arithmetic-bound here)
template <class R, bool TO_CUM, int... I>
__device__ __forceinline__ void cum_convert_seq(R* m, R* c, std::integer_sequence<int, I...>) {
    const R m0 = m[0], im0 = R(1) / m0;
    (cum_one<R, TO_CUM, I + 1>(m, c, m0, im0), ...);  // i = 1..26 in order
}

template <class R, bool TO_CUM>
__device__ __forceinline__ void cum_convert(R* m, R* c) {
    cum_convert_seq<R, TO_CUM>(m, c, std::make_integer_sequence<int, 26>{});
}

// Run() = RunBoundaries() then CollisionMRT (Dynamics.c.Rt:226-290,
// 326-390), without the AVG, IB and SMAG build options.  par = [nu,
// nubuffer, GalileanCorrection, ForceX, ForceY, ForceZ, Omega]; zone =
// [Velocity, Pressure, Turbulence, DensityInObj .. ZMomentumInObj].
// BOUNDARY values: Buffer 1, WVelocityTurbulent 2, NVelocity 3, SVelocity 4,
// NPressure 5, SPressure 6, EPressure 7, EVelocity 8, Solid 9, Wall 10,
// WPressure 11, WVelocity 12.
template <class R, bool GLOB>
__device__ __forceinline__ void cum_boundaries(R* f, unsigned nt, const double* zr, double* g) {
    using L = Lat<MODEL_D3Q27_CUMULANT>;
    q27_symmetry_sel<R, L, 1>(f, (nt & 0x3u) == 1u, (nt & 0x3u) == 2u);  // NSymmetry, SSymmetry
    const R V = (R)zr[0], P = (R)zr[1];
    switch ((nt >> 3) & 0xFu) {
    case 2:   // WVelocityTurbulent: no synthetic turbulence modes, so V = Velocity
    case 12: q27_zouhe<R, L, 0, 1, true>(f, V, P); break;    // WVelocity
    case 7: q27_zouhe<R, L, 0, -1, false>(f, V, P); break;   // EPressure
    case 11: q27_zouhe<R, L, 0, 1, false>(f, V, P); break;   // WPressure
    case 6: q27_zouhe<R, L, 1, 1, false>(f, V, P); break;    // SPressure
    case 5: q27_zouhe<R, L, 1, -1, false>(f, V, P); break;   // NPressure
    case 3: q27_zouhe<R, L, 1, -1, true>(f, V, P); break;    // NVelocity
    case 4: q27_zouhe<R, L, 1, 1, true>(f, V, P); break;     // SVelocity
    case 8: q27_zouhe<R, L, 0, -1, true>(f, V, P); break;    // EVelocity
    case 10: {  // Wall: FullBounceBack; on Body nodes the momentum change
        R p1[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < 27; i++) {
            p1[0] += f[i] * R(L::ex(i)); p1[1] += f[i] * R(L::ey(i)); p1[2] += f[i] * R(L::ez(i));
        }
#pragma unroll
        for (int i = 0; i < 27; i++) {
            const int o = L::mirror(i, 3);
            if (o > i) { const R t = f[i]; f[i] = f[o]; f[o] = t; }
        }
        if (GLOB && (nt & 0x4u)) {
            R p2[3] = {0, 0, 0};
#pragma unroll
            for (int i = 0; i < 27; i++) {
                p2[0] += f[i] * R(L::ex(i)); p2[1] += f[i] * R(L::ey(i)); p2[2] += f[i] * R(L::ez(i));
            }
            g[2] += (double)(-(p2[0] - p1[0])); g[9] += zr[5];   // Drag
            g[4] += (double)(-(p2[1] - p1[1])); g[9] += zr[7];   // Lateral
            g[3] += (double)(-(p2[2] - p1[2])); g[9] += zr[6];   // Lift
        }
        break; }
    default: break;
    }
}

// InterpolatedBounceBack of the IB builds (Dynamics.c.Rt:180-222), run
// first in RunBoundaries on nodes of NodeType IB.  For each direction
// i = 1..26 in order whose link to x + e_i is cut at the fraction
// q/CUT_MAX (cut[(i-1)*N] != NO_CUT), with p = 2q/CUT_MAX and fa = f_i
// of the node itself in the snapshot being read (its post-collision value):
//   p > 1:  f_opp(i) = fa/p + old_opp(i) (p-1)/p
//   p <= 1: f_opp(i) = fa p + f_i (1-p)
// (Bouzidi's linear interpolation); a later direction reads what an earlier
// one wrote.  F = sum of (f_opp(i) - fa) e_i is added to Drag, Lateral and
// Lift on Body nodes.  old(i) loads f_i of the node from the read snapshot.
template <class R, bool GLOB, class OLD>
__device__ __forceinline__ void cum_ib(R* f, unsigned nt, const uint16_t* cut, size_t N, OLD old,
                                       const double* zr, double* g) {
    using L = Lat<MODEL_D3Q27_CUMULANT>;
    R F[3] = {0, 0, 0};
#pragma unroll
    for (int i = 1; i < 27; i++) {
        const unsigned q = cut[(size_t)(i - 1) * N];
        if (q != 65535u) {  // NO_CUT
            const R p = (R)((double)q * (2.0 / 65000.0));  // CUT_MAX = 65000
            const R fa = old(i);
            const int b = L::mirror(i, 3);
            if (p > R(1)) f[b] = fa / p + old(b) * (p - R(1)) / p;
            else f[b] = fa * p + f[i] * (R(1) - p);
            const R d = f[b] - fa;
            if (L::ex(i)) F[0] += L::ex(i) > 0 ? d : -d;
            if (L::ey(i)) F[1] += L::ey(i) > 0 ? d : -d;
            if (L::ez(i)) F[2] += L::ez(i) > 0 ? d : -d;
        }
    }
    if (GLOB && (nt & 0x4u)) {  // Body
        g[2] += (double)F[0]; g[9] += zr[5];   // Drag
        g[4] += (double)F[1]; g[9] += zr[7];   // Lateral
        g[3] += (double)F[2]; g[9] += zr[6];   // Lift
    }
}

// SMAG: the d3q27_cumulant_SMAG build (par[7] = Smag).  AVG: the AVG
// builds; av (null: no averages kept) points at the node's first averaged
// density, the k-th at av[k*NN], in the order avgP, varUX, varUY, varUZ,
// varUXUY, varUXUZ, varUYUZ, avgdxu2, avgdyv2, avgdzw2, avgUX, avgUY,
// avgUZ.  On collision nodes they gain dxu^2, dyv^2, dzw^2 of the collision
// (Dynamics.c.Rt:410-414) and P, u, u_i u_j of the post-collision
// populations (getP, getU_: 0 on Wall nodes; Dynamics.c.Rt:507-518).
template <class R, bool GLOB, bool SMAG = false, bool AVG = false>
__device__ __forceinline__ void cum_node(R* f, unsigned nt, const double* par, const double* zr, double* g,
                                         double* av = nullptr, size_t NN = 0) {
    cum_boundaries<R, GLOB>(f, nt, zr, g);
    if ((nt & 0x80u) == 0) return;  // MRT
    const R Fx = (R)par[3], Fy = (R)par[4], Fz = (R)par[5];
    R m[27], c[27];
#pragma unroll
    for (int i = 0; i < 27; i++) m[i] = f[i];
    cum_fwd(m);
    const R rho = m[0];
    if (GLOB) {
        // getU_: 0 on Wall nodes, else (J + F/2)/rho
        const bool wall = ((nt >> 3) & 0xFu) == 10u;
        const R ux = wall ? R(0) : (m[1] + Fx / R(2)) / rho;
        const R uy = wall ? R(0) : (m[3] + Fy / R(2)) / rho;
        const R uz = wall ? R(0) : (m[9] + Fz / R(2)) / rho;
        g[0] += (double)rho; g[9] += zr[3];
        g[1] += (double)ux; g[9] += zr[4];
        g[5] += (double)rho; g[9] += zr[8];
        g[6] += (double)(rho * ux); g[9] += zr[9];
        g[7] += (double)(rho * uy); g[9] += zr[10];
        g[8] += (double)(rho * uz); g[9] += zr[11];
    }
    R w0 = ((nt >> 3) & 0xFu) != 0 ? R(1) / (R(3) * (R)par[1] + R(0.5))
                                    : R(1) / (R(3) * (R)par[0] + R(0.5));
    const R w1 = R(1);
    c[0] = R(1);
    cum_convert<R, true>(m, c);
    if constexpr (SMAG) {
        // Smagorinsky (Dynamics.c.Rt:362-374): tau_t from the deviation of the
        // second cumulants from 1/3; buffer-layer nodes keep nubuffer
        const R third = R(1) / R(3);
        const R q = sqrt((c[2] - third) * (c[2] - third) + (c[6] - third) * (c[6] - third) +
                         (c[18] - third) * (c[18] - third) + R(2) * (c[4] * c[4]) + R(2) * (c[10] * c[10]) +
                         R(2) * (c[12] * c[12]));
        const R tau0 = R(3) * (R)par[0] + R(0.5);
        const R Smag = (R)par[7];
        const R taut = R(0.5) * (sqrt(tau0 * tau0 + R(18) * Smag * Smag * q) - tau0);
        w0 = R(1) / (tau0 + taut);
        if (((nt >> 3) & 0xFu) != 0) w0 = R(1) / (R(3) * (R)par[1] + R(0.5));
    }
    const R h = R(0.5) / rho;
    const R ux = c[1] + Fx * h, uy = c[3] + Fy * h, uz = c[9] + Fz * h;
    // second order: c200 = c[2], c020 = c[6], c002 = c[18], c110 = c[4],
    // c101 = c[10], c011 = c[12]
    const R dxu = -w0 / R(2) * (R(2) * c[2] - c[6] - c[18]) - w1 / R(2) * (c[2] + c[6] + c[18] - R(1));
    const R dyv = dxu + R(3) * w0 / R(2) * (c[2] - c[6]);
    const R dzw = dxu + R(3) * w0 / R(2) * (c[2] - c[18]);
    if constexpr (AVG) {
        if (av != nullptr) {
            av[7 * NN] += (double)(dxu * dxu);
            av[8 * NN] += (double)(dyv * dyv);
            av[9 * NN] += (double)(dzw * dzw);
        }
    }
    const R gcor1 = R(3) * (R(1) - w0 / R(2)) * (ux * ux * dxu - uy * uy * dyv);
    const R gcor2 = R(3) * (R(1) - w0 / R(2)) * (ux * ux * dxu - uz * uz * dzw);
    const R gcor3 = R(3) * (R(1) - w1 / R(2)) * (ux * ux * dxu + uy * uy * dyv + uz * uz * dzw);
    const R GC = (R)par[2];
    const R a = (R(1) - w0) * (c[2] - c[6]) - gcor1 * GC;
    const R b = (R(1) - w0) * (c[2] - c[18]) - gcor2 * GC;
    const R cc = w1 + (R(1) - w1) * (c[2] + c[6] + c[18]) - gcor3 * GC;
    c[1] += Fx; c[3] += Fy; c[9] += Fz;
    const R third = R(1) / R(3);
    c[2] = (a + b + cc) * third;
    c[6] = (cc - R(2) * a + b) * third;
    c[18] = (cc - R(2) * b + a) * third;
    c[4] *= R(1) - w0;
    c[12] *= R(1) - w0;
    c[10] *= R(1) - w0;
    const R om3 = R(1) - (R)par[6];
#pragma unroll
    for (int i = 0; i < 27; i++) {
        const int o = i % 3 + i / 3 % 3 + i / 9;
        if (o == 3) c[i] *= om3;
        if (o > 3) c[i] = R(0);
    }
    cum_convert<R, false>(m, c);
    cum_inv(m);
#pragma unroll
    for (int i = 0; i < 27; i++) f[i] = m[i];
    if constexpr (AVG) {
        if (av != nullptr) {
            using L = Lat<MODEL_D3Q27_CUMULANT>;
            R d = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
            for (int i = 0; i < 27; i++) {
                d += f[i];
                if (L::ex(i)) jx += L::ex(i) > 0 ? f[i] : -f[i];
                if (L::ey(i)) jy += L::ey(i) > 0 ? f[i] : -f[i];
                if (L::ez(i)) jz += L::ez(i) > 0 ? f[i] : -f[i];
            }
            const bool wall = ((nt >> 3) & 0xFu) == 10u;
            const R vx = wall ? R(0) : (jx + Fx / R(2)) / d;
            const R vy = wall ? R(0) : (jy + Fy / R(2)) / d;
            const R vz = wall ? R(0) : (jz + Fz / R(2)) / d;
            av[0] += (double)((d - R(1)) / R(3));
            av[10 * NN] += (double)vx;
            av[11 * NN] += (double)vy;
            av[12 * NN] += (double)vz;
            av[1 * NN] += (double)(vx * vx);
            av[2 * NN] += (double)(vy * vy);
            av[3 * NN] += (double)(vz * vz);
            av[4 * NN] += (double)(vx * vy);
            av[5 * NN] += (double)(vx * vz);
            av[6 * NN] += (double)(vy * vz);
        }
    }
}

// SetEquilibrum = Req M^-1 for this lattice's digit order
template <class R>
__device__ __forceinline__ void cum_feq(R* f, R rho, R Jx, R Jy, R Jz) {
    const R ir = R(1) / rho;
#pragma unroll
    for (int k = 0; k < 27; k++) f[k] = d3q27_req<R>(k, rho, Jx, Jy, Jz, ir);
    cum_inv(f);
}

// ---------------------------------------------------------- reductions ---
// Sums are carried as unevaluated pairs (s, c), s + c, and combined with
// Knuth's TwoSum, so the result is accurate to a few ulps of the exact sum
// whatever the order: block and launch boundaries do not change it beyond
// that, and cancelling sums keep their relative accuracy.
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}
__device__ __forceinline__ void dd_add(double& s, double& c, double s2, double c2) {
    double e;
    two_sum(s, s2, s, e);
    c = c + c2 + e;
}
__device__ __forceinline__ void wave_sum_dd(double& s, double& c) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double s2 = __shfl_xor(s, o, 64), c2 = __shfl_xor(c, o, 64);
        dd_add(s, c, s2, c2);
    }
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Per-block partials, stored Global-major [G][nblocks][2] so the final pass
// reads them contiguously.  A Global that is zero on every lane of a
// wavefront is skipped (most slice Globals are, away from slice nodes); the
// 64 lanes are summed in double, the wavefronts of the block with TwoSum.
template <int G>
__device__ __forceinline__ void block_globals(const double* g, int sum_globals, double* gpart, int slot,
                                              int nslots) {
    __shared__ double lds[16][G];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int i = 0; i < G; i++) {
        double v = g[i];
        if (__ballot(v != 0.0) != 0) v = i < sum_globals ? wave_sum(v) : wave_max(v);
        if (lane == 0) lds[wave][i] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < G) {
        const int i = threadIdx.x;
        double v = lds[0][i], c = 0.0;
        for (int w = 1; w < nw; w++) {
            if (i < sum_globals) dd_add(v, c, lds[w][i], 0.0);
            else v = fmax(v, lds[w][i]);
        }
        const size_t k = ((size_t)i * nslots + slot) * 2;
        gpart[k] = v;
        gpart[k + 1] = c;
    }
    __syncthreads();  // the LDS is reused by the next tile of a persistent block
}

// ------------------------------------------------------------ autosym ---
// SymmetryAccess of the reference's autosym builds of d2q9
// (src/LatticeAccess.inc.cpp.Rt:233-260, 376-470), applied to the pulled
// populations f of a node with NodeType nt.  SYMX is bits 7-8, SYMY bits
// 9-10 (1 = _plus: the plane on the node's + side, 2 = _minus).
//   autosym = 1 (Symmetry, plane through the node): the populations that
//     crossed the plane take their mirror images' values, X first, then Y.
//   autosym = 2 (SymmetryEdge, plane half a node out): a population pulled
//     from beyond the plane is read from its mirror image's field at the
//     mirrored offset -- X first, then Y on the result (loadX_ -> loadY_).
// load(field, ox, oy) reads a field at an offset from the node.
template <int MODEL, class R, class LOAD>
__device__ __forceinline__ void d2q9_autosym(R* f, unsigned nt, LOAD load) {
    const unsigned sx = (nt >> 7) & 3u, sy = (nt >> 9) & 3u;
    if constexpr (autosym_kind<MODEL> == 1) {
        const bool xp = sx == 1u, xm = sx == 2u, yp = sy == 1u, ym = sy == 2u;
        const R a1 = f[1], a3 = f[3], a5 = f[5], a6 = f[6], a7 = f[7], a8 = f[8];
        f[3] = xp ? a1 : a3; f[6] = xp ? a5 : a6; f[7] = xp ? a8 : a7;
        f[1] = xm ? a3 : a1; f[5] = xm ? a6 : a5; f[8] = xm ? a7 : a8;
        const R b2 = f[2], b4 = f[4], b5 = f[5], b6 = f[6], b7 = f[7], b8 = f[8];
        f[4] = yp ? b2 : b4; f[7] = yp ? b6 : b7; f[8] = yp ? b5 : b8;
        f[2] = ym ? b4 : b2; f[5] = ym ? b8 : b5; f[6] = ym ? b7 : b6;
    } else {
        if ((sx | sy) == 0u) return;
        using L = Lat<MODEL_D2Q9_SRT>;
#pragma unroll
        for (int i = 1; i < 9; i++) {
            int fi = i, ox = -L::ex(i), oy = -L::ey(i);
            bool ch = false;
            if ((ox > 0 && sx == 1u) || (ox < 0 && sx == 2u)) {
                fi = d2q9_symx(fi);
                ox = (ox > 0 ? 1 : -1) - ox;
                ch = true;
            }
            if ((oy > 0 && sy == 1u) || (oy < 0 && sy == 2u)) {
                fi = d2q9_symy(fi);
                oy = (oy > 0 ? 1 : -1) - oy;
                ch = true;
            }
            if (ch) f[i] = load(fi, ox, oy);
        }
    }
}

// ---------------------------------------------------------- step kernel ---
// One node of row `row` (y + ny*z) at x = xc: pull with periodic wrap, Run(),
// store; `active` is false for the lanes past the end of the row (they load
// a valid node but store nothing).
// LANE: `row` differs between the lanes of a wave (flat tiles, the sorted
// list), so y and z come from a double estimate of row / ny, corrected to the
// exact quotient, instead of a per-lane integer division.
template <int MODEL, int ST, bool GLOB, int DISPATCH = 0, bool LANE = false>
__device__ __forceinline__ void node_update(const StepArgs& a, int row, int xc, int xm, int xp, bool active,
                                            double* g) {
    using L = Lat<MODEL>;
    using S = Store<ST>;
    using T = typename S::T;
    using R = typename S::R;
    constexpr int Q = L::Q;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    // a 2D lattice has nz = 1: its row is y
    int y, z;
    if constexpr (L::Q == 9) {
        y = row;
        z = 0;
    } else if constexpr (LANE) {
        z = (int)((double)row * a.inv_ny);
        y = row - z * ny;
        if (y < 0) { z--; y += ny; }
        if (y >= ny) { z++; y -= ny; }
    } else {
        y = row % ny;
        z = row / ny;
    }
    const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
    const int zm = z == 0 ? nz - 1 : z - 1, zp = z == nz - 1 ? 0 : z + 1;

    const unsigned n = (unsigned)xc + (unsigned)nx * ((unsigned)y + (unsigned)ny * (unsigned)z);
    const unsigned nt = *(gptr<const uint16_t>)(a.flags + n);
    int zone = (int)(nt >> L::ZSHIFT);
    if (zone >= a.nzones) zone = 0;
    const double* zr = a.zone + (size_t)zone * a.zcols;
    // the bulk pass leaves every node of another class to the list pass,
    // and loads nothing for it
    if constexpr (DISPATCH == 1) active = active && (nt & a.class_mask) == a.bulk_val;

    // plane base pointers are uniform (scalar registers); the per-lane part
    // of every address is a 32-bit byte offset (a plane is at most 4 GiB:
    // check_lattice_size refuses larger ones)
    const char* fin = (const char*)a.fin;
    const size_t pbytes = (size_t)a.pitch * sizeof(T);
    // d3q19 and the 2D models in a shifted format collide on the stored
    // deviations f - w
    constexpr bool DEV = (is_d3q19<MODEL> || L::Q == 9) && S::shifted;
    R f[Q];
    if (DISPATCH != 1 || active) {
        // all Q pulls are issued before any is used: the scheduler barrier
        // keeps the compiler from pairing each load with its conversion and
        // waiting on it alone (which it did in the batch step's flat-tile
        // loop, serializing the pulls).  Not for the 27-velocity models,
        // whose capped steps would then spill.
        T v[Q];
#pragma unroll
        for (int i = 0; i < Q; i++) {
            // source node x - e_i with periodic wrap
            const int sx = L::ex(i) > 0 ? xm : (L::ex(i) < 0 ? xp : xc);
            const int sy = L::ey(i) > 0 ? ym : (L::ey(i) < 0 ? yp : y);
            const int sz = L::ez(i) > 0 ? zm : (L::ez(i) < 0 ? zp : z);
            const unsigned src = (unsigned)sx + (unsigned)nx * ((unsigned)sy + (unsigned)ny * (unsigned)sz);
            v[i] = *(gptr<const T>)(fin + (size_t)i * pbytes + src * (unsigned)sizeof(T));
        }
        if constexpr (Q < 27) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < Q; i++) f[i] = DEV ? S::raw(v[i]) : S::get(v[i], (R)L::w(i));
    }
    if constexpr (autosym_kind<MODEL> != 0) {
        d2q9_autosym<MODEL>(f, nt, [&](int fi, int ox, int oy) {
            const int sxn = ox > 0 ? xp : (ox < 0 ? xm : xc);
            const int syn = oy > 0 ? yp : (oy < 0 ? ym : y);
            const unsigned src = (unsigned)sxn + (unsigned)nx * (unsigned)syn;
            const T v = *(gptr<const T>)(fin + (size_t)fi * pbytes + src * (unsigned)sizeof(T));
            return DEV ? S::raw(v) : S::get(v, (R)L::w(fi));   // a mirror image has the same w
        });
    }

    if (active) {
        if constexpr (DEV && is_d3q19<MODEL>) {
            d3q19_node_dev<R, GLOB, MODEL == MODEL_D3Q19_LID>(f, nt, a.par, zr, g, a.flags, xc, y, z, nx, ny, nz,
                                                              a.fpar);
        } else if constexpr (is_d3q19<MODEL>) {
            d3q19_node<R, GLOB, MODEL == MODEL_D3Q19_LID>(f, nt, a.par, zr, g, a.flags, xc, y, z, nx, ny, nz,
                                                          a.fpar);
        } else if constexpr (is_d2q9m<MODEL> && DEV) {
            d2q9m_node_dev<GLOB, MODEL == MODEL_D2Q9_DRAG, is_d2q9_bc<MODEL>>(
                f, nt, a.par, zr, g, a.extra ? a.extra + n : nullptr, (size_t)nx * ny * nz);
        } else if constexpr (is_d2q9m<MODEL>) {
            d2q9m_node<R, GLOB, MODEL == MODEL_D2Q9_DRAG, is_d2q9_bc<MODEL>>(
                f, nt, a.par, zr, g, a.extra ? a.extra + n : nullptr, (size_t)nx * ny * nz);
        } else if constexpr (MODEL == MODEL_D3Q27) {
            d3q27_node<R, GLOB>(f, nt, a.par, zr, g);
        } else if constexpr (is_cum<MODEL>) {
            // NodeType IB with cuts present; every cumulant build carries
            // the branch, whose mask is 0 in the builds without IB
            if ((nt & a.ib_mask) && a.cuts != nullptr)
                cum_ib<R, GLOB>(f, nt, a.cuts + n, (size_t)nx * ny * nz, [&](int i) {
                    return S::get(*(gptr<const T>)(fin + (size_t)i * pbytes + n * (unsigned)sizeof(T)), (R)L::w(i));
                }, zr, g);
            if constexpr (is_cum_avg<MODEL>)
                cum_node<R, GLOB, is_cum_smag<MODEL>, true>(f, nt, a.par, zr, g, a.extra ? a.extra + n : nullptr,
                                                            (size_t)nx * ny * nz);
            else
                cum_node<R, GLOB, is_cum_smag<MODEL>>(f, nt, a.par, zr, g);
        } else if constexpr (DEV) {
            d2q9_node_dev<MODEL == MODEL_D2Q9_SRT_DRAG, GLOB>(f, nt, (R)a.par[0], (R)a.par[1], zr, g);
        } else {
            d2q9_node<R, MODEL == MODEL_D2Q9_SRT_DRAG, GLOB>(f, nt, (R)a.par[0], (R)a.par[1], zr, g);
        }
        // streaming stores: nothing reads them back this step
        char* fout = (char*)a.fout;
        if constexpr (S::stochastic) {
            // direction 2j rounds with word j, 2j + 1 with word j << 16
            uint32_t w = sr_word0(n, a.sr_key);
#pragma unroll
            for (int i = 0; i < Q; i++) {
                const uint32_t r = (i & 1) ? w << 16 : w;
                __builtin_nontemporal_store(DEV ? S::unraw_sr(f[i], r) : S::put_sr(f[i], (R)L::w(i), r),
                                            (gptr<T>)(fout + (size_t)i * pbytes + n * (unsigned)sizeof(T)));
                if (i & 1) w = xorshift32(w);
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < Q; i++)
            __builtin_nontemporal_store(DEV ? S::unraw(f[i]) : S::put(f[i], (R)L::w(i)),
                                        (gptr<T>)(fout + (size_t)i * pbytes + n * (unsigned)sizeof(T)));
    }
}

// The rows of one tile: x segment `xb` of rows [rowblk*rpb, +rpb) of the
// lattice `a`; with GLOB the tile's Globals go to partial slot `slot` of
// `nslots`.
// node index n -> (row, x) for flat tiles: a double-precision estimate of
// n / nx corrected to the exact quotient
__device__ __forceinline__ void split_node(unsigned n, int nx, double inv_nx, int& row, int& x) {
    int r = (int)((double)n * inv_nx);
    int xx = (int)n - r * nx;
    if (xx < 0) { r--; xx += nx; }
    if (xx >= nx) { r++; xx -= nx; }
    row = r;
    x = xx;
}

template <int MODEL, int ST, bool GLOB, int DISPATCH = 0>
__device__ __forceinline__ void step_tile(const StepArgs& a, int tile, int rpb, int slot, int nslots) {
    using L = Lat<MODEL>;
    if (a.flat) {
        // runs of blockDim consecutive nodes (rows of the lattice are
        // concatenated, no lane is left idle at the end of a row)
        double g[L::G > 0 ? L::G : 1];
        if (GLOB) {
#pragma unroll
            for (int i = 0; i < L::G; i++) g[i] = 0.0;
        }
        const unsigned N = (unsigned)a.nx * (unsigned)a.ny * (unsigned)a.nz;
        for (int r = 0; r < rpb; r++) {
            const unsigned n = ((unsigned)tile * rpb + r) * blockDim.x + threadIdx.x;
            const bool in = n < N;
            int row, x;
            split_node(in ? n : N - 1, a.nx, a.inv_nx, row, x);
            const int xm = x == 0 ? a.nx - 1 : x - 1, xp = x == a.nx - 1 ? 0 : x + 1;
            node_update<MODEL, ST, GLOB, DISPATCH, true>(a, row, x, xm, xp, in, g);
        }
        if (GLOB) {
            block_globals<L::G>(g, L::SUMG, a.gpart, slot, nslots);
        }
        return;
    }
    const int rowblk = tile / a.bpr;
    const int xb = tile - rowblk * a.bpr;
    const int x = xb * blockDim.x + threadIdx.x;
    const bool xin = x < a.nx;
    const int xc = xin ? x : a.nx - 1;
    const int xm = xc == 0 ? a.nx - 1 : xc - 1, xp = xc == a.nx - 1 ? 0 : xc + 1;
    double g[L::G > 0 ? L::G : 1];
    if (GLOB) {
#pragma unroll
        for (int i = 0; i < L::G; i++) g[i] = 0.0;
    }
    const int nrows = a.ny * a.nz;
    for (int row = rowblk * rpb; row < (rowblk + 1) * rpb && row < nrows; row++)
        node_update<MODEL, ST, GLOB, DISPATCH>(a, row, xc, xm, xp, xin, g);
    if (GLOB) {
        block_globals<L::G>(g, L::SUMG, a.gpart, slot, nslots);
    }
}

// One step of one lattice: block = one tile.  A plain step tile is one x
// segment of one (y,z) row; a Globals step tile walks `rpb` consecutive rows
// so that one block reduction covers them all.
template <int MODEL, int ST, bool GLOB>
__global__ void __launch_bounds__(256) k_step(StepArgs a) {
    step_tile<MODEL, ST, GLOB>(a, blockIdx.x, GLOB ? a.rpb : a.rpt, blockIdx.x, gridDim.x);
}

// The same step with a floor of W waves per SIMD, which caps its registers.
// The cumulant collisions compile to 130-135 VGPRs (three waves) uncapped
// and fit 128 (four waves) without spilling; the AVG_SMAG builds would
// spill and keep the plain kernel.  More waves keep more loads in flight.
// (The d3q19 float steps, held from seven waves to eight, ran no faster:
// they already sit at the 19-plane pull ceiling.)
template <int MODEL, int ST, bool GLOB, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W))) k_step_w(StepArgs a) {
    step_tile<MODEL, ST, GLOB>(a, blockIdx.x, GLOB ? a.rpb : a.rpt, blockIdx.x, gridDim.x);
}

// NodeType-sorted dispatch of a plain step.  The bulk pass walks the tiles
// of the plain step and updates only the nodes of the plain collision class;
// the list pass takes one node of the class-sorted list per thread, so a
// wavefront meets at most two classes only where the list changes class.
// W: waves per SIMD the kernel is held to (0: none).
template <int MODEL, int ST, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W > 0 ? W : 1)))
k_step_bulk(StepArgs a) {
    step_tile<MODEL, ST, false, 1>(a, blockIdx.x, a.rpt, 0, 0);
}

template <int MODEL, int ST, int W>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W > 0 ? W : 1)))
k_step_list(StepArgs a) {
    const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in = i < a.list_n;
    const unsigned n = a.list[in ? i : a.list_n - 1];
    int row, x;
    split_node(n, a.nx, a.inv_nx, row, x);
    const int xm = x == 0 ? a.nx - 1 : x - 1, xp = x == a.nx - 1 ? 0 : x + 1;
    double g[1];
    node_update<MODEL, ST, false, 0, true>(a, row, x, xm, xp, in, g);
}

// ------------------------------------------------- temporally blocked step ---
// 2D models whose node update reads nothing but its own node's pulled
// populations, flags and zone (no mirror nodes, no densities updated in
// place): d2q9_SRT, d2q9_SRT_drag, d2q9 and d2q9_drag.
template <int MODEL>
constexpr bool tb_model = MODEL == MODEL_D2Q9_SRT || MODEL == MODEL_D2Q9_SRT_DRAG || MODEL == MODEL_D2Q9 ||
                          MODEL == MODEL_D2Q9_DRAG;

// The stored form of population i, rounded as the single step rounds its
// store; with stochastic rounding direction 2j takes word j of the node's
// stream and 2j + 1 word j << 16 (w advances after each odd direction).
template <class S, class L, bool DEV>
__device__ __forceinline__ typename S::T tb_stored(typename S::R v, int i, uint32_t& w) {
    using R = typename S::R;
    if constexpr (S::stochastic) {
        const uint32_t r = (i & 1) ? w << 16 : w;
        const typename S::T out = DEV ? S::unraw_sr(v, r) : S::put_sr(v, (R)L::w(i), r);
        if (i & 1) w = xorshift32(w);
        return out;
    } else {
        return DEV ? S::unraw(v) : S::put(v, (R)L::w(i));
    }
}

// K = a.tb_k plain steps of a 2D lattice in one launch.  Block b takes the
// TX x TY tile whose interior is the (TX - 2K) x (TY - 2K) nodes from
// ((b % tb_bx)(TX - 2K), (b / tb_bx)(TY - 2K)), K nodes of halo on every
// side (periodic wrap), loads its populations of the current snapshot into
// LDS and steps the tile K times there.  After step s the nodes at least s
// from the tile's edge hold exact values (their pulls reached only nodes
// that held exact values), so the last step writes the interior to the
// next snapshot.  Every pull of a step is read before any node writes its
// own populations (a barrier between), so one LDS copy of the tile serves
// all K steps.  Each value written to LDS is rounded to the storage format
// exactly as the single step rounds its stores, and the node update is the
// single step's own (stochastic rounding included: step s draws the words
// of its own iteration): K launches of k_step and one of this give the
// same bits.  The halo is computed again by the neighbouring blocks; the
// launch saves K - 1 round trips through memory and K - 1 launch
// latencies, which pays where a lattice is too small for one step to fill
// the GPU.
template <int MODEL, int ST, int TX, int TY>
__global__ void __launch_bounds__(TX * TY) k_step_tb(StepArgs a) {
    using L = Lat<MODEL>;
    using S = Store<ST>;
    using T = typename S::T;
    using R = typename S::R;
    constexpr int Q = 9;
    constexpr bool DEV = S::shifted;
    static_assert(tb_model<MODEL> && L::Q == 9, "2D models");
    __shared__ T tile[Q][TY][TX];
    const int K = a.tb_k;
    const int nx = a.nx, ny = a.ny;
    const int bx = blockIdx.x % a.tb_bx, by = blockIdx.x / a.tb_bx;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    // this thread's tile node: unwrapped coordinates (ux, uy), lattice node (x, y)
    const int ux = bx * (TX - 2 * K) - K + tx, uy = by * (TY - 2 * K) - K + ty;
    int x = ux % nx, y = uy % ny;
    if (x < 0) x += nx;
    if (y < 0) y += ny;
    const unsigned n = (unsigned)x + (unsigned)nx * (unsigned)y;
    const unsigned nt = *(gptr<const uint16_t>)(a.flags + n);
    int zone = (int)(nt >> L::ZSHIFT);
    if (zone >= a.nzones) zone = 0;
    const double* zr = a.zone + (size_t)zone * a.zcols;
    const size_t pbytes = (size_t)a.pitch * sizeof(T);
    const char* fin = (const char*)a.fin;
#pragma unroll
    for (int i = 0; i < Q; i++) tile[i][ty][tx] = *(gptr<const T>)(fin + (size_t)i * pbytes + n * (unsigned)sizeof(T));
    // the interior node this thread writes at the end (each lattice node is
    // the interior node of exactly one tile)
    const bool writes = tx >= K && tx < TX - K && ty >= K && ty < TY - K && ux < nx && uy < ny;
    double g[1];
    for (int s = 1; s <= K; s++) {
        __syncthreads();
        R f[Q];
#pragma unroll
        for (int i = 0; i < Q; i++) {
            // the pull of the tile's edge nodes is clamped into the tile:
            // their values are never used
            const int sx = min(max(tx - L::ex(i), 0), TX - 1);
            const int sy = min(max(ty - L::ey(i), 0), TY - 1);
            const T v = tile[i][sy][sx];
            f[i] = DEV ? S::raw(v) : S::get(v, (R)L::w(i));
        }
        __syncthreads();
        if constexpr (is_d2q9m<MODEL> && DEV)
            d2q9m_node_dev<false, MODEL == MODEL_D2Q9_DRAG, false>(f, nt, a.par, zr, g, nullptr, 0);
        else if constexpr (is_d2q9m<MODEL>)
            d2q9m_node<R, false, MODEL == MODEL_D2Q9_DRAG, false>(f, nt, a.par, zr, g, nullptr, 0);
        else if constexpr (DEV)
            d2q9_node_dev<MODEL == MODEL_D2Q9_SRT_DRAG, false>(f, nt, (R)a.par[0], (R)a.par[1], zr, g);
        else
            d2q9_node<R, MODEL == MODEL_D2Q9_SRT_DRAG, false>(f, nt, (R)a.par[0], (R)a.par[1], zr, g);
        // stochastic rounding: the words of this node at the iteration
        // step s computes, as the single step of that iteration draws them
        uint32_t w = 0;
        if constexpr (S::stochastic) w = sr_word0(n, sr_key(a.sr_step + (unsigned)(s - 1), a.sr_seed));
        if (s < K) {
#pragma unroll
            for (int i = 0; i < Q; i++) tile[i][ty][tx] = tb_stored<S, L, DEV>(f[i], i, w);
        } else if (writes) {
            char* fout = (char*)a.fout;
#pragma unroll
            for (int i = 0; i < Q; i++)
                __builtin_nontemporal_store(tb_stored<S, L, DEV>(f[i], i, w),
                                            (gptr<T>)(fout + (size_t)i * pbytes + n * (unsigned)sizeof(T)));
        }
    }
}

// tile of the temporally blocked step: 32 x 32 nodes, one per thread
constexpr int TB_TX = TB_TILE, TB_TY = TB_TILE;

// waves per SIMD the plain step of MODEL in storage ST is held to (0: none)
template <int MODEL, int ST>
constexpr int step_waves() {
    if constexpr (is_cum<MODEL>) return is_cum_avg<MODEL> && is_cum_smag<MODEL> ? 0 : 4;
    return 0;
}

// Tile of virtual block v (v = blockIdx.x + k gridDim.x) in XCD-contiguous
// order.  The hardware deals the blocks of a launch to the eight XCDs in
// turn, block b to XCD b % 8 (and virtual block v too while gridDim.x is a
// multiple of 8), so this order gives XCD k the k-th contiguous run of the
// tile list: the rows a tile pulls from its neighbours are mostly fetched
// through the same L2.  A bijection of [0, n): the first n % 8 XCDs take
// n/8 + 1 tiles, the others n/8.
__device__ __forceinline__ int xcd_tile(int v, int n) {
    const int q = n >> 3, r = n & 7, x = v & 7, i = v >> 3;
    return x * q + (x < r ? x : r) + i;
}

// One step of a batch of lattices (a ragged arena): a persistent grid walks
// the tile list; tile t = {case, tile within the case}.  Each case has its
// own descriptor (sizes, pitch, snapshots, flags, zone table, parameters).
// Bit 0 of `phase` selects which snapshot of every case is read this step;
// bit 1 walks the tiles in XCD-contiguous order (xcd_tile).  Globals
// partials of case c occupy the slots of its tiles, [tile0, tile0 + ntiles).
template <int MODEL, int ST, bool GLOB>
__global__ void __launch_bounds__(256) k_batch_step(const BatchCase* __restrict__ cases,
                                                    const int2* __restrict__ tiles, int ntiles, int phase,
                                                    double* gpart) {
    for (int v = blockIdx.x; v < ntiles; v += gridDim.x) {
        const int t = (phase & 2) ? xcd_tile(v, ntiles) : v;
        const int2 tl = tiles[t];
        StepArgs a = cases[tl.x].args;
        if ((phase ^ cases[tl.x].parity) & 1) {
            const void* tmp = a.fin;
            a.fin = a.fout;
            a.fout = (void*)tmp;
        }
        a.gpart = gpart;
        if constexpr (Store<ST>::stochastic)
            a.sr_key = sr_key(cases[tl.x].sr_iter + ((unsigned)phase >> 2), cases[tl.x].sr_seed);
        step_tile<MODEL, ST, GLOB>(a, tl.y, GLOB ? a.rpb : a.rpt, t, ntiles);
    }
}

// The same batch step held to eight waves per SIMD.  The 2D float steps
// need under 64 VGPRs, but the per-tile case descriptor pushes them past
// 96 SGPRs, which alone limits them to seven waves.
template <int MODEL, int ST, bool GLOB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8)))
k_batch_step_w8(const BatchCase* __restrict__ cases, const int2* __restrict__ tiles, int ntiles, int phase,
                double* gpart) {
    for (int v = blockIdx.x; v < ntiles; v += gridDim.x) {
        const int t = (phase & 2) ? xcd_tile(v, ntiles) : v;
        const int2 tl = tiles[t];
        StepArgs a = cases[tl.x].args;
        if ((phase ^ cases[tl.x].parity) & 1) {
            const void* tmp = a.fin;
            a.fin = a.fout;
            a.fout = (void*)tmp;
        }
        a.gpart = gpart;
        if constexpr (Store<ST>::stochastic)
            a.sr_key = sr_key(cases[tl.x].sr_iter + ((unsigned)phase >> 2), cases[tl.x].sr_seed);
        step_tile<MODEL, ST, GLOB>(a, tl.y, GLOB ? a.rpb : a.rpt, t, ntiles);
    }
}

// Init(): d2q9 SetEquilibrium(Density, Velocity_x, Velocity_y) per zone
// (Dynamics.c:26-38); d3q19 SetEquilibrum(1, InletVelocity, 0, 0)
// (Dynamics.c.Rt:219-221).
template <int MODEL, int ST>
__global__ void __launch_bounds__(256) k_init(StepArgs a) {
    using L = Lat<MODEL>;
    using S = Store<ST>;
    using T = typename S::T;
    using R = typename S::R;
    constexpr int Q = L::Q;
    const int row = blockIdx.x / a.bpr;
    const int xb = blockIdx.x - row * a.bpr;
    const int y = row % a.ny;
    const int z = row / a.ny;
    const int x = xb * blockDim.x + threadIdx.x;
    if (x >= a.nx) return;
    const unsigned n = (unsigned)x + (unsigned)a.nx * ((unsigned)y + (unsigned)a.ny * (unsigned)z);
    const unsigned nt = a.flags[n];
    int zone = (int)(nt >> L::ZSHIFT);
    if (zone >= a.nzones) zone = 0;
    const double* zr = a.zone + (size_t)zone * a.zcols;
    R f[Q];
    if constexpr (is_d3q19<MODEL>) {
        d3q19_feq<R>(f, R(1.0), (R)a.par[1], R(0.), R(0.));
    } else if constexpr (is_d2q9m<MODEL>) {
        // Init(), models/flow/d2q9/Dynamics.c.Rt:89-102
        R rho = R(1) + (R)zr[2] * R(3);
        R ux = (R)zr[0], uy = (R)zr[1];
        if ((nt & 0xFu) == 10u) { rho = R(1); ux = R(0); uy = R(0); }
        d2q9m_feq<R>(f, rho, ux * rho, uy * rho);
        if constexpr (is_d2q9_bc<MODEL>) {
            // SetEquilibrum of the bc builds: BC = J/rho on BOUNDARY nodes,
            // else 0 (Dynamics.c.Rt:80-87)
            if (a.extra != nullptr) {
                const size_t NN = (size_t)a.nx * a.ny * a.nz;
                const bool bnd = (nt & 0xFu) != 0u;
                a.extra[n] = bnd ? (double)(ux * rho / rho) : 0.0;
                a.extra[NN + n] = bnd ? (double)(uy * rho / rho) : 0.0;
            }
        }
    } else if constexpr (MODEL == MODEL_D3Q27) {
        // Init(), models/flow/d3q27/Dynamics.c.Rt:184-190: SetEquilibrum(1 +
        // 3 Pressure, Velocity + Turbulence*ST.x, ...) with no synthetic
        // turbulence modes (ST = 0)
        d3q27_feq<R>(f, R(1) + (R)zr[1] * R(3), (R)zr[0], R(0), R(0));
    } else if constexpr (is_cum<MODEL>) {
        // Init(), models/flow/d3q27_cumulant/Dynamics.c.Rt:266-275: Wall
        // nodes at rest, the rest at (1 + 3 Pressure, Velocity, 0, 0)
        if (((nt >> 3) & 0xFu) == 10u) cum_feq<R>(f, R(1), R(0), R(0), R(0));
        else cum_feq<R>(f, R(1) + (R)zr[1] * R(3), (R)zr[0], R(0), R(0));
    } else {
        d2q9_feq<R>(f, (R)a.par[1], (R)zr[1], (R)zr[2]);
    }
    T* fout = (T*)a.fout;
#pragma unroll
    for (int i = 0; i < Q; i++) fout[(size_t)i * a.pitch + n] = S::put(f[i], (R)L::w(i));
}

// Quantities on pulled populations (src/LatticeContainer.inc.cpp.Rt:437-473):
// d2q9: U (J/rho + G/2) = 0, Rho = 1; d3q19: P = 0, U (J/rho, 0 on Wall) = 1;
// d3q19_lid reports the wall velocity on MovingWall nodes.
template <int MODEL, int ST>
__global__ void __launch_bounds__(256) k_quantity(StepArgs a, int qid, int x0, int y0, int z0, int rx,
                                                  int ry, int rz, double scale, double* out) {
    using L = Lat<MODEL>;
    using S = Store<ST>;
    using T = typename S::T;
    constexpr int Q = L::Q;
    const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)rx * ry * rz;
    if (tid >= total) return;
    const int lx = (int)(tid % rx);
    const int ly = (int)((tid / rx) % ry);
    const int lz = (int)(tid / ((long long)rx * ry));
    const int x = x0 + lx, y = y0 + ly, z = z0 + lz;
    const int nx = a.nx, ny = a.ny, nz = a.nz;
    const unsigned n = (unsigned)x + (unsigned)nx * ((unsigned)y + (unsigned)ny * (unsigned)z);
    const unsigned nt = a.flags[n];
    const T* fin = (const T*)a.fin;
    double f[Q];
#pragma unroll
    for (int i = 0; i < Q; i++) {
        const int sx = (x - L::ex(i) + nx) % nx;
        const int sy = (y - L::ey(i) + ny) % ny;
        const int sz = (z - L::ez(i) + nz) % nz;
        const unsigned src = (unsigned)sx + (unsigned)nx * ((unsigned)sy + (unsigned)ny * (unsigned)sz);
        f[i] = (double)S::get(fin[(size_t)i * a.pitch + src], (typename S::R)L::w(i));
    }
    if constexpr (autosym_kind<MODEL> != 0) {
        d2q9_autosym<MODEL>(f, nt, [&](int fi, int ox, int oy) {
            const unsigned src = (unsigned)((x + ox + nx) % nx) + (unsigned)nx * (unsigned)((y + oy + ny) % ny);
            return (double)S::get(fin[(size_t)fi * a.pitch + src], (typename S::R)L::w(fi));
        });
    }
    if constexpr (is_d3q19<MODEL>) {
        double rho = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            rho += f[i]; jx += f[i] * L::ex(i); jy += f[i] * L::ey(i); jz += f[i] * L::ez(i);
        }
        if (qid == 0) {
            out[tid] = (rho - 1.0) / 3.0 * scale;
        } else {
            double ux = jx / rho, uy = jy / rho, uz = jz / rho;
            if ((nt & 0x1Cu) == 0x0Cu) ux = uy = uz = 0;
            if (MODEL == MODEL_D3Q19_LID && (nt & 0x1Cu) == 0x1Cu) { ux = a.par[12]; uy = a.par[13]; uz = a.par[14]; }
            out[3 * tid] = ux * scale; out[3 * tid + 1] = uy * scale; out[3 * tid + 2] = uz * scale;
        }
    } else if constexpr (MODEL == MODEL_D3Q27) {
        // getP, getU, getFd (models/flow/d3q27/Dynamics.c.Rt:33-73): P = 0,
        // U = (J + F/2)/rho = 1, Fd = -2 J on Wall nodes (0 elsewhere) = 2
        double d = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
        for (int i = 0; i < Q; i++) {
            d += f[i]; jx += f[i] * L::ex(i); jy += f[i] * L::ey(i); jz += f[i] * L::ez(i);
        }
        if (qid == 0) {
            out[tid] = (d - 1.0) / 3.0 * scale;
        } else if (qid == 1) {
            out[3 * tid] = (jx + a.par[2] * 0.5) / d * scale;
            out[3 * tid + 1] = (jy + a.par[3] * 0.5) / d * scale;
            out[3 * tid + 2] = (jz + a.par[4] * 0.5) / d * scale;
        } else {
            const bool wall = (nt & 0xFu) == 6u;
            out[3 * tid] = wall ? -2.0 * jx * scale : 0.0;
            out[3 * tid + 1] = wall ? -2.0 * jy * scale : 0.0;
            out[3 * tid + 2] = wall ? -2.0 * jz * scale : 0.0;
        }
    } else if constexpr (is_cum<MODEL>) {
        // getP = 0 on the pulled populations; getU = 1 after RunBoundaries()
        // (0 on Wall nodes, else (J + F/2)/rho); getSolid = 2 (1 on Wall)
        // (models/flow/d3q27_cumulant/Dynamics.c.Rt:23-62)
        const bool wall = ((nt >> 3) & 0xFu) == 10u;
        if (qid == 0) {
            double d = 0;
#pragma unroll
            for (int i = 0; i < Q; i++) d += f[i];
            out[tid] = (d - 1.0) / 3.0 * scale;
        } else if (qid == 1) {
            int zone = (int)(nt >> L::ZSHIFT);
            if (zone >= a.nzones) zone = 0;
            if ((nt & a.ib_mask) && a.cuts != nullptr)
                cum_ib<double, false>(f, nt, a.cuts + n, (size_t)nx * ny * nz, [&](int i) {
                    return (double)S::get(fin[(size_t)i * a.pitch + n], (typename S::R)L::w(i));
                }, nullptr, nullptr);
            cum_boundaries<double, false>(f, nt, a.zone + (size_t)zone * a.zcols, nullptr);
            double d = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
            for (int i = 0; i < Q; i++) {
                d += f[i]; jx += f[i] * L::ex(i); jy += f[i] * L::ey(i); jz += f[i] * L::ez(i);
            }
            out[3 * tid] = wall ? 0.0 : (jx + a.par[3] / 2.) / d * scale;
            out[3 * tid + 1] = wall ? 0.0 : (jy + a.par[4] / 2.) / d * scale;
            out[3 * tid + 2] = wall ? 0.0 : (jz + a.par[5] / 2.) / d * scale;
        } else if (qid == 2 || !is_cum_avg<MODEL>) {
            out[tid] = (wall ? 1.0 : 0.0) * scale;
        } else if (a.extra == nullptr) {  // a lattice that keeps no averages
            const bool vec = qid == 4 || qid == 6 || qid == 7;
            for (int k = 0; k < (vec ? 3 : 1); k++) out[(vec ? 3 : 1) * tid + k] = __builtin_nan("");
        } else {
            // the AVG quantities (Dynamics.c.Rt:62-99) over average_iter
            // iterations: KinE = 3, ReStr = 4, Dissipation = 5, avgU = 6,
            // varU = 7, averageP = 8
            const size_t NN = (size_t)nx * ny * nz;
            auto A = [&](int k, int xx, int yy, int zz) {
                xx = (xx + nx) % nx; yy = (yy + ny) % ny; zz = (zz + nz) % nz;
                return a.extra[(size_t)k * NN + (size_t)xx + (size_t)nx * ((size_t)yy + (size_t)ny * zz)];
            };
            const double it = a.avg_n;
            const double mx = A(10, x, y, z) / it, my = A(11, x, y, z) / it, mz = A(12, x, y, z) / it;
            const double vx = A(1, x, y, z) / it - mx * mx, vy = A(2, x, y, z) / it - my * my,
                         vz = A(3, x, y, z) / it - mz * mz;
            if (qid == 3) {
                out[tid] = 0.5 * (vx + vy + vz) * scale;
            } else if (qid == 4) {
                out[3 * tid] = (A(6, x, y, z) / it - my * mz) * scale;
                out[3 * tid + 1] = (A(5, x, y, z) / it - mx * mz) * scale;
                out[3 * tid + 2] = (A(4, x, y, z) / it - mx * my) * scale;
            } else if (qid == 5) {
                // 4*average_iter*average_iter is int arithmetic there: it
                // wraps past 2^31 (average_iter > 23170), as it does here
                const int ni = (int)it;
                const double q4 = (double)(int)(4u * (unsigned)ni * (unsigned)ni);
                const double nu = a.par[0];
                const double gx = A(10, x + 1, y, z) - A(10, x - 1, y, z);
                const double gy = A(11, x, y + 1, z) - A(11, x, y - 1, z);
                const double gz = A(12, x, y, z + 1) - A(12, x, y, z - 1);
                double diss = nu * ((A(7, x, y, z) / it) - gx * gx / q4);
                diss = diss + nu * ((A(8, x, y, z) / it) - gy * gy / q4);
                diss = diss + nu * ((A(9, x, y, z) / it) - gz * gz / q4);
                out[tid] = diss * scale;
            } else if (qid == 6) {
                out[3 * tid] = mx * scale; out[3 * tid + 1] = my * scale; out[3 * tid + 2] = mz * scale;
            } else if (qid == 7) {
                out[3 * tid] = vx * scale; out[3 * tid + 1] = vy * scale; out[3 * tid + 2] = vz * scale;
            } else {
                out[tid] = A(0, x, y, z) / it * scale;
            }
        }
    } else if constexpr (is_d2q9m<MODEL>) {
        // getRho / getU, Dynamics.c.Rt:46-63: Rho = 0, U = J/rho + G/2 = 1
        double d = 0, jx = 0, jy = 0;
#pragma unroll
        for (int i = 0; i < Q; i++) { d += f[i]; jx += f[i] * L::ex(i); jy += f[i] * L::ey(i); }
        if (qid == 0) {
            out[tid] = d * scale;
        } else {
            double ux = jx / d, uy = jy / d;
            if constexpr (is_d2q9_bc<MODEL>) {
                // + BC/2 off the BOUNDARY nodes (Dynamics.c.Rt:49-53)
                if ((nt & 0xFu) == 0u && a.extra != nullptr) {
                    const size_t NN = (size_t)nx * ny * nz;
                    ux += a.extra[n] * 0.5;
                    uy += a.extra[NN + n] * 0.5;
                }
            }
            out[3 * tid] = (ux + a.par[3] * 0.5) * scale;
            out[3 * tid + 1] = (uy + a.par[4] * 0.5) * scale;
            out[3 * tid + 2] = 0.0;
        }
    } else {
        int zone = (int)(nt >> L::ZSHIFT);
        if (zone >= a.nzones) zone = 0;
        const double* zr = a.zone + (size_t)zone * a.zcols;
        double d = f[8]+f[7]+f[6]+f[5]+f[4]+f[3]+f[2]+f[1]+f[0];
        if (qid == 1) {
            out[tid] = d * scale;
        } else {
            out[3 * tid] = (( f[8]-f[7]-f[6]+f[5]-f[3]+f[1] )/d + zr[3]*0.5 ) * scale;
            out[3 * tid + 1] = ((-f[8]-f[7]+f[6]+f[5]-f[4]+f[2] )/d + zr[4]*0.5 ) * scale;
            out[3 * tid + 2] = 0.0;
        }
    }
}

// ------------------------------------------------------ launch templates ---
inline int nblocks_of(const StepArgs& a, int block) { return a.bpr * a.ny * a.nz; }

template <int MODEL, int ST>
hipError_t step_ms(bool glob, const StepArgs& a, int block, hipStream_t s) {
    dim3 grid(step_blocks(a, glob, block)), blk(block);
    if (!glob && a.mode == 3) {
        if constexpr (tb_model<MODEL>) {
            if (a.tb_k < 2 || 2 * a.tb_k >= TB_TX || 2 * a.tb_k >= TB_TY || a.nz != 1) return hipErrorInvalidValue;
            StepArgs b = a;
            b.tb_bx = (a.nx + TB_TX - 2 * a.tb_k - 1) / (TB_TX - 2 * a.tb_k);
            const int by = (a.ny + TB_TY - 2 * a.tb_k - 1) / (TB_TY - 2 * a.tb_k);
            hipLaunchKernelGGL((k_step_tb<MODEL, ST, TB_TX, TB_TY>), dim3(b.tb_bx * by), dim3(TB_TX * TB_TY), 0, s, b);
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    } else if (!glob && a.mode == 1) {
        hipLaunchKernelGGL((k_step_bulk<MODEL, ST, step_waves<MODEL, ST>()>), grid, blk, 0, s, a);
    } else if (!glob && a.mode == 2) {
        if (a.list_n == 0) return hipSuccess;
        hipLaunchKernelGGL((k_step_list<MODEL, ST, step_waves<MODEL, ST>()>),
                           dim3((a.list_n + block - 1) / block), blk, 0, s, a);
    } else if (glob) {
        hipLaunchKernelGGL((k_step<MODEL, ST, true>), grid, blk, 0, s, a);
    } else if constexpr (step_waves<MODEL, ST>() > 0) {
        hipLaunchKernelGGL((k_step_w<MODEL, ST, false, step_waves<MODEL, ST>()>), grid, blk, 0, s, a);
    } else {
        hipLaunchKernelGGL((k_step<MODEL, ST, false>), grid, blk, 0, s, a);
    }
    return hipGetLastError();
}

template <int MODEL>
hipError_t step_m(int st, bool glob, const StepArgs& a, int block, hipStream_t s) {
    switch (st) {
#if TCLBX_HAS_F64
    case ST_F64: return step_ms<MODEL, ST_F64>(glob, a, block, s);
#endif
#if TCLBX_HAS_FLOAT
    case ST_F32: return step_ms<MODEL, ST_F32>(glob, a, block, s);
    case ST_F32S: return step_ms<MODEL, ST_F32S>(glob, a, block, s);
    case ST_F16S:
        return a.sr ? step_ms<MODEL, ST_F16S_SR>(glob, a, block, s) : step_ms<MODEL, ST_F16S>(glob, a, block, s);
    case ST_BF16S:
        return a.sr ? step_ms<MODEL, ST_BF16S_SR>(glob, a, block, s) : step_ms<MODEL, ST_BF16S>(glob, a, block, s);
#endif
    }
    return hipErrorInvalidValue;
}

template <int MODEL, int ST>
hipError_t init_ms(const StepArgs& a, int block, hipStream_t s) {
    hipLaunchKernelGGL((k_init<MODEL, ST>), dim3(nblocks_of(a, block)), dim3(block), 0, s, a);
    return hipGetLastError();
}

template <int MODEL>
hipError_t init_m(int st, const StepArgs& a, int block, hipStream_t s) {
    switch (st) {
#if TCLBX_HAS_F64
    case ST_F64: return init_ms<MODEL, ST_F64>(a, block, s);
#endif
#if TCLBX_HAS_FLOAT
    case ST_F32: return init_ms<MODEL, ST_F32>(a, block, s);
    case ST_F32S: return init_ms<MODEL, ST_F32S>(a, block, s);
    case ST_F16S: return init_ms<MODEL, ST_F16S>(a, block, s);
    case ST_BF16S: return init_ms<MODEL, ST_BF16S>(a, block, s);
#endif
    }
    return hipErrorInvalidValue;
}

template <int MODEL, int ST>
hipError_t quantity_ms(int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry, int rz,
                       double scale, double* out, hipStream_t s) {
    const long long total = (long long)rx * ry * rz;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL((k_quantity<MODEL, ST>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                       a, q, x0, y0, z0, rx, ry, rz, scale, out);
    return hipGetLastError();
}

template <int MODEL>
hipError_t quantity_m(int st, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry, int rz,
                      double scale, double* out, hipStream_t s) {
    switch (st) {
#if TCLBX_HAS_F64
    case ST_F64: return quantity_ms<MODEL, ST_F64>(q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
#endif
#if TCLBX_HAS_FLOAT
    case ST_F32: return quantity_ms<MODEL, ST_F32>(q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case ST_F32S: return quantity_ms<MODEL, ST_F32S>(q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case ST_F16S: return quantity_ms<MODEL, ST_F16S>(q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
    case ST_BF16S: return quantity_ms<MODEL, ST_BF16S>(q, a, x0, y0, z0, rx, ry, rz, scale, out, s);
#endif
    }
    return hipErrorInvalidValue;
}

template <int MODEL, int ST>
hipError_t batch_ms(bool glob, const BatchCase* cases, const int2* tiles, int ntiles, int phase, double* gpart,
                    int grid, int block, hipStream_t s) {
    if (ntiles <= 0) return hipSuccess;
    if (glob) hipLaunchKernelGGL((k_batch_step<MODEL, ST, true>), dim3(grid), dim3(block), 0, s, cases, tiles,
                                 ntiles, phase, gpart);
    else if constexpr (Lat<MODEL>::Q == 9 && ST != ST_F64)
        hipLaunchKernelGGL((k_batch_step_w8<MODEL, ST, false>), dim3(grid), dim3(block), 0, s, cases, tiles, ntiles,
                           phase, gpart);
    else hipLaunchKernelGGL((k_batch_step<MODEL, ST, false>), dim3(grid), dim3(block), 0, s, cases, tiles,
                            ntiles, phase, gpart);
    return hipGetLastError();
}
template <int MODEL>
hipError_t batch_m(int st, bool glob, const BatchCase* cases, const int2* tiles, int ntiles, int phase,
                   double* gpart, int grid, int block, hipStream_t s) {
    switch (st) {
#if TCLBX_HAS_F64
    case ST_F64: return batch_ms<MODEL, ST_F64>(glob, cases, tiles, ntiles, phase, gpart, grid, block, s);
#endif
#if TCLBX_HAS_FLOAT
    case ST_F32: return batch_ms<MODEL, ST_F32>(glob, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case ST_F32S: return batch_ms<MODEL, ST_F32S>(glob, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case ST_F16S: return batch_ms<MODEL, ST_F16S>(glob, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case ST_BF16S: return batch_ms<MODEL, ST_BF16S>(glob, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case ST_F16S_SR: return batch_ms<MODEL, ST_F16S_SR>(glob, cases, tiles, ntiles, phase, gpart, grid, block, s);
    case ST_BF16S_SR: return batch_ms<MODEL, ST_BF16S_SR>(glob, cases, tiles, ntiles, phase, gpart, grid, block, s);
#endif
    }
    return hipErrorInvalidValue;
}

// Dispatch of one family of models (kernels_<family>.hip); hipErrorNotFound
// for a model of another family.  Each family is compiled twice: the _f64
// dispatchers hold the FP64-storage kernels, built with every multiply and
// add rounded separately (the arithmetic the FP64 oracle follows); the _f32
// ones hold the float-collision kernels of the other storage formats, built
// with multiply-adds contracted into FMAs.
#if TCLBX_FP_GROUP == 64
#define TCLBX_FAM(name) name##_f64
#elif TCLBX_FP_GROUP == 32
#define TCLBX_FAM(name) name##_f32
#else
#define TCLBX_FAM(name) name##_all
#endif
hipError_t step_2d_f64(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_2d_f64(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_2d_f64(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_2d_f64(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);
hipError_t step_d3q19_f64(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_d3q19_f64(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_d3q19_f64(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_d3q19_f64(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);
hipError_t step_d3q27_f64(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_d3q27_f64(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_d3q27_f64(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_d3q27_f64(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);
hipError_t step_cum_f64(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_cum_f64(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_cum_f64(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_cum_f64(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);
hipError_t step_cumavg_f64(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_cumavg_f64(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_cumavg_f64(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_cumavg_f64(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);
hipError_t step_2d_f32(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_2d_f32(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_2d_f32(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_2d_f32(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);
hipError_t step_d3q19_f32(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_d3q19_f32(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_d3q19_f32(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_d3q19_f32(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);
hipError_t step_d3q27_f32(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_d3q27_f32(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_d3q27_f32(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_d3q27_f32(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);
hipError_t step_cum_f32(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_cum_f32(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_cum_f32(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_cum_f32(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);
hipError_t step_cumavg_f32(int model, int storage, bool globals, const StepArgs& a, int block, hipStream_t s);
hipError_t init_cumavg_f32(int model, int storage, const StepArgs& a, int block, hipStream_t s);
hipError_t batch_cumavg_f32(int model, int storage, bool globals, const BatchCase* cases, const int2* tiles, int ntiles,
                    int phase, double* gpart, int grid, int block, hipStream_t s);
hipError_t quantity_cumavg_f32(int model, int storage, int q, const StepArgs& a, int x0, int y0, int z0, int rx, int ry,
                       int rz, double scale, double* out, hipStream_t s);

}  // namespace tclbx
