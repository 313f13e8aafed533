// storage.h -- distribution storage formats in HBM and their conversions.
//
// A stored value is (f_i - shift_i) in the storage type; reading adds the
// shift back (src/LatticeAccess.inc.cpp.Rt:14-35, which TCLB applies only to
// models that declare shift=w_i; here every reduced format is shifted by the
// lattice weight w_i, i.e. f_i - w_i*rho0 with rho0 = 1).  F64 keeps double
// arithmetic; the reduced formats compute in float.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "model.h"

namespace tclbx {

// ---- stochastic rounding of the 16-bit formats (an option of a lattice) --
// A value x between the neighbouring representable values lo and hi (|lo| <=
// |x| < |hi|) is stored as hi with probability (|x| - |lo|)/(|hi| - |lo|),
// so the stored value's expectation is x: rounding errors do not add up
// into a drift over many steps.  gfx950 rounds this way in one instruction
// (v_cvt_sr_f16_f32, v_cvt_sr_bf16_f32) given 32 random bits r:
// - binary16: the float's significand aligned to binary16's spacing at |x|
//   (normal or subnormal), the 13 bits below the kept ones (lower bits
//   dropped), plus the top 13 bits of r, carried into the kept bits and
//   truncated; so below binary16's spacing/2^13 (|x| < 2^-37) x becomes 0;
// - bfloat16: the low 16 bits of x plus the top 16 bits of r, carried and
//   truncated.
// The random words of a node's stores come from MurmurHash3's 32-bit
// finalizer (fmix32) of (node, step, seed), forced odd, and Marsaglia's
// xorshift32 after it; direction 2j takes word j, direction 2j + 1 word j
// shifted left by 16, so the two use disjoint bits.  A run is reproducible.
__host__ __device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
// the key of a step's random words (the same for every node; the host
// computes it for the launch)
__host__ __device__ __forceinline__ uint32_t sr_key(uint32_t step, uint32_t seed) {
    return fmix32(step * 0x85ebca77u + seed);
}
// the first random word of `node`'s stores in the step of `key`
__device__ __forceinline__ uint32_t sr_word0(uint32_t node, uint32_t key) {
    return fmix32(node * 0x9e3779b1u ^ key) | 1u;
}
__device__ __forceinline__ uint32_t xorshift32(uint32_t x) {
    x ^= x << 13; x ^= x >> 17; x ^= x << 5;
    return x;
}
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef __bf16 bf162_t __attribute__((ext_vector_type(2)));
__device__ __forceinline__ _Float16 f16_sr(float x, uint32_t r) {
    return __builtin_amdgcn_cvt_sr_f16_f32(half2_t{}, x, r, false)[0];
}
__device__ __forceinline__ __bf16 bf16_sr(float x, uint32_t r) {
    return __builtin_amdgcn_cvt_sr_bf16_f32(bf162_t{}, x, r, false)[0];
}

template <int ST> struct Store;

template <> struct Store<ST_F64> {
    using T = double;
    using R = double;
    static constexpr bool shifted = false;
    static constexpr bool stochastic = false;
    __device__ __forceinline__ static R get(T v, R) { return v; }
    __device__ __forceinline__ static T put(R v, R) { return v; }
    __device__ __forceinline__ static R raw(T v) { return v; }
    __device__ __forceinline__ static T unraw(R v) { return v; }
};

template <> struct Store<ST_F32> {
    using T = float;
    using R = float;
    static constexpr bool shifted = false;
    static constexpr bool stochastic = false;
    __device__ __forceinline__ static R get(T v, R) { return v; }
    __device__ __forceinline__ static T put(R v, R) { return v; }
    __device__ __forceinline__ static R raw(T v) { return v; }
    __device__ __forceinline__ static T unraw(R v) { return v; }
};

template <> struct Store<ST_F32S> {
    using T = float;
    using R = float;
    static constexpr bool shifted = true;
    static constexpr bool stochastic = false;
    __device__ __forceinline__ static R get(T v, R s) { return v + s; }
    __device__ __forceinline__ static T put(R v, R s) { return v - s; }
    // the stored value itself (f - w)
    __device__ __forceinline__ static R raw(T v) { return v; }
    __device__ __forceinline__ static T unraw(R v) { return v; }
};

template <> struct Store<ST_F16S> {
    using T = _Float16;
    using R = float;
    static constexpr bool shifted = true;
    __device__ __forceinline__ static R get(T v, R s) { return (float)v + s; }
    // float -> half conversion rounds to nearest even (v_cvt_f16_f32)
    __device__ __forceinline__ static T put(R v, R s) { return (_Float16)(v - s); }
    __device__ __forceinline__ static R raw(T v) { return (float)v; }
    __device__ __forceinline__ static T unraw(R v) { return (_Float16)v; }
    // the same stores, rounded stochastically with random bits r (used by
    // the ST_F16S_SR instantiation)
    static constexpr bool stochastic = false;
    __device__ __forceinline__ static T put_sr(R v, R s, uint32_t r) { return f16_sr(v - s, r); }
    __device__ __forceinline__ static T unraw_sr(R v, uint32_t r) { return f16_sr(v, r); }
};

template <> struct Store<ST_BF16S> {
    using T = __bf16;
    using R = float;
    static constexpr bool shifted = true;
    __device__ __forceinline__ static R get(T v, R s) { return (float)v + s; }
    // float -> bfloat16 rounds to nearest even (v_cvt_pk_bf16_f32)
    __device__ __forceinline__ static T put(R v, R s) { return (__bf16)(v - s); }
    __device__ __forceinline__ static R raw(T v) { return (float)v; }
    __device__ __forceinline__ static T unraw(R v) { return (__bf16)v; }
    static constexpr bool stochastic = false;
    __device__ __forceinline__ static T put_sr(R v, R s, uint32_t r) { return bf16_sr(v - s, r); }
    __device__ __forceinline__ static T unraw_sr(R v, uint32_t r) { return bf16_sr(v, r); }
};

// the step kernels of a lattice with stochastic rounding on
template <> struct Store<ST_F16S_SR> : Store<ST_F16S> { static constexpr bool stochastic = true; };
template <> struct Store<ST_BF16S_SR> : Store<ST_BF16S> { static constexpr bool stochastic = true; };

}  // namespace tclbx
