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
// into a drift over many steps.  The random bits come from MurmurHash3's
// 32-bit finalizer (fmix32) over (node, direction, step, seed), so a run is
// reproducible.
__device__ __forceinline__ uint32_t fmix32(uint32_t h) {
    h ^= h >> 16; h *= 0x85ebca6bu;
    h ^= h >> 13; h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}
__device__ __forceinline__ uint32_t sr_bits(uint32_t node, uint32_t q, uint32_t step, uint32_t seed) {
    return fmix32(node * 0x9e3779b1u ^ fmix32(step * 0x85ebca77u + q * 0xc2b2ae3du + seed));
}
// float -> half: k = the bits of x's float significand below binary16's
// spacing at |x| (13 where binary16 is normal, -1 - e for a float exponent
// e in [-24, -15], where it is subnormal); k random bits added to those
// bits carry into the kept ones with probability (low bits)/2^k, and the
// truncated value converts to binary16 exactly.  Below binary16's smallest
// subnormal (e <= -25) x becomes +-2^-24 with probability |x| 2^24, else
// +-0, drawn from the top 24 random bits.
__device__ __forceinline__ _Float16 f16_sr(float x, uint32_t r) {
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    const uint32_t a = u & 0x7fffffffu;
    if (a >= 0x7f800000u) return (_Float16)x;
    const int e = (int)(a >> 23) - 127;
    if (e <= -25) {
        const bool up = (float)(r >> 8) * 5.9604644775390625e-08f < __builtin_bit_cast(float, a) * 16777216.0f;
        return __builtin_bit_cast(_Float16, (uint16_t)(((u >> 16) & 0x8000u) | (up ? 1u : 0u)));
    }
    const uint32_t k = e >= -14 ? 13u : (uint32_t)(-1 - e);
    const uint32_t mask = (1u << k) - 1u;
    return (_Float16)__builtin_bit_cast(float, (u + (r & mask)) & ~mask);
}
// float -> bfloat16: the low 16 bits plus 16 random bits carry into the
// kept half with probability (low bits)/2^16
__device__ __forceinline__ __bf16 bf16_sr(float x, uint32_t r) {
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7fffffffu) >= 0x7f800000u) return (__bf16)x;
    return __builtin_bit_cast(__bf16, (uint16_t)((u + (r & 0xffffu)) >> 16));
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
    // the same stores, rounded stochastically with random bits r
    static constexpr bool stochastic = true;
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
    static constexpr bool stochastic = true;
    __device__ __forceinline__ static T put_sr(R v, R s, uint32_t r) { return bf16_sr(v - s, r); }
    __device__ __forceinline__ static T unraw_sr(R v, uint32_t r) { return bf16_sr(v, r); }
};

}  // namespace tclbx
