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

template <int ST> struct Store;

template <> struct Store<ST_F64> {
    using T = double;
    using R = double;
    static constexpr bool shifted = false;
    __device__ __forceinline__ static R get(T v, R) { return v; }
    __device__ __forceinline__ static T put(R v, R) { return v; }
    __device__ __forceinline__ static R raw(T v) { return v; }
    __device__ __forceinline__ static T unraw(R v) { return v; }
};

template <> struct Store<ST_F32> {
    using T = float;
    using R = float;
    static constexpr bool shifted = false;
    __device__ __forceinline__ static R get(T v, R) { return v; }
    __device__ __forceinline__ static T put(R v, R) { return v; }
    __device__ __forceinline__ static R raw(T v) { return v; }
    __device__ __forceinline__ static T unraw(R v) { return v; }
};

template <> struct Store<ST_F32S> {
    using T = float;
    using R = float;
    static constexpr bool shifted = true;
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
};

}  // namespace tclbx
