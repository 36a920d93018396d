// Lattice constants and storage-precision traits shared by every plbpd HIP
// kernel.  Directions, weights and the reverse map follow LBModel.cpp:23-165
// (D2Q9 lives in the x-z plane with c_y == 0, as in the reference).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plb {

template <int Q>
struct Lat;

template <>
struct Lat<9> {
    static constexpr int q = 9;
    __host__ __device__ static constexpr int cx(int k) {
        constexpr int v[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
        return v[k];
    }
    __host__ __device__ static constexpr int cy(int k) { return (void)k, 0; }
    __host__ __device__ static constexpr int cz(int k) {
        constexpr int v[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
        return v[k];
    }
    __host__ __device__ static constexpr double w(int k) {
        return k == 0 ? 4.0 / 9.0 : (k < 5 ? 1.0 / 9.0 : 1.0 / 36.0);
    }
};

template <>
struct Lat<19> {
    static constexpr int q = 19;
    __host__ __device__ static constexpr int cx(int k) {
        constexpr int v[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
        return v[k];
    }
    __host__ __device__ static constexpr int cy(int k) {
        constexpr int v[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
        return v[k];
    }
    __host__ __device__ static constexpr int cz(int k) {
        constexpr int v[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
        return v[k];
    }
    __host__ __device__ static constexpr double w(int k) {
        return k == 0 ? 1.0 / 3.0 : (k < 7 ? 1.0 / 18.0 : 1.0 / 36.0);
    }
};

template <>
struct Lat<27> {
    static constexpr int q = 27;
    __host__ __device__ static constexpr int cx(int k) {
        constexpr int v[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1,
                               0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
        return v[k];
    }
    __host__ __device__ static constexpr int cy(int k) {
        constexpr int v[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0,
                               1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
        return v[k];
    }
    __host__ __device__ static constexpr int cz(int k) {
        constexpr int v[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1,
                               1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
        return v[k];
    }
    __host__ __device__ static constexpr double w(int k) {
        return k == 0 ? 8.0 / 27.0 : (k < 7 ? 2.0 / 27.0 : (k < 19 ? 1.0 / 54.0 : 1.0 / 216.0));
    }
};

// Every lattice here pairs direction k with its reverse as (1,2), (3,4), ...
template <int Q>
__host__ __device__ constexpr int rev(int k) {
    return k == 0 ? 0 : ((k & 1) ? k + 1 : k - 1);
}

// ------------------------------------------------------ storage variants --
// Variant ids match plb_variant in include/plbpd_hip.h.
//   load/store       stored value <-> f (or the deviation g = f - w_k)
//   load_raw/...     stored value <-> kScale * g, the form the fused and
//                    batched steps compute in (kScale is a power of two, so
//                    scaling is exact and commutes with rounding: the step's
//                    results are bit for bit those of computing on g)
enum Variant { kF64 = 0, kF32 = 1, kF64A32 = 2, kF32S = 3, kF16S = 4 };

template <int V>
struct Prec;

template <>
struct Prec<kF64> {
    using S = double;
    using T = double;
    static constexpr bool shifted = false;
    __device__ static T load(S s) { return s; }
    __device__ static S store(T t) { return t; }
    static constexpr T kScale = 1.0;
    __device__ static T load_raw(S s) { return s; }
    __device__ static S store_raw(T t) { return t; }
};
template <>
struct Prec<kF64A32> {
    using S = double;
    using T = float;
    static constexpr bool shifted = false;
    __device__ static T load(S s) { return (float)s; }
    __device__ static S store(T t) { return (double)t; }
    static constexpr T kScale = 1.0f;
    __device__ static T load_raw(S s) { return (float)s; }
    __device__ static S store_raw(T t) { return (double)t; }
};
template <>
struct Prec<kF32> {
    using S = float;
    using T = float;
    static constexpr bool shifted = false;
    __device__ static T load(S s) { return s; }
    __device__ static S store(T t) { return t; }
    static constexpr T kScale = 1.0f;
    __device__ static T load_raw(S s) { return s; }
    __device__ static S store_raw(T t) { return t; }
};
template <>
struct Prec<kF32S> {
    using S = float;
    using T = float;
    static constexpr bool shifted = true;
    __device__ static T load(S s) { return s; }
    __device__ static S store(T t) { return t; }
    static constexpr T kScale = 1.0f;
    __device__ static T load_raw(S s) { return s; }
    __device__ static S store_raw(T t) { return t; }
};
// Shifted deviations g = f - w_k stored as IEEE half of g * 2^15: the scale is
// a power of two (exact), and maps |g| from about 2e-9 up to 1 into half's
// normal range (6.1e-5 .. 65504).
template <>
struct Prec<kF16S> {
    using S = _Float16;
    using T = float;
    static constexpr bool shifted = true;
    __device__ static T load(S s) { return (float)s * 0x1.0p-15f; }
    __device__ static S store(T t) { return (_Float16)(t * 0x1.0p15f); }
    static constexpr T kScale = 0x1.0p15f;
    __device__ static T load_raw(S s) { return (float)s; }
    __device__ static S store_raw(T t) { return (_Float16)t; }
};

inline size_t storage_bytes(int v) {
    switch (v) {
        case kF64: case kF64A32: return 8;
        case kF32: case kF32S: return 4;
        case kF16S: return 2;
        default: return 0;
    }
}

}  // namespace plb
