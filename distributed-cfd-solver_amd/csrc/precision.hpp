// Storage types of the precision forms: FP64, FP32 and bfloat16 block values.
//
// bfloat16 is the upper half of an IEEE binary32; float -> bf16 rounds to
// nearest, ties to even (NaN stays a quiet NaN).  A double is first rounded to
// float, then to bf16.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

namespace cfdg {

struct bf16 {
    std::uint16_t u;
};

__host__ __device__ __forceinline__ std::uint32_t f32_bits(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __float_as_uint(f);
#else
    std::uint32_t u;
    std::memcpy(&u, &f, sizeof(u));
    return u;
#endif
}

__host__ __device__ __forceinline__ float f32_from_bits(std::uint32_t u) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __uint_as_float(u);
#else
    float f;
    std::memcpy(&f, &u, sizeof(f));
    return f;
#endif
}

__host__ __device__ __forceinline__ bf16 bf16_from_float(float f) {
    const std::uint32_t x = f32_bits(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return bf16{static_cast<std::uint16_t>((x >> 16) | 0x0040u)};
    const std::uint32_t r = x + 0x7fffu + ((x >> 16) & 1u);
    return bf16{static_cast<std::uint16_t>(r >> 16)};
}

__host__ __device__ __forceinline__ float bf16_to_float(bf16 h) {
    return f32_from_bits(static_cast<std::uint32_t>(h.u) << 16);
}

// Layout of one stored 5x5 block: entry (a, b) at a * row + b of a block of
// `block` values.  bfloat16 rows are padded to 6 values (12 bytes, 4-byte
// aligned) in 64-byte blocks, so that a block row is one 12-byte load rather
// than five 2-byte loads.
template <class TS>
struct BlockLayout {
    static constexpr int block = 25, row = 5;
};
template <>
struct BlockLayout<bf16> {
    static constexpr int block = 32, row = 6;
};

// Conversions between the storage and arithmetic types.
template <class To>
__host__ __device__ __forceinline__ To cvt(double v);
template <>
__host__ __device__ __forceinline__ double cvt<double>(double v) { return v; }
template <>
__host__ __device__ __forceinline__ float cvt<float>(double v) { return static_cast<float>(v); }
template <>
__host__ __device__ __forceinline__ bf16 cvt<bf16>(double v) { return bf16_from_float(static_cast<float>(v)); }

__host__ __device__ __forceinline__ double to_f64(double v) { return v; }
__host__ __device__ __forceinline__ double to_f64(float v) { return static_cast<double>(v); }
__host__ __device__ __forceinline__ double to_f64(bf16 v) { return static_cast<double>(bf16_to_float(v)); }
__host__ __device__ __forceinline__ float to_f32(double v) { return static_cast<float>(v); }
__host__ __device__ __forceinline__ float to_f32(float v) { return v; }
__host__ __device__ __forceinline__ float to_f32(bf16 v) { return bf16_to_float(v); }

// Load a stored value into the arithmetic type C.
template <class C, class S>
__host__ __device__ __forceinline__ C load_as(S v);
template <>
__host__ __device__ __forceinline__ double load_as<double, double>(double v) { return v; }
template <>
__host__ __device__ __forceinline__ double load_as<double, float>(float v) { return to_f64(v); }
template <>
__host__ __device__ __forceinline__ double load_as<double, bf16>(bf16 v) { return to_f64(v); }
template <>
__host__ __device__ __forceinline__ float load_as<float, float>(float v) { return v; }
template <>
__host__ __device__ __forceinline__ float load_as<float, bf16>(bf16 v) { return to_f32(v); }
template <>
__host__ __device__ __forceinline__ float load_as<float, double>(double v) { return to_f32(v); }

}  // namespace cfdg
