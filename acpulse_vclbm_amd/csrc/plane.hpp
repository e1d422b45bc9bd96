// Addressing of the structure-of-arrays population planes (plane q of array F
// starts at F + q*N) and the compile-time shape of the bulk kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

// Workgroup width along x and the minimum number of waves per SIMD the
// register allocation must allow.
#ifndef ACP_BULK_BLOCK
#define ACP_BULK_BLOCK 512
#endif
#ifndef ACP_BULK_WAVES
#define ACP_BULK_WAVES 1
#endif

namespace acp {

constexpr int kBulkBlock = ACP_BULK_BLOCK;

// NARROW: a plane spans at most 4 GiB, so every access is a uniform (scalar)
// plane base plus a 32-bit byte offset, which global loads and stores take
// directly with an SGPR base; this keeps 64-bit address arithmetic out of the
// vector registers.  Otherwise 64-bit element indices.
template <bool NARROW>
struct Plane;

template <>
struct Plane<true> {
    using idx = uint32_t;
    template <typename T>
    static __device__ __forceinline__ T& at(T* base, int q, idx l, int64_t N) {
        using Byte = std::conditional_t<std::is_const_v<T>, const char, char>;
        return *reinterpret_cast<T*>(reinterpret_cast<Byte*>(base + q * N) + (uint32_t)(l * (uint32_t)sizeof(T)));
    }
};

template <>
struct Plane<false> {
    using idx = int64_t;
    template <typename T>
    static __device__ __forceinline__ T& at(T* base, int q, idx l, int64_t N) {
        return base[q * N + l];
    }
};

// true when the narrow form can address every plane of N elements of `bytes`
inline bool narrow_planes(int64_t N, int bytes) { return N * (int64_t)bytes <= ((int64_t)1 << 32); }

}  // namespace acp
