// Addressing of the population arrays and the compile-time shape of the bulk
// kernels.
//
// Population layout.  The 19 populations of a node live in blocks of B
// consecutive nodes: block b holds, plane after plane, population q of nodes
// [bB, (b+1)B), so element (q, l) sits at
//     (l >> log2 B) * 19B + q * B + (l & (B - 1)).
// One block spanning the whole grid is the plain structure of arrays (plane q
// starts at q*N).  Blocks of 8192 nodes are 64 KiB per plane segment in float64.
// They keep the node update's 19 read and 19 write streams inside a few
// contiguous regions of the array, which the memory system serves measurably
// faster than 19 streams a whole plane (N elements) apart. The last block is padded
// to B nodes; the padding is zero and never written.
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
#ifndef ACP_HRR_WAVES
#define ACP_HRR_WAVES 4
#endif

namespace acp {

constexpr int kBulkBlock = ACP_BULK_BLOCK;

struct PopLayout {
    int64_t N = 0;      // nodes
    int64_t mask = -1;  // nodes per block - 1; all bits for a single block
    int64_t plane = 0;  // distance between populations q and q + 1 of a node
    int64_t block = 0;  // distance between consecutive blocks (0: a single block)
    int64_t elems = 0;  // elements of one population array, padding included
    int32_t shift = 0;  // log2(nodes per block)
};

__host__ __device__ __forceinline__ int64_t pidx(const PopLayout& L, int q, int64_t l) {
    return (l >> L.shift) * L.block + (int64_t)q * L.plane + (l & L.mask);
}

// Layout of N nodes in blocks of `block_nodes` (a power of two; 0, or a grid
// of at most one block, gives the plain structure of arrays).
inline PopLayout make_layout(int64_t N, int64_t block_nodes) {
    PopLayout L;
    L.N = N;
    if (block_nodes <= 0 || N <= block_nodes) {
        L.mask = -1;
        L.plane = N;
        L.block = 0;
        L.shift = 0;
        L.elems = 19 * N;
        return L;
    }
    int s = 0;
    while (((int64_t)1 << s) < block_nodes) ++s;
    L.shift = s;
    L.mask = ((int64_t)1 << s) - 1;
    L.plane = (int64_t)1 << s;
    L.block = 19 * L.plane;
    L.elems = ((N + L.mask) >> s) * L.block;
    return L;
}

// NARROW: the per-node part of every population offset (the node's block and
// its position in the block, or its index in a whole plane) and every macro
// plane stay below 4 GiB, so each access is a uniform base (array plus plane
// offset) and a 32-bit byte offset, which global loads and stores take
// directly with an SGPR base; this keeps 64-bit address arithmetic out of the
// vector registers.  Otherwise 64-bit element indices.
template <bool NARROW>
struct Plane;

template <>
struct Plane<true> {
    using idx = uint32_t;
    // population q of node l in layout L: the plane offset q * plane is uniform
    // (64-bit, scalar), the node's block and position in it a 32-bit offset
    template <typename T>
    static __device__ __forceinline__ T& at(T* base, int q, idx l, const PopLayout& L) {
        using Byte = std::conditional_t<std::is_const_v<T>, const char, char>;
        const uint32_t e = (l >> L.shift) * (uint32_t)L.block + (l & (uint32_t)L.mask);
        return *reinterpret_cast<T*>(reinterpret_cast<Byte*>(base + q * L.plane) + e * (uint32_t)sizeof(T));
    }
    // plane v of a structure-of-arrays buffer with planes of N elements
    template <typename T>
    static __device__ __forceinline__ T& soa(T* base, int v, idx l, int64_t N) {
        using Byte = std::conditional_t<std::is_const_v<T>, const char, char>;
        return *reinterpret_cast<T*>(reinterpret_cast<Byte*>(base + v * N) + (uint32_t)(l * (uint32_t)sizeof(T)));
    }
};

template <>
struct Plane<false> {
    using idx = int64_t;
    template <typename T>
    static __device__ __forceinline__ T& at(T* base, int q, idx l, const PopLayout& L) {
        return base[pidx(L, q, l)];
    }
    template <typename T>
    static __device__ __forceinline__ T& soa(T* base, int v, idx l, int64_t N) {
        return base[v * N + l];
    }
};

// Streaming accesses.  A step reads each population of A once and writes each
// population of B once.  When the two arrays are larger than the 256 MB
// last-level cache (Infinity Cache) they cannot stay in it from one kernel to
// the next, and the bulk kernels issue these accesses with the non-temporal
// hint (NT = true): on a copy of one float64 step of a 512^3 grid in this
// layout (19 planes in blocks of 8192 nodes, one node per thread, 512-thread
// workgroups) it reaches 6.18 TB/s against 5.72 TB/s without (6.01 TB/s with
// the hint on stores only).  Smaller cases keep plain accesses: there the hint
// evicts what the LODI pass and the next step would find in the cache (a
// 151 x 151 x 21 HRR case, 145 MB of populations, took 12 % longer per step
// with it).  Arrays whose offsets need 64-bit indices are always far larger.
constexpr int64_t kCacheBytes = (int64_t)256 << 20;
inline bool stream_populations(const PopLayout& L, int elem_bytes) {
    return 2 * L.elems * (int64_t)elem_bytes > kCacheBytes;
}
template <bool NT, typename T>
__device__ __forceinline__ T ld_pop(const T* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename T>
__device__ __forceinline__ void st_pop(T* p, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

// true when the narrow form addresses population arrays of `pop_bytes` per
// element and macro planes of `plane_bytes` per element: the per-node part of
// every offset (the node's block and position, or its index in a whole plane)
// stays below 4 GiB
inline bool narrow_planes(const PopLayout& L, int pop_bytes, int plane_bytes) {
    const int64_t node_part = L.block ? (((L.N - 1) >> L.shift) * L.block + L.mask + 1) : L.N;
    return node_part * (int64_t)pop_bytes <= ((int64_t)1 << 32) && L.N * (int64_t)plane_bytes <= ((int64_t)1 << 32);
}

// Addressing and streaming form of a bulk-kernel launch.  By default
// (BULK_AUTO) it follows the arrays' size: 64-bit offsets when the narrow
// form cannot address them, the non-temporal hint when they exceed the
// last-level cache.  Tests force the forms that only large grids take onto
// small grids, so that every instantiation the large cases run is compared
// with the oracle; a narrow form is never forced onto arrays it cannot address.
enum BulkVariant : int {
    BULK_AUTO = 0,
    BULK_WIDE_NT = 1,
    BULK_NARROW_NT = 2,
    BULK_NARROW_PLAIN = 3,
    BULK_MFMA = 4,
    BULK_RELATIVE_NT = 5
};
// rel: block-relative 32-bit offsets (see bulk64): for arrays the narrow form
// cannot address whose workgroups each stay within a window that 32-bit
// offsets reach; the float64 bulk kernel takes it in place of the 64-bit form
struct BulkForm {
    bool narrow, nt, rel;
};
// Blocks a workgroup's accesses reach beyond its row's first block: `reach`
// nodes (the farthest neighbour index a row's pushes take, both ways: a plane
// and two rows, two planes when the y wrap of a periodic grid's last row lands
// a plane back), rounded up to blocks, plus one.
inline int64_t rel_span_blocks(int64_t reach, const PopLayout& L) {
    return L.block ? ((reach + L.mask) >> L.shift) + 1 : 0;
}
// whether block-relative offsets reach every access of a workgroup: a window
// of 2 span + 2 blocks of pop_bytes populations, node indices below 2^32 and
// macro planes of plane_bytes addressable with 32-bit offsets
inline bool relative_fits(const PopLayout& L, int64_t reach, int pop_bytes, int plane_bytes) {
    if (!L.block) return false;
    const int64_t window = (2 * rel_span_blocks(reach, L) + 2) * L.block * (int64_t)pop_bytes;
    return window <= ((int64_t)1 << 32) && L.N < ((int64_t)1 << 32) &&
           L.N * (int64_t)plane_bytes <= ((int64_t)1 << 32);
}
inline BulkForm bulk_form(const PopLayout& L, int pop_bytes, int plane_bytes, int variant, int64_t reach = 0) {
    const bool narrow = narrow_planes(L, pop_bytes, plane_bytes);
    const bool rel_ok = reach > 0 && relative_fits(L, reach, pop_bytes, plane_bytes);
    switch (variant) {
        case BULK_WIDE_NT: return {false, true, false};
        case BULK_NARROW_NT: return {narrow, true, false};
        case BULK_NARROW_PLAIN: return {narrow, !narrow, false};
        case BULK_RELATIVE_NT: return {false, true, rel_ok};
        default: return {narrow, !narrow || stream_populations(L, pop_bytes), !narrow && rel_ok};
    }
}

}  // namespace acp
