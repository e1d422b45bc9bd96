// Host-side launchers of the float64 kernels (k64_bulk.hip, k64_aux.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "phys64.hpp"
#include "plane.hpp"

namespace acp {

struct KShift;

// bulk-kernel operator codes
enum BulkOp : int { OP_SRT = 0, OP_MRT_GS = 1, OP_MRT_RM = 2, OP_RR = 3, OP_HRR = 4, OP_STREAM = 5 };
// The MRT-RM collision on the matrix cores (bulk variant BULK_MFMA; periodic
// float64 grids only, see k64_bulk.hip): a kernel form, never a case's operator.
constexpr int OP_MRT_RM_MFMA = 6;
// Ct of that form: the case's collision matrix, transposed and zero-padded.
constexpr int kMfmaRows = 20, kMfmaCols = 32;

// One boundary node handled by the LODI kernel (Node.cpp:229-504): its index, the
// first and second normal neighbours (read from the B array), the four tangent
// neighbours (A array, -1 when absent), the inward normal direction and the
// corner flag of Lattice::setConnectivity.
struct LodiNode {
    int64_t node, n1, n2, t[4];
    int32_t dir, corner;
};
static_assert(sizeof(LodiNode) == 64, "one LODI record per 64 bytes");

// For grids wrapped in y and z (the kernels' PERIODIC form) each x face is
// wrapped, open or LODI; bit 0 of xopen / xlodi is the face i = 0, bit 1 the
// face i = nx - 1 (both flags 0: x wraps too).  An open face is not wrapped:
// its nodes collide like interior nodes and links leaving the grid are
// dropped (the faces of a refined case's fine level that the coarse level
// sets).  A LODI face holds boundary nodes as in the reference: they do not
// collide, push only their inward x links, and the LODI pass updates them.
// L: layout of the population arrays (plane.hpp).
//
// Nodes strictly inside the box [b0, b1] x [b2, b3] x [b4, b5] are no part of
// the lattice: the hole of the reference's two-level shell geometry, whose
// volume the coarse level holds.  The bulk kernels and the macro pre-pass skip
// them; nothing reads them (the fine level's couplings, restriction and output
// touch the box surface and the nodes outside).  The all-zero box holds no node.
struct Hole {
    int b[6];
    __host__ __device__ __forceinline__ bool inside(int i, int j, int k) const {
        return i > b[0] && i < b[1] && j > b[2] && j < b[3] && k > b[4] && k < b[5];
    }
    // the x row (j, k) crosses the hole: its nodes b0 + 1 .. b1 - 1 are skipped
    __host__ __device__ __forceinline__ bool crosses(int j, int k) const {
        return b[1] - b[0] > 1 && j > b[2] && j < b[3] && k > b[4] && k < b[5];
    }
    bool empty() const { return b[1] - b[0] < 2 || b[3] - b[2] < 2 || b[5] - b[4] < 2; }
};

struct Grid {
    int nx, ny, nz;
    int64_t N;
    int xopen = 0;
    PopLayout L;
    int xlodi = 0;
    int variant = BULK_AUTO;  // addressing and streaming form of the bulk kernels (plane.hpp)
    Hole hole{};
};

hipError_t launch_bulk64(int op, int order, bool cubic, bool periodic, const K64* k, const double* A, double* B,
                         double* psi, const double* macro, const Grid& g, hipStream_t s);
// The matrix-core form of the relaxed MRT-RM bulk kernel on a periodic grid
// (nx a multiple of 64): Ct [kMfmaRows][kMfmaCols] as written by
// mrt_rm_collision_matrix.
hipError_t launch_bulk64_mfma(const K64* k, const double* A, double* B, const double* Ct, const Grid& g,
                              hipStream_t s);
// C^T of f* = f + C (feq - f), the relaxed MRT-RM collision as one matrix
// (C = M^-1 S M), zero-padded to [kMfmaRows][kMfmaCols] (host arithmetic).
void mrt_rm_collision_matrix(const K64& k, double* Ct);
hipError_t launch_macro64(const K64* k, const double* A, double* macro, const Grid& g, bool periodic,
                          hipStream_t s);
// Workgroup size of a LODI phase: a phase of fewer than 16 Ki nodes (the x
// faces of a quasi-2D grid, edges, corners) fills few CUs with 256-thread
// workgroups; 64-thread workgroups spread it over four times as many.
inline int lodi_block(int64_t count) { return count >= ((int64_t)1 << 14) ? 256 : 64; }

hipError_t launch_lodi64(int order, bool cubic, bool transverse, const K64* k, const double* A, double* B,
                         const double* psi, const LodiNode* nodes, int64_t count, const PopLayout& L, hipStream_t s);
hipError_t launch_init_random64(int order, const K64* k, double* A, double* B, const Grid& g, uint64_t seed,
                               double rho0, double rho_amp, double u_amp, hipStream_t s);
hipError_t launch_probe64(const double* A, const PopLayout& L, const int64_t* nodes, int nprobe, double* out,
                          int32_t* slot, int32_t cap, hipStream_t s);
hipError_t launch_checksum(const void* a, int elem_bytes, int64_t n, unsigned long long* partial, int nblocks,
                           hipStream_t s);
hipError_t launch_mass64(const double* A, const PopLayout& L, double* partial, int nblocks, hipStream_t s);
// one population plane q of the node range [l0, l0 + n) between a contiguous
// float64 buffer and the arrays of a case of any tier (converted to and from
// the shifted storage for those tiers; dB may be null)
hipError_t launch_put_plane(int prec, const KShift* k, const double* src, void* dA, void* dB, const PopLayout& L,
                            int q, int64_t l0, int64_t n, hipStream_t s);
hipError_t launch_get_plane(int prec, const KShift* k, const void* src, double* dst, const PopLayout& L, int q,
                            int64_t l0, int64_t n, hipStream_t s);
hipError_t launch_macros_out64(const K64* k, const double* A, const Grid& g, double* rho, double* u,
                               hipStream_t s);

}  // namespace acp
