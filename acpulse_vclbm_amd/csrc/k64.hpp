// Host-side launchers of the float64 kernels (k64_bulk.hip, k64_aux.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "phys64.hpp"

namespace acp {

// bulk-kernel operator codes
enum BulkOp : int { OP_SRT = 0, OP_MRT_GS = 1, OP_MRT_RM = 2, OP_RR = 3, OP_HRR = 4, OP_STREAM = 5 };

// One boundary node handled by the LODI kernel (Node.cpp:229-504): its index, the
// first and second normal neighbours (read from the B array), the four tangent
// neighbours (A array, -1 when absent), the inward normal direction and the
// corner flag of Lattice::setConnectivity.
struct LodiNode {
    int64_t node, n1, n2, t[4];
    int32_t dir, corner;
};

// xopen (periodic grids only): the x faces are not wrapped; their nodes
// collide like interior nodes and links leaving the grid are dropped (the
// fine level of a refined case, whose x faces the coarse level sets).
struct Grid {
    int nx, ny, nz;
    int64_t N;
    int xopen = 0;
};

hipError_t launch_bulk64(int op, int order, bool cubic, bool periodic, const K64* k, const double* A, double* B,
                         double* psi, const double* macro, const Grid& g, hipStream_t s);
hipError_t launch_macro64(const K64* k, const double* A, double* macro, const Grid& g, bool periodic,
                          hipStream_t s);
hipError_t launch_lodi64(int order, bool cubic, bool transverse, const K64* k, const double* A, double* B,
                         const double* psi, const LodiNode* nodes, int64_t count, int64_t N, hipStream_t s);
hipError_t launch_init_random64(int order, const K64* k, double* A, double* B, const Grid& g, uint64_t seed,
                               double rho0, double rho_amp, double u_amp, hipStream_t s);
hipError_t launch_probe64(const double* A, int64_t N, const int64_t* nodes, int nprobe, double* out,
                          int32_t* slot, int32_t cap, hipStream_t s);
hipError_t launch_checksum(const void* a, int elem_bytes, int64_t n, unsigned long long* partial, int nblocks,
                           hipStream_t s);
hipError_t launch_mass64(const double* A, int64_t N, double* partial, int nblocks, hipStream_t s);
hipError_t launch_macros_out64(const K64* k, const double* A, const Grid& g, double* rho, double* u,
                               hipStream_t s);

}  // namespace acp
