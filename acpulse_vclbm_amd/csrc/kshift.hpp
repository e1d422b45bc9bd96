// Shifted-population tiers: populations stored as s*(f_i/rho0 - w_i) in float32
// (s = 1) or float16 (s = per-lattice scale), collided in float32 in lattice
// units (rho0 = 1, dx = dt = 1, cs^2 = 1/3).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "acp.h"
#include "k64.hpp"

namespace acp {

struct KShift {
    float omega;       // collision frequency (dimensionless, same as the float64 tier)
    float s, inv_s;    // storage scale and its inverse
    float sigma;       // HRR hybridisation
    float hrr_pre;     // -(2 cs^2 / omega): (1-sigma) * this * rho * S enters a2
    float k1, ma, ma2; // LODI relaxation and Mach number
    float inv_len[3];  // 1 / (nodes along x, y, z): cs/L of the LODI relaxation uses L in nodes
    float k6;          // 1 / (6 c^2), c = sqrt(3) cs in m/s: the fourth-order terms' mixing factor (physf.hpp)
    double rho0;       // SI density scale for conversions
    double c_si;       // molecular velocity cs*sqrt(3) for SI velocity output
};

// storage element type per tier
template <int PREC>
struct StoreT;
template <>
struct StoreT<ACP_PREC_F32_SHIFTED> {
    using type = float;
};
template <>
struct StoreT<ACP_PREC_F16_SHIFTED> {
    using type = _Float16;
};

hipError_t launch_bulk_shift(int prec, int op, int order, bool periodic, const KShift* k, const void* A, void* B,
                             const float* macro, const Grid& g, hipStream_t s);
hipError_t launch_macro_shift(int prec, const KShift* k, const void* A, float* macro, const Grid& g, hipStream_t s);
hipError_t launch_lodi_shift(int prec, int order, bool transverse, const KShift* k, const void* A, void* B,
                             const LodiNode* nodes, int64_t count, const PopLayout& L, hipStream_t s);
// density and velocity (SI, float64) of every node: rho [N], u [3][N]
hipError_t launch_macros_out_shift(int prec, const KShift* k, const void* A, const Grid& g, double* rho, double* u,
                                   hipStream_t s);
hipError_t launch_init_random_shift(int prec, int order, const KShift* k, void* A, void* B, const Grid& g,
                                    uint64_t seed, double rho_amp, double u_amp, hipStream_t s);
// density (SI, float64) at probe nodes and per-block mass partial sums (float64)
hipError_t launch_probe_shift(int prec, const KShift* k, const void* A, const PopLayout& L, const int64_t* nodes,
                              int nprobe, double* out, int32_t* slot, int32_t cap, hipStream_t s);
hipError_t launch_mass_shift(int prec, const KShift* k, const void* A, const PopLayout& L, double* partial,
                             int nblocks, hipStream_t s);

// Equilibrium state from a density band rho_rows[(j - j0) * nx + i] (u = 0)
// into both arrays of any tier (k64_aux.hip, float64 arithmetic).
hipError_t launch_init_rows(int prec, int order, const K64* k, const KShift* ks, const double* rho_rows, int j0,
                            int rows, const Grid& g, void* A, void* B, hipStream_t s);
// The same from one density per node of the range [l0, l0 + n).
hipError_t launch_init_nodes(int prec, int order, const K64* k, const KShift* ks, const double* rho, int64_t l0,
                             int64_t n, const Grid& g, void* A, void* B, hipStream_t s);

}  // namespace acp
