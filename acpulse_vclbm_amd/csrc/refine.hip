// Two-level vertex-centred grid refinement (ratio 2), float64.
//
// The reference's shipped case file selects the standard coupling
// (Coupling-Approach STD: Filippova & Hanel 1998, doi 10.1006/jcph.1998.6089)
// with one coarse overlap cell, every fine link replaced from the coarse level
// (Coarse-to-Fine fAll), every coarse link replaced from the fine level
// (Fine-to-Coarse cAll), no fine-to-coarse filter (F2C-Filter NOFI) and
// population-based scaling (Scaling-Style SRTSTYLE)
// (BoundaryConditions.txt:52-90).  Its coupling arithmetic lives in
// GridCoupling.cpp, which the reference does not contain; the declarations
// (GridCoupling.h:28-46: interpolTimeScalingCoarseToFine, interpolSpaceMidNode,
// interpolSpaceCenterNode, scalingFineToCoarse) and the nested stepping of
// Lattice.cpp:4354-5770 fix the sequence, and the published scheme the rest:
//
//   one coarse step:  coarse collide + stream            t -> t+1
//                     fine collide + stream              t -> t+1/2
//                     coarse-to-fine at t+1/2            (time-interpolated)
//                     fine collide + stream              t+1/2 -> t+1
//                     coarse-to-fine at t+1
//                     fine-to-coarse at t+1
//
// Geometry: the coarse level is an nx x ny x nz grid, periodic or with LODI x
// faces (ACP_BC_LODI_X); the fine level covers coarse x in [x0, x1] (every
// y, z) at half the spacing, i.e. 2 (x1 - x0) + 1 by 2 ny by 2 nz nodes,
// periodic in y and z, with open x faces (ACP_BC_OPEN_X), or, when the region
// reaches the high-x face of a LODI_X coarse level, an open face at x0 and a
// LODI face at the other end (ACP_BC_OPEN_LODI_X).  Coarse node (I, J, K)
// coincides with fine node (2 (I - x0), 2 J, 2 K).
//
// Coarse-to-fine on the open fine x faces (all 19 links): coincident nodes take
// the coarse populations of the interface column, linearly interpolated in
// time (f = 0.5 (f(t) + f(t+1)) at t+1/2), with the non-equilibrium part
// scaled:  f_fine = feq(rho, u) + a (f - feq(rho, u)),  a = tau_f / (2 tau_c);
// the other face nodes average their coincident neighbours along y and z
// (mid nodes: 2, centre nodes: 4).
// Fine-to-coarse on every coarse node strictly inside the fine region (the
// overlap cell x0 + 1, x1 - 1 and, so that the whole coarse field stays a
// valid solution, the nodes beyond it):  f_coarse = feq + (1 / a) (f_fine - feq)
// of the coincident fine node.  When the fine region reaches the coarse
// level's high-x LODI face, the coarse node on that face keeps its own LODI
// update (the fine level's LODI face lies under it).
// HRR hands strain rates across the interface as the reference's partner
// functions compute them (Node.cpp:1892-2490, called at Lattice.cpp:12142,
// 12884 and 13211): the fine face nodes collide with their coarse partners'
// strain rate (at t and at t + 1/2), the coarse grid-coupling column (x0 + 1,
// x1 - 1) with its fine partners', instead of their own stencils.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <initializer_list>
#include <type_traits>
#include <vector>

#include "acp.h"
#include "engine.hpp"
#include "physf.hpp"

namespace acp {

namespace refine {

// Scaling-Style of the standard coupling (BoundaryConditions.txt:89, chosen at
// Lattice.cpp:4634-4704): the non-equilibrium part n = f - feq of a
// population set handed between the levels is scaled by one factor a
// (SRTSTYLE: tau_f / (2 tau_c) coarse to fine, its inverse fine to coarse) or,
// MRTSTYLE, moment by moment in the Gram-Schmidt basis of the MRT-GS operator
// by m[r]: tau_f / (2 tau_c) for the moments that operator relaxes with omega,
// 1/2 for those it relaxes at rate 1, 0 for the conserved ones, the inverses
// fine to coarse (the oracle's scale_neq; the reference's
// interpolTimeScalingCoarseToFineMRT_GSbasis and scalingFineToCoarseMRT*_GSbasis
// are declared at GridCoupling.h:30-36 in the absent GridCoupling.cpp).
// Passed to the kernels by value.
struct NeqScale {
    int mrt;
    double a;
    double m[19];
};

__device__ __forceinline__ void scale_neq(const K64& k, double n[19], const NeqScale& s) {
    if (!s.mrt) {
#pragma unroll
        for (int q = 0; q < 19; ++q) n[q] = s.a * n[q];
        return;
    }
    double m[19];
    gs_moments64(k, n, m);
#pragma unroll
    for (int r = 0; r < 19; ++r) m[r] = m[r] * s.m[r];
    gs_inverse64(k, m, n);
}

// density and velocity (SI) and feq of one node's populations; the
// non-equilibrium part rescaled
template <int ORDER>
__device__ __forceinline__ void rescale(const K64& k, double f[19], const NeqScale& s) {
    double rho, ux, uy, uz, eq[19], n[19];
    moments64(k, f, rho, ux, uy, uz);
    feq64<ORDER>(k, rho, ux, uy, uz, eq);
#pragma unroll
    for (int q = 0; q < 19; ++q) n[q] = f[q] - eq[q];
    scale_neq(k, n, s);
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = eq[q] + n[q];
}

// The populations a coincident fine face node takes from coarse node lc:
// time-interpolated (half: at t + 1/2) and rescaled.
// Both time levels are loaded unconditionally, all 38 loads ahead of the
// arithmetic, so that one memory latency covers them.
template <int ORDER>
__device__ __forceinline__ void c2f_value(const K64& k, const double* __restrict__ Cprev,
                                          const double* __restrict__ Ccur, const PopLayout& CN, int64_t lc, int half,
                                          const NeqScale& a, double f[19]) {
    const int64_t o = (lc >> CN.shift) * CN.block + (lc & CN.mask);
    double prev[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) {
        f[q] = Ccur[o + q * CN.plane];
        prev[q] = Cprev[o + q * CN.plane];
    }
    if (half) {
#pragma unroll
        for (int q = 0; q < 19; ++q) f[q] = 0.5 * (prev[q] + f[q]);
    }
    rescale<ORDER>(k, f, a);
}

// Coarse-to-fine on the coupled fine x faces (nfaces: 2 both, 1 the face at
// x0 only), in two launches.  Face 0 is fine column 0 under coarse column
// ci0, face 1 fine column fnx - 1 under ci1.
// First the coincident nodes (even j and k, one thread per coarse (J, K) of
// either face) take their coarse node's value; then the other face nodes
// average their 2 (mid nodes) or 4 (centre nodes) coincident neighbours.
// Link sets of the standard coupling (BoundaryConditions.txt:73-86), bit q
// for direction q: the links the fine face nodes at coarse positions take from
// the coarse level (Coarse-to-Fine fAll / MfP / Mf) and those the other face
// nodes average (Hanging-Node-Reconstruction), per face; the coarse columns
// that take only some restricted links (Fine-to-Coarse Mc: the grid-coupling
// columns x0 + 1 and x1 - 1, -1 for none) and those links.  Every other
// restricted node takes all 19.
constexpr uint32_t kAllLinks = (1u << 19) - 1;
struct StdLinks {
    uint32_t face[2], hang[2];
    int gc[2];
    uint32_t gcm[2];
};
__device__ __forceinline__ uint32_t gc_mask(const StdLinks& lk, int I) {
    return I == lk.gc[0] ? lk.gcm[0] : (I == lk.gc[1] ? lk.gcm[1] : kAllLinks);
}

template <int ORDER>
__global__ __launch_bounds__(256) void c2f_coincident(const K64* __restrict__ kp, const double* __restrict__ Cprev,
                                                      const double* __restrict__ Ccur, int half, const NeqScale a, int cnx,
                                                      int cny, int cnz, int ci0, int ci1, double* __restrict__ F,
                                                      int fnx, int fny, const PopLayout CN, const PopLayout FN,
                                                      int nfaces, const StdLinks lk) {
    __shared__ K64 sk;
    const K64& k = stage_k64(kp, sk);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = cny * cnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int J = r % cny, K = r / cny;
    const int ci = face ? ci1 : ci0, fi = face ? fnx - 1 : 0;
    const int64_t lc = ci + (int64_t)cnx * (J + (int64_t)cny * K);
    const int64_t lf = fi + (int64_t)fnx * (2 * J + (int64_t)fny * (2 * K));
    double f[19];
    c2f_value<ORDER>(k, Cprev, Ccur, CN, lc, half, a, f);
    const uint32_t m = lk.face[face];
#pragma unroll
    for (int q = 0; q < 19; ++q)
        if ((m >> q) & 1u) F[pidx(FN, q, lf)] = f[q];
}

// c2f_between: Fc and F are the same array; the kernel reads only coincident nodes (even j
// and k) through Fc and writes only the others through F, so no element is
// both read and written and both pointers may be restrict-qualified, which
// lets every load of a thread be issued before its first store.
__global__ __launch_bounds__(256) void c2f_between(const double* __restrict__ Fc, double* __restrict__ F, int fnx,
                                                   int fny, int fnz, const PopLayout FN, int nfaces,
                                                   const StdLinks lk) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = fny * fnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int j = r % fny, k = r / fny;
    if (!(j & 1) && !(k & 1)) return;
    const int fi = face ? fnx - 1 : 0;
    auto at = [&](int jj, int kk) {
        jj = (jj + fny) % fny;
        kk = (kk + fnz) % fnz;
        return fi + (int64_t)fnx * (jj + (int64_t)fny * kk);
    };
    const int64_t l = at(j, k);
    const uint32_t m = lk.hang[face];
    if ((j & 1) && (k & 1)) {
        const int64_t a = at(j - 1, k - 1), b = at(j + 1, k - 1), c = at(j - 1, k + 1), d = at(j + 1, k + 1);
#pragma unroll
        for (int q = 0; q < 19; ++q)
            if ((m >> q) & 1u)
                F[pidx(FN, q, l)] =
                    ((Fc[pidx(FN, q, a)] + Fc[pidx(FN, q, b)]) + (Fc[pidx(FN, q, c)] + Fc[pidx(FN, q, d)])) * 0.25;
    } else {
        const int64_t a = (j & 1) ? at(j - 1, k) : at(j, k - 1);
        const int64_t b = (j & 1) ? at(j + 1, k) : at(j, k + 1);
#pragma unroll
        for (int q = 0; q < 19; ++q)
            if ((m >> q) & 1u) F[pidx(FN, q, l)] = (Fc[pidx(FN, q, a)] + Fc[pidx(FN, q, b)]) * 0.5;
    }
}

// F2C-Filter LAG (Lagrava et al. 2012, doi 10.1016/j.jcp.2012.03.015; the
// reference calls it at Lattice.cpp:5190-5199): the coarse node receives the
// coincident fine node's equilibrium plus (1/a) times the mean non-equilibrium
// of that node and its 18 D3Q19 neighbours, each against its own equilibrium,
// summed in the order: the node, then directions 0..17.  The coincident fine
// node of coarse node (x0 + 1 + di, J, K) is (2 (1 + di), 2 J, 2 K); its x
// neighbours exist (x in [2, fnx - 3]), y and z wrap.
// The 19 non-equilibria of a coarse node's mean are spread over 19 lanes:
// lane s evaluates fine node s of the mean (s = 0 the coincident node, s = 1 + j
// its neighbour in direction j) into LDS, and after the barrier lane q sums
// direction q over s = 0, 1, ..., 18 in that order.  (One thread per coarse
// node chaining the 19 equilibria ran at 256 VGPRs and under one wave per SIMD.)
// Four waves per SIMD: 39.5 KB of LDS per workgroup lets four workgroups share
// a CU's 160 KB, so the case constants are read from memory rather than from a
// copy in LDS, and registers are held to 128 (orders 2 and 3; order 4 needs
// 130 and runs at three waves, where four would spill).
constexpr int kLagNodes = 13;  // coarse nodes per 256-thread workgroup (247 lanes)

//
// F2C-Filter TOU (TOU = true; Touil, Ricot and Leveque 2014, doi
// 10.1016/j.jcp.2013.07.037, selected at Lattice.cpp:5179-5211) shares the
// first half: each population is filtered along its own lattice direction,
//   nbar_q = (1/4 n_q(x - e_q) + 1/2 n_q(x)) + 1/4 n_q(x + e_q),  nbar_18 = n_18(x),
// n_q(x - e_q) being lane 1 + OPP[q]'s and n_q(x + e_q) lane 1 + q's; after a
// second barrier every lane sums the filtered part's mass M and momentum J
// over q = 0..18 and clears them, nbar_q - w_q (M + xi_q . J / cs^2), so that
// the coarse node keeps its coincident fine node's density and velocity.
// LIST: the coarse nodes list[t] (the shell geometry's restricted nodes) with
// their masks lmask[t] in place of the columns x0 + 1 .. x1 - 1.
template <int ORDER, bool TOU, bool LIST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ORDER < 4 ? 4 : 3))) void f2c_lag(const K64* __restrict__ kp, const double* __restrict__ F, int fnx,
                                               int fny, int fnz, const NeqScale ainv, double* __restrict__ C, int cnx,
                                               int cny, int cnz, int x0, int width, const PopLayout CN,
                                               const PopLayout FN, const StdLinks lk,
                                               const int64_t* __restrict__ list, const uint32_t* __restrict__ lmask,
                                               int64_t nlist) {
    __shared__ double neq[kLagNodes][19][19];  // [node][fine node s][direction]
    __shared__ double eq0[kLagNodes][19];      // the coincident node's equilibrium
    const K64& k = *kp;
    const int node = (int)threadIdx.x / 19, s = (int)threadIdx.x % 19;
    const int64_t t = (int64_t)blockIdx.x * kLagNodes + node;
    const bool active = node < kLagNodes && t < (LIST ? nlist : (int64_t)width * cny * cnz);
    int di = 0, J = 0, K = 0, Il = 0;
    if (active) {
        if constexpr (LIST) {
            const int64_t lc = list[t];
            Il = (int)(lc % cnx);
            J = (int)((lc / cnx) % cny);
            K = (int)(lc / ((int64_t)cnx * cny));
        } else {
            di = (int)(t % width);
            const int64_t r = t / width;
            J = (int)(r % cny);
            K = (int)(r / cny);
        }
        int dx = 0, dy = 0, dz = 0;
        static_for<0, 18>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            if (s == j + 1) {
                dx = EX[j];
                dy = EY[j];
                dz = EZ[j];
            }
        });
        const int fx = LIST ? 2 * Il : 2 * (1 + di), fy = 2 * J, fz = 2 * K;
        const int y = (fy + dy + fny) % fny, z = (fz + dz + fnz) % fnz;
        const int64_t l = fx + dx + (int64_t)fnx * (y + (int64_t)fny * z);
        double g[19], e[19], rho, ux, uy, uz;
#pragma unroll
        for (int q = 0; q < 19; ++q) g[q] = F[pidx(FN, q, l)];
        moments64(k, g, rho, ux, uy, uz);
        feq64<ORDER>(k, rho, ux, uy, uz, e);
#pragma unroll
        for (int q = 0; q < 19; ++q) neq[node][s][q] = g[q] - e[q];
        if (s == 0) {
#pragma unroll
            for (int q = 0; q < 19; ++q) eq0[node][q] = e[q];
        }
    }
    __syncthreads();
    const int q = s;
    const int I = LIST ? Il : (x0 + 1 + di) % cnx;
    const int64_t lc = I + (int64_t)cnx * (J + (int64_t)cny * K);
    // Fine-to-Coarse link set of this column (or listed node)
    const bool take = ((LIST ? (active ? lmask[t] : 0u) : gc_mask(lk, I)) >> q) & 1u;
    double nq = 0.;  // the non-equilibrium part lane q hands to population q
    if constexpr (TOU) {
        __shared__ double nbar[kLagNodes][19];
        if (active) {
            double nb = neq[node][0][q];
            static_for<0, 18>([&](auto qc) {
                constexpr int qq = decltype(qc)::value;
                if (q == qq) nb = (0.25 * neq[node][1 + OPP[qq]][qq] + 0.5 * neq[node][0][qq]) + 0.25 * neq[node][1 + qq][qq];
            });
            nbar[node][q] = nb;
        }
        __syncthreads();
        if (active) {
            double M = 0., Jx = 0., Jy = 0., Jz = 0.;
            static_for<0, 19>([&](auto qc) {
                constexpr int qq = decltype(qc)::value;
                const double v = nbar[node][qq];
                M += v;
                if constexpr (EX[qq] != 0) Jx += (EX[qq] > 0 ? k.c : -k.c) * v;
                if constexpr (EY[qq] != 0) Jy += (EY[qq] > 0 ? k.c : -k.c) * v;
                if constexpr (EZ[qq] != 0) Jz += (EZ[qq] > 0 ? k.c : -k.c) * v;
            });
            double sv = M, w = 0.;
            static_for<0, 19>([&](auto qc) {
                constexpr int qq = decltype(qc)::value;
                if (q == qq) {
                    if constexpr (EX[qq] != 0) sv += k.rcs2 * ((EX[qq] > 0 ? k.c : -k.c) * Jx);
                    if constexpr (EY[qq] != 0) sv += k.rcs2 * ((EY[qq] > 0 ? k.c : -k.c) * Jy);
                    if constexpr (EZ[qq] != 0) sv += k.rcs2 * ((EZ[qq] > 0 ? k.c : -k.c) * Jz);
                    w = weight(qq);
                }
            });
            nq = nbar[node][q] - w * sv;
        }
    } else if (active) {
        double acc = neq[node][0][q];
#pragma unroll
        for (int m = 1; m < 19; ++m) acc += neq[node][m][q];
        nq = acc / 19.;
    }
    if (!ainv.mrt) {
        if (active && take) C[pidx(CN, q, lc)] = eq0[node][q] + ainv.a * nq;
        return;
    }
    // MRTSTYLE: the node's 19 parts gathered in LDS, lane r forms moment r
    // (gs_moments64's order), scales it, and lane q the population q of the
    // scaled moments (gs_inverse64's order)
    __shared__ double part[kLagNodes][19];
    if (active) part[node][q] = nq;
    __syncthreads();
    double mr = 0.;
    if (active) {
        static_for<0, 19>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if (q == r) {
                double sum = 0.;
                static_for<0, 19>([&](auto qc) {
                    constexpr int qq = decltype(qc)::value;
                    constexpr int cf = gs_coef(r, qq);
                    if constexpr (cf == 1) sum += part[node][qq];
                    else if constexpr (cf == -1) sum -= part[node][qq];
                    else if constexpr (cf == 2) sum += 2. * part[node][qq];
                    else if constexpr (cf == -2) sum -= 2. * part[node][qq];
                });
                mr = (r == 0 ? sum : sum * k.gs_cp[GS_PW[r]]) * ainv.m[r];
            }
        });
    }
    __syncthreads();
    if (active) part[node][q] = mr;
    __syncthreads();
    if (!active) return;
    double out = 0.;
    static_for<0, 19>([&](auto qc) {
        constexpr int qq = decltype(qc)::value;
        if (q == qq) {
            double sum = 0.;
            bool first = true;
            static_for<0, 19>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                if constexpr (gs_coef(r, qq) != 0) {
                    const double t = k.gs_pre[r][qq] * part[node][r];
                    sum = first ? t : sum + t;
                    first = false;
                }
            });
            out = sum;
        }
    });
    if (take) C[pidx(CN, q, lc)] = eq0[node][q] + out;
}

// Coarse nodes with x in [x0 + 1, x1 - 1] from their coincident fine nodes
// (F2C-Filter NOFI; LAG is f2c_lag).
template <int ORDER>
__global__ __launch_bounds__(256) void f2c(const K64* __restrict__ kp, const double* __restrict__ F, int fnx,
                                           int fny, int fnz, const NeqScale ainv, double* __restrict__ C, int cnx,
                                           int cny, int cnz, int x0, int width, const PopLayout CN,
                                           const PopLayout FN, const StdLinks lk) {
    __shared__ K64 sk;
    const K64& k = stage_k64(kp, sk);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)width * cny * cnz) return;
    const int di = (int)(t % width);
    const int64_t r = t / width;
    const int J = (int)(r % cny), K = (int)(r / cny);
    const int I = (x0 + 1 + di) % cnx;
    const int64_t lc = I + (int64_t)cnx * (J + (int64_t)cny * K);
    double f[19];
    const int64_t lf = 2 * (1 + di) + (int64_t)fnx * (2 * J + (int64_t)fny * (2 * K));
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = F[pidx(FN, q, lf)];
    rescale<ORDER>(k, f, ainv);
    const uint32_t m = gc_mask(lk, I);
#pragma unroll
    for (int q = 0; q < 19; ++q)
        if ((m >> q) & 1u) C[pidx(CN, q, lc)] = f[q];
}

// ---- shifted tiers: stored g = s (f / rho0 - w) ---------------------------
// The coupling arithmetic runs in float64 on SI populations (F64 = true; the
// stored values converted exactly as the population upload and download
// convert them) or in float32 on lattice-unit deviations (the tier's own
// moments and equilibrium).
template <typename T>
__device__ __forceinline__ double g_to_f(const KShift& ks, T g, int q) {
    return ks.rho0 * (weight(q) + (double)g / (double)ks.s);
}
template <typename T>
__device__ __forceinline__ T f_to_g(const KShift& ks, double f, int q) {
    return (T)((double)ks.s * (f / ks.rho0 - weight(q)));
}

// The stored values (tier T) a node receives from node lc of another level:
// time-interpolated (half) and rescaled by a, in float64 on SI populations
// (F64) or in float32 on lattice-unit deviations.
// The 19 stored values of node l (any storage type), loaded before any of
// them is used.
template <typename T>
__device__ __forceinline__ void load_raw19(const T* __restrict__ A, const PopLayout& L, int64_t l, T out[19]) {
    const int64_t o = (l >> L.shift) * L.block + (l & L.mask);
#pragma unroll
    for (int q = 0; q < 19; ++q) out[q] = A[o + q * L.plane];
}

template <typename T, int ORDER, bool F64>
__device__ __forceinline__ void couple_value(const K64& k, const KShift& ks, const T* Cprev, const T* Ccur,
                                             const PopLayout& CN, int64_t lc, int half, const NeqScale& a, T out[19]) {
    // every load ahead of the arithmetic (one memory latency for all of them)
    T cur[19], prev[19];
    load_raw19(Ccur, CN, lc, cur);
    if (half) {
        load_raw19(Cprev, CN, lc, prev);
    } else {
#pragma unroll
        for (int q = 0; q < 19; ++q) prev[q] = cur[q];
    }
    if constexpr (F64) {
        double f[19];
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const double fc = g_to_f(ks, cur[q], q);
            f[q] = half ? 0.5 * (g_to_f(ks, prev[q], q) + fc) : fc;
        });
        rescale<ORDER>(k, f, a);
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            out[q] = f_to_g<T>(ks, f[q], q);
        });
    } else {
        float d[19], eq[19], dr, ux, uy, uz;
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const float dc = (float)cur[q] * ks.inv_s;
            d[q] = half ? 0.5f * ((float)prev[q] * ks.inv_s + dc) : dc;
        });
        moments_dev(d, dr, ux, uy, uz);
        feq_dev<ORDER>(dr, ux, uy, uz, eq, ks.k6);
        const float af = (float)a.a;
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            out[q] = (T)((eq[q] + af * (d[q] - eq[q])) * ks.s);
        });
    }
}

template <typename T, int ORDER, bool F64>
__device__ __forceinline__ void couple_node(const K64& k, const KShift& ks, const T* Cprev, const T* Ccur,
                                            const PopLayout& CN, int64_t lc, int half, const NeqScale& a, T* F,
                                            const PopLayout& FN, int64_t lf, uint32_t m) {
    T v[19];
    couple_value<T, ORDER, F64>(k, ks, Cprev, Ccur, CN, lc, half, a, v);
#pragma unroll
    for (int q = 0; q < 19; ++q)
        if ((m >> q) & 1u) F[pidx(FN, q, lf)] = v[q];
}

// LAG restriction in the shifted tiers: the same mean as f2c_lag, in float64
// on SI populations (F64) or in float32 on lattice-unit deviations.
template <typename T, int ORDER, bool F64>
__device__ __forceinline__ void lag_node_s(const K64& k, const KShift& ks, const T* __restrict__ F, int fnx, int fny,
                                           int fnz, int fx, int fy, int fz, const NeqScale& ainv, T* __restrict__ C,
                                           const PopLayout& CN, int64_t lc, const PopLayout& FN, uint32_t m) {
    const int64_t l0 = fx + (int64_t)fnx * (fy + (int64_t)fny * fz);
    T raw[19];
    load_raw19(F, FN, l0, raw);
    if constexpr (F64) {
        double v[19], eq[19], acc[19], rho, ux, uy, uz;
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            v[q] = g_to_f(ks, raw[q], q);
        });
        moments64(k, v, rho, ux, uy, uz);
        feq64<ORDER>(k, rho, ux, uy, uz, eq);
#pragma unroll
        for (int q = 0; q < 19; ++q) acc[q] = v[q] - eq[q];
        static_for<0, 18>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const int y = (fy + EY[j] + fny) % fny, z = (fz + EZ[j] + fnz) % fnz;
            const int64_t l = fx + EX[j] + (int64_t)fnx * (y + (int64_t)fny * z);
            double g[19], e[19], r, x1, x2, x3;
            T rn[19];
            load_raw19(F, FN, l, rn);
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                g[q] = g_to_f(ks, rn[q], q);
            });
            moments64(k, g, r, x1, x2, x3);
            feq64<ORDER>(k, r, x1, x2, x3, e);
#pragma unroll
            for (int q = 0; q < 19; ++q) acc[q] += g[q] - e[q];
        });
#pragma unroll
        for (int q = 0; q < 19; ++q) acc[q] = acc[q] / 19.;
        scale_neq(k, acc, ainv);
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if ((m >> q) & 1u) C[pidx(CN, q, lc)] = f_to_g<T>(ks, eq[q] + acc[q], q);
        });
    } else {
        float d[19], eq[19], acc[19], dr, ux, uy, uz;
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            d[q] = (float)raw[q] * ks.inv_s;
        });
        moments_dev(d, dr, ux, uy, uz);
        feq_dev<ORDER>(dr, ux, uy, uz, eq, ks.k6);
#pragma unroll
        for (int q = 0; q < 19; ++q) acc[q] = d[q] - eq[q];
        static_for<0, 18>([&](auto jc) {
            constexpr int j = decltype(jc)::value;
            const int y = (fy + EY[j] + fny) % fny, z = (fz + EZ[j] + fnz) % fnz;
            const int64_t l = fx + EX[j] + (int64_t)fnx * (y + (int64_t)fny * z);
            float g[19], e[19], r, x1, x2, x3;
            T rn[19];
            load_raw19(F, FN, l, rn);
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                g[q] = (float)rn[q] * ks.inv_s;
            });
            moments_dev(g, r, x1, x2, x3);
            feq_dev<ORDER>(r, x1, x2, x3, e, ks.k6);
#pragma unroll
            for (int q = 0; q < 19; ++q) acc[q] += g[q] - e[q];
        });
        const float af = (float)ainv.a;
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if ((m >> q) & 1u) C[pidx(CN, q, lc)] = (T)((eq[q] + af * (acc[q] * (1.f / 19.f))) * ks.s);
        });
    }
}

// TOU restriction in the shifted tiers: the same filter as f2c_lag's TOU form,
// one thread per coarse node, in float64 on SI populations (F64) or in float32
// on lattice-unit deviations.
template <typename T, int ORDER, bool F64>
__device__ __forceinline__ void tou_node_s(const K64& k, const KShift& ks, const T* __restrict__ F, int fnx, int fny,
                                           int fnz, int fx, int fy, int fz, const NeqScale& ainv, T* __restrict__ C,
                                           const PopLayout& CN, int64_t lc, const PopLayout& FN, uint32_t m) {
    using R = std::conditional_t<F64, double, float>;
    const int64_t l0 = fx + (int64_t)fnx * (fy + (int64_t)fny * fz);
    T raw[19];
    load_raw19(F, FN, l0, raw);
    R v[19], eq[19], n0[19], up[19], dn[19], nb[19];
    auto convert = [&](const T in[19], R out[19]) {
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (F64) out[q] = g_to_f(ks, in[q], q);
            else out[q] = (float)in[q] * ks.inv_s;
        });
    };
    auto equilibrium = [&](const R f[19], R e[19]) {
        if constexpr (F64) {
            double r, x1, x2, x3;
            moments64(k, f, r, x1, x2, x3);
            feq64<ORDER>(k, r, x1, x2, x3, e);
        } else {
            float r, x1, x2, x3;
            moments_dev(f, r, x1, x2, x3);
            feq_dev<ORDER>(r, x1, x2, x3, e, ks.k6);
        }
    };
    convert(raw, v);
    equilibrium(v, eq);
#pragma unroll
    for (int q = 0; q < 19; ++q) n0[q] = v[q] - eq[q];
    static_for<0, 18>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const int y = (fy + EY[j] + fny) % fny, z = (fz + EZ[j] + fnz) % fnz;
        const int64_t l = fx + EX[j] + (int64_t)fnx * (y + (int64_t)fny * z);
        T rn[19];
        R g[19], e[19];
        load_raw19(F, FN, l, rn);
        convert(rn, g);
        equilibrium(g, e);
        up[j] = g[j] - e[j];
        dn[OPP[j]] = g[OPP[j]] - e[OPP[j]];
    });
    const R quarter = (R)0.25, half = (R)0.5;
#pragma unroll
    for (int q = 0; q < 18; ++q) nb[q] = (quarter * dn[q] + half * n0[q]) + quarter * up[q];
    nb[18] = n0[18];
    // lattice speed and 1 / cs^2: SI (float64) or lattice units (float32)
    const R c = F64 ? (R)k.c : (R)1, rcs2 = F64 ? (R)k.rcs2 : (R)3;
    R M = 0, Jx = 0, Jy = 0, Jz = 0;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        M += nb[q];
        if constexpr (EX[q] != 0) Jx += (EX[q] > 0 ? c : -c) * nb[q];
        if constexpr (EY[q] != 0) Jy += (EY[q] > 0 ? c : -c) * nb[q];
        if constexpr (EZ[q] != 0) Jz += (EZ[q] > 0 ? c : -c) * nb[q];
    });
    R n[19];
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        R sv = M;
        if constexpr (EX[q] != 0) sv += rcs2 * ((EX[q] > 0 ? c : -c) * Jx);
        if constexpr (EY[q] != 0) sv += rcs2 * ((EY[q] > 0 ? c : -c) * Jy);
        if constexpr (EZ[q] != 0) sv += rcs2 * ((EZ[q] > 0 ? c : -c) * Jz);
        n[q] = nb[q] - (R)weight(q) * sv;
    });
    if constexpr (F64) {
        scale_neq(k, n, ainv);
    } else {
        const R ai = (R)ainv.a;
#pragma unroll
        for (int q = 0; q < 19; ++q) n[q] = ai * n[q];
    }
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const R out = eq[q] + n[q];
        if (!((m >> q) & 1u)) return;
        if constexpr (F64) C[pidx(CN, q, lc)] = f_to_g<T>(ks, out, q);
        else C[pidx(CN, q, lc)] = (T)(out * ks.s);
    });
}

// The coarse-to-fine kernels of the shifted tiers (see c2f_coincident and
// c2f_between); the averages are formed from the coincident values as
// stored, in float64 (F64 coupling) or float32.
template <typename T, int ORDER, bool F64>
__global__ __launch_bounds__(256) void c2f_coincident_s(const K64* __restrict__ kp, const KShift* __restrict__ ksp,
                                                        const T* __restrict__ Cprev, const T* __restrict__ Ccur,
                                                        int half, const NeqScale a, int cnx, int cny, int cnz, int ci0,
                                                        int ci1, T* __restrict__ F, int fnx, int fny,
                                                        const PopLayout CN, const PopLayout FN, int nfaces,
                                                        const StdLinks lk) {
    __shared__ K64 sk;
    const K64& k = F64 ? stage_k64(kp, sk) : *kp;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = cny * cnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int J = r % cny, K = r / cny;
    const int ci = face ? ci1 : ci0, fi = face ? fnx - 1 : 0;
    const int64_t lc = ci + (int64_t)cnx * (J + (int64_t)cny * K);
    const int64_t lf = fi + (int64_t)fnx * (2 * J + (int64_t)fny * (2 * K));
    couple_node<T, ORDER, F64>(k, *ksp, Cprev, Ccur, CN, lc, half, a, F, FN, lf, lk.face[face]);
}

template <typename T, bool F64>
__global__ __launch_bounds__(256) void c2f_between_s(const KShift* __restrict__ ksp, const T* __restrict__ Fc,
                                                     T* __restrict__ F, int fnx, int fny, int fnz, const PopLayout FN,
                                                     int nfaces, const StdLinks lk) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = fny * fnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int j = r % fny, k = r / fny;
    if (!(j & 1) && !(k & 1)) return;
    const int fi = face ? fnx - 1 : 0;
    const KShift& ks = *ksp;
    auto at = [&](int jj, int kk) {
        jj = (jj + fny) % fny;
        kk = (kk + fnz) % fnz;
        return fi + (int64_t)fnx * (jj + (int64_t)fny * kk);
    };
    const int64_t l = at(j, k);
    int64_t nb[4];
    int n;
    if ((j & 1) && (k & 1)) {
        nb[0] = at(j - 1, k - 1), nb[1] = at(j + 1, k - 1), nb[2] = at(j - 1, k + 1), nb[3] = at(j + 1, k + 1);
        n = 4;
    } else {
        nb[0] = (j & 1) ? at(j - 1, k) : at(j, k - 1);
        nb[1] = (j & 1) ? at(j + 1, k) : at(j, k + 1);
        n = 2;
    }
    const uint32_t m = lk.hang[face];
    T r0[19], r1[19], r2[19], r3[19];
    load_raw19(Fc, FN, nb[0], r0);
    load_raw19(Fc, FN, nb[1], r1);
    if (n == 4) {
        load_raw19(Fc, FN, nb[2], r2);
        load_raw19(Fc, FN, nb[3], r3);
    }
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if (!((m >> q) & 1u)) return;
        if constexpr (F64) {
            double v = n == 4 ? ((g_to_f(ks, r0[q], q) + g_to_f(ks, r1[q], q)) +
                                 (g_to_f(ks, r2[q], q) + g_to_f(ks, r3[q], q))) * 0.25
                              : (g_to_f(ks, r0[q], q) + g_to_f(ks, r1[q], q)) * 0.5;
            F[pidx(FN, q, l)] = f_to_g<T>(ks, v, q);
        } else {
            float v = n == 4 ? (((float)r0[q] + (float)r1[q]) + ((float)r2[q] + (float)r3[q])) * 0.25f
                             : ((float)r0[q] + (float)r1[q]) * 0.5f;
            F[pidx(FN, q, l)] = (T)v;
        }
    });
}

template <typename T, int ORDER, bool F64, int FILT>
__global__ __launch_bounds__(256) void f2c_s(const K64* __restrict__ kp, const KShift* __restrict__ ksp,
                                             const T* __restrict__ F, int fnx, int fny, int fnz, const NeqScale ainv,
                                             T* __restrict__ C, int cnx, int cny, int cnz, int x0, int width,
                                             const PopLayout CN, const PopLayout FN, const StdLinks lk) {
    __shared__ K64 sk;
    const K64& k = F64 ? stage_k64(kp, sk) : *kp;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)width * cny * cnz) return;
    const int di = (int)(t % width);
    const int64_t r = t / width;
    const int J = (int)(r % cny), K = (int)(r / cny);
    const int I = (x0 + 1 + di) % cnx;
    const int64_t lc = I + (int64_t)cnx * (J + (int64_t)cny * K);
    const uint32_t m = gc_mask(lk, I);
    if constexpr (FILT == ACP_F2C_LAG) {
        lag_node_s<T, ORDER, F64>(k, *ksp, F, fnx, fny, fnz, 2 * (1 + di), 2 * J, 2 * K, ainv, C, CN, lc, FN, m);
    } else if constexpr (FILT == ACP_F2C_TOU) {
        tou_node_s<T, ORDER, F64>(k, *ksp, F, fnx, fny, fnz, 2 * (1 + di), 2 * J, 2 * K, ainv, C, CN, lc, FN, m);
    } else {
        const int64_t lf = 2 * (1 + di) + (int64_t)fnx * (2 * J + (int64_t)fny * (2 * K));
        couple_node<T, ORDER, F64>(k, *ksp, F, F, FN, lf, 0, ainv, C, CN, lc, m);
    }
}

// ---- HRR strain-rate transfer between the levels (Lattice.cpp:12142,
// 12884, 13211; Node.cpp:1892-2490).  A grid-coupling node collides with the
// strain rate handed to it (Node.cpp:1795-1798): the fine face nodes with
// their coarse partners' (coincident nodes their partner's, mid nodes the
// mean of two, centre nodes of four, as the populations' coarse-to-fine), the
// coarse grid-coupling column (one coarse cell inside the fine region) with
// its fine partner's.  Strain rates are float64 in SI units, planes of six
// components over the column's (J, K) or the face's (j, k), in every tier
// (the shifted tiers' populations are read as SI populations).

// density and velocity (SI) of node l of a level, any storage type
template <typename T>
__device__ __forceinline__ void macros_si(const K64& k, const KShift* ks, const T* __restrict__ A, const PopLayout& L,
                                          int64_t l, double m4[4]) {
    T raw[19];
    load_raw19(A, L, l, raw);
    double f[19];
    if constexpr (std::is_same_v<T, double>) {
#pragma unroll
        for (int q = 0; q < 19; ++q) f[q] = raw[q];
    } else {
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            f[q] = g_to_f(*ks, raw[q], q);
        });
    }
    moments64(k, f, m4[0], m4[1], m4[2], m4[3]);
}

// The strain rate of neighbour macros nb into entry r of the planes o (nv of
// `per` entries): S, and with the cubic-Mach terms (nv = 25) psi from the same
// stencil, as the partner functions evaluate both (Node.cpp:1892-2490).
__device__ __forceinline__ void store_handed(const K64& k, const double nb[6][4], int nv, double* __restrict__ o,
                                             int per, int r) {
    double S[6];
    if (nv > 6) {
        double ps[19];
        strain_psi64<true>(k, nb, S, ps);
#pragma unroll
        for (int q = 0; q < 19; ++q) o[(6 + q) * per + r] = ps[q];
    } else {
        strain_psi64<false>(k, nb, S, nullptr);
    }
#pragma unroll
    for (int v = 0; v < 6; ++v) o[v * per + r] = S[v];
}

// The coarse partners of the fine face nodes: coarse column ci0 (face 0) and
// ci1 (face 1), their grid-coupling neighbour (+x for face 0, -x for face 1)
// read through its fine partner (column fi0 / fi1 of the fine level, array F).
// HALF: the other neighbours as the mean of their states at t (Cprev) and
// t + 1 (Ccur), Node.cpp:2244-2380; else their current states (Ccur),
// Node.cpp:2114-2240.  Distances are the coarse spacing (k.dist).
template <typename T, bool HALF>
__global__ __launch_bounds__(64) void partner_coarse(const K64* __restrict__ kcp, const KShift* __restrict__ kcs,
                                                     const KShift* __restrict__ kfs, const T* __restrict__ Cprev,
                                                     const T* __restrict__ Ccur, const PopLayout CL, int cnx, int cny,
                                                     int cnz, int ci0, int ci1, const T* __restrict__ F,
                                                     const PopLayout FL, int fnx, int fny, int fi0, int fi1,
                                                     int nfaces, int nv, double* __restrict__ Sout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = cny * cnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int J = r % cny, K = r / cny;
    const int ci = face ? ci1 : ci0, gcm = face ? 2 : 0, fi = face ? fi1 : fi0;
    const K64& k = *kcp;
    double nb[6][4];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        if (m == gcm) {
            const int64_t lf = fi + (int64_t)fnx * (2 * J + (int64_t)fny * (2 * K));
            macros_si(k, kfs, F, FL, lf, nb[m]);
            continue;
        }
        const int a = (ci + EX[m] + cnx) % cnx, b = (J + EY[m] + cny) % cny, d = (K + EZ[m] + cnz) % cnz;
        const int64_t lc = a + (int64_t)cnx * (b + (int64_t)cny * d);
        if constexpr (HALF) {
            double p4[4], c4[4];
            macros_si(k, kcs, Cprev, CL, lc, p4);
            macros_si(k, kcs, Ccur, CL, lc, c4);
            nb[m][0] = 0.5 * (p4[0] + c4[0]);
#pragma unroll
            for (int v = 1; v < 4; ++v) nb[m][v] = 0.5 * (p4[v] + c4[v]);
        } else {
            macros_si(k, kcs, Ccur, CL, lc, nb[m]);
        }
    }
    store_handed(k, nb, nv, Sout + (int64_t)face * nv * per, per, r);
}

// The fine partners of the coarse grid-coupling columns: fine node (fi, 2J,
// 2K) from its six fine neighbours' current states at the fine spacing
// (Node.cpp:2383-2490).
template <typename T>
__global__ __launch_bounds__(64) void partner_fine(const K64* __restrict__ kfp, const KShift* __restrict__ kfs,
                                                   const T* __restrict__ F, const PopLayout FL, int fnx, int fny,
                                                   int fnz, int fi0, int fi1, int cny, int cnz, int nfaces, int nv,
                                                   double* __restrict__ Sout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = cny * cnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int J = r % cny, K = r / cny;
    const int fi = face ? fi1 : fi0;
    const K64& k = *kfp;
    double nb[6][4];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const int a = fi + EX[m], b = (2 * J + EY[m] + fny) % fny, d = (2 * K + EZ[m] + fnz) % fnz;
        macros_si(k, kfs, F, FL, a + (int64_t)fnx * (b + (int64_t)fny * d), nb[m]);
    }
    store_handed(k, nb, nv, Sout + (int64_t)face * nv * per, per, r);
}

// HRR under the direct coupling at t + 1/2 (Node.cpp:1892-2111 on level 1,
// called at Lattice.cpp:14064-14108): the fine face nodes at coarse positions,
// (fi, 2J, 2K), from the fine nodes two fine nodes away along the face and into
// the fine region (array F, after the exchange at t + 1/2), and for the
// missing outer neighbour the coarse partner's neighbour (column ci - 1 for
// face 0, ci + 1 for face 1) with its density and velocity averaged over t
// (Cprev) and t + 1 (Ccur); every distance twice the fine spacing, the coarse
// spacing of kcp.  Planes per face [6][cny cnz].
template <typename T>
__global__ __launch_bounds__(64) void dc_strain_fine(const K64* __restrict__ kcp, const KShift* __restrict__ kcs,
                                                     const KShift* __restrict__ kfs, const T* __restrict__ Cprev,
                                                     const T* __restrict__ Ccur, const PopLayout CL, int cnx, int cny,
                                                     int cnz, int ci0, int ci1, const T* __restrict__ F,
                                                     const PopLayout FL, int fnx, int fny, int fnz, int nfaces,
                                                     int nv, double* __restrict__ Sout) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = cny * cnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int J = r % cny, K = r / cny;
    const int ci = face ? ci1 : ci0, fi = face ? fnx - 1 : 0, mm = face ? 0 : 2;
    const K64& k = *kcp;
    double nb[6][4];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        if (m == mm) {
            const int64_t lc = (ci + EX[m] + cnx) % cnx + (int64_t)cnx * (J + (int64_t)cny * K);
            double p4[4], c4[4];
            macros_si(k, kcs, Cprev, CL, lc, p4);
            macros_si(k, kcs, Ccur, CL, lc, c4);
            nb[m][0] = 0.5 * (p4[0] + c4[0]);
#pragma unroll
            for (int v = 1; v < 4; ++v) nb[m][v] = 0.5 * (p4[v] + c4[v]);
        } else {
            const int a = fi + 2 * EX[m], b = (2 * J + 2 * EY[m] + fny) % fny, d = (2 * K + 2 * EZ[m] + fnz) % fnz;
            macros_si(k, kfs, F, FL, a + (int64_t)fnx * (b + (int64_t)fny * d), nb[m]);
        }
    }
    store_handed(k, nb, nv, Sout + (int64_t)face * nv * per, per, r);
}

// The fine face nodes' strain rates (and psi) from their coarse partners'
// (coarse column planes Scol, per face [nv][cny cnz]) into per-face [nv][fny fnz].
__global__ __launch_bounds__(256) void face_strain(const double* __restrict__ Scol, int cny, int cnz, int fny, int fnz,
                                                   int nfaces, int nv, double* __restrict__ Sface) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = fny * fnz, cper = cny * cnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int j = r % fny, kk = r / fny;
    const int J = j / 2, K = kk / 2, J1 = (J + 1) % cny, K1 = (K + 1) % cnz;
    const double* S = Scol + (int64_t)face * nv * cper;
    double* o = Sface + (int64_t)face * nv * per;
    const int a = J + cny * K, b = J1 + cny * K, c = J + cny * K1, d = J1 + cny * K1;
    for (int v = 0; v < nv; ++v) {
        const double* P = S + v * cper;
        double x;
        if (!(j & 1) && !(kk & 1)) x = P[a];
        else if ((j & 1) && (kk & 1)) x = ((P[a] + P[b]) + (P[c] + P[d])) * 0.25;
        else if (j & 1) x = (P[a] + P[b]) * 0.5;
        else x = (P[a] + P[c]) * 0.5;
        o[v * per + r] = x;
    }
}

// ---- direct coupling (Coupling-Approach DC1 / DC2, Astoul et al. 2021, doi
// 10.1016/j.jcp.2021.110667; called at Lattice.cpp:14018-14058 at t + 1/2 and
// 14699-14739 at t + 1, declared at GridCoupling.h:28-31, 45, its arithmetic
// in the absent GridCoupling.cpp).  The coarse interface column and the fine
// face column coincide and each level misses the links arriving from the
// other level's side.  For a coincident pair (coarse populations cv, fine fv)
// the set C of links the coarse node takes from the fine side (e_x = -n, with
// the common e_x = 0 links under McP) gives the pair's density and momentum
// X = (rho, rho u) as those of the coarse node completed with the fine-side
// links rescaled to the coarse non-equilibrium:
//   X = sum_{q not in C} cv_q (1, xi_q)
//     + sum_{q in C} [feq_q(X) + (1/a) (fv_q - feq_q(X))] (1, xi_q),
// solved by chord steps X -= M^-1 (X - b S(X) - R) with b = 1 - 1/a, S the
// mass and momentum of feq(X) over C and M^-1 the inverse of the equation's
// derivative at u = 0 (per face, from the host), starting from the moments of
// the unscaled mix.  The fine node's links of its set then take
// feq + a (cv - feq), at t + 1 the coarse node's links of its set
// feq + (1/a) (fv - feq); the other fine face nodes average their coincident
// neighbours for the hanging-node link set.
struct DcFace {
    double m[6];   // chord matrix: (rho, rho u_x) block by rows, then y, z (SI)
    float mf[6];   // the same on lattice-unit deviations (float32 coupling)
    uint32_t fine, coarse, hang;  // link sets, bit q for direction q
};
struct DcArgs {
    DcFace face[2];
    double a, ainv;
};
constexpr int kDcMaxIters = 16;          // float64: steps until converged, at most this many
constexpr double kDcTol = 4. * 2.220446049250313e-16;
constexpr int kDcItersF32 = 4;           // float32: a fixed count (each gains a factor O(|u| / c_s))

// mass and momentum over the links of `mask` (inside) or the others
template <typename R>
__device__ __forceinline__ void dc_sums(R c, const R v[19], uint32_t mask, bool inside, R out[4]) {
    R r = 0, x = 0, y = 0, z = 0;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if ((((mask >> q) & 1u) != 0u) == inside) {
            r += v[q];
            if constexpr (EX[q] != 0) x += (EX[q] > 0 ? c : -c) * v[q];
            if constexpr (EY[q] != 0) y += (EY[q] > 0 ? c : -c) * v[q];
            if constexpr (EZ[q] != 0) z += (EZ[q] > 0 ? c : -c) * v[q];
        }
    });
    out[0] = r, out[1] = x, out[2] = y, out[3] = z;
}

template <int ORDER>
__device__ __forceinline__ void dc_state64(const K64& k, double ainv, const double m[6], uint32_t C,
                                           const double cv[19], const double fv[19], double e[19]) {
    const double b = 1. - ainv;
    double g[19], X[4], Rc[4], Rf[4], R[4], S[4];
#pragma unroll
    for (int q = 0; q < 19; ++q) g[q] = ((C >> q) & 1u) ? fv[q] : cv[q];
    dc_sums<double>(k.c, g, C, true, Rf);
    dc_sums<double>(k.c, g, C, false, Rc);
#pragma unroll
    for (int i = 0; i < 4; ++i) X[i] = Rc[i] + Rf[i], R[i] = Rc[i] + ainv * Rf[i];
    for (int it = 0; it < kDcMaxIters; ++it) {
        feq64<ORDER>(k, X[0], X[1] / X[0], X[2] / X[0], X[3] / X[0], e);
        dc_sums<double>(k.c, e, C, true, S);
        double F[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) F[i] = (X[i] - b * S[i]) - R[i];
        const double d0 = m[0] * F[0] + m[1] * F[1], d1 = m[2] * F[0] + m[3] * F[1];
        const double d2 = m[4] * F[2], d3 = m[5] * F[3];
        X[0] -= d0, X[1] -= d1, X[2] -= d2, X[3] -= d3;
        if (fabs(d0) <= kDcTol * fabs(X[0]) && ((fabs(d1) + fabs(d2)) + fabs(d3)) <= kDcTol * (k.c * fabs(X[0])))
            break;
    }
    feq64<ORDER>(k, X[0], X[1] / X[0], X[2] / X[0], X[3] / X[0], e);
}

// the same on lattice-unit deviations d = f / rho0 - w in float32 (X = (d rho, rho u))
template <int ORDER>
__device__ __forceinline__ void dc_state32(float k6, float ainv, const float m[6], uint32_t C, const float cv[19],
                                           const float fv[19], float e[19]) {
    const float b = 1.f - ainv;
    float g[19], X[4], Rc[4], Rf[4], R[4], S[4];
#pragma unroll
    for (int q = 0; q < 19; ++q) g[q] = ((C >> q) & 1u) ? fv[q] : cv[q];
    dc_sums<float>(1.f, g, C, true, Rf);
    dc_sums<float>(1.f, g, C, false, Rc);
#pragma unroll
    for (int i = 0; i < 4; ++i) X[i] = Rc[i] + Rf[i], R[i] = Rc[i] + ainv * Rf[i];
    for (int it = 0; it < kDcItersF32; ++it) {
        const float ir = 1.f / (1.f + X[0]);
        feq_dev<ORDER>(X[0], X[1] * ir, X[2] * ir, X[3] * ir, e, k6);
        dc_sums<float>(1.f, e, C, true, S);
        float F[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) F[i] = (X[i] - b * S[i]) - R[i];
        const float d0 = m[0] * F[0] + m[1] * F[1], d1 = m[2] * F[0] + m[3] * F[1];
        X[0] -= d0, X[1] -= d1, X[2] -= m[4] * F[2], X[3] -= m[5] * F[3];
    }
    const float ir = 1.f / (1.f + X[0]);
    feq_dev<ORDER>(X[0], X[1] * ir, X[2] * ir, X[3] * ir, e, k6);
}

// One thread per coincident pair of the coupled faces.  MODE 0: float64
// populations; 1: shifted tier, coupling in float64 on SI populations; 2:
// shifted tier, coupling in float32 on lattice-unit deviations.  C and F are
// read and written by the same thread at the same node (not restrict).
template <typename T, int ORDER, int MODE>
__global__ __launch_bounds__(64) void dc_coincident(const K64* __restrict__ kp, const KShift* __restrict__ ksp,
                                                    const T* __restrict__ Cprev, T* C, int half, const DcArgs args,
                                                    int cnx, int cny, int cnz, int ci0, int ci1, T* F, int fnx,
                                                    int fny, const PopLayout CN, const PopLayout FN, int nfaces) {
    __shared__ K64 sk;
    const K64& k = MODE < 2 ? stage_k64(kp, sk) : *kp;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = cny * cnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int J = r % cny, K = r / cny;
    const int ci = face ? ci1 : ci0, fi = face ? fnx - 1 : 0;
    const int64_t lc = ci + (int64_t)cnx * (J + (int64_t)cny * K);
    const int64_t lf = fi + (int64_t)fnx * (2 * J + (int64_t)fny * (2 * K));
    const uint32_t fm = face ? args.face[1].fine : args.face[0].fine;
    const uint32_t cm = face ? args.face[1].coarse : args.face[0].coarse;
    T cur[19], prev[19], fr[19];
    load_raw19<T>(C, CN, lc, cur);
    if (half) load_raw19<T>(Cprev, CN, lc, prev);
    load_raw19<T>(F, FN, lf, fr);
    if constexpr (MODE < 2) {
        double m[6], cv[19], fv[19], e[19];
#pragma unroll
        for (int i = 0; i < 6; ++i) m[i] = face ? args.face[1].m[i] : args.face[0].m[i];
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (MODE == 0) {
                cv[q] = half ? 0.5 * ((double)prev[q] + (double)cur[q]) : (double)cur[q];
                fv[q] = (double)fr[q];
            } else {
                const double c1 = g_to_f(*ksp, cur[q], q);
                cv[q] = half ? 0.5 * (g_to_f(*ksp, prev[q], q) + c1) : c1;
                fv[q] = g_to_f(*ksp, fr[q], q);
            }
        });
        dc_state64<ORDER>(k, args.ainv, m, cm, cv, fv, e);
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if ((fm >> q) & 1u) {
                const double v = e[q] + args.a * (cv[q] - e[q]);
                if constexpr (MODE == 0) F[pidx(FN, q, lf)] = (T)v;
                else F[pidx(FN, q, lf)] = f_to_g<T>(*ksp, v, q);
            }
            if (!half && ((cm >> q) & 1u)) {
                const double v = e[q] + args.ainv * (fv[q] - e[q]);
                if constexpr (MODE == 0) C[pidx(CN, q, lc)] = (T)v;
                else C[pidx(CN, q, lc)] = f_to_g<T>(*ksp, v, q);
            }
        });
    } else {
        const KShift& ks = *ksp;
        float m[6], cv[19], fv[19], e[19];
#pragma unroll
        for (int i = 0; i < 6; ++i) m[i] = face ? args.face[1].mf[i] : args.face[0].mf[i];
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const float c1 = (float)cur[q] * ks.inv_s;
            cv[q] = half ? 0.5f * ((float)prev[q] * ks.inv_s + c1) : c1;
            fv[q] = (float)fr[q] * ks.inv_s;
        });
        const float af = (float)args.a, aif = (float)args.ainv;
        dc_state32<ORDER>(ks.k6, aif, m, cm, cv, fv, e);
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if ((fm >> q) & 1u) F[pidx(FN, q, lf)] = (T)((e[q] + af * (cv[q] - e[q])) * ks.s);
            if (!half && ((cm >> q) & 1u)) C[pidx(CN, q, lc)] = (T)((e[q] + aif * (fv[q] - e[q])) * ks.s);
        });
    }
}

// The other fine face nodes: the mean of their 2 (mid nodes) or 4 (centre
// nodes) coincident neighbours for the links of the hanging-node set, formed
// as c2f_between / c2f_between_s form it.  Reads only coincident nodes, writes
// only the others.
template <typename T, int MODE>
__global__ __launch_bounds__(64) void dc_between(const KShift* __restrict__ ksp, const T* __restrict__ Fc,
                                                 T* __restrict__ F, int fnx, int fny, int fnz, const PopLayout FN,
                                                 int nfaces, uint32_t hang0, uint32_t hang1) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = fny * fnz;
    if (t >= nfaces * per) return;
    const int face = t / per, r = t - face * per;
    const int j = r % fny, k = r / fny;
    if (!(j & 1) && !(k & 1)) return;
    const int fi = face ? fnx - 1 : 0;
    const uint32_t hm = face ? hang1 : hang0;
    auto at = [&](int jj, int kk) {
        jj = (jj + fny) % fny;
        kk = (kk + fnz) % fnz;
        return fi + (int64_t)fnx * (jj + (int64_t)fny * kk);
    };
    const int64_t l = at(j, k);
    const bool centre = (j & 1) && (k & 1);
    int64_t nb[4];
    if (centre) {
        nb[0] = at(j - 1, k - 1), nb[1] = at(j + 1, k - 1), nb[2] = at(j - 1, k + 1), nb[3] = at(j + 1, k + 1);
    } else {
        nb[0] = (j & 1) ? at(j - 1, k) : at(j, k - 1);
        nb[1] = (j & 1) ? at(j + 1, k) : at(j, k + 1);
        nb[2] = nb[0], nb[3] = nb[1];
    }
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if (!((hm >> q) & 1u)) return;
        const T v0 = Fc[pidx(FN, q, nb[0])], v1 = Fc[pidx(FN, q, nb[1])];
        const T v2 = Fc[pidx(FN, q, nb[2])], v3 = Fc[pidx(FN, q, nb[3])];
        if constexpr (MODE == 0) {
            F[pidx(FN, q, l)] = centre ? ((v0 + v1) + (v2 + v3)) * 0.25 : (v0 + v1) * 0.5;
        } else if constexpr (MODE == 1) {
            const KShift& ks = *ksp;
            const double v = centre ? ((g_to_f(ks, v0, q) + g_to_f(ks, v1, q)) + (g_to_f(ks, v2, q) + g_to_f(ks, v3, q))) * 0.25
                                    : (g_to_f(ks, v0, q) + g_to_f(ks, v1, q)) * 0.5;
            F[pidx(FN, q, l)] = f_to_g<T>(ks, v, q);
        } else {
            const float v = centre ? (((float)v0 + (float)v1) + ((float)v2 + (float)v3)) * 0.25f
                                   : ((float)v0 + (float)v1) * 0.5f;
            F[pidx(FN, q, l)] = (T)v;
        }
    });
}

}  // namespace refine

// The nodes of x column `col` of a float64 level collide with the strain rate
// S (planes [6][ny nz] over (j, k)) and push as the bulk kernel pushes them:
// all 18 links, y and z wrapped, x wrapped when the level is periodic in x and
// otherwise dropped when they leave the grid.  Launched right after the bulk
// kernel, it overwrites what that kernel pushed from these nodes.  KIND 0:
// HRR; with the cubic-Mach terms (KIND 1 HRR, 2 RR) the planes continue with
// the handed psi ([19][ny nz]), which the collision takes and which replaces
// the node's psi state (written there by the bulk kernel).
template <int ORDER, int KIND>
__global__ __launch_bounds__(64) void fix64(const K64* __restrict__ kp, const double* __restrict__ A,
                                            double* __restrict__ B, double* __restrict__ psi, int nx, int ny, int nz,
                                            const PopLayout L, int col, int xwrap, const double* __restrict__ S) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = ny * nz;
    if (t >= per) return;
    const int j = t % ny, kz = t / ny;
    const int64_t l = col + (int64_t)nx * (j + (int64_t)ny * kz);
    double f[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = A[pidx(L, q, l)];
    double s6[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) s6[v] = S[v * per + t];
    if constexpr (KIND == 0) {
        collide_hrr64_strain<ORDER>(*kp, f, s6);
    } else {
        double ps[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) ps[q] = S[(6 + q) * per + t];
        collide_rr64_handed<ORDER, KIND == 1>(*kp, f, ps, s6);
#pragma unroll
        for (int q = 0; q < 19; ++q) psi[q * L.N + l] = ps[q];
    }
    static_for<0, 18>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        int a = col + EX[q];
        if (xwrap) a = (a + nx) % nx;
        else if (a < 0 || a >= nx) return;
        const int b = (j + EY[q] + ny) % ny, d = (kz + EZ[q] + nz) % nz;
        B[pidx(L, q, a + (int64_t)nx * (b + (int64_t)ny * d))] = f[q];
    });
    B[pidx(L, 18, l)] = f[18];
}

// fix64 over a list of nodes (list[t], none on a face of the grid, so every
// push stays inside it), handed values P [nv][n] over the list.
template <int ORDER, int KIND>
__global__ __launch_bounds__(64) void fix64_list(const K64* __restrict__ kp, const double* __restrict__ A,
                                                 double* __restrict__ B, double* __restrict__ psi, int nx, int ny,
                                                 const PopLayout L, const int64_t* __restrict__ list, int64_t n,
                                                 const double* __restrict__ P) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t l = list[t];
    const int64_t sxy = (int64_t)nx * ny;
    double f[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = A[pidx(L, q, l)];
    double s6[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) s6[v] = P[v * n + t];
    if constexpr (KIND == 0) {
        collide_hrr64_strain<ORDER>(*kp, f, s6);
    } else {
        double ps[19];
#pragma unroll
        for (int q = 0; q < 19; ++q) ps[q] = P[(6 + q) * n + t];
        collide_rr64_handed<ORDER, KIND == 1>(*kp, f, ps, s6);
#pragma unroll
        for (int q = 0; q < 19; ++q) psi[q * L.N + l] = ps[q];
    }
    static_for<0, 18>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        B[pidx(L, q, l + EX[q] + (int64_t)nx * EY[q] + sxy * EZ[q])] = f[q];
    });
    B[pidx(L, 18, l)] = f[18];
}

hipError_t launch_fix64_list(int op, int order, bool cubic, const K64* k, const double* A, double* B, double* psi,
                             const Grid& g, const int64_t* list, int64_t n, const double* P, hipStream_t s) {
    const dim3 grid((unsigned)((n + 63) / 64)), block(64);
    const int kind = !cubic ? 0 : (op == OP_HRR ? 1 : 2);
    if (kind == 0 && op != OP_HRR) return hipErrorInvalidValue;
#define FIXL(O, KD) hipLaunchKernelGGL((fix64_list<O, KD>), grid, block, 0, s, k, A, B, psi, g.nx, g.ny, g.L, list, n, P)
#define FIXLK(O) \
    if (kind == 0) FIXL(O, 0); \
    else if (kind == 1) FIXL(O, 1); \
    else FIXL(O, 2)
    if (order == 2) { FIXLK(2); }
    else if (order == 3) { FIXLK(3); }
    else { FIXLK(4); }
#undef FIXLK
#undef FIXL
    return hipGetLastError();
}

namespace refine {

// ---- the reference's two-level geometry (MaxLevel 1; createLattice,
// Lattice.cpp:6-213): a fine shell around a coarse core.  The fine level
// covers the domain (2 nx - 1 nodes per row), the coarse level too; the fine
// grid-coupling nodes are the surface of the hole box hb, the coarse
// grid-coupling nodes the surface of the core box cb, one coarse cell outside
// it (Overlap 1).  The kernels run over node lists built on the host: the fine
// grid-coupling nodes (coincident ones first) with their Coarse-to-Fine and
// hanging link masks, the restricted coarse nodes (the core box's surface
// first) with their Fine-to-Coarse masks.
struct Box6 {
    int b[6];
};
__device__ __forceinline__ bool box_surface(const Box6& B, int i, int j, int k) {
    const bool in = i >= B.b[0] && i <= B.b[1] && j >= B.b[2] && j <= B.b[3] && k >= B.b[4] && k <= B.b[5];
    const bool inner = i > B.b[0] && i < B.b[1] && j > B.b[2] && j < B.b[3] && k > B.b[4] && k < B.b[5];
    return in && !inner;
}

// coincident fine grid-coupling nodes from their coarse node (time
// interpolation, rescaling), the links of their Coarse-to-Fine mask
template <int ORDER>
__global__ __launch_bounds__(64) void shell_c2f_co(const K64* __restrict__ kp, const double* __restrict__ Cprev,
                                                   const double* __restrict__ Ccur, int half, const NeqScale a, int cnx,
                                                   int cny, const PopLayout CL, double* __restrict__ F, int fnx,
                                                   int fny, const PopLayout FL, const int64_t* __restrict__ fs,
                                                   const uint32_t* __restrict__ mask, int64_t nfc) {
    __shared__ K64 sk;
    const K64& k = stage_k64(kp, sk);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfc) return;
    const int64_t l = fs[t];
    const int i = (int)(l % fnx), j = (int)((l / fnx) % fny), kk = (int)(l / ((int64_t)fnx * fny));
    const int64_t lc = i / 2 + (int64_t)cnx * (j / 2 + (int64_t)cny * (kk / 2));
    double f[19];
    c2f_value<ORDER>(k, Cprev, Ccur, CL, lc, half, a, f);
    const uint32_t m = mask[t];
#pragma unroll
    for (int q = 0; q < 19; ++q)
        if ((m >> q) & 1u) F[pidx(FL, q, l)] = f[q];
}

// the other fine grid-coupling nodes: the mean of their two (one odd index)
// or four (two odd indices) coincident neighbours along the surface, in the
// order (-,-), (+,-), (-,+), (+,+) of their odd axes, over their hanging mask;
// Fc and F the same array (coincident nodes read, the others written)
__global__ __launch_bounds__(64) void shell_c2f_between(const double* __restrict__ Fc, double* __restrict__ F, int fnx,
                                                        int fny, const PopLayout FL, const int64_t* __restrict__ fs,
                                                        const uint32_t* __restrict__ hmask, int64_t nfc, int64_t nfs) {
    const int64_t t = nfc + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfs) return;
    const int64_t l = fs[t];
    const int x[3] = {(int)(l % fnx), (int)((l / fnx) % fny), (int)(l / ((int64_t)fnx * fny))};
    const int64_t st[3] = {1, fnx, (int64_t)fnx * fny};
    int odd[2] = {0, 0}, no = 0;
    for (int d = 0; d < 3; ++d)
        if (x[d] & 1) odd[no++] = d;
    const uint32_t m = hmask[t];
    if (no == 1) {
        const int64_t a = l - st[odd[0]], b = l + st[odd[0]];
#pragma unroll
        for (int q = 0; q < 19; ++q)
            if ((m >> q) & 1u) F[pidx(FL, q, l)] = (Fc[pidx(FL, q, a)] + Fc[pidx(FL, q, b)]) * 0.5;
    } else {
        const int64_t a = l - st[odd[0]] - st[odd[1]], b = l + st[odd[0]] - st[odd[1]];
        const int64_t c = l - st[odd[0]] + st[odd[1]], d = l + st[odd[0]] + st[odd[1]];
#pragma unroll
        for (int q = 0; q < 19; ++q)
            if ((m >> q) & 1u)
                F[pidx(FL, q, l)] =
                    ((Fc[pidx(FL, q, a)] + Fc[pidx(FL, q, b)]) + (Fc[pidx(FL, q, c)] + Fc[pidx(FL, q, d)])) * 0.25;
    }
}

// the restricted coarse nodes from their coincident fine node (F2C-Filter
// NOFI; LAG and TOU are f2c_lag over the same list), over their masks
template <int ORDER>
__global__ __launch_bounds__(256) void shell_f2c(const K64* __restrict__ kp, const double* __restrict__ F, int fnx,
                                                 int fny, const NeqScale ainv, double* __restrict__ C, int cnx,
                                                 int cny, const PopLayout CL, const PopLayout FL,
                                                 const int64_t* __restrict__ cr, const uint32_t* __restrict__ mask,
                                                 int64_t ncr) {
    __shared__ K64 sk;
    const K64& k = stage_k64(kp, sk);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncr) return;
    const int64_t lc = cr[t];
    const int I = (int)(lc % cnx), J = (int)((lc / cnx) % cny), K = (int)(lc / ((int64_t)cnx * cny));
    const int64_t lf = 2 * I + (int64_t)fnx * (2 * J + (int64_t)fny * (2 * K));
    double f[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) f[q] = F[pidx(FL, q, lf)];
    rescale<ORDER>(k, f, ainv);
    const uint32_t m = mask[t];
#pragma unroll
    for (int q = 0; q < 19; ++q)
        if ((m >> q) & 1u) C[pidx(CL, q, lc)] = f[q];
}

// HRR (and the cubic-Mach terms): the coarse partners of the coincident fine
// grid-coupling nodes (Node.cpp:2114-2380), their neighbours on the core box's
// surface read through their fine partners; HALF at t + 1/2 (the other
// neighbours the mean of t and t + 1).  Into S [nv][nfs] at the node's list
// position.
template <bool HALF>
__global__ __launch_bounds__(64) void shell_partner_coarse(const K64* __restrict__ kcp, const double* __restrict__ Cprev,
                                                           const double* __restrict__ Ccur, const PopLayout CL, int cnx,
                                                           int cny, const double* __restrict__ F, const PopLayout FL,
                                                           int fnx, int fny, const int64_t* __restrict__ fs,
                                                           int64_t nfc, int64_t nfs, const Box6 cb, int nv,
                                                           double* __restrict__ S) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfc) return;
    const int64_t l = fs[t];
    const int P[3] = {(int)(l % fnx) / 2, (int)((l / fnx) % fny) / 2, (int)(l / ((int64_t)fnx * fny)) / 2};
    const K64& k = *kcp;
    double nb[6][4];
#pragma unroll
    for (int m = 0; m < 6; ++m) {
        const int a = P[0] + EX[m], b = P[1] + EY[m], d = P[2] + EZ[m];
        if (box_surface(cb, a, b, d)) {
            macros_si<double>(k, nullptr, F, FL, 2 * a + (int64_t)fnx * (2 * b + (int64_t)fny * (2 * d)), nb[m]);
            continue;
        }
        const int64_t lc = a + (int64_t)cnx * (b + (int64_t)cny * d);
        if constexpr (HALF) {
            double p4[4], c4[4];
            macros_si<double>(k, nullptr, Cprev, CL, lc, p4);
            macros_si<double>(k, nullptr, Ccur, CL, lc, c4);
#pragma unroll
            for (int v = 0; v < 4; ++v) nb[m][v] = 0.5 * (p4[v] + c4[v]);
        } else {
            macros_si<double>(k, nullptr, Ccur, CL, lc, nb[m]);
        }
    }
    store_handed(k, nb, nv, S, (int)nfs, (int)t);
}

// the other fine grid-coupling nodes: the mean of their coincident
// neighbours' handed values (list positions nbpos [4] per node, -1 unused)
__global__ __launch_bounds__(256) void shell_spread(const int32_t* __restrict__ nbpos, int64_t nfc, int64_t nfs, int nv,
                                                    double* __restrict__ S) {
    const int64_t t = nfc + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nfs) return;
    const int32_t* p = nbpos + 4 * (t - nfc);
    for (int v = 0; v < nv; ++v) {
        const double* V = S + (int64_t)v * nfs;
        S[(int64_t)v * nfs + t] = p[2] < 0 ? (V[p[0]] + V[p[1]]) * 0.5 : ((V[p[0]] + V[p[1]]) + (V[p[2]] + V[p[3]])) * 0.25;
    }
}

// the fine partners of the core box's surface (Node.cpp:2383-2490): fine node
// 2 Q from its six fine neighbours; into S [nv][ncs]
__global__ __launch_bounds__(64) void shell_partner_fine(const K64* __restrict__ kfp, const double* __restrict__ F,
                                                         const PopLayout FL, int fnx, int fny, int cnx, int cny,
                                                         const int64_t* __restrict__ cr, int64_t ncs, int nv,
                                                         double* __restrict__ S) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncs) return;
    const int64_t lc = cr[t];
    const int I = (int)(lc % cnx), J = (int)((lc / cnx) % cny), K = (int)(lc / ((int64_t)cnx * cny));
    const K64& k = *kfp;
    double nb[6][4];
#pragma unroll
    for (int m = 0; m < 6; ++m)
        macros_si<double>(k, nullptr, F, FL,
                          2 * I + EX[m] + (int64_t)fnx * (2 * J + EY[m] + (int64_t)fny * (2 * K + EZ[m])), nb[m]);
    store_handed(k, nb, nv, S, (int)ncs, (int)t);
}

}  // namespace refine

hipError_t launch_fix64(int op, int order, bool cubic, const K64* k, const double* A, double* B, double* psi,
                        const Grid& g, int col, const double* S, hipStream_t s) {
    const int per = g.ny * g.nz;
    const dim3 grid((unsigned)((per + 63) / 64)), block(64);
    const int xwrap = !g.xopen && !g.xlodi;
    const int kind = !cubic ? 0 : (op == OP_HRR ? 1 : 2);
    if (kind == 0 && op != OP_HRR) return hipErrorInvalidValue;
#define FIX(O, KD) \
    hipLaunchKernelGGL((fix64<O, KD>), grid, block, 0, s, k, A, B, psi, g.nx, g.ny, g.nz, g.L, col, xwrap, S)
#define FIXK(O) \
    if (kind == 0) FIX(O, 0); \
    else if (kind == 1) FIX(O, 1); \
    else FIX(O, 2)
    if (order == 2) { FIXK(2); }
    else if (order == 3) { FIXK(3); }
    else { FIXK(4); }
#undef FIXK
#undef FIX
    return hipGetLastError();
}

}  // namespace acp

struct acp_refined {
    acp_case* coarse = nullptr;
    acp_case* fine = nullptr;
    int x0 = 0, x1 = 0;
    double a = 1.0;        // tau_f / (2 tau_c)
    bool mrt_style = false;  // Scaling-Style MRTSTYLE (else SRTSTYLE)
    bool f64_coupling = true;  // shifted tiers: coupling arithmetic in float64 (else float32)
    int filter = ACP_F2C_NOFI;  // fine-to-coarse filter (F2C-Filter)
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int32_t graph_chunk = 0;              // coarse steps per captured graph (0: plain launches)
    hipGraphExec_t gexec[4] = {nullptr, nullptr, nullptr, nullptr};  // by coarse + 2 * fine parity
    uint64_t epoch[2] = {0, 0};  // the levels' graph epochs the captured graphs hold
    // HRR strain-rate transfer, with the cubic-Mach terms also for RR (per
    // coupled face, float64 planes of nv components: S[6], then psi[19] with
    // the cubic-Mach terms): s_col at the coarse partners of the fine face
    // nodes [nf][nv][cny cnz], s_face at the fine face nodes [nf][nv][fny fnz],
    // s_gc at the coarse grid-coupling columns [nf][nv][cny cnz]; null when
    // nothing is handed
    int nv = 6;
    double* s_col = nullptr;
    double* s_face = nullptr;
    double* s_gc = nullptr;
    // Coupling-Approach: ACP_COUPLING_STD or ACP_COUPLING_DC with its link
    // sets (ACP_LINKS_*: Coarse-to-Fine, Fine-to-Coarse, Hanging-Node-Reconstruction)
    int scheme = ACP_COUPLING_STD;
    int c2f_set = ACP_LINKS_ALL, f2c_set = ACP_LINKS_ALL, hang_set = ACP_LINKS_ALL;
    // the reference's geometry, a fine shell around a coarse core
    // (acp_refined_create_shell): the fine hole box hb and the coarse core box
    // cb; on the host the fine grid-coupling nodes (coincident first) and the
    // restricted coarse nodes (the core box's surface first), on the device the
    // same lists with their link masks, the hanging nodes' coincident
    // neighbours (list positions) and the handed values [nv][nfs] + [nv][ncs]
    bool shell = false;
    acp::refine::Box6 hb{}, cb{};
    std::vector<int64_t> h_fs, h_cr;
    int64_t nfs = 0, nfc = 0, ncr = 0, ncs = 0;
    int64_t* d_fs = nullptr;
    uint32_t* d_fs_mask = nullptr;  // [2][nfs]: Coarse-to-Fine, then hanging
    int32_t* d_fs_nb = nullptr;     // [nfs - nfc][4]
    int64_t* d_cr = nullptr;
    uint32_t* d_cr_mask = nullptr;
    double* d_sh = nullptr;
};

namespace {

using acp::fail;

// The fine faces the coarse level sets: both, or only the one at x0 when the
// fine level's high-x face is a LODI face.
int coupled_faces(const acp_refined* r) { return (r->fine->g.xlodi & 2) ? 1 : 2; }

// The Scaling-Style factors of the standard coupling, coarse to fine or fine
// to coarse (the Gram-Schmidt moments the MRT-GS operator relaxes with omega,
// Node.cpp:2815-2825, take the SRTSTYLE factor; the conserved ones 0; the
// others 1/2, or 2 fine to coarse).
acp::refine::NeqScale neq_scale(const acp_refined* r, bool f2c) {
    acp::refine::NeqScale s{};
    s.mrt = r->mrt_style ? 1 : 0;
    s.a = f2c ? 1.0 / r->a : r->a;
    for (int m = 0; m < 19; ++m) {
        const bool conserved = m == 0 || m == 3 || m == 5 || m == 7;
        const bool omega = m == 1 || m == 9 || m == 11 || m == 13 || m == 14 || m == 15;
        s.m[m] = conserved ? 0. : (omega ? s.a : (f2c ? 2.0 : 0.5));
    }
    return s;
}

bool dc_link(int set, int q, int n);

// The standard coupling's link sets (see StdLinks); every link under the
// direct coupling, whose pair exchange hands the interface column its links.
acp::refine::StdLinks std_links(const acp_refined* r) {
    acp::refine::StdLinks l{};
    const bool std_ = r->scheme == ACP_COUPLING_STD;
    for (int s = 0; s < 2; ++s) {
        const int n = s == 0 ? 1 : -1;
        for (int q = 0; q < 19; ++q) {
            if (!std_ || dc_link(r->c2f_set, q, n)) l.face[s] |= 1u << q;
            if (!std_ || dc_link(r->hang_set, q, n)) l.hang[s] |= 1u << q;
            if (dc_link(r->f2c_set, q, -n)) l.gcm[s] |= 1u << q;
        }
        l.gc[s] = -1;
    }
    if (std_ && r->f2c_set != ACP_LINKS_ALL) {
        const int cnx = r->coarse->g.nx;
        l.gc[0] = (r->x0 + 1) % cnx;
        if (coupled_faces(r) == 2) l.gc[1] = (r->x1 - 1) % cnx;
    }
    return l;
}

// Coarse-to-fine on the coupled fine faces: coincident nodes, then the
// others.  These are a few thousand latency-bound threads each: 64-thread
// workgroups spread them over more CUs.
template <typename T, bool F64>
hipError_t launch_c2f_s(acp_refined* r, int half) {
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    const int cnx = c->g.nx, cny = c->g.ny, cnz = c->g.nz;
    const int fnx = f->g.nx, fny = f->g.ny, fnz = f->g.nz;
    const int nf = coupled_faces(r);
    const acp::refine::StdLinks lk = std_links(r);
    const dim3 b(64), g1((unsigned)((nf * cny * cnz + 63) / 64)), g2((unsigned)((nf * fny * fnz + 63) / 64));
#define C2FS(O)                                                                                                  \
    hipLaunchKernelGGL((acp::refine::c2f_coincident_s<T, O, F64>), g1, b, 0, r->stream, c->dk, c->dks,            \
                       (const T*)c->SB, (const T*)c->SA, half, neq_scale(r, false), cnx, cny, cnz, r->x0 % cnx,   \
                       r->x1 % cnx,                                                                               \
                       (T*)f->SA, fnx, fny, c->g.L, f->g.L, nf, lk)
    if (c->d.order == 2) C2FS(2);
    else if (c->d.order == 3) C2FS(3);
    else C2FS(4);
#undef C2FS
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((acp::refine::c2f_between_s<T, F64>), g2, b, 0, r->stream, f->dks, (const T*)f->SA,
                       (T*)f->SA, fnx, fny, fnz, f->g.L, nf, lk);
    return hipGetLastError();
}

template <typename T, bool F64>
hipError_t launch_f2c_s(acp_refined* r) {
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    const int width = r->x1 - r->x0 - 1;
    const int64_t n = (int64_t)width * c->g.ny * c->g.nz;
    const acp::refine::StdLinks lk = std_links(r);
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
#define F2CS(O, FI)                                                                                              \
    hipLaunchKernelGGL((acp::refine::f2c_s<T, O, F64, FI>), g, b, 0, r->stream, c->dk, c->dks, (const T*)f->SA, f->g.nx, \
                       f->g.ny, f->g.nz, neq_scale(r, true), (T*)c->SA, c->g.nx, c->g.ny, c->g.nz, r->x0, width,   \
                       c->g.L, f->g.L, lk)
#define F2CS_O(FI)                       \
    if (c->d.order == 2) F2CS(2, FI);    \
    else if (c->d.order == 3) F2CS(3, FI); \
    else F2CS(4, FI)
    if (r->filter == ACP_F2C_LAG) {
        F2CS_O(ACP_F2C_LAG);
    } else if (r->filter == ACP_F2C_TOU) {
        F2CS_O(ACP_F2C_TOU);
    } else {
        F2CS_O(ACP_F2C_NOFI);
    }
#undef F2CS_O
#undef F2CS
    return hipGetLastError();
}

// ---- the shell geometry's launches (float64)
hipError_t launch_shell_c2f(acp_refined* r, int half) {
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    const int fnx = f->g.nx, fny = f->g.ny;
    const dim3 b(64), g1((unsigned)((r->nfc + 63) / 64)), g2((unsigned)((r->nfs - r->nfc + 63) / 64));
#define SC2F(O)                                                                                                   \
    hipLaunchKernelGGL((acp::refine::shell_c2f_co<O>), g1, b, 0, r->stream, c->dk, c->B, c->A, half,             \
                       neq_scale(r, false), c->g.nx, c->g.ny, c->g.L, f->A, fnx, fny, f->g.L, r->d_fs, r->d_fs_mask, \
                       r->nfc)
    if (c->d.order == 2) SC2F(2);
    else if (c->d.order == 3) SC2F(3);
    else SC2F(4);
#undef SC2F
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || r->nfs == r->nfc) return e;
    hipLaunchKernelGGL(acp::refine::shell_c2f_between, g2, b, 0, r->stream, (const double*)f->A, f->A, fnx, fny,
                       f->g.L, r->d_fs, r->d_fs_mask + r->nfs, r->nfc, r->nfs);
    return hipGetLastError();
}

hipError_t launch_shell_f2c(acp_refined* r) {
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    const int64_t n = r->ncr;
    if (n == 0) return hipSuccess;
    const acp::refine::StdLinks lk{};
    const dim3 b(256);
#define SF2C(O)                                                                                                   \
    hipLaunchKernelGGL((acp::refine::shell_f2c<O>), dim3((unsigned)((n + 255) / 256)), b, 0, r->stream, c->dk,    \
                       f->A, f->g.nx, f->g.ny, neq_scale(r, true), c->A, c->g.nx, c->g.ny, c->g.L, f->g.L, r->d_cr, \
                       r->d_cr_mask, n)
#define SLAG(O, TOU)                                                                                              \
    hipLaunchKernelGGL((acp::refine::f2c_lag<O, TOU, true>), dim3((unsigned)((n + acp::refine::kLagNodes - 1) /  \
                                                                              acp::refine::kLagNodes)),          \
                       b, 0, r->stream, c->dk, f->A, f->g.nx, f->g.ny, f->g.nz, neq_scale(r, true), c->A, c->g.nx, \
                       c->g.ny, c->g.nz, 0, 0, c->g.L, f->g.L, lk, r->d_cr, r->d_cr_mask, n)
    const int o = c->d.order;
    if (r->filter == ACP_F2C_NOFI) {
        if (o == 2) SF2C(2);
        else if (o == 3) SF2C(3);
        else SF2C(4);
    } else if (r->filter == ACP_F2C_LAG) {
        if (o == 2) SLAG(2, false);
        else if (o == 3) SLAG(3, false);
        else SLAG(4, false);
    } else {
        if (o == 2) SLAG(2, true);
        else if (o == 3) SLAG(3, true);
        else SLAG(4, true);
    }
#undef SLAG
#undef SF2C
    return hipGetLastError();
}

hipError_t launch_shell_transfer(acp_refined* r, int half) {
    if (!r->d_sh) return hipSuccess;
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    const int fnx = f->g.nx, fny = f->g.ny, cnx = c->g.nx, cny = c->g.ny;
    double* Sface = r->d_sh;
    double* Sgc = r->d_sh + (int64_t)r->nv * r->nfs;
    const dim3 b64(64), b256(256), g1((unsigned)((r->nfc + 63) / 64));
    if (half)
        hipLaunchKernelGGL((acp::refine::shell_partner_coarse<true>), g1, b64, 0, r->stream, c->dk, c->B, c->A, c->g.L,
                           cnx, cny, f->A, f->g.L, fnx, fny, r->d_fs, r->nfc, r->nfs, r->cb, r->nv, Sface);
    else
        hipLaunchKernelGGL((acp::refine::shell_partner_coarse<false>), g1, b64, 0, r->stream, c->dk, c->A, c->A,
                           c->g.L, cnx, cny, f->A, f->g.L, fnx, fny, r->d_fs, r->nfc, r->nfs, r->cb, r->nv, Sface);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && r->nfs > r->nfc) {
        hipLaunchKernelGGL(acp::refine::shell_spread, dim3((unsigned)((r->nfs - r->nfc + 255) / 256)), b256, 0,
                           r->stream, r->d_fs_nb, r->nfc, r->nfs, r->nv, Sface);
        e = hipGetLastError();
    }
    if (e != hipSuccess || half || r->ncs == 0) return e;
    hipLaunchKernelGGL(acp::refine::shell_partner_fine, dim3((unsigned)((r->ncs + 63) / 64)), b64, 0, r->stream, f->dk,
                       f->A, f->g.L, fnx, fny, cnx, cny, r->d_cr, r->ncs, r->nv, Sgc);
    return hipGetLastError();
}

hipError_t launch_c2f(acp_refined* r, int half) {
    if (r->shell) return launch_shell_c2f(r, half);
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    if (c->precision == ACP_PREC_F32_SHIFTED)
        return r->f64_coupling ? launch_c2f_s<float, true>(r, half) : launch_c2f_s<float, false>(r, half);
    if (c->precision == ACP_PREC_F16_SHIFTED)
        return r->f64_coupling ? launch_c2f_s<_Float16, true>(r, half) : launch_c2f_s<_Float16, false>(r, half);
    const int cnx = c->g.nx, cny = c->g.ny, cnz = c->g.nz;
    const int fnx = f->g.nx, fny = f->g.ny, fnz = f->g.nz;
    const int nf = coupled_faces(r);
    const acp::refine::StdLinks lk = std_links(r);
    const dim3 b(64), g1((unsigned)((nf * cny * cnz + 63) / 64)), g2((unsigned)((nf * fny * fnz + 63) / 64));
#define C2F(O)                                                                                                    \
    hipLaunchKernelGGL((acp::refine::c2f_coincident<O>), g1, b, 0, r->stream, c->dk, c->B, c->A, half,           \
                       neq_scale(r, false), cnx,                                                                  \
                       cny, cnz, r->x0 % cnx, r->x1 % cnx, f->A, fnx, fny, c->g.L, f->g.L, nf, lk)
    if (c->d.order == 2) C2F(2);
    else if (c->d.order == 3) C2F(3);
    else C2F(4);
#undef C2F
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(acp::refine::c2f_between, g2, b, 0, r->stream, (const double*)f->A, f->A, fnx, fny, fnz,
                       f->g.L, nf, lk);
    return hipGetLastError();
}

hipError_t launch_f2c(acp_refined* r) {
    if (r->shell) return launch_shell_f2c(r);
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    const int width = r->x1 - r->x0 - 1;
    if (width > 0 && c->precision == ACP_PREC_F32_SHIFTED)
        return r->f64_coupling ? launch_f2c_s<float, true>(r) : launch_f2c_s<float, false>(r);
    if (width > 0 && c->precision == ACP_PREC_F16_SHIFTED)
        return r->f64_coupling ? launch_f2c_s<_Float16, true>(r) : launch_f2c_s<_Float16, false>(r);
    if (width <= 0) return hipSuccess;
    const int64_t n = (int64_t)width * c->g.ny * c->g.nz;
    const acp::refine::StdLinks lk = std_links(r);
    const dim3 g((unsigned)((n + 255) / 256)), b(256);
#define F2C(O)                                                                                                  \
    hipLaunchKernelGGL((acp::refine::f2c<O>), g, b, 0, r->stream, c->dk, f->A, f->g.nx, f->g.ny, f->g.nz,       \
                       neq_scale(r, true),                                                                        \
                       c->A, c->g.nx, c->g.ny, c->g.nz, r->x0, width, c->g.L, f->g.L, lk)
#define F2CLAG(O, TOU)                                                                                           \
    hipLaunchKernelGGL((acp::refine::f2c_lag<O, TOU, false>), dim3((unsigned)((n + acp::refine::kLagNodes - 1) / \
                                                                               acp::refine::kLagNodes)),         \
                       b, 0, r->stream, c->dk, f->A, f->g.nx, f->g.ny, f->g.nz, neq_scale(r, true), c->A, c->g.nx, \
                       c->g.ny,                                                                                   \
                       c->g.nz, r->x0, width, c->g.L, f->g.L, lk, nullptr, nullptr, (int64_t)0)
    if (r->filter != ACP_F2C_NOFI) {
        const bool tou = r->filter == ACP_F2C_TOU;
        if (c->d.order == 2) {
            if (tou) F2CLAG(2, true);
            else F2CLAG(2, false);
        } else if (c->d.order == 3) {
            if (tou) F2CLAG(3, true);
            else F2CLAG(3, false);
        } else {
            if (tou) F2CLAG(4, true);
            else F2CLAG(4, false);
        }
    } else {
        if (c->d.order == 2) F2C(2);
        else if (c->d.order == 3) F2C(3);
        else F2C(4);
    }
#undef F2C
#undef F2CLAG
    return hipGetLastError();
}

// ---- direct coupling: per-face link sets and chord matrices
// link q belongs to `set` on the face whose links arriving from outside the
// fine region have e_x = n
bool dc_link(int set, int q, int n) {
    return set == ACP_LINKS_ALL || acp::EX[q] == n || (set == ACP_LINKS_MISSING_COMMON && acp::EX[q] == 0);
}

// The inverse of I - b J, J the derivative at u = 0 of the mass and momentum
// of feq over the coarse node's fine-side links with respect to (rho, rho u),
// for lattice speed c and squared sound speed cs2 (the set is symmetric in y
// and z, so J couples rho with rho u_x only).
void dc_matrix(double c, double cs2, double a, int set, int n, double m[6]) {
    const double b = 1. - 1. / a;
    double W = 0., V = 0., Tx = 0., Ty = 0., Tz = 0.;
    for (int q = 0; q < 19; ++q) {
        if (!dc_link(set, q, -n)) continue;
        const double w = acp::weight(q);
        const double x = acp::EX[q] == 0 ? 0. : (acp::EX[q] > 0 ? c : -c);
        const double y = acp::EY[q] == 0 ? 0. : (acp::EY[q] > 0 ? c : -c);
        const double z = acp::EZ[q] == 0 ? 0. : (acp::EZ[q] > 0 ? c : -c);
        W += w;
        V += w * x;
        Tx += w * (x * x);
        Ty += w * (y * y);
        Tz += w * (z * z);
    }
    const double a00 = 1. - b * W, a01 = -(b * V) / cs2, a10 = -(b * V), a11 = 1. - (b * Tx) / cs2;
    const double det = a00 * a11 - a01 * a10;
    m[0] = a11 / det, m[1] = -a01 / det, m[2] = -a10 / det, m[3] = a00 / det;
    m[4] = 1. / (1. - (b * Ty) / cs2);
    m[5] = 1. / (1. - (b * Tz) / cs2);
}

acp::refine::DcArgs dc_args(const acp_refined* r) {
    acp::refine::DcArgs d{};
    d.a = r->a;
    d.ainv = 1.0 / r->a;
    for (int s = 0; s < 2; ++s) {
        const int n = s == 0 ? 1 : -1;
        acp::refine::DcFace& f = d.face[s];
        for (int q = 0; q < 19; ++q) {
            if (dc_link(r->c2f_set, q, n)) f.fine |= 1u << q;
            if (dc_link(r->f2c_set, q, -n)) f.coarse |= 1u << q;
            if (dc_link(r->hang_set, q, n)) f.hang |= 1u << q;
        }
        dc_matrix(r->coarse->hk.c, r->coarse->hk.cs2, r->a, r->f2c_set, n, f.m);
        double ml[6];
        dc_matrix(1., 1. / 3., r->a, r->f2c_set, n, ml);
        for (int i = 0; i < 6; ++i) f.mf[i] = (float)ml[i];
    }
    return d;
}

// The direct coupling on the coupled faces: coincident pairs (half: at
// t + 1/2, only the fine nodes receive), then the hanging fine face nodes.
template <typename T, int MODE>
hipError_t launch_dc_t(acp_refined* r, int half) {
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    const bool shifted = c->precision != ACP_PREC_F64;
    T* Ccur = (T*)(shifted ? c->SA : (void*)c->A);
    const T* Cprev = (const T*)(shifted ? c->SB : (const void*)c->B);
    T* F = (T*)(shifted ? f->SA : (void*)f->A);
    const int cnx = c->g.nx, cny = c->g.ny, cnz = c->g.nz, fnx = f->g.nx, fny = f->g.ny, fnz = f->g.nz;
    const int nf = coupled_faces(r);
    const acp::refine::DcArgs args = dc_args(r);
    const dim3 b(64), g1((unsigned)((nf * cny * cnz + 63) / 64)), g2((unsigned)((nf * fny * fnz + 63) / 64));
#define DC(O)                                                                                                     \
    hipLaunchKernelGGL((acp::refine::dc_coincident<T, O, MODE>), g1, b, 0, r->stream, c->dk, c->dks, Cprev, Ccur, half, \
                       args, cnx, cny, cnz, r->x0 % cnx, r->x1 % cnx, F, fnx, fny, c->g.L, f->g.L, nf)
    if (c->d.order == 2) DC(2);
    else if (c->d.order == 3) DC(3);
    else DC(4);
#undef DC
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL((acp::refine::dc_between<T, MODE>), g2, b, 0, r->stream, c->dks, (const T*)F, F, fnx, fny, fnz,
                       f->g.L, nf, args.face[0].hang, args.face[1].hang);
    return hipGetLastError();
}

hipError_t launch_dc(acp_refined* r, int half) {
    switch (r->coarse->precision) {
        case ACP_PREC_F64: return launch_dc_t<double, 0>(r, half);
        case ACP_PREC_F32_SHIFTED:
            return r->f64_coupling ? launch_dc_t<float, 1>(r, half) : launch_dc_t<float, 2>(r, half);
        default: return r->f64_coupling ? launch_dc_t<_Float16, 1>(r, half) : launch_dc_t<_Float16, 2>(r, half);
    }
}

// The strain rates the interface columns collide with.  half = 0 at a coarse
// step's start (states at t), half = 1 after the coarse step and the first
// fine step (coarse at t and t + 1, fine at t + 1/2): the coarse partners'
// strain rates and their interpolation onto the fine faces; at the start also
// the fine partners' for the coarse grid-coupling columns (standard coupling;
// under the direct coupling the coarse interface columns collide with the
// coarse partners' own, Node.cpp:1892-2111 on level 0 reading the missing
// inner neighbour through the fine partner's second neighbour as the
// coarse-partner function does).  Under the direct coupling half = 1 runs
// after the exchange at t + 1/2 and takes the level-1 stencil.
template <typename T>
hipError_t launch_transfer_t(acp_refined* r, int half) {
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    const bool shifted = c->precision != ACP_PREC_F64;
    const T* Ccur = (const T*)(shifted ? c->SA : (const void*)c->A);
    const T* Cprev = (const T*)(shifted ? c->SB : (const void*)c->B);
    const T* F = (const T*)(shifted ? f->SA : (const void*)f->A);
    const int cnx = c->g.nx, cny = c->g.ny, cnz = c->g.nz, fnx = f->g.nx, fny = f->g.ny, fnz = f->g.nz;
    const int nf = coupled_faces(r);
    const dim3 b64(64), gc((unsigned)((nf * cny * cnz + 63) / 64));
    const dim3 gface((unsigned)((nf * fny * fnz + 255) / 256)), b256(256);
    if (half && r->scheme == ACP_COUPLING_DC) {
        // the level-1 stencil into the planes s_gc (no grid-coupling column
        // inside the fine region under the direct coupling), spread over the faces
        hipLaunchKernelGGL((acp::refine::dc_strain_fine<T>), gc, b64, 0, r->stream, c->dk, c->dks, f->dks, Cprev, Ccur,
                           c->g.L, cnx, cny, cnz, r->x0 % cnx, r->x1 % cnx, F, f->g.L, fnx, fny, fnz, nf, r->nv,
                           r->s_gc);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(acp::refine::face_strain, gface, b256, 0, r->stream, (const double*)r->s_gc, cny, cnz, fny,
                           fnz, nf, r->nv, r->s_face);
        return hipGetLastError();
    }
    if (half)
        hipLaunchKernelGGL((acp::refine::partner_coarse<T, true>), gc, b64, 0, r->stream, c->dk, c->dks, f->dks, Cprev,
                           Ccur, c->g.L, cnx, cny, cnz, r->x0 % cnx, r->x1 % cnx, F, f->g.L, fnx, fny, 2, fnx - 3, nf,
                           r->nv, r->s_col);
    else
        hipLaunchKernelGGL((acp::refine::partner_coarse<T, false>), gc, b64, 0, r->stream, c->dk, c->dks, f->dks, Ccur,
                           Ccur, c->g.L, cnx, cny, cnz, r->x0 % cnx, r->x1 % cnx, F, f->g.L, fnx, fny, 2, fnx - 3, nf,
                           r->nv, r->s_col);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(acp::refine::face_strain, gface, b256, 0, r->stream, (const double*)r->s_col, cny, cnz, fny, fnz,
                       nf, r->nv, r->s_face);
    e = hipGetLastError();
    // under the direct coupling the coarse interface column collides with the
    // coarse partners' strain rates themselves (s_col)
    if (e != hipSuccess || half || r->scheme == ACP_COUPLING_DC) return e;
    hipLaunchKernelGGL((acp::refine::partner_fine<T>), gc, b64, 0, r->stream, f->dk, f->dks, F, f->g.L, fnx, fny, fnz, 2,
                       fnx - 3, cny, cnz, nf, r->nv, r->s_gc);
    return hipGetLastError();
}

hipError_t launch_transfer(acp_refined* r, int half) {
    if (r->shell) return launch_shell_transfer(r, half);
    if (!r->s_col) return hipSuccess;
    switch (r->coarse->precision) {
        case ACP_PREC_F64: return launch_transfer_t<double>(r, half);
        case ACP_PREC_F32_SHIFTED: return launch_transfer_t<float>(r, half);
        default: return launch_transfer_t<_Float16>(r, half);
    }
}

// Whether the interface columns collide with values handed by the other
// level: HRR's strain rate, and with the cubic-Mach terms psi (RR and HRR).
bool handed(const acp_case_desc& d) { return d.op == ACP_OP_HRR || (d.op == ACP_OP_RR && d.cubic); }
// planes per interface node: S[6], then psi[19] with the cubic-Mach terms
int handed_planes(const acp_case_desc& d) { return d.cubic ? 25 : 6; }

// Bytes of the transfer buffers of a refined case (every buffer a multiple of
// 256 bytes; none when nothing is handed).
int64_t transfer_bytes(const acp_case_desc& d, const acp_case_desc& fd) {
    if (!handed(d)) return 0;
    const int64_t nf = fd.boundary == ACP_BC_OPEN_LODI_X ? 1 : 2, nv = handed_planes(d);
    const int64_t col = acp::round_alloc(nf * nv * d.ny * d.nz * 8), face = acp::round_alloc(nf * nv * fd.ny * fd.nz * 8);
    return 2 * col + face;
}

}  // namespace

namespace {

// The coarse description's checks and the fine level's description.
int fine_desc(const acp_case_desc& d, int32_t x0, int32_t x1, acp_case_desc& fd) {
    if (d.boundary != ACP_BC_PERIODIC && d.boundary != ACP_BC_LODI_X)
        return fail(ACP_E_UNSUPPORTED, "refined cases need a periodic or LODI_X coarse level");
    if (d.cubic && d.op != ACP_OP_RR && d.op != ACP_OP_HRR)
        return fail(ACP_E_UNSUPPORTED, "refined cases: the cubic-Mach terms belong to RR and HRR");
    if (d.boundary == ACP_BC_PERIODIC && (x0 < 1 || x1 <= x0 + 1 || x1 > d.nx || x1 - x0 >= d.nx - 1))
        return fail(ACP_E_ARG, "fine region [x0, x1] must satisfy 1 <= x0, x0 + 2 <= x1 <= nx, x1 - x0 < nx - 1");
    if (d.boundary == ACP_BC_LODI_X && (x0 < 1 || x1 <= x0 + 1 || x1 > d.nx - 1))
        return fail(ACP_E_ARG, "fine region [x0, x1] of a LODI_X coarse level must satisfy 1 <= x0, x0 + 2 <= x1 <= nx - 1");
    fd = d;
    fd.nx = 2 * (x1 - x0) + 1;
    fd.ny = 2 * d.ny;
    fd.nz = 2 * d.nz;
    fd.spacing = d.spacing / 2.;
    fd.domain[0] = (fd.nx - 0.5) * fd.spacing;
    fd.domain[1] = (fd.ny - 0.5) * fd.spacing;
    fd.domain[2] = (fd.nz - 0.5) * fd.spacing;
    fd.boundary = (d.boundary == ACP_BC_LODI_X && x1 == d.nx - 1) ? ACP_BC_OPEN_LODI_X : ACP_BC_OPEN_X;
    if (handed(d) && fd.boundary == ACP_BC_OPEN_X && x1 - x0 < 3)
        return fail(ACP_E_ARG, "HRR (or cubic-Mach) refined cases with two coupled faces need x1 - x0 >= 3 (one "
                               "grid-coupling column inside each face)");
    return ACP_OK;
}

void drop_graphs(acp_refined* r) {
    for (hipGraphExec_t& g : r->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

int parity(const acp_case* c) {
    const void* cur = c->precision == ACP_PREC_F64 ? (const void*)c->A : c->SA;
    return cur == c->arr0 ? 0 : 1;
}

int refined_create(const acp_case_desc* coarse, int32_t x0, int32_t x1, void* stream, char* region, int64_t bytes,
                   acp_refined** out) {
    if (!coarse || !out) return fail(ACP_E_ARG, "acp_refined_create: null argument");
    *out = nullptr;
    const acp_case_desc& d = *coarse;
    acp_case_desc fd;
    if (int rc = fine_desc(d, x0, x1, fd)) return rc;
    int64_t cb = 0, fb = 0;
    if (region) {
        if (!stream) return fail(ACP_E_ARG, "acp_refined_create_in: a placed refined case needs the caller's stream");
        if (int rc = acp_case_footprint(&d, 0, 0, &cb)) return rc;
        if (int rc = acp_case_footprint(&fd, 0, 0, &fb)) return rc;
        cb += transfer_bytes(d, fd);
        if (fb % 256 != 0 || cb + fb > bytes)
            return fail(ACP_E_ARG, "acp_refined_create_in: region smaller than acp_refined_footprint");
    }
    acp_refined* r = new acp_refined();
    r->x0 = x0;
    r->x1 = x1;
    if (stream) {
        r->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete r;
            return acp::hip_fail(e, "hipStreamCreate");
        }
        r->own_stream = true;
    }
    int rc;
    if (region) {
        // the fine level first (its footprint is a multiple of 256 bytes, so the
        // coarse part stays aligned); the coarse level takes the rest, where
        // probes set on it later find the room acp_refined_footprint counted
        rc = acp_case_create_in(&fd, r->stream, region, fb, &r->fine);
        if (rc == ACP_OK) rc = acp_case_create_in(&d, r->stream, region + fb, bytes - fb, &r->coarse);
    } else {
        rc = acp_case_create(&d, r->stream, &r->coarse);
        if (rc == ACP_OK) rc = acp_case_create(&fd, r->stream, &r->fine);
    }
    if (rc != ACP_OK) {
        acp_case_destroy(r->coarse);
        acp_case_destroy(r->fine);
        if (r->own_stream) (void)hipStreamDestroy(r->stream);
        delete r;
        return rc;
    }
    // A fine LODI face relaxes towards rho0 over the level-0 domain length,
    // with its own spacing and time step (Node.cpp:237-238, 311: every level's
    // LODIpressureOutlet takes getNumberOfNodesX() * getSpacing()); the shifted
    // tiers hold that length in the fine level's lattice units.
    if (r->fine->g.xlodi) {
        acp_case* f = r->fine;
        const int n[3] = {d.nx, d.ny, d.nz};
        for (int a = 0; a < 3; ++a) {
            f->hk.dom_len[a] = n[a] * d.spacing;
            f->hks.inv_len[a] = 1.0f / (float)(2 * n[a]);
        }
        hipError_t e = hipMemcpyAsync(f->dk, &f->hk, sizeof(acp::K64), hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess && f->dks)
            e = hipMemcpyAsync(f->dks, &f->hks, sizeof(acp::KShift), hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess) {
            acp_case_destroy(r->coarse);
            acp_case_destroy(r->fine);
            if (r->own_stream) (void)hipStreamDestroy(r->stream);
            delete r;
            return acp::hip_fail(e, "fine-level constants");
        }
    }
    // HRR (and RR with the cubic-Mach terms): the transfer buffers, on the
    // coarse level (counted in its footprint), and the interface columns that
    // collide with them: the coarse grid-coupling columns x0 + 1 and x1 - 1,
    // the fine face columns
    r->nv = handed_planes(d);
    if (handed(d)) {
        const int nf = coupled_faces(r), nv = r->nv;
        const int64_t cp = (int64_t)d.ny * d.nz, fp = (int64_t)fd.ny * fd.nz;
        void *pc = nullptr, *pg = nullptr, *pf = nullptr;
        int arc = acp::case_alloc_extra(r->coarse, nf * nv * cp * 8, &pc);
        if (arc == ACP_OK) arc = acp::case_alloc_extra(r->coarse, nf * nv * cp * 8, &pg);
        if (arc == ACP_OK) arc = acp::case_alloc_extra(r->coarse, nf * nv * fp * 8, &pf);
        if (arc != ACP_OK) {
            acp_case_destroy(r->coarse);
            acp_case_destroy(r->fine);
            if (r->own_stream) (void)hipStreamDestroy(r->stream);
            delete r;
            return arc;
        }
        r->s_col = (double*)pc;
        r->s_gc = (double*)pg;
        r->s_face = (double*)pf;
        r->coarse->nfix = nf;
        r->fine->nfix = nf;
        for (int s = 0; s < nf; ++s) {
            r->coarse->fix_col[s] = s == 0 ? x0 + 1 : (x1 - 1) % d.nx;
            r->coarse->fix_S[s] = r->s_gc + (int64_t)s * nv * cp;
            r->fine->fix_col[s] = s == 0 ? 0 : fd.nx - 1;
            r->fine->fix_S[s] = r->s_face + (int64_t)s * nv * fp;
        }
    }
    // tau = 1 / omega of each level (Control.cpp:54 with each level's time step)
    r->a = (1. / r->fine->hk.omega) / (2. * (1. / r->coarse->hk.omega));
    *out = r;
    acp::clear_error();
    return ACP_OK;
}


// ---- the shell geometry on the host: node lists (in the oracle's order) and
// link masks
bool box_in_h(const acp::refine::Box6& B, int i, int j, int k) {
    return i >= B.b[0] && i <= B.b[1] && j >= B.b[2] && j <= B.b[3] && k >= B.b[4] && k <= B.b[5];
}
bool box_inner_h(const acp::refine::Box6& B, int i, int j, int k) {
    return i > B.b[0] && i < B.b[1] && j > B.b[2] && j < B.b[3] && k > B.b[4] && k < B.b[5];
}

// the link masks of the current link sets, uploaded (Coarse-to-Fine and
// hanging: the links arriving from inside the hole, with MfP / hMfP also those
// arriving along its surface; Fine-to-Coarse Mc: the core box's surface takes
// the links arriving from outside the box)
int shell_upload_masks(acp_refined* r) {
    const int fnx = r->fine->g.nx, fny = r->fine->g.ny, cnx = r->coarse->g.nx, cny = r->coarse->g.ny;
    std::vector<uint32_t> fm(2 * r->nfs), cm(r->ncr);
    for (int64_t t = 0; t < r->nfs; ++t) {
        const int64_t l = r->h_fs[t];
        const int i = (int)(l % fnx), j = (int)((l / fnx) % fny), k = (int)(l / ((int64_t)fnx * fny));
        uint32_t miss = 0, common = 0;
        for (int q = 0; q < 19; ++q) {
            const int a = i - acp::EX[q], b = j - acp::EY[q], d = k - acp::EZ[q];
            if (box_inner_h(r->hb, a, b, d)) miss |= 1u << q;
            else if (box_in_h(r->hb, a, b, d)) common |= 1u << q;
        }
        const uint32_t all = acp::refine::kAllLinks;
        fm[t] = r->c2f_set == ACP_LINKS_ALL ? all : (r->c2f_set == ACP_LINKS_MISSING_COMMON ? (miss | common) : miss);
        fm[r->nfs + t] =
            r->hang_set == ACP_LINKS_ALL ? all : (r->hang_set == ACP_LINKS_MISSING_COMMON ? (miss | common) : miss);
    }
    for (int64_t t = 0; t < r->ncr; ++t) {
        uint32_t m = acp::refine::kAllLinks;
        if (t < r->ncs && r->f2c_set != ACP_LINKS_ALL) {
            const int64_t l = r->h_cr[t];
            const int I = (int)(l % cnx), J = (int)((l / cnx) % cny), K = (int)(l / ((int64_t)cnx * cny));
            m = 0;
            for (int q = 0; q < 19; ++q)
                if (!box_in_h(r->cb, I - acp::EX[q], J - acp::EY[q], K - acp::EZ[q])) m |= 1u << q;
        }
        cm[t] = m;
    }
    hipError_t e = hipMemcpy(r->d_fs_mask, fm.data(), sizeof(uint32_t) * fm.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && !cm.empty())
        e = hipMemcpy(r->d_cr_mask, cm.data(), sizeof(uint32_t) * cm.size(), hipMemcpyHostToDevice);
    return e == hipSuccess ? ACP_OK : acp::hip_fail(e, "shell link masks");
}

void shell_free(acp_refined* r) {
    for (void* p : {(void*)r->d_fs, (void*)r->d_fs_mask, (void*)r->d_fs_nb, (void*)r->d_cr, (void*)r->d_cr_mask,
                    (void*)r->d_sh})
        if (p) (void)hipFree(p);
    r->d_fs = nullptr, r->d_fs_mask = nullptr, r->d_fs_nb = nullptr, r->d_cr = nullptr, r->d_cr_mask = nullptr;
    r->d_sh = nullptr;
}

// the lists, their device copies and the levels' fix-node lists
int shell_build(acp_refined* r) {
    acp_case* c = r->coarse;
    acp_case* f = r->fine;
    const int fnx = f->g.nx, fny = f->g.ny, cnx = c->g.nx, cny = c->g.ny, cnz = c->g.nz;
    for (int pass = 0; pass < 2; ++pass) {
        for (int64_t l = 0; l < f->g.N; ++l) {
            const int i = (int)(l % fnx), j = (int)((l / fnx) % fny), k = (int)(l / ((int64_t)fnx * fny));
            if (!box_in_h(r->hb, i, j, k) || box_inner_h(r->hb, i, j, k)) continue;
            const bool co = !(i & 1) && !(j & 1) && !(k & 1);
            if (co == (pass == 0)) r->h_fs.push_back(l);
        }
        if (pass == 0) r->nfc = (int64_t)r->h_fs.size();
    }
    r->nfs = (int64_t)r->h_fs.size();
    for (int pass = 0; pass < 2; ++pass) {
        for (int64_t l = 0; l < c->g.N; ++l) {
            const int I = (int)(l % cnx), J = (int)((l / cnx) % cny), K = (int)(l / ((int64_t)cnx * cny));
            const bool face = I == 0 || I == cnx - 1 || J == 0 || J == cny - 1 || K == 0 || K == cnz - 1;
            const bool take = pass == 0 ? (box_in_h(r->cb, I, J, K) && !box_inner_h(r->cb, I, J, K))
                                        : (!box_in_h(r->cb, I, J, K) && !face);
            if (take) r->h_cr.push_back(l);
        }
        if (pass == 0) r->ncs = (int64_t)r->h_cr.size();
    }
    r->ncr = (int64_t)r->h_cr.size();
    // the hanging nodes' coincident neighbours along the surface, as list positions
    std::vector<int32_t> pos(f->g.N, -1);
    for (int64_t t = 0; t < r->nfc; ++t) pos[r->h_fs[t]] = (int32_t)t;
    std::vector<int32_t> nbp(4 * (r->nfs - r->nfc), -1);
    const int64_t st[3] = {1, fnx, (int64_t)fnx * fny};
    for (int64_t t = r->nfc; t < r->nfs; ++t) {
        const int64_t l = r->h_fs[t];
        const int x[3] = {(int)(l % fnx), (int)((l / fnx) % fny), (int)(l / ((int64_t)fnx * fny))};
        int odd[2] = {0, 0}, no = 0;
        for (int d = 0; d < 3; ++d)
            if (x[d] & 1) odd[no++] = d;
        int32_t* p = &nbp[4 * (t - r->nfc)];
        if (no == 1) {
            p[0] = pos[l - st[odd[0]]];
            p[1] = pos[l + st[odd[0]]];
        } else {
            p[0] = pos[l - st[odd[0]] - st[odd[1]]];
            p[1] = pos[l + st[odd[0]] - st[odd[1]]];
            p[2] = pos[l - st[odd[0]] + st[odd[1]]];
            p[3] = pos[l + st[odd[0]] + st[odd[1]]];
        }
        for (int e = 0; e < (no == 1 ? 2 : 4); ++e)
            if (p[e] < 0) return fail(ACP_E_STATE, "shell geometry: a hanging node without coincident neighbours");
    }
    const bool handed = c->d.op == ACP_OP_HRR || (c->d.op == ACP_OP_RR && c->d.cubic);
    hipError_t e = hipMalloc((void**)&r->d_fs, sizeof(int64_t) * r->nfs);
    if (e == hipSuccess) e = hipMalloc((void**)&r->d_fs_mask, sizeof(uint32_t) * 2 * r->nfs);
    if (e == hipSuccess) e = hipMalloc((void**)&r->d_fs_nb, sizeof(int32_t) * std::max<size_t>(nbp.size(), 1));
    if (e == hipSuccess) e = hipMalloc((void**)&r->d_cr, sizeof(int64_t) * std::max<int64_t>(r->ncr, 1));
    if (e == hipSuccess) e = hipMalloc((void**)&r->d_cr_mask, sizeof(uint32_t) * std::max<int64_t>(r->ncr, 1));
    if (e == hipSuccess && handed) {
        e = hipMalloc((void**)&r->d_sh, sizeof(double) * r->nv * (r->nfs + r->ncs));
        if (e == hipSuccess) e = hipMemset(r->d_sh, 0, sizeof(double) * r->nv * (r->nfs + r->ncs));
    }
    if (e == hipSuccess) e = hipMemcpy(r->d_fs, r->h_fs.data(), sizeof(int64_t) * r->nfs, hipMemcpyHostToDevice);
    if (e == hipSuccess && !nbp.empty())
        e = hipMemcpy(r->d_fs_nb, nbp.data(), sizeof(int32_t) * nbp.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && r->ncr)
        e = hipMemcpy(r->d_cr, r->h_cr.data(), sizeof(int64_t) * r->ncr, hipMemcpyHostToDevice);
    if (e != hipSuccess) return acp::hip_fail(e, "shell geometry lists");
    if (handed) {
        f->fix_list = r->d_fs;
        f->fix_list_n = r->nfs;
        f->fix_list_P = r->d_sh;
        c->fix_list = r->d_cr;
        c->fix_list_n = r->ncs;
        c->fix_list_P = r->d_sh + (int64_t)r->nv * r->nfs;
    }
    return shell_upload_masks(r);
}

// One refined coarse step as plain launches on the case's stream.
int step_once(acp_refined* r) {
    hipError_t e = launch_transfer(r, 0);
    if (e != hipSuccess) return acp::hip_fail(e, "acp_refined_step launch");
    int rc = acp::case_step_launch(r->coarse, 1);
    if (rc == ACP_OK) rc = acp::case_step_launch(r->fine, 1);
    if (rc != ACP_OK) return rc;
    if (r->scheme == ACP_COUPLING_DC) {
        // direct coupling: the fine nodes at t + 1/2 (then the HRR strain
        // rates of t + 1/2), the second fine step, the restriction of the
        // coarse nodes inside the fine region, then the pair exchange at t + 1
        // (the restriction reads no face node, the exchange writes only face
        // nodes and the coarse interface columns)
        e = launch_dc(r, 1);
        if (e == hipSuccess) e = launch_transfer(r, 1);
        if (e == hipSuccess) {
            rc = acp::case_step_launch(r->fine, 1);
            if (rc != ACP_OK) return rc;
            e = launch_f2c(r);
        }
        if (e == hipSuccess) e = launch_dc(r, 0);
        if (e != hipSuccess) return acp::hip_fail(e, "acp_refined_step launch");
        return ACP_OK;
    }
    e = launch_transfer(r, 1);
    if (e == hipSuccess) e = launch_c2f(r, 1);
    if (e == hipSuccess) {
        rc = acp::case_step_launch(r->fine, 1);
        if (rc != ACP_OK) return rc;
        e = launch_c2f(r, 0);
    }
    if (e == hipSuccess) e = launch_f2c(r);
    if (e != hipSuccess) return acp::hip_fail(e, "acp_refined_step launch");
    return ACP_OK;
}

// graph_chunk coarse steps (an even count: both levels end on the arrays they
// started from) replayed from a graph captured on first use for each parity.
int step_graph_chunk(acp_refined* r) {
    const int p = parity(r->coarse) + 2 * parity(r->fine);
    acp_case* lv[2] = {r->coarse, r->fine};
    if (!r->gexec[p]) {
        int32_t steps0[2], count0[2];
        for (int i = 0; i < 2; ++i) {
            steps0[i] = lv[i]->steps;
            count0[i] = lv[i]->probe_count;
        }
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(r->stream, hipStreamCaptureModeRelaxed);
        if (e != hipSuccess) return acp::hip_fail(e, "refined graph capture");
        int rc = ACP_OK;
        for (int32_t s = 0; s < r->graph_chunk && rc == ACP_OK; ++s) rc = step_once(r);
        e = hipStreamEndCapture(r->stream, &g);
        for (int i = 0; i < 2; ++i) {
            lv[i]->steps = steps0[i];
            lv[i]->probe_count = count0[i];
        }
        if (rc != ACP_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&r->gexec[p], g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            r->gexec[p] = nullptr;
            return acp::hip_fail(e, "refined graph capture");
        }
    }
    const hipError_t e = hipGraphLaunch(r->gexec[p], r->stream);
    if (e != hipSuccess) return acp::hip_fail(e, "refined graph launch");
    for (int i = 0; i < 2; ++i) {
        const int32_t k = (i == 0 ? 1 : 2) * r->graph_chunk;
        lv[i]->steps += k;
        if (lv[i]->nprobe > 0) lv[i]->probe_count += k;
    }
    return ACP_OK;
}

}  // namespace

extern "C" int acp_refined_create(const acp_case_desc* coarse, int32_t x0, int32_t x1, void* stream,
                                  acp_refined** out) {
    return refined_create(coarse, x0, x1, stream, nullptr, 0, out);
}

extern "C" int acp_refined_fine_desc(const acp_case_desc* coarse, int32_t x0, int32_t x1, acp_case_desc* fine) {
    if (!coarse || !fine) return fail(ACP_E_ARG, "acp_refined_fine_desc: null argument");
    return fine_desc(*coarse, x0, x1, *fine);
}

extern "C" int acp_refined_footprint(const acp_case_desc* coarse, int32_t x0, int32_t x1, int32_t nprobe,
                                     int32_t probe_capacity, int64_t* bytes) {
    if (!coarse || !bytes) return fail(ACP_E_ARG, "acp_refined_footprint: null argument");
    acp_case_desc fd;
    if (int rc = fine_desc(*coarse, x0, x1, fd)) return rc;
    int64_t cb = 0, fb = 0;
    if (int rc = acp_case_footprint(coarse, nprobe, probe_capacity, &cb)) return rc;
    if (int rc = acp_case_footprint(&fd, 0, 0, &fb)) return rc;
    *bytes = fb + cb + transfer_bytes(*coarse, fd);
    return ACP_OK;
}

extern "C" int acp_refined_create_in(const acp_case_desc* coarse, int32_t x0, int32_t x1, void* stream,
                                     void* region, int64_t region_bytes, acp_refined** out) {
    if (!region || region_bytes <= 0 || ((uintptr_t)region % 256) != 0)
        return fail(ACP_E_ARG, "acp_refined_create_in: region must be non-null, 256-byte aligned and non-empty");
    return refined_create(coarse, x0, x1, stream, (char*)region, region_bytes, out);
}

extern "C" int acp_refined_enable_graphs(acp_refined* r, int32_t chunk) {
    if (!r) return fail(ACP_E_ARG, "acp_refined_enable_graphs: null argument");
    if (chunk < 0 || chunk == 1 || (chunk % 2) != 0)
        return fail(ACP_E_ARG, "graph chunk must be 0 (off) or an even number of coarse steps");
    const hipError_t e = hipStreamSynchronize(r->stream);
    if (e != hipSuccess) return acp::hip_fail(e, "acp_refined_enable_graphs");
    drop_graphs(r);
    r->graph_chunk = chunk;
    return ACP_OK;
}

extern "C" int acp_refined_create_shell(const acp_case_desc* coarse, int32_t layer_yz, int32_t layer_x, void* stream,
                                        acp_refined** out) {
    if (!coarse || !out) return fail(ACP_E_ARG, "acp_refined_create_shell: null argument");
    *out = nullptr;
    const acp_case_desc& d = *coarse;
    if (d.boundary != ACP_BC_LODI)
        return fail(ACP_E_UNSUPPORTED, "shell geometry: the coarse level has LODI faces (ACP_BC_LODI)");
    if (d.precision != ACP_PREC_F64) return fail(ACP_E_UNSUPPORTED, "shell geometry: float64 only");
    if (d.cubic && d.op != ACP_OP_RR && d.op != ACP_OP_HRR)
        return fail(ACP_E_UNSUPPORTED, "refined cases: the cubic-Mach terms belong to RR and HRR");
    if (layer_yz < 2 || layer_x < 2)
        return fail(ACP_E_ARG, "shell geometry: RefinementLayerYZ and RefinementLayerX of at least 2 (the core box's "
                               "grid-coupling nodes off the domain faces)");
    const int hc[6] = {layer_x, d.nx - 1 - layer_yz, layer_yz, d.ny - 1 - layer_yz, layer_yz, d.nz - 1 - layer_yz};
    if (hc[1] - hc[0] < 2 || hc[3] - hc[2] < 2 || hc[5] - hc[4] < 2)
        return fail(ACP_E_ARG, "shell geometry: the refinement layers leave no coarse core");
    acp_case_desc fd = d;
    fd.nx = 2 * d.nx - 1;
    fd.ny = 2 * d.ny - 1;
    fd.nz = 2 * d.nz - 1;
    fd.spacing = d.spacing / 2.;
    for (int a = 0; a < 3; ++a) fd.domain[a] = ((a == 0 ? fd.nx : (a == 1 ? fd.ny : fd.nz)) - 0.5) * fd.spacing;
    acp_refined* r = new acp_refined();
    r->shell = true;
    for (int a = 0; a < 6; ++a) {
        r->hb.b[a] = 2 * hc[a];
        r->cb.b[a] = hc[a] + (a % 2 ? 1 : -1);
    }
    r->nv = handed_planes(d);
    if (stream) {
        r->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete r;
            return acp::hip_fail(e, "hipStreamCreate");
        }
        r->own_stream = true;
    }
    int rc = acp_case_create(&d, r->stream, &r->coarse);
    if (rc == ACP_OK) rc = acp_case_create(&fd, r->stream, &r->fine);
    if (rc == ACP_OK) {
        // the fine LODI faces relax over the level-0 domain length (Node.cpp:237-238, 311)
        acp_case* f = r->fine;
        const int n[3] = {d.nx, d.ny, d.nz};
        for (int a = 0; a < 3; ++a) f->hk.dom_len[a] = n[a] * d.spacing;
        hipError_t e = hipMemcpyAsync(f->dk, &f->hk, sizeof(acp::K64), hipMemcpyHostToDevice, r->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(r->stream);
        if (e != hipSuccess) rc = acp::hip_fail(e, "fine-level constants");
    }
    if (rc == ACP_OK) rc = shell_build(r);
    if (rc != ACP_OK) {
        shell_free(r);
        acp_case_destroy(r->coarse);
        acp_case_destroy(r->fine);
        if (r->own_stream) (void)hipStreamDestroy(r->stream);
        delete r;
        return rc;
    }
    r->a = (1. / r->fine->hk.omega) / (2. * (1. / r->coarse->hk.omega));
    // the fine nodes strictly inside the core box are no part of the shell
    // (the reference's level-1 parts leave them out): the fine level's bulk
    // kernels skip them
    for (int a = 0; a < 6; ++a) r->fine->g.hole.b[a] = r->hb.b[a];
    *out = r;
    acp::clear_error();
    return ACP_OK;
}

extern "C" void acp_refined_destroy(acp_refined* r) {
    if (!r) return;
    if (r->graph_chunk > 0) (void)hipStreamSynchronize(r->stream);
    drop_graphs(r);
    if (r->shell) (void)hipStreamSynchronize(r->stream);
    shell_free(r);
    acp_case_destroy(r->fine);
    acp_case_destroy(r->coarse);
    if (r->own_stream && r->stream) (void)hipStreamDestroy(r->stream);
    delete r;
}

extern "C" int acp_refined_set_coupling_precision(acp_refined* r, int32_t precision) {
    if (!r) return fail(ACP_E_ARG, "acp_refined_set_coupling_precision: null argument");
    if (precision != ACP_PREC_F64 && precision != ACP_PREC_F32_SHIFTED)
        return fail(ACP_E_ARG, "coupling precision: ACP_PREC_F64 or ACP_PREC_F32_SHIFTED");
    if (r->coarse->precision == ACP_PREC_F64 && precision != ACP_PREC_F64)
        return fail(ACP_E_UNSUPPORTED, "a float64 refined case couples in float64");
    if (r->mrt_style && precision != ACP_PREC_F64)
        return fail(ACP_E_UNSUPPORTED, "Scaling-Style MRTSTYLE needs the coupling arithmetic in float64");
    if (r->graph_chunk > 0) (void)hipStreamSynchronize(r->stream);
    drop_graphs(r);
    r->f64_coupling = precision == ACP_PREC_F64;
    return ACP_OK;
}

extern "C" int acp_refined_set_scaling(acp_refined* r, int32_t style) {
    if (!r) return fail(ACP_E_ARG, "acp_refined_set_scaling: null argument");
    if (style != ACP_SCALING_SRT && style != ACP_SCALING_MRT)
        return fail(ACP_E_ARG, "Scaling-Style: ACP_SCALING_SRT or ACP_SCALING_MRT");
    if (style == ACP_SCALING_MRT && !r->f64_coupling)
        return fail(ACP_E_UNSUPPORTED, "Scaling-Style MRTSTYLE needs the coupling arithmetic in float64");
    if (r->graph_chunk > 0) (void)hipStreamSynchronize(r->stream);
    drop_graphs(r);
    r->mrt_style = style == ACP_SCALING_MRT;
    return ACP_OK;
}

extern "C" int acp_refined_set_filter(acp_refined* r, int32_t filter) {
    if (!r) return fail(ACP_E_ARG, "acp_refined_set_filter: null argument");
    if (filter != ACP_F2C_NOFI && filter != ACP_F2C_LAG && filter != ACP_F2C_TOU)
        return fail(ACP_E_UNSUPPORTED, "fine-to-coarse filter: ACP_F2C_NOFI, ACP_F2C_LAG or ACP_F2C_TOU");
    if (r->graph_chunk > 0) (void)hipStreamSynchronize(r->stream);
    drop_graphs(r);
    r->filter = filter;
    return ACP_OK;
}

extern "C" int acp_refined_set_coupling(acp_refined* r, int32_t scheme, int32_t c2f, int32_t f2c, int32_t hanging) {
    if (!r) return fail(ACP_E_ARG, "acp_refined_set_coupling: null argument");
    auto valid = [](int32_t v) { return v == ACP_LINKS_MISSING || v == ACP_LINKS_MISSING_COMMON || v == ACP_LINKS_ALL; };
    if (!valid(c2f) || !valid(f2c) || !valid(hanging)) return fail(ACP_E_ARG, "coupling link sets: ACP_LINKS_*");
    if (r->shell) {
        if (scheme != ACP_COUPLING_STD)
            return fail(ACP_E_UNSUPPORTED, "shell geometry: the standard coupling (Overlap 1) only");
        if (f2c == ACP_LINKS_MISSING_COMMON)
            return fail(ACP_E_UNSUPPORTED, "Fine-to-Coarse McP belongs to the direct coupling (the standard "
                                           "coupling takes Mc or cAll)");
        if (r->graph_chunk > 0) (void)hipStreamSynchronize(r->stream);
        drop_graphs(r);
        (void)hipStreamSynchronize(r->stream);
        r->c2f_set = c2f;
        r->f2c_set = f2c;
        r->hang_set = hanging;
        return shell_upload_masks(r);
    }
    if (scheme == ACP_COUPLING_STD) {
        if (f2c == ACP_LINKS_MISSING_COMMON)
            return fail(ACP_E_UNSUPPORTED, "Fine-to-Coarse McP belongs to the direct coupling (the standard "
                                           "coupling takes Mc or cAll)");
        if (f2c == ACP_LINKS_MISSING && coupled_faces(r) == 2 && r->x1 - r->x0 < 3)
            return fail(ACP_E_ARG, "Fine-to-Coarse Mc with two coupled faces needs x1 - x0 >= 3 (one grid-coupling "
                                   "column inside each face)");
    } else if (scheme == ACP_COUPLING_DC) {
        if (c2f == ACP_LINKS_ALL || f2c == ACP_LINKS_ALL)
            return fail(ACP_E_ARG, "direct coupling: Coarse-to-Fine Mf or MfP, Fine-to-Coarse Mc or McP");
        if (c2f == ACP_LINKS_MISSING_COMMON && f2c == ACP_LINKS_MISSING_COMMON)
            return fail(ACP_E_ARG, "direct coupling: either MfP-Mc or Mf-McP");
    } else {
        return fail(ACP_E_ARG, "coupling scheme: ACP_COUPLING_STD or ACP_COUPLING_DC");
    }
    if (r->graph_chunk > 0) (void)hipStreamSynchronize(r->stream);
    drop_graphs(r);
    r->scheme = scheme;
    // HRR (and RR with the cubic-Mach terms): the coarse column that collides
    // with handed values is the grid-coupling column inside the fine region
    // (standard coupling) or the interface column itself (direct coupling)
    const int64_t cp = (int64_t)r->coarse->g.ny * r->coarse->g.nz * r->nv;
    for (int s = 0; s < r->coarse->nfix; ++s) {
        const int cnx = r->coarse->g.nx;
        if (scheme == ACP_COUPLING_STD) {
            r->coarse->fix_col[s] = s == 0 ? r->x0 + 1 : (r->x1 - 1) % cnx;
            r->coarse->fix_S[s] = r->s_gc + (int64_t)s * cp;
        } else {
            r->coarse->fix_col[s] = s == 0 ? r->x0 % cnx : r->x1 % cnx;
            r->coarse->fix_S[s] = r->s_col + (int64_t)s * cp;
        }
    }
    r->c2f_set = c2f;
    r->f2c_set = f2c;
    r->hang_set = hanging;
    return ACP_OK;
}

extern "C" int acp_refined_restrict(acp_refined* r) {
    if (!r) return fail(ACP_E_ARG, "acp_refined_restrict: null argument");
    if (!r->coarse->initialized || !r->fine->initialized) return fail(ACP_E_STATE, "refined case not initialised");
    const hipError_t e = launch_f2c(r);
    if (e != hipSuccess) return acp::hip_fail(e, "acp_refined_restrict launch");
    return ACP_OK;
}

extern "C" int acp_refined_level(acp_refined* r, int32_t level, acp_case** out) {
    if (!r || !out || (level != 0 && level != 1)) return fail(ACP_E_ARG, "acp_refined_level: level 0 or 1");
    *out = level == 0 ? r->coarse : r->fine;
    return ACP_OK;
}

extern "C" int acp_refined_init_pulse(acp_refined* r, double amp, double x0, double y0, double b) {
    if (!r) return fail(ACP_E_ARG, "acp_refined_init_pulse: null argument");
    int rc = acp_case_init_pulse(r->coarse, amp, x0, y0, b);
    // the same field at the fine nodes: x = x0_region + i/2, y = j/2 in coarse node units
    if (rc == ACP_OK) rc = acp_case_init_pulse(r->fine, amp, 2. * (x0 - r->x0), 2. * y0, 2. * b);
    return rc;
}

extern "C" int acp_refined_init_reference(acp_refined* r) {
    if (!r) return fail(ACP_E_ARG, "acp_refined_init_reference: null argument");
    int rc = acp::init_reference_level(r->coarse, 0, 0, r->coarse->d);
    // the fine level's first node is coarse node x0, fine index 2 x0 of the domain
    if (rc == ACP_OK) rc = acp::init_reference_level(r->fine, 1, 2 * r->x0, r->coarse->d);
    return rc;
}

extern "C" int acp_refined_write_vtk(acp_refined* r, const char* stem) {
    if (!r || !stem) return fail(ACP_E_ARG, "acp_refined_write_vtk: null argument");
    if (!r->coarse->initialized || !r->fine->initialized) return fail(ACP_E_STATE, "refined case not initialised");
    std::vector<double> arr[2][2];  // [level][current, other]
    acp_case* lv[2] = {r->coarse, r->fine};
    for (int l = 0; l < 2; ++l) {
        const size_t n = (size_t)19 * lv[l]->g.N;
        arr[l][0].resize(n);
        if (int rc = acp_case_get_populations(lv[l], arr[l][0].data())) return rc;
        if (lv[l]->d.nonreflecting) {
            arr[l][1].resize(n);
            if (int rc = acp_case_get_previous_populations(lv[l], arr[l][1].data())) return rc;
        }
    }
    auto other = [&](int l) { return arr[l][1].empty() ? arr[l][0].data() : arr[l][1].data(); };
    // the cubic-Mach state enters the printed density and velocity (calcDensity, Node.cpp:11-30)
    std::vector<double> psi[2];
    if (r->coarse->d.cubic)
        for (int l = 0; l < 2; ++l) {
            psi[l].resize((size_t)19 * lv[l]->g.N);
            if (int rc = acp_case_get_psi(lv[l], psi[l].data())) return rc;
        }
    const double* cpsi = psi[0].empty() ? nullptr : psi[0].data();
    const double* fpsi = psi[1].empty() ? nullptr : psi[1].data();
    if (r->shell) {
        const int ly = r->hb.b[2] / 2, lx = r->hb.b[0] / 2;
        return acp::write_vtk_shell_host(r->coarse->d, r->coarse->hk, arr[0][0].data(), other(0), cpsi, r->fine->d,
                                         r->fine->hk, arr[1][0].data(), other(1), fpsi, ly, lx, stem);
    }
    return acp::write_vtk2_host(r->coarse->d, r->coarse->hk, arr[0][0].data(), other(0), r->fine->d, r->fine->hk,
                                arr[1][0].data(), other(1), r->x0, r->x1, stem, cpsi, fpsi);
}

extern "C" int acp_refined_step(acp_refined* r, int32_t n) {
    if (!r) return fail(ACP_E_ARG, "acp_refined_step: null argument");
    if (n < 0) return fail(ACP_E_ARG, "negative step count");
    if (!r->coarse->initialized || !r->fine->initialized) return fail(ACP_E_STATE, "refined case not initialised");
    for (acp_case* c : {r->coarse, r->fine})
        if (c->nprobe > 0 && c->probe_count + (c == r->fine ? 2 : 1) * n > c->probe_cap)
            return fail(ACP_E_STATE, "probe buffer full: read the probes first");
    // a level whose probes (or kernel form) changed since the graphs were
    // captured: the graphs hold its old buffers, so capture them anew
    if (r->coarse->graph_epoch != r->epoch[0] || r->fine->graph_epoch != r->epoch[1]) {
        if (r->graph_chunk > 0) {
            const hipError_t e = hipStreamSynchronize(r->stream);
            if (e != hipSuccess) return acp::hip_fail(e, "acp_refined_step");
        }
        drop_graphs(r);
        r->epoch[0] = r->coarse->graph_epoch;
        r->epoch[1] = r->fine->graph_epoch;
    }
    while (r->graph_chunk > 0 && n >= r->graph_chunk) {
        if (int rc = step_graph_chunk(r)) return rc;
        n -= r->graph_chunk;
    }
    for (int32_t s = 0; s < n; ++s)
        if (int rc = step_once(r)) return rc;
    return ACP_OK;
}
