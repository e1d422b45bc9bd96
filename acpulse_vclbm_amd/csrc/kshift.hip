// Shifted-population tiers: float32 collision in lattice units on stored
// deviations g_q = s (f_q/rho0 - w_q), float32 (s = 1) or float16 storage.
//
// The kernels follow the float64 tier step for step (collide + push with the
// same push rule, the same LODI pass order and phases) but read/write 4 or 2
// bytes per population: 2*19*4 = 152 and 2*19*2 = 76 bytes per node and step.
// Global reductions (mass, probes) and conversions to SI are done in float64.
#include <algorithm>

#include "kshift.hpp"
#include "pack.hpp"
#include "physf.hpp"
#include "plane.hpp"

namespace acp {

template <typename T>
__device__ __forceinline__ float ld(const T* __restrict__ p) {
    return (float)(*p);
}

template <typename T>
__device__ __forceinline__ void load_dev(const T* __restrict__ F, const PopLayout& L, int64_t l, float inv_s,
                                         float d[19]) {
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        d[q] = (float)F[pidx(L, q, l)] * inv_s;
    });
}

// A deviation (float32, storage scale) converted to the storage type.  For
// float16 the product that made v is first rounded to float32 in every
// kernel: otherwise the compiler folds a multiply with a single use into the
// conversion (one rounding, v_fma_mixlo_f16) in some kernels and not in
// others, and the same node update would round differently in a case
// stepped alone and in a pack.  The empty asm only pins v to a register.
template <typename T>
__device__ __forceinline__ T to_store(float v) {
    if constexpr (!std::is_same_v<T, float>) __asm__ volatile("" : "+v"(v));
    return (T)v;
}

// Collision of one node in lattice units on its stored deviations g (in place):
// the operators of the float64 tier rewritten for deviations.  o[6] are the
// +x, +y, -x, -y, +z, -z neighbour offsets (HRR reads their macros from M).
template <int OP, int ORDER, bool NARROW>
__device__ __forceinline__ void collide_shift(const KShift& k, float g[19], const float* __restrict__ M,
                                              typename Plane<NARROW>::idx l,
                                              const typename Plane<NARROW>::idx o[6], int64_t N, int ghost) {
    using P = Plane<NARROW>;
    using I = typename P::idx;
    float d[19];
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        d[q] = g[q] * k.inv_s;
    });
    float dr, ux, uy, uz;
    moments_dev(d, dr, ux, uy, uz);
    float eq[19];
    if constexpr (OP == OP_MRT_RM) rm_feq_dev(dr, ux, uy, uz, eq);
    else if constexpr (OP != OP_MRT_GS) feq_dev<ORDER>(dr, ux, uy, uz, eq, k.k6);
    const float om1 = 1.f - k.omega;
    if constexpr (OP == OP_SRT) {
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            d[q] = eq[q] + om1 * (d[q] - eq[q]);
        });
    } else if constexpr (OP == OP_RR || OP == OP_HRR) {
        float fneq[19], a2[6], f1[19];
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            fneq[q] = d[q] - eq[q];
        });
        pi1_dev(fneq, a2);
        if constexpr (OP == OP_HRR) {
            // FD strain of the neighbours' pre-collision velocity (Node.cpp:1750-1889), dx = 1
                        float v[6][3];
#pragma unroll
            for (int m = 0; m < 6; ++m)
#pragma unroll
                for (int a = 0; a < 3; ++a) v[m][a] = P::soa(M, a + 1, (I)(l + o[m]), N);
            if (ghost) {
                // open x face (ghost = +1: no +x neighbour, -1: no -x neighbour):
                // quadratic extrapolation of the missing neighbour's velocity
                const int m = ghost > 0 ? 0 : 2;
                const I s1 = ghost > 0 ? (I)-1 : (I)1;
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    v[m][a] = (3.f * P::soa(M, a + 1, l, N) - 3.f * P::soa(M, a + 1, (I)(l + s1), N)) +
                              P::soa(M, a + 1, (I)(l + s1 + s1), N);
            }
            const float S[6] = {0.5f * (v[0][0] - v[2][0]), 0.5f * (v[1][1] - v[3][1]),
                                0.5f * (v[4][2] - v[5][2]),
                                0.5f * (0.5f * (v[0][1] - v[2][1]) + 0.5f * (v[1][0] - v[3][0])),
                                0.5f * (0.5f * (v[0][2] - v[2][2]) + 0.5f * (v[4][0] - v[5][0])),
                                0.5f * (0.5f * (v[1][2] - v[3][2]) + 0.5f * (v[4][1] - v[5][1]))};
            const float c = (1.f - k.sigma) * k.hrr_pre * (1.f + dr);
#pragma unroll
            for (int a = 0; a < 6; ++a) a2[a] = k.sigma * a2[a] + c * S[a];
        }
        offeq_dev<ORDER>(a2, ux, uy, uz, f1, k.k6);
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            d[q] = eq[q] + om1 * f1[q];
        });
    } else if constexpr (OP == OP_MRT_RM) {
        // feq_RM + (1-omega) M^-1 (non-equilibrium of N_xz, N_yz, Pi_xy, Pi_xz, Pi_yz);
        // every other raw moment relaxes with rate 1 (Node.cpp:2573-2600)
        float n5 = 0.f, n6 = 0.f, nxy = 0.f, nxz = 0.f, nyz = 0.f;
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const float f = d[q] - eq[q];
            constexpr int c5 = EX[q] * EX[q] - EZ[q] * EZ[q], c6 = EY[q] * EY[q] - EZ[q] * EZ[q];
            if constexpr (c5 > 0) n5 += f;
            if constexpr (c5 < 0) n5 -= f;
            if constexpr (c6 > 0) n6 += f;
            if constexpr (c6 < 0) n6 -= f;
            if constexpr (EX[q] * EY[q] > 0) nxy += f;
            if constexpr (EX[q] * EY[q] < 0) nxy -= f;
            if constexpr (EX[q] * EZ[q] > 0) nxz += f;
            if constexpr (EX[q] * EZ[q] < 0) nxz -= f;
            if constexpr (EY[q] * EZ[q] > 0) nyz += f;
            if constexpr (EY[q] * EZ[q] < 0) nyz -= f;
        });
        const float t3 = 1.f / 3.f, t6 = 1.f / 6.f;
        const float ax = om1 * (t3 * n5 - t6 * n6), ay = om1 * (t3 * n6 - t6 * n5), az = -om1 * t6 * (n5 + n6);
        const float dxy = 0.25f * om1 * nxy, dxz = 0.25f * om1 * nxz, dyz = 0.25f * om1 * nyz;
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            float add = 0.f;
            if constexpr (EX[q] != 0 && EY[q] == 0 && EZ[q] == 0) add = ax;
            if constexpr (EY[q] != 0 && EX[q] == 0 && EZ[q] == 0) add = ay;
            if constexpr (EZ[q] != 0 && EX[q] == 0 && EY[q] == 0) add = az;
            if constexpr (EX[q] * EY[q] != 0) add = (float)(EX[q] * EY[q]) * dxy;
            if constexpr (EX[q] * EZ[q] != 0) add = (float)(EX[q] * EZ[q]) * dxz;
            if constexpr (EY[q] * EZ[q] != 0) add = (float)(EY[q] * EZ[q]) * dyz;
            d[q] = eq[q] + add;
        });
    } else if constexpr (OP == OP_MRT_GS) {
        // Gram-Schmidt moments of the deviations (the rest state has only m0 = 1),
        // relaxation (Node.cpp:2798-2835) and inverse transform f += M^-1 dm
        float m[19];
        static_for<0, 19>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            float s = 0.f;
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int a = gsf_coef(r, q);
                if constexpr (a != 0) s += (float)a * d[q];
            });
            m[r] = s;
        });
        const float rho = 1.f + m[0];
        const float ir = 1.f / rho, jx = m[3], jy = m[5], jz = m[7];
        const float xx = jx * jx, yy = jy * jy, zz = jz * jz;
        float meq[19] = {0.f};
        meq[1] = (xx + yy + zz) * ir;
        meq[9] = (2.f * xx - yy - zz) * ir;
        meq[11] = (yy - zz) * ir;
        meq[13] = jx * jy * ir;
        meq[14] = jy * jz * ir;
        meq[15] = jx * jz * ir;
        float dm[19];
        static_for<0, 19>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if constexpr (r == 0 || r == 3 || r == 5 || r == 7) dm[r] = 0.f;
            else if constexpr (r == 1 || r == 9 || r == 11 || r == 13 || r == 14 || r == 15)
                dm[r] = k.omega * (meq[r] - m[r]);
            else dm[r] = meq[r] - m[r];
        });
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            float s = 0.f;
            static_for<0, 19>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                constexpr int a = gsf_coef(r, q);
                if constexpr (a != 0) {
                    constexpr float coef = wf(q) * (float)a / gsf_norm(r);
                    s += coef * dm[r];
                }
            });
            d[q] += s;
        });
    }
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        g[q] = d[q] * k.s;
    });
}

// HRR with a given strain rate S (lattice units of the level: 1 / time
// step): the collision of a grid-coupling node of a two-level step, which
// keeps the strain rate handed to it (Node.cpp:1795-1798); otherwise the HRR
// branch of collide_shift.
template <int ORDER>
__device__ __forceinline__ void collide_shift_hrr_strain(const KShift& k, float g[19], const float S[6]) {
    float d[19];
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        d[q] = g[q] * k.inv_s;
    });
    float dr, ux, uy, uz;
    moments_dev(d, dr, ux, uy, uz);
    float eq[19], fneq[19], a2[6], f1[19];
    feq_dev<ORDER>(dr, ux, uy, uz, eq, k.k6);
    const float om1 = 1.f - k.omega;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        fneq[q] = d[q] - eq[q];
    });
    pi1_dev(fneq, a2);
    const float c = (1.f - k.sigma) * k.hrr_pre * (1.f + dr);
#pragma unroll
    for (int a = 0; a < 6; ++a) a2[a] = k.sigma * a2[a] + c * S[a];
    offeq_dev<ORDER>(a2, ux, uy, uz, f1, k.k6);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        g[q] = (eq[q] + om1 * f1[q]) * k.s;
    });
}

// The nodes of x column `col` of a shifted level collide with the strain rate
// S_SI (float64 planes [6][ny nz], 1/s) times the level's time step and push
// as the bulk kernel pushes them (y, z wrapped; x wrapped on a level periodic
// in x, else links leaving the grid dropped), overwriting what the bulk
// kernel pushed from these nodes.
template <typename T, int ORDER>
__global__ __launch_bounds__(64) void fix_hrr_shift(const KShift* __restrict__ kp, const T* __restrict__ A,
                                                    T* __restrict__ B, int nx, int ny, int nz, const PopLayout L,
                                                    int col, int xwrap, const double* __restrict__ S, double dt) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int per = ny * nz;
    if (t >= per) return;
    const int j = t % ny, kz = t / ny;
    const int64_t l = col + (int64_t)nx * (j + (int64_t)ny * kz);
    float g[19];
#pragma unroll
    for (int q = 0; q < 19; ++q) g[q] = (float)A[pidx(L, q, l)];
    float s6[6];
#pragma unroll
    for (int v = 0; v < 6; ++v) s6[v] = (float)(S[v * per + t] * dt);
    collide_shift_hrr_strain<ORDER>(*kp, g, s6);
    static_for<0, 18>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        int a = col + EX[q];
        if (xwrap) a = (a + nx) % nx;
        else if (a < 0 || a >= nx) return;
        const int b = (j + EY[q] + ny) % ny, d = (kz + EZ[q] + nz) % nz;
        B[pidx(L, q, a + (int64_t)nx * (b + (int64_t)ny * d))] = (T)g[q];
    });
    B[pidx(L, 18, l)] = (T)g[18];
}

hipError_t launch_fix_hrr_shift(int prec, int order, const KShift* k, const void* A, void* B, const Grid& g, int col,
                                const double* S, double dt, hipStream_t s) {
    const int per = g.ny * g.nz;
    const dim3 grid((unsigned)((per + 63) / 64)), block(64);
    const int xwrap = !g.xopen && !g.xlodi;
#define FIXS(T, O)                                                                                              \
    hipLaunchKernelGGL((fix_hrr_shift<T, O>), grid, block, 0, s, k, (const T*)A, (T*)B, g.nx, g.ny, g.nz, g.L, \
                       col, xwrap, S, dt)
#define FIXT(T)                  \
    if (order == 2) FIXS(T, 2);  \
    else if (order == 3) FIXS(T, 3); \
    else FIXS(T, 4)
    if (prec == ACP_PREC_F32_SHIFTED) {
        FIXT(float);
    } else {
        FIXT(_Float16);
    }
#undef FIXT
#undef FIXS
    return hipGetLastError();
}

// ------------------------------------------------------------------ bulk collide + push
// The update of node (i, j, kz); `active` false: a thread beyond a row held
// by one workgroup, which only takes part in its LDS exchange (see bulk64).
template <typename T, int OP, int ORDER, bool PERIODIC, bool NARROW, bool NT>
__device__ __forceinline__ void bulk_shift_node(const KShift* __restrict__ kp, const T* __restrict__ A,
                                                T* __restrict__ B, const float* __restrict__ M, int nx, int ny, int nz,
                                                const PopLayout& L, int xopen, int row_lds, int xlodi, int i, int j,
                                                int kz, bool active) {
    using P = Plane<NARROW>;
    using I = typename P::idx;
    const I sxy = (I)nx * (I)ny;
    const I l = (I)i + (I)nx * (I)j + sxy * (I)kz;
    const KShift& k = *kp;
    float g[19];
    if (active) {
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            g[q] = (float)ld_pop<NT>(&P::at(A, q, l, L));
        });
    }
    int bi = 0, bj = 0, bk = 0;
    if constexpr (!PERIODIC) {
        bi = i == 0 ? 1 : (i == nx - 1 ? -1 : 0);
        bj = j == 0 ? 1 : (j == ny - 1 ? -1 : 0);
        bk = kz == 0 ? 1 : (kz == nz - 1 ? -1 : 0);
    } else {
        bi = (i == 0 && (xlodi & 1)) ? 1 : ((i == nx - 1 && (xlodi & 2)) ? -1 : 0);
    }
    const bool boundary = (bi | bj | bk) != 0;
    // neighbour offsets; unsigned wrap-around in the narrow form is exact because
    // every final index is in range
    I oxp = 1, oxm = (I)-1, oyp = (I)nx, oym = (I)0 - (I)nx, ozp = sxy, ozm = (I)0 - sxy;
    if constexpr (PERIODIC) {
        if (i == nx - 1) oxp = (I)0 - (I)(nx - 1);
        if (i == 0) oxm = (I)(nx - 1);
        if (j == ny - 1) oyp = (I)0 - (I)nx * (I)(ny - 1);
        if (j == 0) oym = (I)nx * (I)(ny - 1);
        if (kz == nz - 1) ozp = (I)0 - sxy * (I)(nz - 1);
        if (kz == 0) ozm = sxy * (I)(nz - 1);
    }
    if constexpr (OP != OP_STREAM) {
        if (active && !boundary) {
            const I o[6] = {oxp, oyp, oxm, oym, ozp, ozm};
            const int ghost = PERIODIC ? ((i == nx - 1 && (xopen & 2)) ? 1 : ((i == 0 && (xopen & 1)) ? -1 : 0)) : 0;
            collide_shift<OP, ORDER, NARROW>(k, g, M, l, o, L.N, ghost);
        }
    }
    // rows held by one workgroup: x-shifted links realigned through LDS, each
    // position written only if its source node pushes the link (see bulk64)
    if (row_lds) {
        __shared__ T row[kBulkBlock];
        static_for<0, 18>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            I base = l - (I)i;
            if constexpr (EY[q] > 0) base += oyp;
            if constexpr (EY[q] < 0) base += oym;
            if constexpr (EZ[q] > 0) base += ozp;
            if constexpr (EZ[q] < 0) base += ozm;
            if constexpr (EX[q] == 0) {
                if (active && row_source_pushes<PERIODIC>(q, i, nx, bj, bk, xopen, xlodi))
                    st_pop<NT>(&P::at(B, q, (I)(base + (I)i), L), to_store<T>(g[q]));
            } else {
                __syncthreads();
                if (active) row[(i + EX[q] + nx) % nx] = to_store<T>(g[q]);
                __syncthreads();
                if (active && row_source_pushes<PERIODIC>(q, i - EX[q], nx, bj, bk, xopen, xlodi))
                    st_pop<NT>(&P::at(B, q, (I)(base + (I)i), L), row[i]);
            }
        });
        if (active) st_pop<NT>(&P::at(B, 18, l, L), to_store<T>(g[18]));
        return;
    }

    static_for<0, 18>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (!PERIODIC) {
            if (boundary && !boundary_pushes(q, bi, bj, bk)) return;
        } else {
            if (boundary && EX[q] != bi) return;  // LODI x face: inward x links only
            if constexpr (EX[q] > 0) {
                if ((xopen & 2) && i == nx - 1) return;
            }
            if constexpr (EX[q] < 0) {
                if ((xopen & 1) && i == 0) return;
            }
        }
        I dd = l;
        if constexpr (EX[q] > 0) dd += oxp;
        if constexpr (EX[q] < 0) dd += oxm;
        if constexpr (EY[q] > 0) dd += oyp;
        if constexpr (EY[q] < 0) dd += oym;
        if constexpr (EZ[q] > 0) dd += ozp;
        if constexpr (EZ[q] < 0) dd += ozm;
        st_pop<NT>(&P::at(B, q, dd, L), to_store<T>(g[q]));
    });
    st_pop<NT>(&P::at(B, 18, l, L), to_store<T>(g[18]));
}

template <typename T, int OP, int ORDER, bool PERIODIC, bool NARROW, bool NT>
__global__ __launch_bounds__(kBulkBlock) __attribute__((amdgpu_waves_per_eu(ACP_BULK_WAVES))) void bulk_shift(
    const KShift* __restrict__ kp, const T* __restrict__ A, T* __restrict__ B, const float* __restrict__ M, int nx,
    int ny, int nz, const PopLayout L, int xopen, int row_lds, int xlodi) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < nx;
    if (!active && !row_lds) return;
    bulk_shift_node<T, OP, ORDER, PERIODIC, NARROW, NT>(kp, A, B, M, nx, ny, nz, L, xopen, row_lds, xlodi, i,
                                                        (int)blockIdx.y, (int)blockIdx.z, active);
}

// Packed form (pack.hpp): one persistent grid over the 256-node chunks of
// every member of a pack with operator OP (one-level grids with LODI faces,
// narrow addressing, plain loads and stores), each node as bulk_shift.
template <typename T, int OP, int ORDER>
__global__ __launch_bounds__(kPackChunk) __attribute__((amdgpu_waves_per_eu(ACP_BULK_WAVES))) void bulk_shift_packed(
    const PackTable* __restrict__ t) {
    const PackList& pl = t->bulk[OP];
    const int total = pl.prefix[pl.n];
    int m = 0;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        m = pack_member(pl, w, m);
        const PackSlot& s = t->slot[pl.slot[m]];
        if (s.steps_left <= 0) continue;
        const uint32_t l = (uint32_t)(w - pl.prefix[m]) * kPackChunk + threadIdx.x;
        if (l >= (uint32_t)s.L.N) continue;
        const uint32_t nx = (uint32_t)s.nx, nxy = nx * (uint32_t)s.ny;
        const uint32_t kz = l / nxy, r = l - kz * nxy, j = r / nx, i = r - j * nx;
        bulk_shift_node<T, OP, ORDER, false, true, false>((const KShift*)s.k, (const T*)s.A, (T*)s.B,
                                                          (const float*)s.M, s.nx, s.ny, s.nz, s.L, 0, 0, 0, (int)i,
                                                          (int)j, (int)kz, true);
    }
}

// Periodic rows of nx = 2 * blockDim.x nodes (nx a multiple of 128, no open
// faces), blockDim.y consecutive rows per workgroup: two nodes per thread, so
// that each load and store of a wave moves 64 pairs of consecutive
// populations; each node collides exactly as in bulk_shift and the x-shifted
// links are realigned through LDS (one buffer row per grid row) before the
// pair stores.
template <typename T>
struct alignas(2 * sizeof(T)) Pair2 {
    T x, y;
};

// A pair as a two-element vector, so that it moves with a single access (with
// the non-temporal hint when NT).
template <typename T>
using PairVec = T __attribute__((ext_vector_type(2)));

template <bool NT, typename T>
__device__ __forceinline__ Pair2<T> ld_pair(const T* p) {
    const PairVec<T> u = ld_pop<NT>(reinterpret_cast<const PairVec<T>*>(p));
    Pair2<T> v;
    v.x = u.x;
    v.y = u.y;
    return v;
}

template <bool NT, typename T>
__device__ __forceinline__ void st_pair(T* p, const Pair2<T>& v) {
    PairVec<T> u;
    u.x = v.x;
    u.y = v.y;
    st_pop<NT>(reinterpret_cast<PairVec<T>*>(p), u);
}

// ROWS: blockDim.y rows per workgroup (else one, blockDim.y = 1, compiled
// without the checks for rows beyond the grid).
template <typename T, int OP, int ORDER, bool NARROW, bool NT, bool ROWS>
__global__ __launch_bounds__(kBulkBlock) __attribute__((amdgpu_waves_per_eu(ACP_BULK_WAVES))) void bulk_shift2(
    const KShift* __restrict__ kp, const T* __restrict__ A, T* __restrict__ B, const float* __restrict__ M, int nx,
    int ny, int nz, const PopLayout L) {
    using P = Plane<NARROW>;
    using I = typename P::idx;
    const int t = threadIdx.x;
    // blockDim.y consecutive rows per workgroup; a row beyond the grid only
    // takes part in the LDS exchange
    const int j = ROWS ? (int)(blockIdx.y * blockDim.y + threadIdx.y) : (int)blockIdx.y, kz = blockIdx.z;
    const bool active = !ROWS || j < ny;
    const I sxy = (I)nx * (I)ny;
    const I rs = (I)nx * (I)j + sxy * (I)kz;
    const KShift& k = *kp;
    float g[2][19];
    if (active) {
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const Pair2<T> v = ld_pair<NT>(&P::at(A, q, (I)(rs + (I)(2 * t)), L));
            g[0][q] = (float)v.x;
            g[1][q] = (float)v.y;
        });
    }
    I oyp = (I)nx, oym = (I)0 - (I)nx, ozp = sxy, ozm = (I)0 - sxy;
    if (j == ny - 1) oyp = (I)0 - (I)nx * (I)(ny - 1);
    if (j == 0) oym = (I)nx * (I)(ny - 1);
    if (kz == nz - 1) ozp = (I)0 - sxy * (I)(nz - 1);
    if (kz == 0) ozm = sxy * (I)(nz - 1);
    if constexpr (OP != OP_STREAM) {
        if (active) {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int i = 2 * t + u;
                const I oxp = i == nx - 1 ? (I)0 - (I)(nx - 1) : (I)1;
                const I oxm = i == 0 ? (I)(nx - 1) : (I)-1;
                const I o[6] = {oxp, oyp, oxm, oym, ozp, ozm};
                collide_shift<OP, ORDER, NARROW>(k, g[u], M, (I)(rs + (I)i), o, L.N, 0);
            }
        }
    }
    __shared__ T rows[2 * kBulkBlock];
    T* row = ROWS ? rows + threadIdx.y * nx : rows;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        I base = rs;
        if constexpr (EY[q] > 0) base += oyp;
        if constexpr (EY[q] < 0) base += oym;
        if constexpr (EZ[q] > 0) base += ozp;
        if constexpr (EZ[q] < 0) base += ozm;
        Pair2<T> v;
        if constexpr (EX[q] == 0) {
            v.x = (T)g[0][q];
            v.y = (T)g[1][q];
        } else {
            __syncthreads();
            if (active) {
                row[(2 * t + EX[q] + nx) % nx] = (T)g[0][q];
                row[(2 * t + 1 + EX[q] + nx) % nx] = (T)g[1][q];
            }
            __syncthreads();
            v.x = row[2 * t];
            v.y = row[2 * t + 1];
        }
        if (active) st_pair<NT>(&P::at(B, q, (I)(base + (I)(2 * t)), L), v);
    });
}

template <typename T, int OP, int ORDER>
static hipError_t bulk_go(bool periodic, const KShift* k, const void* A, void* B, const float* M, const Grid& g,
                          hipStream_t s) {
    const int bx = g.nx >= kBulkBlock ? kBulkBlock : ((g.nx + 63) / 64) * 64;
    dim3 block(bx, 1, 1), grid((g.nx + bx - 1) / bx, g.ny, g.nz);
    // the macro array M is float32 like a float32 population plane
    const BulkForm form = bulk_form(g.L, sizeof(T), sizeof(float), g.variant);
    const bool narrow = form.narrow, nt = form.nt;
    if (periodic && !g.xopen && !g.xlodi && g.nx % 128 == 0 && g.nx / 2 <= kBulkBlock) {
        // float32 without HRR: as many whole rows per workgroup as fit
        // kBulkBlock threads (512-node rows: 36.2 -> 38.9 GLUPS on 512^3); HRR
        // and float16 keep one row (two rows: HRR 22.6 -> 21.7, float16 75 -> 66)
        constexpr bool rows = sizeof(T) == 4 && OP != OP_HRR;
        const int ry = rows ? std::max(1, kBulkBlock / (g.nx / 2)) : 1;
        const dim3 b2(g.nx / 2, ry, 1), g2(1, (g.ny + ry - 1) / ry, g.nz);
#define BS2(NAR, NT)                                                                                                \
    hipLaunchKernelGGL((bulk_shift2<T, OP, ORDER, NAR, NT, rows>), g2, b2, 0, s, k, (const T*)A, (T*)B, M, g.nx, g.ny, \
                       g.nz, g.L)
        if (!narrow) BS2(false, true);
        else if (nt) BS2(true, true);
        else BS2(true, false);
#undef BS2
        return hipGetLastError();
    }
    // the whole row in one workgroup of whole waves: rows of 64 to 512 nodes (a
    // multiple of 64), whose realigned stores cover whole cache lines.  With
    // boundary nodes in the rows (LODI, open x faces) only on grids of at least
    // 4096 rows: on fewer the kernel is latency-bound and the exchange's
    // barriers cost more than the aligned stores gain (measured: a 256 x 128 x 3
    // duct 17.3 against 20.7 us per step, a 256^3 LODI cube 1.18 against 1.32 ms).
    const bool fully_periodic = periodic && !g.xopen && !g.xlodi;
    const int row_lds = g.nx == bx && (fully_periodic || (int64_t)g.ny * g.nz >= 4096);
#define BS(PER, NAR, NT)                                                                                             \
    hipLaunchKernelGGL((bulk_shift<T, OP, ORDER, PER, NAR, NT>), grid, block, 0, s, k, (const T*)A, (T*)B, M, g.nx, \
                       g.ny, g.nz, g.L, g.xopen, row_lds, g.xlodi)
    if (periodic) {
        if (!narrow) BS(true, false, true);
        else if (nt) BS(true, true, true);
        else BS(true, true, false);
    } else {
        if (!narrow) BS(false, false, true);
        else if (nt) BS(false, true, true);
        else BS(false, true, false);
    }
#undef BS
    return hipGetLastError();
}

template <typename T>
static hipError_t bulk_t(int op, int order, bool periodic, const KShift* k, const void* A, void* B, const float* M,
                         const Grid& g, hipStream_t s) {
#define BO(OPC)                                                                  \
    switch (order) {                                                             \
        case 2: return bulk_go<T, OPC, 2>(periodic, k, A, B, M, g, s);           \
        case 3: return bulk_go<T, OPC, 3>(periodic, k, A, B, M, g, s);           \
        case 4: return bulk_go<T, OPC, 4>(periodic, k, A, B, M, g, s);           \
        default: return hipErrorInvalidValue;                                    \
    }
    switch (op) {
        case OP_SRT: BO(OP_SRT)
        case OP_RR: BO(OP_RR)
        case OP_HRR: BO(OP_HRR)
        case OP_MRT_RM: return bulk_go<T, OP_MRT_RM, 3>(periodic, k, A, B, M, g, s);
        case OP_MRT_GS: return bulk_go<T, OP_MRT_GS, 3>(periodic, k, A, B, M, g, s);
        case OP_STREAM: return bulk_go<T, OP_STREAM, 3>(periodic, k, A, B, M, g, s);
        default: return hipErrorInvalidValue;
    }
#undef BO
}

template <typename T>
static hipError_t bulk_packed_t(int op, int order, const PackTable* t, hipStream_t s) {
    const dim3 grid(kPackBulkGrid), block(kPackChunk);
#define BP(OPC, O) hipLaunchKernelGGL((bulk_shift_packed<T, OPC, O>), grid, block, 0, s, t)
#define BPO(OPC)                                  \
    switch (order) {                              \
        case 2: BP(OPC, 2); break;                \
        case 3: BP(OPC, 3); break;                \
        case 4: BP(OPC, 4); break;                \
        default: return hipErrorInvalidValue;     \
    }                                             \
    break
    switch (op) {
        case OP_SRT: BPO(OP_SRT);
        case OP_RR: BPO(OP_RR);
        case OP_HRR: BPO(OP_HRR);
        case OP_MRT_RM: BP(OP_MRT_RM, 3); break;
        case OP_MRT_GS: BP(OP_MRT_GS, 3); break;
        case OP_STREAM: BP(OP_STREAM, 3); break;
        default: return hipErrorInvalidValue;
    }
#undef BPO
#undef BP
    return hipGetLastError();
}

hipError_t launch_bulk_shift_packed(int prec, int op, int order, const PackTable* t, hipStream_t s) {
    if (prec == ACP_PREC_F32_SHIFTED) return bulk_packed_t<float>(op, order, t, s);
    return bulk_packed_t<_Float16>(op, order, t, s);
}

hipError_t launch_bulk_shift(int prec, int op, int order, bool periodic, const KShift* k, const void* A, void* B,
                             const float* macro, const Grid& g, hipStream_t s) {
    if (prec == ACP_PREC_F32_SHIFTED) return bulk_t<float>(op, order, periodic, k, A, B, macro, g, s);
    return bulk_t<_Float16>(op, order, periodic, k, A, B, macro, g, s);
}

// ------------------------------------------------------------------ HRR pre-pass
template <typename T>
__device__ __forceinline__ void macro_shift_node(const KShift& k, const T* __restrict__ A, float* __restrict__ M,
                                                 const PopLayout& L, int64_t l) {
    const int64_t N = L.N;
    float d[19], dr, ux, uy, uz;
    load_dev(A, L, l, k.inv_s, d);
    moments_dev(d, dr, ux, uy, uz);
    M[l] = dr;
    M[N + l] = ux;
    M[2 * N + l] = uy;
    M[3 * N + l] = uz;
}

template <typename T>
__global__ __launch_bounds__(256) void macro_shift(const KShift* __restrict__ kp, const T* __restrict__ A,
                                                   float* __restrict__ M, const PopLayout L) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L.N) return;
    macro_shift_node(*kp, A, M, L, l);
}

// Packed form: the pre-pass of every HRR member of a pack.
template <typename T>
__global__ __launch_bounds__(kPackChunk) void macro_shift_packed(const PackTable* __restrict__ t) {
    const PackList& pl = t->bulk[OP_HRR];
    const int total = pl.prefix[pl.n];
    int m = 0;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        m = pack_member(pl, w, m);
        const PackSlot& s = t->slot[pl.slot[m]];
        if (s.steps_left <= 0) continue;
        const int64_t l = (int64_t)(w - pl.prefix[m]) * kPackChunk + threadIdx.x;
        if (l >= s.L.N) continue;
        macro_shift_node(*(const KShift*)s.k, (const T*)s.A, (float*)s.M, s.L, l);
    }
}

hipError_t launch_macro_shift_packed(int prec, const PackTable* t, hipStream_t s) {
    const dim3 grid(kPackBulkGrid), block(kPackChunk);
    if (prec == ACP_PREC_F32_SHIFTED) hipLaunchKernelGGL(macro_shift_packed<float>, grid, block, 0, s, t);
    else hipLaunchKernelGGL(macro_shift_packed<_Float16>, grid, block, 0, s, t);
    return hipGetLastError();
}

hipError_t launch_macro_shift(int prec, const KShift* k, const void* A, float* macro, const Grid& g, hipStream_t s) {
    dim3 grid((unsigned)((g.N + 255) / 256));
    if (prec == ACP_PREC_F32_SHIFTED)
        hipLaunchKernelGGL(macro_shift<float>, grid, dim3(256), 0, s, k, (const float*)A, macro, g.L);
    else
        hipLaunchKernelGGL(macro_shift<_Float16>, grid, dim3(256), 0, s, k, (const _Float16*)A, macro, g.L);
    return hipGetLastError();
}

// ------------------------------------------------------------------ LODI
// Node.cpp:229-504 in lattice units on deviations: (rho_b - rho0) and the density
// differences of the one-sided gradient are formed from the stored deviations.
template <typename T, int ORDER>
__device__ __forceinline__ void lodi_shift_node(const KShift& k, const T* __restrict__ A, T* __restrict__ B,
                                                const LodiNode& ln, const PopLayout& L, int transverse) {
    float d[19];
    float rb, vb[3], r1, v1[3], r2, v2[3];
    load_dev(A, L, ln.node, k.inv_s, d);
    moments_dev(d, rb, vb[0], vb[1], vb[2]);
    load_dev(B, L, ln.n1, k.inv_s, d);
    moments_dev(d, r1, v1[0], v1[1], v1[2]);
    load_dev(B, L, ln.n2, k.inv_s, d);
    moments_dev(d, r2, v2[0], v2[1], v2[2]);
    float dr = (4.f * r1 - r2 - 3.f * rb) * -0.5f;
    float vg[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) vg[a] = (4.f * v1[a] - v2[a] - 3.f * vb[a]) * -0.5f;
    float dT[2] = {0.f, 0.f};
    float vgT[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
    if (!ln.corner && transverse) {
        float rt[4], vt[4][3];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            load_dev(A, L, ln.t[t], k.inv_s, d);
            moments_dev(d, rt[t], vt[t][0], vt[t][1], vt[t][2]);
        }
        dT[0] = 0.5f * (rt[0] - rt[1]);
        dT[1] = 0.5f * (rt[2] - rt[3]);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vgT[0][a] = 0.5f * (vt[0][a] - vt[1][a]);
            vgT[1][a] = 0.5f * (vt[2][a] - vt[3][a]);
        }
    }
    const int dir = ln.dir;
    const bool neg = (dir == 0 || dir == 1 || dir == 4);
    if (neg) {
        dr = -dr;
        vg[0] = -vg[0];
        vg[1] = -vg[1];
        vg[2] = -vg[2];
    }
    const int axis = (dir == 0 || dir == 2) ? 0 : ((dir == 1 || dir == 3) ? 1 : 2);
    const int p0 = axis, p1 = axis == 0 ? 1 : 0, p2 = axis == 2 ? 1 : 2;
    const float svg[3] = {vg[p0], vg[p1], vg[p2]};
    const float svb[3] = {vb[p0], vb[p1], vb[p2]};
    const float svgT[2][3] = {{vgT[0][p0], vgT[0][p1], vgT[0][p2]}, {vgT[1][p0], vgT[1][p1], vgT[1][p2]}};
    const float cs = 0.57735026918962576f, cs2 = 1.f / 3.f;
    const float csDivL = cs * k.inv_len[axis];
    const float rho_b = 1.f + rb;
    const float T1 = -(cs2 * (rho_b * (svgT[0][1] + svgT[1][2]) + svb[1] * dT[0] + svb[2] * dT[1]) -
                       rho_b * cs * (svb[1] * svgT[0][0] + svb[2] * svgT[1][0]));
    const float T5 = -(cs2 * (rho_b * (svgT[0][1] + svgT[1][2]) + svb[1] * dT[0] + svb[2] * dT[1]) +
                       rho_b * cs * (svb[1] * svgT[0][0] + svb[2] * svgT[1][0]));
    const float Lin = k.k1 * (1.f - k.ma2) * csDivL * cs2 * rb;
    float L1, L5;
    if (neg) {
        L5 = Lin - k.ma * T5 + T5;
        L1 = (svb[0] - cs) * cs * (cs * dr - rho_b * svg[0]);
    } else {
        L1 = Lin - k.ma * T1 + T1;
        L5 = (svb[0] + cs) * cs * (cs * dr + rho_b * svg[0]);
    }
    const float L3 = svb[0] * svg[1], L4 = svb[0] * svg[2];
    const float T3 = -(svb[1] * svgT[0][1] + svb[2] * svgT[1][1]) - (cs2 / rho_b) * dT[0];
    const float T4 = -(svb[1] * svgT[0][2] + svb[2] * svgT[1][2]) - (cs2 / rho_b) * dT[1];
    const float dr_new = rb - 1.5f * (L5 + L1) + 1.5f * (T5 + T1);
    const float c2 = 1.f / (2.f * rho_b * cs);
    float u[3];
    u[p0] = svb[0] - c2 * (L5 - L1) + c2 * (T5 - T1);
    u[p1] = svb[1] - L3 + T3;
    u[p2] = svb[2] - L4 + T4;
    float eq[19];
    feq_dev<ORDER>(dr_new, u[0], u[1], u[2], eq, k.k6);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        B[pidx(L, q, ln.node)] = to_store<T>(eq[q] * k.s);
    });
}

template <typename T, int ORDER>
__global__ __launch_bounds__(256) void lodi_shift(const KShift* __restrict__ kp, const T* __restrict__ A,
                                                  T* __restrict__ B, const LodiNode* __restrict__ nodes, int64_t count,
                                                  const PopLayout L, int transverse) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= count) return;
    const LodiNode ln = nodes[s];
    lodi_shift_node<T, ORDER>(*kp, A, B, ln, L, transverse);
}

// Packed form: launch `phase` of every member's LODI pass.
template <typename T, int ORDER>
__global__ __launch_bounds__(kPackChunk) void lodi_shift_packed(const PackTable* __restrict__ t, int phase) {
    const PackList& pl = t->lodi[phase];
    const int total = pl.prefix[pl.n];
    int m = 0;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        m = pack_member(pl, w, m);
        const PackSlot& s = t->slot[pl.slot[m]];
        if (s.steps_left <= 0) continue;
        const int64_t r = s.phase_start[phase] + (int64_t)(w - pl.prefix[m]) * kPackChunk + threadIdx.x;
        if (r >= s.phase_start[phase + 1]) continue;
        const LodiNode ln = s.lodi[r];
        lodi_shift_node<T, ORDER>(*(const KShift*)s.k, (const T*)s.A, (T*)s.B, ln, s.L, s.transverse);
    }
}

hipError_t launch_lodi_shift_packed(int prec, int order, int phase, const PackTable* t, hipStream_t s) {
    const dim3 grid(kPackLodiGrid), block(kPackChunk);
#define LSP(T, O) hipLaunchKernelGGL((lodi_shift_packed<T, O>), grid, block, 0, s, t, phase)
    if (prec == ACP_PREC_F32_SHIFTED) {
        if (order == 2) LSP(float, 2);
        else if (order == 3) LSP(float, 3);
        else LSP(float, 4);
    } else {
        if (order == 2) LSP(_Float16, 2);
        else if (order == 3) LSP(_Float16, 3);
        else LSP(_Float16, 4);
    }
#undef LSP
    return hipGetLastError();
}

hipError_t launch_lodi_shift(int prec, int order, bool transverse, const KShift* k, const void* A, void* B,
                             const LodiNode* nodes, int64_t count, const PopLayout& L, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int bs = lodi_block(count);
    dim3 grid((unsigned)((count + bs - 1) / bs)), block(bs);
    const int tr = transverse ? 1 : 0;
#define LS(T, O) hipLaunchKernelGGL((lodi_shift<T, O>), grid, block, 0, s, k, (const T*)A, (T*)B, nodes, count, L, tr)
    if (prec == ACP_PREC_F32_SHIFTED) {
        if (order == 2) LS(float, 2);
        else if (order == 3) LS(float, 3);
        else LS(float, 4);
    } else {
        if (order == 2) LS(_Float16, 2);
        else if (order == 3) LS(_Float16, 3);
        else LS(_Float16, 4);
    }
#undef LS
    return hipGetLastError();
}

// density and velocity in float64 (SI) from the stored deviations
template <typename T>
__device__ __forceinline__ void macro_f64(const KShift& k, const T* __restrict__ A, const PopLayout& L, int64_t l, double& rho,
                                          double& ux, double& uy, double& uz) {
    double r = 0., jx = 0., jy = 0., jz = 0.;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double v = (double)A[pidx(L, q, l)];
        r += v;
        if constexpr (EX[q] != 0) jx += EX[q] * v;
        if constexpr (EY[q] != 0) jy += EY[q] * v;
        if constexpr (EZ[q] != 0) jz += EZ[q] * v;
    });
    const double is = 1.0 / (double)k.s;
    const double rl = 1.0 + r * is;
    rho = k.rho0 * rl;
    ux = jx * is / rl * k.c_si;
    uy = jy * is / rl * k.c_si;
    uz = jz * is / rl * k.c_si;
}

template <typename T>
__global__ __launch_bounds__(256) void macros_out_shift(const KShift* __restrict__ kp, const T* __restrict__ A,
                                                        double* __restrict__ rho, double* __restrict__ u,
                                                        const PopLayout L) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N = L.N;
    if (l >= N) return;
    double r, ux, uy, uz;
    macro_f64(*kp, A, L, l, r, ux, uy, uz);
    rho[l] = r;
    u[l] = ux;
    u[N + l] = uy;
    u[2 * N + l] = uz;
}

hipError_t launch_macros_out_shift(int prec, const KShift* k, const void* A, const Grid& g, double* rho, double* u,
                                   hipStream_t s) {
    dim3 grid((unsigned)((g.N + 255) / 256));
    if (prec == ACP_PREC_F32_SHIFTED)
        hipLaunchKernelGGL(macros_out_shift<float>, grid, dim3(256), 0, s, k, (const float*)A, rho, u, g.L);
    else
        hipLaunchKernelGGL(macros_out_shift<_Float16>, grid, dim3(256), 0, s, k, (const _Float16*)A, rho, u, g.L);
    return hipGetLastError();
}

// ------------------------------------------------------------------ random initial condition
__device__ __forceinline__ void philox(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += W0;
        k1 += W1;
    }
}

template <typename T, int ORDER>
__global__ __launch_bounds__(256) void init_random_shift(const KShift* __restrict__ kp, T* __restrict__ A,
                                                         T* __restrict__ B, const PopLayout L, uint64_t seed,
                                                         double rho_amp, double u_amp) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L.N) return;
    uint32_t c[4] = {(uint32_t)l, (uint32_t)((uint64_t)l >> 32), 0u, 0u};
    philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double uu[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) uu[a] = ((double)c[a] + 0.5) * (1.0 / 4294967296.0);
    const double two_pi = 6.283185307179586476925286766559;
    const double r0 = sqrt(-2.0 * log(uu[0])), r1 = sqrt(-2.0 * log(uu[2]));
    const double n0 = r0 * cos(two_pi * uu[1]), n1 = r0 * sin(two_pi * uu[1]);
    const double n2 = r1 * cos(two_pi * uu[3]), n3 = r1 * sin(two_pi * uu[3]);
    float eq[19];
    feq_dev<ORDER>((float)(rho_amp * n0), (float)(u_amp * n1), (float)(u_amp * n2), (float)(u_amp * n3), eq, kp->k6);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const T v = (T)(eq[q] * kp->s);
        A[pidx(L, q, l)] = v;
        B[pidx(L, q, l)] = v;
    });
}

hipError_t launch_init_random_shift(int prec, int order, const KShift* k, void* A, void* B, const Grid& g,
                                    uint64_t seed, double rho_amp, double u_amp, hipStream_t s) {
    dim3 grid((unsigned)((g.N + 255) / 256)), block(256);
#define IR(T, O) hipLaunchKernelGGL((init_random_shift<T, O>), grid, block, 0, s, k, (T*)A, (T*)B, g.L, seed, rho_amp, u_amp)
    if (prec == ACP_PREC_F32_SHIFTED) {
        if (order == 2) IR(float, 2);
        else if (order == 3) IR(float, 3);
        else IR(float, 4);
    } else {
        if (order == 2) IR(_Float16, 2);
        else if (order == 3) IR(_Float16, 3);
        else IR(_Float16, 4);
    }
#undef IR
    return hipGetLastError();
}

// ------------------------------------------------------------------ probes and mass (float64)
// One block; the row index is kept on the device (see probe64).
template <typename T>
__device__ __forceinline__ void probe_shift_row(const KShift& k, const T* __restrict__ A, const PopLayout& L,
                                                const int64_t* __restrict__ nodes, int nprobe,
                                                double* __restrict__ out, int32_t* __restrict__ slot, int32_t cap) {
    const int32_t row = *slot;
    for (int p = threadIdx.x; p < nprobe; p += blockDim.x) {
        double r, ux, uy, uz;
        macro_f64(k, A, L, nodes[p], r, ux, uy, uz);
        if (row < cap) out[(int64_t)row * nprobe + p] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) *slot = row + 1;
}

template <typename T>
__global__ __launch_bounds__(1024) void probe_shift(const KShift* __restrict__ kp, const T* __restrict__ A, const PopLayout L,
                                                    const int64_t* __restrict__ nodes, int nprobe,
                                                    double* __restrict__ out, int32_t* __restrict__ slot,
                                                    int32_t cap) {
    probe_shift_row(*kp, A, L, nodes, nprobe, out, slot, cap);
}

// Packed form: the probe row of every member (see probe64_packed).
template <typename T>
__global__ __launch_bounds__(256) void probe_shift_packed(const PackTable* __restrict__ t) {
    for (int i = 0; i < t->nslot; ++i) {
        const PackSlot& s = t->slot[i];
        if (s.steps_left <= 0 || s.nprobe <= 0) continue;
        probe_shift_row(*(const KShift*)s.k, (const T*)s.B, s.L, s.probe_nodes, s.nprobe, s.probe_buf, s.probe_slot,
                        s.probe_cap);
    }
}

hipError_t launch_probe_shift_packed(int prec, const PackTable* t, hipStream_t s) {
    if (prec == ACP_PREC_F32_SHIFTED) hipLaunchKernelGGL(probe_shift_packed<float>, dim3(1), dim3(256), 0, s, t);
    else hipLaunchKernelGGL(probe_shift_packed<_Float16>, dim3(1), dim3(256), 0, s, t);
    return hipGetLastError();
}

hipError_t launch_probe_shift(int prec, const KShift* k, const void* A, const PopLayout& L, const int64_t* nodes,
                              int nprobe, double* out, int32_t* slot, int32_t cap, hipStream_t s) {
    if (nprobe <= 0) return hipSuccess;
    dim3 grid(1), block(nprobe >= 1024 ? 1024 : ((nprobe + 63) / 64) * 64);
    if (prec == ACP_PREC_F32_SHIFTED)
        hipLaunchKernelGGL(probe_shift<float>, grid, block, 0, s, k, (const float*)A, L, nodes, nprobe, out, slot,
                           cap);
    else
        hipLaunchKernelGGL(probe_shift<_Float16>, grid, block, 0, s, k, (const _Float16*)A, L, nodes, nprobe, out,
                           slot, cap);
    return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void mass_shift(const KShift* __restrict__ kp, const T* __restrict__ A,
                                                  const PopLayout L, double* __restrict__ partial) {
    __shared__ double sh[256];
    double acc = 0.;
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < L.N; l += (int64_t)gridDim.x * blockDim.x) {
        double r = 0.;
#pragma unroll
        for (int q = 0; q < 19; ++q) r += (double)A[pidx(L, q, l)];
        acc += r;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        // sum over this block's nodes of rho0 (1 + sum_q g_q / s): the node count part is
        // added on the host from the grid size, here only the deviation part
        partial[blockIdx.x] = kp->rho0 * sh[0] / (double)kp->s;
    }
}

hipError_t launch_mass_shift(int prec, const KShift* k, const void* A, const PopLayout& L, double* partial,
                             int nblocks, hipStream_t s) {
    if (prec == ACP_PREC_F32_SHIFTED)
        hipLaunchKernelGGL(mass_shift<float>, dim3(nblocks), dim3(256), 0, s, k, (const float*)A, L, partial);
    else
        hipLaunchKernelGGL(mass_shift<_Float16>, dim3(nblocks), dim3(256), 0, s, k, (const _Float16*)A, L, partial);
    return hipGetLastError();
}

}  // namespace acp
