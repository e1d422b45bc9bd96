// Fused collide + push kernel, float64 tier.
//
// One thread per node, one workgroup per grid row (x is the fastest index), so
// that every cache line of the destination arrays is written by a single
// workgroup: the x-shifted stores of a row stay inside that row.  A (time t) is
// read once, B (time t+1) written once: 2*19*8 = 304 bytes per node and step,
// with the non-temporal hint when the arrays exceed the cache (NT, plane.hpp).
// Populations are addressed through the case's block layout (plane.hpp).
//
// Semantics follow Lattice::timeStepping* (Lattice.cpp:1524-1580): interior
// nodes collide in place and push all 18 moving populations to their neighbours
// (FluidNode::transport, FluidNode.h:22-32); boundary nodes do not collide and
// push only the links that point inward in every boundary dimension
// (BoundaryNode::transport with the pointers of Lattice::setConnectivity); every
// node copies its rest population.  B entries that nothing pushes keep their
// previous contents, as in the reference.  In periodic mode every node is
// interior and the neighbours wrap.
#include "k64.hpp"
#include "pack.hpp"
#include "plane.hpp"

namespace acp {

// The pre-collision (rho, ux, uy, uz) of the six neighbours along links 0..5
// from the macro planes M.  On an open x face (a fine level's coupled face,
// PERIODIC form with xopen) the missing neighbour's macros come by quadratic
// extrapolation, u(-1) = 3 u(0) - 3 u(1) + u(2), which makes the central
// difference the second-order one-sided one.
template <typename P, bool PERIODIC, typename I>
__device__ __forceinline__ void load_neighbour_macros(const double* __restrict__ M, I l, const PopLayout& L, int i,
                                                      int nx, int xopen, I oxp, I oyp, I oxm, I oym, I ozp, I ozm,
                                                      double nb[6][4]) {
    const I o[6] = {oxp, oyp, oxm, oym, ozp, ozm};
#pragma unroll
    for (int m = 0; m < 6; ++m)
#pragma unroll
        for (int v = 0; v < 4; ++v) nb[m][v] = P::soa(M, v, (I)(l + o[m]), L.N);
    if constexpr (PERIODIC) {
        if (((xopen & 1) && i == 0) || ((xopen & 2) && i == nx - 1)) {
            const int m = i == 0 ? 2 : 0;
            const I s1 = i == 0 ? (I)1 : (I)-1;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                nb[m][v] = (3. * P::soa(M, v, l, L.N) - 3. * P::soa(M, v, (I)(l + s1), L.N)) +
                           P::soa(M, v, (I)(l + s1 + s1), L.N);
        }
    }
}

// ---------------------------------------------------------------- MFMA trial
// The relaxed MRT-RM collision (collide_mrt_rm64) as one 19 x 19 matrix per
// case:  f* = f + C (feq - f),  C = M^-1 S M, with M the raw-moment transform,
// S the rates (omega for moments 5-9, 1 for the others) and M^-1 its inverse;
// in exact arithmetic the same update.  The 64 nodes of a wave form one block
// X = feq - f (64 x 19), staged through LDS (rows padded to 65 doubles against
// bank conflicts), and X C^T runs on the matrix cores: 4 x 2 output tiles of
// 16 nodes x 16 populations, each five steps of v_mfma_f64_16x16x4f64 over the
// 19 populations padded to 20.  Operand layout of that instruction for lane l
// (checked on the device): A[l % 16][l / 16], B[l / 16][l % 16], and D rows
// l / 16 + 4 v, column l % 16, v = 0..3.  Every thread of the workgroup
// must take part (the launcher admits only grids whose rows fill whole
// workgroups).  Ct: C^T, [kMfmaRows][kMfmaCols], zero beyond 19 x 19.
typedef double mfma_d4 __attribute__((ext_vector_type(4)));
constexpr int kMfmaBlock = 256;  // threads of a workgroup of the trial kernel (LDS: 4 waves x 20 x 65 doubles)

__device__ __forceinline__ void collide_mrt_rm_mfma(const K64& k, double f[19], const double* __restrict__ Ct) {
    __shared__ double X[kMfmaBlock / 64][kMfmaRows][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double m[19], eq[19];
    rm_moments64(k, f, m);
    rm_feq64(k, m[0], m[1] / m[0], m[2] / m[0], m[3] / m[0], eq);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        X[w][q][lane] = eq[q] - f[q];
    });
    X[w][19][lane] = 0.;
    __syncthreads();
    const int r = lane >> 4, c = lane & 15;
    mfma_d4 acc[2][4];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            acc[nt][mt] = mfma_d4{0., 0., 0., 0.};
#pragma unroll
            for (int kk = 0; kk < 5; ++kk) {
                const double a = X[w][4 * kk + r][16 * mt + c];
                const double b = Ct[(4 * kk + r) * kMfmaCols + 16 * nt + c];
                acc[nt][mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[nt][mt], 0, 0, 0);
            }
        }
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < 2; ++nt)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int v = 0; v < 4; ++v)
                if (16 * nt + c < 19) X[w][16 * nt + c][16 * mt + r + 4 * v] = acc[nt][mt][v];
    __syncthreads();
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] += X[w][q][lane];
    });
}

// HRR without the cubic terms, orders 2 and 3, is held to 128 registers (four
// waves per SIMD): at 512^3 order 3 runs 9.4 GLUPS there, with a few bytes of
// spill, against 8.2 at the 130 registers and three waves it takes
// unconstrained.  Order 4 would spill about 100 bytes and ran slower (7.6
// against 7.8 GLUPS), the cubic forms several hundred.
// The update of node (i, j, kz): collide and push.  `active` is false for the
// threads beyond a row held by one workgroup (row_lds), which keep only their
// part in the row's LDS exchange; row_lds is uniform per workgroup.
template <int OP, int ORDER, bool CUBIC, bool PERIODIC, bool NARROW, bool NT>
__device__ __forceinline__ void bulk64_node(const K64* kp, const double* A, double* B, double* psi, const double* M,
                                            int nx, int ny, int nz, const PopLayout& L,
                                            int xopen, int row_lds, int xlodi, int i, int j, int kz, bool active,
                                            typename Plane<NARROW>::idx l0 = 0) {
    using P = Plane<NARROW>;
    using I = typename P::idx;
    const I sxy = (I)nx * (I)ny;
    // l0: the first node of the block the arrays are addressed from (0, or
    // the workgroup's block-relative base; see bulk64)
    const I l = (I)i + (I)nx * (I)j + sxy * (I)kz - l0;
    const K64& k = *kp;

    int bi = 0, bj = 0, bk = 0;
    if constexpr (!PERIODIC) {
        bi = i == 0 ? 1 : (i == nx - 1 ? -1 : 0);
        bj = j == 0 ? 1 : (j == ny - 1 ? -1 : 0);
        bk = kz == 0 ? 1 : (kz == nz - 1 ? -1 : 0);
    } else {
        bi = (i == 0 && (xlodi & 1)) ? 1 : ((i == nx - 1 && (xlodi & 2)) ? -1 : 0);
    }
    const bool boundary = (bi | bj | bk) != 0;

    // neighbour offsets (wrapped in periodic mode); unsigned wrap-around in the
    // narrow form lands on the right index because the final index is in range
    I oxp = 1, oxm = (I)-1, oyp = (I)nx, oym = (I)0 - (I)nx, ozp = sxy, ozm = (I)0 - sxy;
    if constexpr (PERIODIC) {
        if (i == nx - 1) oxp = (I)0 - (I)(nx - 1);
        if (i == 0) oxm = (I)(nx - 1);
        if (j == ny - 1) oyp = (I)0 - (I)nx * (I)(ny - 1);
        if (j == 0) oym = (I)nx * (I)(ny - 1);
        if (kz == nz - 1) ozp = (I)0 - sxy * (I)(nz - 1);
        if (kz == 0) ozm = sxy * (I)(nz - 1);
    }

    // HRR without the cubic terms: the strain rate from the neighbours'
    // macros first, so that their 24 values and the node's 19 populations are
    // never live together (the schedule barrier keeps the population loads
    // below); collide_hrr64_strain then equals collide_rr64's hybrid form
    double S[6] = {0., 0., 0., 0., 0., 0.};
    if constexpr (OP == OP_HRR && !CUBIC) {
        if (active && !boundary) {
            double nb[6][4];
            load_neighbour_macros<P, PERIODIC>(M, l, L, i, nx, xopen, oxp, oyp, oxm, oym, ozp, ozm, nb);
            strain_psi64<false>(k, nb, S, nullptr);
        }
        __builtin_amdgcn_sched_barrier(0);
    }

    double f[19];
    if (active) {
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            f[q] = ld_pop<NT>(&P::at(A, q, l, L));
        });
    }

    if (active && !boundary) {
        if constexpr (OP == OP_SRT) {
            collide_srt64<ORDER, false>(k, f, nullptr);
        } else if constexpr (OP == OP_HRR && !CUBIC) {
            collide_hrr64_strain<ORDER>(k, f, S);
        } else if constexpr (OP == OP_RR || OP == OP_HRR) {
            double ps[19];
            if constexpr (CUBIC) {
                static_for<0, 19>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    ps[q] = ld_pop<NT>(&P::soa(psi, q, l, L.N));
                });
            }
            double nb[6][4];
            if constexpr (OP == OP_HRR || CUBIC)
                load_neighbour_macros<P, PERIODIC>(M, l, L, i, nx, xopen, oxp, oyp, oxm, oym, ozp, ozm, nb);
            collide_rr64<ORDER, OP == OP_HRR, CUBIC>(k, f, ps, nb);
            if constexpr (OP == OP_HRR || CUBIC) {
                if (psi != nullptr) {
                    static_for<0, 19>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        st_pop<NT>(&P::soa(psi, q, l, L.N), ps[q]);
                    });
                }
            }
        } else if constexpr (OP == OP_MRT_RM) {
            collide_mrt_rm64(k, f);
        } else if constexpr (OP == OP_MRT_RM_MFMA) {
            collide_mrt_rm_mfma(k, f, M);  // M carries the collision matrix in this form
        } else if constexpr (OP == OP_MRT_GS) {
            collide_mrt_gs64(k, f);
        }
    }

    // Rows held by one workgroup (row_lds, uniform per block, no thread left
    // early): the x-shifted links are realigned through LDS so that every store
    // of a wave covers whole cache lines.  Position i of the destination row
    // takes the value of node i - e_x, and is written only if that node pushes
    // the link (row_source_pushes: boundary nodes, open and LODI x faces);
    // positions nothing pushes keep their contents, as with direct stores.
    if (row_lds) {
        __shared__ double row[kBulkBlock];
        static_for<0, 18>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            I base = l - (I)i;
            if constexpr (EY[q] > 0) base += oyp;
            if constexpr (EY[q] < 0) base += oym;
            if constexpr (EZ[q] > 0) base += ozp;
            if constexpr (EZ[q] < 0) base += ozm;
            if constexpr (EX[q] == 0) {
                if (active && row_source_pushes<PERIODIC>(q, i, nx, bj, bk, xopen, xlodi))
                    st_pop<NT>(&P::at(B, q, (I)(base + (I)i), L), f[q]);
            } else {
                __syncthreads();
                if (active) row[(i + EX[q] + nx) % nx] = f[q];
                __syncthreads();
                if (active && row_source_pushes<PERIODIC>(q, i - EX[q], nx, bj, bk, xopen, xlodi))
                    st_pop<NT>(&P::at(B, q, (I)(base + (I)i), L), row[i]);
            }
        });
        if (active) st_pop<NT>(&P::at(B, 18, l, L), f[18]);
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
        I d = l;
        if constexpr (EX[q] > 0) d += oxp;
        if constexpr (EX[q] < 0) d += oxm;
        if constexpr (EY[q] > 0) d += oyp;
        if constexpr (EY[q] < 0) d += oym;
        if constexpr (EZ[q] > 0) d += ozp;
        if constexpr (EZ[q] < 0) d += ozm;
        st_pop<NT>(&P::at(B, q, d, L), f[q]);
    });
    st_pop<NT>(&P::at(B, 18, l, L), f[18]);
}

// Block-relative addressing (REL, with the narrow form): on arrays too large
// for 32-bit offsets from the array start, every access of a workgroup still
// lies within `rel_span` blocks of the workgroup's own block (its row's
// neighbours are at most one plane and a row away, or two planes on a grid
// wrapped in y, whose last row pushes to the first row of the next and
// previous planes), so the arrays are addressed from a
// uniform base `rel_span` blocks before that block, with 32-bit offsets from
// it: half the address registers of the 64-bit form.  Rows whose z
// neighbours wrap to the far end of a periodic grid run in a separate launch
// of the 64-bit form.  Planes kz = kz0 + kzstep * blockIdx.z.
template <int OP, int ORDER, bool CUBIC, bool PERIODIC, bool NARROW, bool NT, bool REL = false>
__global__ __launch_bounds__(kBulkBlock) __attribute__((amdgpu_waves_per_eu(OP == OP_HRR && !CUBIC && ORDER < 4 ? ACP_HRR_WAVES : ACP_BULK_WAVES))) void bulk64(const K64* __restrict__ kp, const double* __restrict__ A,
                                                     double* __restrict__ B, double* __restrict__ psi,
                                                     const double* __restrict__ M, int nx, int ny, int nz,
                                                     const PopLayout L, int xopen, int row_lds, int xlodi,
                                                     int kz0, int kzstep, int rel_span, const Hole hole) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y, kz = kz0 + kzstep * (int)blockIdx.z;
    // a row crossing a hole (never with row_lds: the launcher takes the plain
    // rows) packs the nodes on either side of it into the first lanes, so the
    // waves over the hole exit at once instead of running part-filled
    if (!row_lds && hole.crosses(j, kz) && i > hole.b[0]) i += hole.b[1] - hole.b[0] - 1;
    // a row held by one workgroup (row_lds) keeps the threads beyond it for the
    // LDS exchange of its stores; they load, compute and store nothing
    const bool active = i < nx;
    if (!active && !row_lds) return;
    if constexpr (REL) {
        static_assert(NARROW, "block-relative addressing takes 32-bit offsets");
        const int64_t first = (int64_t)nx * ((int64_t)j + (int64_t)ny * kz);
        const int64_t b = first >> L.shift;
        const int64_t b0 = b > rel_span ? b - rel_span : 0;
        const int64_t l0 = b0 << L.shift;
        bulk64_node<OP, ORDER, CUBIC, PERIODIC, true, NT>(kp, A + b0 * L.block, B + b0 * L.block,
                                                          psi ? psi + l0 : nullptr, M ? M + l0 : nullptr, nx, ny, nz,
                                                          L, xopen, row_lds, xlodi, i, j, kz, active, (uint32_t)l0);
    } else {
        bulk64_node<OP, ORDER, CUBIC, PERIODIC, NARROW, NT>(kp, A, B, psi, M, nx, ny, nz, L, xopen, row_lds, xlodi, i,
                                                            j, kz, active);
    }
}

// Packed form (pack.hpp): one persistent grid over the 256-node chunks of
// every member of a pack with operator OP.  Members are one-level grids with
// boundary nodes on every face (LODI), addressed in the narrow form with
// plain loads and stores; each node is updated exactly as bulk64 updates it.
template <int OP, int ORDER>
__global__ __launch_bounds__(kPackChunk) __attribute__((amdgpu_waves_per_eu(OP == OP_HRR && ORDER < 4 ? ACP_HRR_WAVES : ACP_BULK_WAVES))) void bulk64_packed(const PackTable* __restrict__ t) {
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
        bulk64_node<OP, ORDER, false, false, true, false>((const K64*)s.k, (const double*)s.A, (double*)s.B, nullptr,
                                                          (const double*)s.M, s.nx, s.ny, s.nz, s.L, 0, 0, 0, (int)i,
                                                          (int)j, (int)kz, true);
    }
}

template <int OP, int ORDER, bool CUBIC, bool PERIODIC>
static hipError_t go(const K64* k, const double* A, double* B, double* psi, const double* M, const Grid& g,
                     hipStream_t s) {
    const int bx = g.nx >= kBulkBlock ? kBulkBlock : ((g.nx + 63) / 64) * 64;
    dim3 block(bx, 1, 1), grid((g.nx + bx - 1) / bx, g.ny, g.nz);
    // the whole row in one workgroup of whole waves: rows of 64 to 512 nodes (a
    // multiple of 64), whose realigned stores cover whole cache lines.  With
    // boundary nodes in the rows (LODI, open x faces) only on grids of at least
    // 4096 rows: on fewer the kernel is latency-bound and the exchange's
    // barriers cost more than the aligned stores gain (measured: a 256 x 128 x 3
    // duct 17.3 against 20.7 us per step, a 256^3 LODI cube 1.18 against 1.32 ms).
    const bool fully_periodic = PERIODIC && !g.xopen && !g.xlodi;
    const int row_lds = g.nx == bx && g.hole.empty() && (fully_periodic || (int64_t)g.ny * g.nz >= 4096);
#define B64(NAR, NT, REL, GZ, KZ0, KZS)                                                                          \
    hipLaunchKernelGGL((bulk64<OP, ORDER, CUBIC, PERIODIC, NAR, NT, REL>), dim3(grid.x, grid.y, GZ), block, 0, s, k, A, \
                       B, psi, M, g.nx, g.ny, g.nz, g.L, g.xopen, row_lds, g.xlodi, KZ0, KZS, span, \
                       g.hole)
    // a row's accesses reach one plane and a row (z and y neighbours), and on
    // grids wrapped in y (the periodic form) the wrapped row of the next or
    // previous plane: nearly two planes
    const int64_t plane = (int64_t)g.nx * g.ny;
    const int64_t reach = (PERIODIC ? 2 * plane : plane) + 2 * (int64_t)g.nx;
    const BulkForm form = bulk_form(g.L, sizeof(double), sizeof(double), g.variant, reach);
    const int span = (int)rel_span_blocks(reach, g.L);
    if (form.rel) {
        // planes whose z neighbours wrap (periodic grids) in the 64-bit form
        if (PERIODIC && g.nz > 2) {
            B64(true, true, true, g.nz - 2, 1, 1);
            B64(false, true, false, 2, 0, g.nz - 1);
        } else if (!PERIODIC) {
            B64(true, true, true, g.nz, 0, 1);
        } else {
            B64(false, true, false, g.nz, 0, 1);
        }
    } else if (!form.narrow) {
        B64(false, true, false, g.nz, 0, 1);
    } else if (form.nt) {
        B64(true, true, false, g.nz, 0, 1);
    } else {
        B64(true, false, false, g.nz, 0, 1);
    }
#undef B64
    return hipGetLastError();
}

template <int OP, int ORDER, bool CUBIC>
static hipError_t go_p(bool periodic, const K64* k, const double* A, double* B, double* psi, const double* M,
                       const Grid& g, hipStream_t s) {
    return periodic ? go<OP, ORDER, CUBIC, true>(k, A, B, psi, M, g, s)
                    : go<OP, ORDER, CUBIC, false>(k, A, B, psi, M, g, s);
}

template <int OP, bool CUBIC>
static hipError_t go_o(int order, bool periodic, const K64* k, const double* A, double* B, double* psi,
                       const double* M, const Grid& g, hipStream_t s) {
    switch (order) {
        case 2: return go_p<OP, 2, CUBIC>(periodic, k, A, B, psi, M, g, s);
        case 3: return go_p<OP, 3, CUBIC>(periodic, k, A, B, psi, M, g, s);
        case 4: return go_p<OP, 4, CUBIC>(periodic, k, A, B, psi, M, g, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_bulk64(int op, int order, bool cubic, bool periodic, const K64* k, const double* A, double* B,
                         double* psi, const double* macro, const Grid& g, hipStream_t s) {
    switch (op) {
        case OP_SRT: return go_o<OP_SRT, false>(order, periodic, k, A, B, psi, macro, g, s);
        case OP_RR:
            return cubic ? go_o<OP_RR, true>(order, periodic, k, A, B, psi, macro, g, s)
                         : go_o<OP_RR, false>(order, periodic, k, A, B, psi, macro, g, s);
        case OP_HRR:
            return cubic ? go_o<OP_HRR, true>(order, periodic, k, A, B, psi, macro, g, s)
                         : go_o<OP_HRR, false>(order, periodic, k, A, B, psi, macro, g, s);
        case OP_MRT_RM: return go_p<OP_MRT_RM, 3, false>(periodic, k, A, B, psi, macro, g, s);
        case OP_MRT_GS: return go_p<OP_MRT_GS, 3, false>(periodic, k, A, B, psi, macro, g, s);
        case OP_STREAM: return go_p<OP_STREAM, 3, false>(periodic, k, A, B, psi, macro, g, s);
        default: return hipErrorInvalidValue;
    }
}

// The trial kernel: the bulk kernel's periodic form with the matrix-core
// collision, 256-thread workgroups along x.
template <bool NARROW, bool NT>
__global__ __launch_bounds__(kMfmaBlock) void bulk64_mfma(const K64* __restrict__ kp, const double* __restrict__ A,
                                                          double* __restrict__ B, const double* __restrict__ Ct,
                                                          int nx, int ny, int nz, const PopLayout L) {
    const int i = blockIdx.x * kMfmaBlock + threadIdx.x;
    bulk64_node<OP_MRT_RM_MFMA, 3, false, true, NARROW, NT>(kp, A, B, nullptr, Ct, nx, ny, nz, L, 0, 0, 0, i,
                                                            (int)blockIdx.y, (int)blockIdx.z, true);
}

hipError_t launch_bulk64_mfma(const K64* k, const double* A, double* B, const double* Ct, const Grid& g,
                              hipStream_t s) {
    if (g.nx % kMfmaBlock != 0 || g.xopen || g.xlodi) return hipErrorInvalidValue;
    const dim3 grid(g.nx / kMfmaBlock, g.ny, g.nz), block(kMfmaBlock);
    const BulkForm form = bulk_form(g.L, sizeof(double), sizeof(double), BULK_AUTO);
    if (!form.narrow) hipLaunchKernelGGL((bulk64_mfma<false, true>), grid, block, 0, s, k, A, B, Ct, g.nx, g.ny, g.nz, g.L);
    else if (form.nt) hipLaunchKernelGGL((bulk64_mfma<true, true>), grid, block, 0, s, k, A, B, Ct, g.nx, g.ny, g.nz, g.L);
    else hipLaunchKernelGGL((bulk64_mfma<true, false>), grid, block, 0, s, k, A, B, Ct, g.nx, g.ny, g.nz, g.L);
    return hipGetLastError();
}

void mrt_rm_collision_matrix(const K64& k, double* Ct) {
    // M (moments of the unit populations), M^-1 (populations of the unit
    // moments), S; C = M^-1 S M
    double Mm[19][19], Mi[19][19];
    for (int q = 0; q < 19; ++q) {
        double e[19] = {0.}, m[19];
        e[q] = 1.;
        rm_moments64(k, e, m);
        for (int r = 0; r < 19; ++r) Mm[r][q] = m[r];
    }
    for (int r = 0; r < 19; ++r) {
        double e[19] = {0.}, f[19];
        e[r] = 1.;
        rm_inverse64(k, e, f);
        for (int q = 0; q < 19; ++q) Mi[q][r] = f[q];
    }
    for (int i = 0; i < kMfmaRows * kMfmaCols; ++i) Ct[i] = 0.;
    for (int q = 0; q < 19; ++q)        // output population
        for (int p = 0; p < 19; ++p) {  // input population
            double s = 0.;
            for (int r = 0; r < 19; ++r) s += Mi[q][r] * ((r >= 5 && r <= 9) ? k.omega : 1.) * Mm[r][p];
            Ct[p * kMfmaCols + q] = s;
        }
}

template <int OP>
static hipError_t packed_o(int order, const PackTable* t, hipStream_t s) {
    const dim3 grid(kPackBulkGrid), block(kPackChunk);
    switch (order) {
        case 2: hipLaunchKernelGGL((bulk64_packed<OP, 2>), grid, block, 0, s, t); break;
        case 3: hipLaunchKernelGGL((bulk64_packed<OP, 3>), grid, block, 0, s, t); break;
        case 4: hipLaunchKernelGGL((bulk64_packed<OP, 4>), grid, block, 0, s, t); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bulk64_packed(int op, int order, const PackTable* t, hipStream_t s) {
    switch (op) {
        case OP_SRT: return packed_o<OP_SRT>(order, t, s);
        case OP_RR: return packed_o<OP_RR>(order, t, s);
        case OP_HRR: return packed_o<OP_HRR>(order, t, s);
        case OP_MRT_RM: return packed_o<OP_MRT_RM>(3, t, s);
        case OP_MRT_GS: return packed_o<OP_MRT_GS>(3, t, s);
        case OP_STREAM: return packed_o<OP_STREAM>(3, t, s);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace acp
