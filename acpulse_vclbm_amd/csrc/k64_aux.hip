// Float64 auxiliary kernels: LODI boundary pass, HRR macro pre-pass, device
// initial condition, probes and the mass reduction.
#include "k64.hpp"
#include "kshift.hpp"
#include "pack.hpp"
#include "plane.hpp"

namespace acp {

__device__ __forceinline__ void load19(const double* __restrict__ F, const PopLayout& L, int64_t l, double f[19]) {
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = F[pidx(L, q, l)];
    });
}

// 19 planes of N values (the cubic-Mach terms psi)
__device__ __forceinline__ void load19_soa(const double* __restrict__ F, int64_t N, int64_t l, double f[19]) {
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = F[q * N + l];
    });
}

// ------------------------------------------------------------------ LODI
// Node::LODIpressureOutlet + LODIpressureCalcVelocity (Node.cpp:229-504) for one
// boundary node: own state from A (time t), normal neighbours from B (time t+1,
// plus their cubic-Mach terms), tangent neighbours from A; writes feq_N of the
// characteristic update into the node's B entries.
template <int ORDER, bool CUBIC>
__device__ __forceinline__ void lodi64_node(const K64& k, const double* __restrict__ A, double* __restrict__ B,
                                            const double* __restrict__ psi, const LodiNode& ln, const PopLayout& PL,
                                            int transverse) {
    double f[19], ps[19];
    double rb, vb[3], r1, v1[3], r2, v2[3];
    load19(A, PL, ln.node, f);
    moments64(k, f, rb, vb[0], vb[1], vb[2]);
    load19(B, PL, ln.n1, f);
    if constexpr (CUBIC) {
        load19_soa(psi, PL.N, ln.n1, ps);
        moments64_psi(k, f, ps, r1, v1[0], v1[1], v1[2]);
    } else {
        moments64(k, f, r1, v1[0], v1[1], v1[2]);
    }
    load19(B, PL, ln.n2, f);
    if constexpr (CUBIC) {
        load19_soa(psi, PL.N, ln.n2, ps);
        moments64_psi(k, f, ps, r2, v2[0], v2[1], v2[2]);
    } else {
        moments64(k, f, r2, v2[0], v2[1], v2[2]);
    }
    const double ls = k.ls;
    double dr = (4. * r1 - r2 - 3. * rb) / (-2. * ls);
    double vg[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) vg[a] = (4. * v1[a] - v2[a] - 3. * vb[a]) / (-2. * ls);
    double dT[2] = {0., 0.};
    double vgT[2][3] = {{0., 0., 0.}, {0., 0., 0.}};
    if (!ln.corner && transverse) {
        double rt[4], vt[4][3];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            load19(A, PL, ln.t[t], f);
            moments64(k, f, rt[t], vt[t][0], vt[t][1], vt[t][2]);
        }
        dT[0] = (rt[0] - rt[1]) / (2. * ls);
        dT[1] = (rt[2] - rt[3]) / (2. * ls);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            vgT[0][a] = (vt[0][a] - vt[1][a]) / (2. * ls);
            vgT[1][a] = (vt[2][a] - vt[3][a]) / (2. * ls);
        }
    }
    const int dir = ln.dir;
    const bool neg = (dir == 0 || dir == 1 || dir == 4);
    if (neg) {
        dr *= -1.;
        vg[0] *= -1.;
        vg[1] *= -1.;
        vg[2] *= -1.;
    }
    // local frame with the normal axis first (Node.cpp:303-413)
    const int axis = (dir == 0 || dir == 2) ? 0 : ((dir == 1 || dir == 3) ? 1 : 2);
    const int p0 = axis, p1 = axis == 0 ? 1 : 0, p2 = axis == 2 ? 1 : 2;
    const double svg[3] = {vg[p0], vg[p1], vg[p2]};
    const double svb[3] = {vb[p0], vb[p1], vb[p2]};
    const double svgT[2][3] = {{vgT[0][p0], vgT[0][p1], vgT[0][p2]}, {vgT[1][p0], vgT[1][p1], vgT[1][p2]}};
    const double L = k.dom_len[axis];
    const double cs = k.cs, cs2 = k.lodi_cs2;
    const double csDivL = cs / L;
    const double K1 = k.k1, K2 = k.ma, MaSq = k.ma2;
    const double rout = k.rho0;
    const double dtl = k.dt_l;
    const double Tex = 0.;
    const double T1 = -(cs2 * (rb * (svgT[0][1] + svgT[1][2]) + svb[1] * dT[0] + svb[2] * dT[1]) -
                        rb * cs * (svb[1] * svgT[0][0] + svb[2] * svgT[1][0]));
    const double T5 = -(cs2 * (rb * (svgT[0][1] + svgT[1][2]) + svb[1] * dT[0] + svb[2] * dT[1]) +
                        rb * cs * (svb[1] * svgT[0][0] + svb[2] * svgT[1][0]));
    double L1, L5;
    if (neg) {
        L5 = K1 * (1 - MaSq) * csDivL * cs2 * (rb - rout) - K2 * (T5 - Tex) + T5;
        L1 = (svb[0] - cs) * cs * (cs * dr - rb * svg[0]);
    } else {
        L1 = K1 * (1 - MaSq) * csDivL * cs2 * (rb - rout) - K2 * (T1 - Tex) + T1;
        L5 = (svb[0] + cs) * cs * (cs * dr + rb * svg[0]);
    }
    const double L3 = svb[0] * svg[1];
    const double L4 = svb[0] * svg[2];
    const double T3 = -(svb[1] * svgT[0][1] + svb[2] * svgT[1][1]) - (cs2 / rb) * dT[0];
    const double T4 = -(svb[1] * svgT[0][2] + svb[2] * svgT[1][2]) - (cs2 / rb) * dT[1];
    const double rho = rb - dtl / (2. * cs2) * (L5 + L1) + dtl / (2. * cs2) * (T5 + T1);
    const double lv0 = svb[0] - dtl / (2. * rb * cs) * (L5 - L1) + dtl / (2. * rb * cs) * (T5 - T1);
    const double lv1 = svb[1] - dtl * L3 + dtl * T3;
    const double lv2 = svb[2] - dtl * L4 + dtl * T4;
    double u[3];
    u[p0] = lv0;
    u[p1] = lv1;
    u[p2] = lv2;
    double eq[19];
    feq64<ORDER>(k, rho, u[0], u[1], u[2], eq);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        B[pidx(PL, q, ln.node)] = eq[q];
    });
}

template <int ORDER, bool CUBIC>
__global__ __launch_bounds__(256) void lodi64(const K64* __restrict__ kp, const double* __restrict__ A,
                                              double* __restrict__ B, const double* __restrict__ psi,
                                              const LodiNode* __restrict__ nodes, int64_t count, const PopLayout PL,
                                              int transverse) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= count) return;
    const LodiNode ln = nodes[s];
    lodi64_node<ORDER, CUBIC>(*kp, A, B, psi, ln, PL, transverse);
}

// Packed form (pack.hpp): launch `phase` of every member's LODI pass, one
// persistent grid over 256-record chunks.
template <int ORDER>
__global__ __launch_bounds__(kPackChunk) void lodi64_packed(const PackTable* __restrict__ t, int phase) {
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
        lodi64_node<ORDER, false>(*(const K64*)s.k, (const double*)s.A, (double*)s.B, nullptr, ln, s.L,
                                  s.transverse);
    }
}

hipError_t launch_lodi64_packed(int order, int phase, const PackTable* t, hipStream_t s) {
    const dim3 grid(kPackLodiGrid), block(kPackChunk);
    switch (order) {
        case 2: hipLaunchKernelGGL(lodi64_packed<2>, grid, block, 0, s, t, phase); break;
        case 3: hipLaunchKernelGGL(lodi64_packed<3>, grid, block, 0, s, t, phase); break;
        case 4: hipLaunchKernelGGL(lodi64_packed<4>, grid, block, 0, s, t, phase); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_lodi64(int order, bool cubic, bool transverse, const K64* k, const double* A, double* B,
                         const double* psi, const LodiNode* nodes, int64_t count, const PopLayout& L, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    const int bs = lodi_block(count);
    dim3 grid((unsigned)((count + bs - 1) / bs));
    const int tr = transverse ? 1 : 0;
#define L64(O, C) hipLaunchKernelGGL((lodi64<O, C>), grid, dim3(bs), 0, s, k, A, B, psi, nodes, count, L, tr)
    if (cubic) {
        if (order == 2) L64(2, true);
        else if (order == 3) L64(3, true);
        else L64(4, true);
    } else {
        if (order == 2) L64(2, false);
        else if (order == 3) L64(3, false);
        else L64(4, false);
    }
#undef L64
    return hipGetLastError();
}

// ------------------------------------------------------------------ macro pre-pass
// Pre-collision density and velocity of every node (calcDensityPreCol /
// calcVelocityPreCol, Node.cpp:32-51, 121-154) for the HRR strain stencil.
// Populations larger than the last-level cache are read with the
// non-temporal hint (NT), as the bulk kernels read them; the macro planes are
// stored plainly, since the bulk kernel reads each of them from six nodes.
template <bool NT>
__device__ __forceinline__ void macro64_node(const K64& k, const double* __restrict__ A, double* __restrict__ M,
                                             const PopLayout& L, int64_t l) {
    const int64_t N = L.N;
    double f[19], r, ux, uy, uz;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] = ld_pop<NT>(&A[pidx(L, q, l)]);
    });
    moments64(k, f, r, ux, uy, uz);
    M[l] = r;
    M[N + l] = ux;
    M[2 * N + l] = uy;
    M[3 * N + l] = uz;
}

template <bool NT>
__global__ __launch_bounds__(256) void macro64(const K64* __restrict__ kp, const double* __restrict__ A,
                                               double* __restrict__ M, const PopLayout L, const Hole hole, int nx,
                                               int ny) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L.N) return;
    if (hole.b[1] > hole.b[0]) {
        const int64_t r = l / nx;
        if (hole.inside((int)(l - r * nx), (int)(r % ny), (int)(r / ny))) return;
    }
    macro64_node<NT>(*kp, A, M, L, l);
}

// Packed form: the pre-pass of every HRR member of a pack.
__global__ __launch_bounds__(kPackChunk) void macro64_packed(const PackTable* __restrict__ t) {
    const PackList& pl = t->bulk[OP_HRR];
    const int total = pl.prefix[pl.n];
    int m = 0;
    for (int w = blockIdx.x; w < total; w += gridDim.x) {
        m = pack_member(pl, w, m);
        const PackSlot& s = t->slot[pl.slot[m]];
        if (s.steps_left <= 0) continue;
        const int64_t l = (int64_t)(w - pl.prefix[m]) * kPackChunk + threadIdx.x;
        if (l >= s.L.N) continue;
        macro64_node<false>(*(const K64*)s.k, (const double*)s.A, (double*)s.M, s.L, l);
    }
}

hipError_t launch_macro64_packed(const PackTable* t, hipStream_t s) {
    hipLaunchKernelGGL(macro64_packed, dim3(kPackBulkGrid), dim3(kPackChunk), 0, s, t);
    return hipGetLastError();
}

hipError_t launch_macro64(const K64* k, const double* A, double* macro, const Grid& g, bool, hipStream_t s) {
    const int bs = 256;
    const dim3 grid((unsigned)((g.N + bs - 1) / bs));
    if (stream_populations(g.L, (int)sizeof(double)))
        hipLaunchKernelGGL(macro64<true>, grid, dim3(bs), 0, s, k, A, macro, g.L, g.hole, g.nx, g.ny);
    else
        hipLaunchKernelGGL(macro64<false>, grid, dim3(bs), 0, s, k, A, macro, g.L, g.hole, g.nx, g.ny);
    return hipGetLastError();
}

// macros in the output layout rho[N], u[3][N]
__global__ __launch_bounds__(256) void macros_out64(const K64* __restrict__ kp, const double* __restrict__ A,
                                                    double* __restrict__ rho, double* __restrict__ u, const PopLayout L) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t N = L.N;
    if (l >= N) return;
    double f[19], r, ux, uy, uz;
    load19(A, L, l, f);
    moments64(*kp, f, r, ux, uy, uz);
    rho[l] = r;
    u[l] = ux;
    u[N + l] = uy;
    u[2 * N + l] = uz;
}

hipError_t launch_macros_out64(const K64* k, const double* A, const Grid& g, double* rho, double* u,
                               hipStream_t s) {
    const int bs = 256;
    hipLaunchKernelGGL(macros_out64, dim3((unsigned)((g.N + bs - 1) / bs)), dim3(bs), 0, s, k, A, rho, u, g.L);
    return hipGetLastError();
}

// ------------------------------------------------------------------ random initial condition
// Philox-4x32-10 (Salmon et al., SC'11): counter = node index, key = seed.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += W0;
        k1 += W1;
    }
}

template <int ORDER>
__global__ __launch_bounds__(256) void init_random64(const K64* __restrict__ kp, double* __restrict__ A,
                                                     double* __restrict__ B, const PopLayout L, uint64_t seed, double rho0,
                                                     double rho_amp, double u_amp) {
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L.N) return;
    uint32_t c[4] = {(uint32_t)l, (uint32_t)((uint64_t)l >> 32), 0u, 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double uu[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) uu[a] = ((double)c[a] + 0.5) * (1.0 / 4294967296.0);
    const double two_pi = 6.283185307179586476925286766559;
    const double r0 = sqrt(-2.0 * log(uu[0])), r1 = sqrt(-2.0 * log(uu[2]));
    const double n0 = r0 * cos(two_pi * uu[1]), n1 = r0 * sin(two_pi * uu[1]);
    const double n2 = r1 * cos(two_pi * uu[3]), n3 = r1 * sin(two_pi * uu[3]);
    const K64& k = *kp;
    const double rho = rho0 * (1.0 + rho_amp * n0);
    const double us = u_amp * k.c;
    double eq[19];
    feq64<ORDER>(k, rho, us * n1, us * n2, us * n3, eq);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        A[pidx(L, q, l)] = eq[q];
        B[pidx(L, q, l)] = eq[q];
    });
}

hipError_t launch_init_random64(int order, const K64* k, double* A, double* B, const Grid& g, uint64_t seed,
                               double rho0, double rho_amp, double u_amp, hipStream_t s) {
    const int bs = 256;
    dim3 grid((unsigned)((g.N + bs - 1) / bs));
    if (order == 2) hipLaunchKernelGGL(init_random64<2>, grid, dim3(bs), 0, s, k, A, B, g.L, seed, rho0, rho_amp, u_amp);
    else if (order == 3) hipLaunchKernelGGL(init_random64<3>, grid, dim3(bs), 0, s, k, A, B, g.L, seed, rho0, rho_amp, u_amp);
    else hipLaunchKernelGGL(init_random64<4>, grid, dim3(bs), 0, s, k, A, B, g.L, seed, rho0, rho_amp, u_amp);
    return hipGetLastError();
}

// Equilibrium initial state from a density field given per (i, j) for a band
// of rows [j0, j0 + rows), u = 0, every k: feq_N(rho, 0) in float64 with the
// host's association (this translation unit has FMA contraction off), written
// to both arrays, float64 or converted to the shifted storage exactly as the
// population upload converts it, (T)(s * (f / rho0 - w_q)).
template <typename T, int ORDER>
__global__ __launch_bounds__(256) void init_rows(const K64* __restrict__ kp, const KShift* __restrict__ ksp,
                                                 const double* __restrict__ rho_rows, int j0, int rows, int nx, int ny,
                                                 int nz, T* __restrict__ A, T* __restrict__ B, const PopLayout L) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t band = (int64_t)rows * nx;
    if (t >= band * nz) return;
    const int64_t kz = t / band, r = t - kz * band;
    const int64_t l = (int64_t)j0 * nx + r + kz * (int64_t)nx * ny;
    double eq[19];
    feq64<ORDER>(*kp, rho_rows[r], 0., 0., 0., eq);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        T v;
        if constexpr (sizeof(T) == 8) v = eq[q];
        else v = (T)((double)ksp->s * (eq[q] / ksp->rho0 - weight(q)));
        A[pidx(L, q, l)] = v;
        B[pidx(L, q, l)] = v;
    });
}

// The same for a contiguous node range [l0, l0 + n) with one density per node.
template <typename T, int ORDER>
__global__ __launch_bounds__(256) void init_nodes(const K64* __restrict__ kp, const KShift* __restrict__ ksp,
                                                  const double* __restrict__ rho, int64_t l0, int64_t n,
                                                  T* __restrict__ A, T* __restrict__ B, const PopLayout L) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int64_t l = l0 + t;
    double eq[19];
    feq64<ORDER>(*kp, rho[t], 0., 0., 0., eq);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        T v;
        if constexpr (sizeof(T) == 8) v = eq[q];
        else v = (T)((double)ksp->s * (eq[q] / ksp->rho0 - weight(q)));
        A[pidx(L, q, l)] = v;
        B[pidx(L, q, l)] = v;
    });
}

hipError_t launch_init_nodes(int prec, int order, const K64* k, const KShift* ks, const double* rho, int64_t l0,
                             int64_t n, const Grid& g, void* A, void* B, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
#define IN(T, O) hipLaunchKernelGGL((init_nodes<T, O>), grid, block, 0, s, k, ks, rho, l0, n, (T*)A, (T*)B, g.L)
#define INO(T)                     \
    if (order == 2) IN(T, 2);      \
    else if (order == 3) IN(T, 3); \
    else IN(T, 4)
    if (prec == ACP_PREC_F64) INO(double);
    else if (prec == ACP_PREC_F32_SHIFTED) INO(float);
    else INO(_Float16);
#undef INO
#undef IN
    return hipGetLastError();
}

hipError_t launch_init_rows(int prec, int order, const K64* k, const KShift* ks, const double* rho_rows, int j0,
                            int rows, const Grid& g, void* A, void* B, hipStream_t s) {
    const int64_t n = (int64_t)rows * g.nx * g.nz;
    if (n <= 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
#define IR(T, O)                                                                                                  \
    hipLaunchKernelGGL((init_rows<T, O>), grid, block, 0, s, k, ks, rho_rows, j0, rows, g.nx, g.ny, g.nz, (T*)A, \
                       (T*)B, g.L)
#define IRO(T)                     \
    if (order == 2) IR(T, 2);      \
    else if (order == 3) IR(T, 3); \
    else IR(T, 4)
    if (prec == ACP_PREC_F64) IRO(double);
    else if (prec == ACP_PREC_F32_SHIFTED) IRO(float);
    else IRO(_Float16);
#undef IRO
#undef IR
    return hipGetLastError();
}

// ------------------------------------------------------------------ probes and mass
// One block: every thread reads the slot counter, writes its probes into row
// `slot` of the series (if below the capacity), and thread 0 advances the
// counter after the barrier.  Keeping the row index on the device lets a
// captured graph of several steps write a new row on every replay.
__device__ __forceinline__ void probe64_row(const double* __restrict__ A, const PopLayout& L,
                                            const int64_t* __restrict__ nodes, int nprobe, double* __restrict__ out,
                                            int32_t* __restrict__ slot, int32_t cap) {
    const int32_t row = *slot;
    for (int p = threadIdx.x; p < nprobe; p += blockDim.x) {
        const int64_t l = nodes[p];
        double r = 0.;
#pragma unroll
        for (int q = 0; q < 19; ++q) r += A[pidx(L, q, l)];
        if (row < cap) out[(int64_t)row * nprobe + p] = r;
    }
    __syncthreads();
    if (threadIdx.x == 0) *slot = row + 1;
}

__global__ __launch_bounds__(1024) void probe64(const double* __restrict__ A, const PopLayout L,
                                                const int64_t* __restrict__ nodes, int nprobe,
                                                double* __restrict__ out, int32_t* __restrict__ slot, int32_t cap) {
    probe64_row(A, L, nodes, nprobe, out, slot, cap);
}

// Packed form: the probe row of every member, one member after the other in
// one workgroup (the step's new populations are B, before the swap).
__global__ __launch_bounds__(256) void probe64_packed(const PackTable* __restrict__ t) {
    for (int i = 0; i < t->nslot; ++i) {
        const PackSlot& s = t->slot[i];
        if (s.steps_left <= 0 || s.nprobe <= 0) continue;
        probe64_row((const double*)s.B, s.L, s.probe_nodes, s.nprobe, s.probe_buf, s.probe_slot, s.probe_cap);
    }
}

hipError_t launch_probe64_packed(const PackTable* t, hipStream_t s) {
    hipLaunchKernelGGL(probe64_packed, dim3(1), dim3(256), 0, s, t);
    return hipGetLastError();
}

// Last kernel of a packed step: the arrays of every member with steps left
// change roles and its count goes down by one.
__global__ __launch_bounds__(64) void pack_advance(PackTable* __restrict__ t) {
    const int i = threadIdx.x;
    if (i >= t->nslot) return;
    PackSlot& s = t->slot[i];
    if (s.steps_left <= 0) return;
    void* a = s.A;
    s.A = s.B;
    s.B = a;
    s.steps_left -= 1;
}

hipError_t launch_pack_advance(PackTable* t, hipStream_t s) {
    static_assert(kPackSlots <= 64, "one thread per pack slot");
    hipLaunchKernelGGL(pack_advance, dim3(1), dim3(64), 0, s, t);
    return hipGetLastError();
}

hipError_t launch_probe64(const double* A, const PopLayout& L, const int64_t* nodes, int nprobe, double* out,
                          int32_t* slot, int32_t cap, hipStream_t s) {
    if (nprobe <= 0) return hipSuccess;
    const int bs = nprobe >= 1024 ? 1024 : ((nprobe + 63) / 64) * 64;
    hipLaunchKernelGGL(probe64, dim3(1), dim3(bs), 0, s, A, L, nodes, nprobe, out, slot, cap);
    return hipGetLastError();
}

// Order-independent 64-bit fingerprint of an array of n elements of T: the
// wrapping sum over elements of a mix of (bit pattern, index).  Integer sums
// are associative, so the result does not depend on the reduction order.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

template <typename T>
__global__ __launch_bounds__(256) void checksum_k(const T* __restrict__ a, int64_t n,
                                                  unsigned long long* __restrict__ partial) {
    __shared__ unsigned long long sh[256];
    uint64_t acc = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t bits = 0;
        if constexpr (sizeof(T) == 8) bits = __builtin_bit_cast(uint64_t, a[i]);
        else if constexpr (sizeof(T) == 4) bits = __builtin_bit_cast(uint32_t, a[i]);
        else bits = __builtin_bit_cast(uint16_t, a[i]);
        acc += mix64(bits ^ ((uint64_t)i * 0x9e3779b97f4a7c15ull));
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

hipError_t launch_checksum(const void* a, int elem_bytes, int64_t n, unsigned long long* partial, int nblocks,
                           hipStream_t s) {
    if (elem_bytes == 8)
        hipLaunchKernelGGL(checksum_k<double>, dim3(nblocks), dim3(256), 0, s, (const double*)a, n, partial);
    else if (elem_bytes == 4)
        hipLaunchKernelGGL(checksum_k<float>, dim3(nblocks), dim3(256), 0, s, (const float*)a, n, partial);
    else
        hipLaunchKernelGGL(checksum_k<_Float16>, dim3(nblocks), dim3(256), 0, s, (const _Float16*)a, n, partial);
    return hipGetLastError();
}

// per-block partial sums of sum_q f_q in a fixed order (deterministic)
__global__ __launch_bounds__(256) void mass64(const double* __restrict__ A, const PopLayout L,
                                              double* __restrict__ partial) {
    __shared__ double sh[256];
    double acc = 0.;
    for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < L.N; l += (int64_t)gridDim.x * blockDim.x) {
        double r = 0.;
#pragma unroll
        for (int q = 0; q < 19; ++q) r += A[pidx(L, q, l)];
        acc += r;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

hipError_t launch_mass64(const double* A, const PopLayout& L, double* partial, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(mass64, dim3(nblocks), dim3(256), 0, s, A, L, partial);
    return hipGetLastError();
}

// ------------------------------------------------------------------ host transfers
// Population plane q of nodes [l0, l0 + n) between a contiguous float64 buffer
// (SI values) and a case's arrays: float64 copied, the shifted tiers converted
// as (T)(s * (f / rho0 - w_q)) and back as rho0 * (w_q + g / s).
template <typename T>
__global__ __launch_bounds__(256) void put_plane(const KShift* __restrict__ kp, const double* __restrict__ src,
                                                 T* __restrict__ dA, T* __restrict__ dB, const PopLayout L, int q,
                                                 int64_t l0, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    T v;
    if constexpr (sizeof(T) == 8) v = src[t];
    else v = (T)((double)kp->s * (src[t] / kp->rho0 - weight(q)));
    const int64_t e = pidx(L, q, l0 + t);
    dA[e] = v;
    if (dB) dB[e] = v;
}

template <typename T>
__global__ __launch_bounds__(256) void get_plane(const KShift* __restrict__ kp, const T* __restrict__ src,
                                                 double* __restrict__ dst, const PopLayout L, int q, int64_t l0,
                                                 int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const T v = src[pidx(L, q, l0 + t)];
    if constexpr (sizeof(T) == 8) dst[t] = v;
    else dst[t] = kp->rho0 * (weight(q) + (double)v / (double)kp->s);
}

hipError_t launch_put_plane(int prec, const KShift* k, const double* src, void* dA, void* dB, const PopLayout& L,
                            int q, int64_t l0, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (prec == ACP_PREC_F64)
        hipLaunchKernelGGL(put_plane<double>, grid, block, 0, s, k, src, (double*)dA, (double*)dB, L, q, l0, n);
    else if (prec == ACP_PREC_F32_SHIFTED)
        hipLaunchKernelGGL(put_plane<float>, grid, block, 0, s, k, src, (float*)dA, (float*)dB, L, q, l0, n);
    else
        hipLaunchKernelGGL(put_plane<_Float16>, grid, block, 0, s, k, src, (_Float16*)dA, (_Float16*)dB, L, q, l0,
                           n);
    return hipGetLastError();
}

hipError_t launch_get_plane(int prec, const KShift* k, const void* src, double* dst, const PopLayout& L, int q,
                            int64_t l0, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const dim3 grid((unsigned)((n + 255) / 256)), block(256);
    if (prec == ACP_PREC_F64)
        hipLaunchKernelGGL(get_plane<double>, grid, block, 0, s, k, (const double*)src, dst, L, q, l0, n);
    else if (prec == ACP_PREC_F32_SHIFTED)
        hipLaunchKernelGGL(get_plane<float>, grid, block, 0, s, k, (const float*)src, dst, L, q, l0, n);
    else
        hipLaunchKernelGGL(get_plane<_Float16>, grid, block, 0, s, k, (const _Float16*)src, dst, L, q, l0, n);
    return hipGetLastError();
}

}  // namespace acp
