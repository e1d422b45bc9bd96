// Float64 node physics with the reference solver's floating-point association.
//
// Every function evaluates the same sums and products, in the same order, as
// the reference's Node.cpp does, so that with FMA contraction disabled the
// results are bit-identical to the reference on the same inputs.  Terms whose
// lattice factor is an exact zero are dropped at compile time (adding +-0 does
// not change an IEEE sum with a non-zero term).  The lattice constants come
// from K64, filled on the host exactly as the reference derives them.
#pragma once

#include "lattice.hpp"
#include "static_for.hpp"

namespace acp {

struct K64 {
    double c, dt, omega, cs, cs2;
    double rcs2, c2, c3, c4;        // 1/cs^2, (1/2)/cs^4, (1/6)/cs^6, (1/4)/cs^8
    double hA, hB;                   // xi*xi - cs^2 for a non-zero / zero component
    double pA, pB;                   // pow(xi, 2.) - cs^2 (Node.cpp:2514-2516)
    double cc;                       // c*c, the magnitude of xi_a*xi_b
    // Third-order orthogonal Hermite values (Node.cpp:1024-1037): h3[a][q] is
    // a sign times one of four magnitudes, fixed per (a, q) at compile time
    // (h3_of), since xi = 0 or +-c and IEEE products and sums are symmetric in
    // sign: PB + PB, PA + PB, PA + PA, PA - PB with PA = (c c - cs^2) c and
    // PB = (0 0 - cs^2) c.  Holding 4 values instead of 114 keeps the tables
    // in scalar registers of the collision kernels.
    double h3m[4];
    // fourth-order fully orthogonal Hermite values (Node.cpp:1040-1052), with
    // their weights 8/7, 56/45, 40/27 already applied, per pattern of
    // non-zero components (bit 0 x, bit 1 y, bit 2 z)
    double h4e[8][3];
    double half_dt;                  // 0.5*dt
    double rec_cs4_strain;           // 1/(cs2*cs2) (Node.cpp:1758)
    double inv_pre;                  // -((dt*cs2)/(omega*nu)) (Node.cpp:1449)
    double nu, sigma;
    double dist;                     // neighbour distance of the strain stencil
    double lodi_cs2, dom_len[3], ma, ma2, k1, rho0, ls, dt_l;
    double rm_cp[5];                 // 1, c, pow(c,2.0), pow(c,3.0), pow(c,4.0) (Node.cpp:2695-2698)
    double rm_ix[5];                 // 1, 1/c, pow(1/c, 2..4) (Node.cpp:2768-2771)
    double gs_cp[5];                 // 1, c, pow(c,2.), pow(c,3.), pow(c,4.) (Node.cpp:2936-2939)
    double gs_pre[19][19];           // inverse Gram-Schmidt transform (Node.cpp:2878-2897)
};

// Copy of the case constants in LDS, for kernels of a few hundred to a few
// tens of thousands of threads (the coupling passes of a refined case).  Such
// a kernel is latency-bound, and its
// equilibria read the Hermite tables through a chain of dependent
// scalar-cache loads; read from LDS they cost one load latency for the whole
// table.  The first wave copies; every thread of the workgroup must call this
// (it holds a barrier) before any of them returns.
__device__ __forceinline__ const K64& stage_k64(const K64* __restrict__ kp, K64& sk) {
    constexpr int n = (int)(sizeof(K64) / sizeof(double));
    static_assert(sizeof(K64) % sizeof(double) == 0, "K64 holds doubles only");
    const double* src = reinterpret_cast<const double*>(kp);
    double* dst = reinterpret_cast<double*>(&sk);
    if (threadIdx.x < 64) {
        double v[(n + 63) / 64];
#pragma unroll
        for (int c = 0; c < (n + 63) / 64; ++c) {
            const int i = c * 64 + (int)threadIdx.x;
            v[c] = i < n ? src[i] : 0.;
        }
#pragma unroll
        for (int c = 0; c < (n + 63) / 64; ++c) {
            const int i = c * 64 + (int)threadIdx.x;
            if (i < n) dst[i] = v[c];
        }
    }
    __syncthreads();
    return sk;
}

// ---------------------------------------------------------------- moments (a3)
// Node.cpp:11-30 and 85-118
__host__ __device__ __forceinline__ void moments64(const K64& k, const double f[19], double& rho, double& ux,
                                                    double& uy, double& uz) {
    double r = 0., vx = 0., vy = 0., vz = 0.;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        r += f[q];
    });
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (EX[q] != 0) vx += (EX[q] > 0 ? k.c : -k.c) * f[q];
        if constexpr (EY[q] != 0) vy += (EY[q] > 0 ? k.c : -k.c) * f[q];
        if constexpr (EZ[q] != 0) vz += (EZ[q] > 0 ? k.c : -k.c) * f[q];
    });
    rho = r;
    ux = vx / r;
    uy = vy / r;
    uz = vz / r;
}

// the same with the 0.5*dt*psi terms of the cubic-Mach correction
__host__ __device__ __forceinline__ void moments64_psi(const K64& k, const double f[19], const double psi[19],
                                                        double& rho, double& ux, double& uy, double& uz) {
    double r = 0., vx = 0., vy = 0., vz = 0.;
    const double h = k.half_dt;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        r += f[q] + h * psi[q];
    });
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double xx = EX[q] == 0 ? 0. : (EX[q] > 0 ? k.c : -k.c);
        const double xy = EY[q] == 0 ? 0. : (EY[q] > 0 ? k.c : -k.c);
        const double xz = EZ[q] == 0 ? 0. : (EZ[q] > 0 ? k.c : -k.c);
        vx += xx * f[q] + h * xx * psi[q];
        vy += xy * f[q] + h * xy * psi[q];
        vz += xz * f[q] + h * xz * psi[q];
    });
    rho = r;
    ux = vx / r;
    uy = vy / r;
    uz = vz / r;
}

// sign (0: the value is zero) and magnitude index into K64::h3m of h3[a][q]
struct H3Ref {
    int sign, idx;
};
__host__ __device__ constexpr H3Ref h3_of(int a, int q) {
    const int ex = EX[q], ey = EY[q], ez = EZ[q];
    const int nx = ex != 0, ny = ey != 0, nz = ez != 0;
    switch (a) {
        case 0: return {ey, nx + nz};                           // xx y + zz y
        case 1: return {ez, nx + ny};                           // xx z + yy z
        case 2: return {ex, ny + nz};                           // yy x + zz x
        case 3: return {nz == nx ? 0 : ey * (nz ? 1 : -1), 3};  // zz y - xx y
        case 4: return {ny == nx ? 0 : ez * (ny ? 1 : -1), 3};  // yy z - xx z
        default: return {nz == ny ? 0 : ex * (nz ? 1 : -1), 3}; // zz x - yy x
    }
}
__host__ __device__ constexpr int h4_pattern(int q) { return (EX[q] != 0) | ((EY[q] != 0) << 1) | ((EZ[q] != 0) << 2); }

// sum_a h3[a][q] * v[a] * (a < 3 ? 3 : 1), left to right, as Node.cpp sums it;
// terms whose Hermite value is zero add +-0 to a non-zero partial sum or to a
// sum that only ever meets non-zero values, so they are left out
template <int Q>
__host__ __device__ __forceinline__ double h3_sum(const K64& k, const double v[6]) {
    double s = 0.;
    bool first = true;
    static_for<0, 6>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        constexpr H3Ref h = h3_of(a, Q);
        if constexpr (h.sign != 0) {
            const double hv = h.sign > 0 ? k.h3m[h.idx] : -k.h3m[h.idx];
            const double t = a < 3 ? hv * v[a] * 3. : hv * v[a];
            s = first ? t : s + t;
            first = false;
        }
    });
    return s;
}

// ---------------------------------------------------------------- equilibrium (a4)
// Node.cpp:739-813 (N=2), 816-949 (N=3), 952-1132 (N=4)
template <int ORDER>
__host__ __device__ __forceinline__ void feq64(const K64& k, double rho, double ux, double uy, double uz,
                                                double out[19]) {
    const double xx = ux * ux, yy = uy * uy, zz = uz * uz, xy = ux * uy, xz = ux * uz, yz = uy * uz;
    const double XXY = uy * xx, XYY = ux * yy, XXZ = uz * xx, XZZ = ux * zz, YYZ = uz * yy, YZZ = uy * zz;
    const double a3[6] = {XXY + YZZ, XXZ + YYZ, XYY + XZZ, YZZ - XXY, YYZ - XXZ, XZZ - XYY};
    const double o6 = 1. / 6., t7 = 2. / 7., t5 = 2. / 5.;
    double A = 0., B = 0., C = 0.;
    if constexpr (ORDER >= 4) {
        const double XXYY = uy * XXY, XXZZ = uz * XXZ, YYZZ = uy * YZZ;
        A = XXYY + o6 * zz;
        B = XXZZ + o6 * yy + t7 * A;
        C = YYZZ + o6 * xx + t7 * A + t5 * B;
    }
    const double e7 = 8. / 7., e45 = 56. / 45., e27 = 40. / 27.;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        double s1 = 0.;
        if constexpr (EX[q] != 0) s1 = (EX[q] > 0 ? k.c : -k.c) * ux;
        if constexpr (EY[q] != 0) {
            if constexpr (EX[q] != 0) s1 = s1 + (EY[q] > 0 ? k.c : -k.c) * uy;
            else s1 = (EY[q] > 0 ? k.c : -k.c) * uy;
        }
        if constexpr (EZ[q] != 0) {
            if constexpr (EX[q] != 0 || EY[q] != 0) s1 = s1 + (EZ[q] > 0 ? k.c : -k.c) * uz;
            else s1 = (EZ[q] > 0 ? k.c : -k.c) * uz;
        }
        const double hxx = EX[q] ? k.hA : k.hB, hyy = EY[q] ? k.hA : k.hB, hzz = EZ[q] ? k.hA : k.hB;
        double s2 = hxx * xx + hyy * yy + hzz * zz;
        if constexpr (EX[q] * EY[q] != 0) s2 = s2 + (EX[q] * EY[q] > 0 ? k.cc : -k.cc) * xy * 2.;
        if constexpr (EX[q] * EZ[q] != 0) s2 = s2 + (EX[q] * EZ[q] > 0 ? k.cc : -k.cc) * xz * 2.;
        if constexpr (EY[q] * EZ[q] != 0) s2 = s2 + (EY[q] * EZ[q] > 0 ? k.cc : -k.cc) * yz * 2.;
        double x = 1. * 1. + k.rcs2 * s1 + k.c2 * s2;
        if constexpr (ORDER >= 3) x = x + k.c3 * h3_sum<q>(k, a3);
        if constexpr (ORDER >= 4) {
            constexpr int p = h4_pattern(q);
            const double s4 = k.h4e[p][0] * A + k.h4e[p][1] * B + k.h4e[p][2] * C;
            x = x + k.c4 * s4;
        }
        out[q] = rho * weight(q) * x;
    });
}

// ---------------------------------------------------------------- Pi1 (a6)
// Node.cpp:2493-2558; pi = {xx, yy, zz, xy, xz, yz}
template <bool PSI>
__host__ __device__ __forceinline__ void pi1_64(const K64& k, const double f[19], const double eq[19],
                                                 const double* psi, double pi[6]) {
    double xx = 0., yy = 0., zz = 0., xy = 0., xz = 0., yz = 0.;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        double d;
        if constexpr (PSI) d = f[q] - eq[q] + k.half_dt * psi[q];
        else d = f[q] - eq[q];
        xx += (EX[q] ? k.pA : k.pB) * d;
        yy += (EY[q] ? k.pA : k.pB) * d;
        zz += (EZ[q] ? k.pA : k.pB) * d;
        if constexpr (EX[q] * EY[q] != 0) xy += (EX[q] * EY[q] > 0 ? k.cc : -k.cc) * d;
        if constexpr (EX[q] * EZ[q] != 0) xz += (EX[q] * EZ[q] > 0 ? k.cc : -k.cc) * d;
        if constexpr (EY[q] * EZ[q] != 0) yz += (EY[q] * EZ[q] > 0 ? k.cc : -k.cc) * d;
    });
    pi[0] = xx; pi[1] = yy; pi[2] = zz; pi[3] = xy; pi[4] = xz; pi[5] = yz;
}

// ---------------------------------------------------------------- regularised off-equilibrium (a6, a7)
// calcRRoffEquilibrium{2,3,4}_GHbasis (Node.cpp:1136-1427) with a2 = Pi1 and
// calcHRRoffEquilibrium{2,3,4}_GHbasis (Node.cpp:1430-1747) with the hybrid a2.
template <int ORDER>
__host__ __device__ __forceinline__ void offeq64(const K64& k, const double a2[6], double ux, double uy, double uz,
                                                  double out[19]) {
    const double XX = a2[0], YY = a2[1], ZZ = a2[2], XY = a2[3], XZ = a2[4], YZ = a2[5];
    double b3[6] = {0., 0., 0., 0., 0., 0.};
    if constexpr (ORDER >= 3) {
        const double tXXY = 2. * ux * XY + uy * XX, tXYY = 2. * uy * XY + ux * YY, tXXZ = 2. * ux * XZ + uz * XX;
        const double tXZZ = 2. * uz * XZ + ux * ZZ, tYYZ = 2. * uy * YZ + uz * YY, tYZZ = 2. * uz * YZ + uy * ZZ;
        b3[0] = tXXY + tYZZ; b3[1] = tXXZ + tYYZ; b3[2] = tXYY + tXZZ;
        b3[3] = tYZZ - tXXY; b3[4] = tYYZ - tXXZ; b3[5] = tXZZ - tXYY;
    }
    double A = 0., B = 0., C = 0.;
    if constexpr (ORDER >= 4) {
        const double o6 = 1. / 6., t7 = 2. / 7., t5 = 2. / 5.;
        const double fXXYY = uy * uy * XX + ux * ux * YY + 4. * ux * uy * XY;
        const double fXXZZ = uz * uz * XX + ux * ux * ZZ + 4. * ux * uz * XZ;
        const double fYYZZ = uy * uy * ZZ + uz * uz * YY + 4. * uy * uz * YZ;
        A = fXXYY + o6 * ZZ;
        B = fXXZZ + o6 * YY + t7 * A;
        C = fYYZZ + o6 * XX + t7 * A + t5 * B;
    }
    const double e7 = 8. / 7., e45 = 56. / 45., e27 = 40. / 27.;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double hxx = EX[q] ? k.hA : k.hB, hyy = EY[q] ? k.hA : k.hB, hzz = EZ[q] ? k.hA : k.hB;
        double s2 = hxx * XX + hyy * YY + hzz * ZZ;
        if constexpr (EX[q] * EY[q] != 0) s2 = s2 + (EX[q] * EY[q] > 0 ? k.cc : -k.cc) * XY * 2.;
        if constexpr (EX[q] * EZ[q] != 0) s2 = s2 + (EX[q] * EZ[q] > 0 ? k.cc : -k.cc) * XZ * 2.;
        if constexpr (EY[q] * EZ[q] != 0) s2 = s2 + (EY[q] * EZ[q] > 0 ? k.cc : -k.cc) * YZ * 2.;
        double x = k.c2 * s2;
        if constexpr (ORDER >= 3) x = x + k.c3 * h3_sum<q>(k, b3);
        if constexpr (ORDER >= 4) {
            constexpr int p = h4_pattern(q);
            const double s4 = k.h4e[p][0] * A + k.h4e[p][1] * B + k.h4e[p][2] * C;
            x = x + k.c4 * s4;
        }
        out[q] = weight(q) * x;
    });
}

// ---------------------------------------------------------------- FD strain rate and cubic-Mach terms (a7)
// Node.cpp:1750-1889.  nb[m] = pre-collision (rho, ux, uy, uz) of the neighbour along
// link m (0:+x 1:+y 2:-x 3:-y 4:+z 5:-z).  S = {xx, yy, zz, xy, xz, yz}.
template <bool WANT_PSI>
__host__ __device__ __forceinline__ void strain_psi64(const K64& k, const double nb[6][4], double S[6],
                                                       double psi[19]) {
    const double d = k.dist + k.dist;
    if constexpr (WANT_PSI) {
        double P[10][6];
        for (int m = 0; m < 6; ++m) {
            const double r = nb[m][0], vx = nb[m][1], vy = nb[m][2], vz = nb[m][3];
            P[0][m] = r * vx * vx * vx;
            P[1][m] = 0.5 * r * vy * vz * vz;
            P[2][m] = 0.5 * r * vz * vy * vy;
            P[3][m] = 0.5 * r * vx * vz * vz;
            P[4][m] = r * vy * vy * vy;
            P[5][m] = 0.5 * r * vz * vx * vx;
            P[6][m] = 0.5 * r * vx * vy * vy;
            P[7][m] = 0.5 * r * vy * vx * vx;
            P[8][m] = r * vz * vz * vz;
            P[9][m] = r * vx * vy * vz;
        }
        enum { XXX, XXY, XXZ, XYY, YYY, YYZ, XZZ, YZZ, ZZZ, XYZ };
        const double XXX_dx = (P[XXX][0] - P[XXX][2]) / d, XYY_dx = (P[XYY][0] - P[XYY][2]) / d;
        const double XZZ_dx = (P[XZZ][0] - P[XZZ][2]) / d, XXY_dx = (P[XXY][0] - P[XXY][2]) / d;
        const double XXZ_dx = (P[XXZ][0] - P[XXZ][2]) / d, XYZ_dx = (P[XYZ][0] - P[XYZ][2]) / d;
        const double XXY_dy = (P[XXY][1] - P[XXY][3]) / d, YYY_dy = (P[YYY][1] - P[YYY][3]) / d;
        const double YZZ_dy = (P[YZZ][1] - P[YZZ][3]) / d, XYY_dy = (P[XYY][1] - P[XYY][3]) / d;
        const double XYZ_dy = (P[XYZ][1] - P[XYZ][3]) / d, YYZ_dy = (P[YYZ][1] - P[YYZ][3]) / d;
        const double XXZ_dz = (P[XXZ][4] - P[XXZ][5]) / d, YYZ_dz = (P[YYZ][4] - P[YYZ][5]) / d;
        const double ZZZ_dz = (P[ZZZ][4] - P[ZZZ][5]) / d, XYZ_dz = (P[XYZ][4] - P[XYZ][5]) / d;
        const double XZZ_dz = (P[XZZ][4] - P[XZZ][5]) / d, YZZ_dz = (P[YZZ][4] - P[YZZ][5]) / d;
        const double gxx = XXX_dx + XXY_dy + XXZ_dz, gyy = XYY_dx + YYY_dy + YYZ_dz, gzz = XZZ_dx + YZZ_dy + ZZZ_dz;
        const double gxy = XXY_dx + XYY_dy + XYZ_dz, gxz = XXZ_dx + XYZ_dy + XZZ_dz, gyz = XYZ_dx + YYZ_dy + YZZ_dz;
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            double s = (EX[q] ? k.pA : k.pB) * gxx + (EY[q] ? k.pA : k.pB) * gyy + (EZ[q] ? k.pA : k.pB) * gzz;
            if constexpr (EX[q] * EY[q] != 0) s = s + (EX[q] * EY[q] > 0 ? k.cc : -k.cc) * gxy * 2.;
            if constexpr (EX[q] * EZ[q] != 0) s = s + (EX[q] * EZ[q] > 0 ? k.cc : -k.cc) * gxz * 2.;
            if constexpr (EY[q] * EZ[q] != 0) s = s + (EY[q] * EZ[q] > 0 ? k.cc : -k.cc) * gyz * 2.;
            psi[q] = 0.5 * weight(q) * k.rec_cs4_strain * s;
        });
    }
    S[0] = (nb[0][1] - nb[2][1]) / d;
    S[1] = (nb[1][2] - nb[3][2]) / d;
    S[2] = (nb[4][3] - nb[5][3]) / d;
    S[3] = 0.5 * ((nb[0][2] - nb[2][2]) / d + (nb[1][1] - nb[3][1]) / d);
    S[4] = 0.5 * ((nb[0][3] - nb[2][3]) / d + (nb[4][1] - nb[5][1]) / d);
    S[5] = 0.5 * ((nb[1][3] - nb[3][3]) / d + (nb[4][2] - nb[5][2]) / d);
}

// ---------------------------------------------------------------- collisions
// BGK, Node.cpp:511-557
template <int ORDER, bool CUBIC>
__host__ __device__ __forceinline__ void collide_srt64(const K64& k, double f[19], const double* psi) {
    double rho, ux, uy, uz, eq[19];
    if constexpr (CUBIC) moments64_psi(k, f, psi, rho, ux, uy, uz);
    else moments64(k, f, rho, ux, uy, uz);
    feq64<ORDER>(k, rho, ux, uy, uz, eq);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] += k.omega * (eq[q] - f[q]);
    });
}

// The part of RR / HRR after the strain rate (Node.cpp:560-650, 653-736):
// Pi1, the hybrid mix with the strain rate S (xx, yy, zz, xy, xz, yz), the
// regularised off-equilibrium and the relaxation, on the pre-collision
// equilibrium eq and velocity of the node.
template <int ORDER, bool HYBRID, bool CUBIC>
__host__ __device__ __forceinline__ void rr64_relax(const K64& k, double f[19], const double psi[19],
                                                    const double S[6], const double eq[19], double ux, double uy,
                                                    double uz) {
    double a2[6], f1[19];
    pi1_64<CUBIC>(k, f, eq, psi, a2);
    if constexpr (HYBRID) {
        double dens = 0.;
        if constexpr (CUBIC) {
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                dens += f[q] + k.half_dt * psi[q];
            });
        } else {
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                dens += f[q];
            });
        }
        const double dyn = k.nu * dens;
        for (int i = 0; i < 6; ++i) a2[i] = k.sigma * a2[i] + (1. - k.sigma) * k.inv_pre * (2. * dyn * S[i]);
    }
    offeq64<ORDER>(k, a2, ux, uy, uz, f1);
    const double om = k.omega;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (CUBIC) f[q] = eq[q] + (1 - om) * f1[q] + k.half_dt * psi[q];
        else f[q] = eq[q] + (1 - om) * f1[q];
    });
}

// RR (Node.cpp:560-650) and HRR (Node.cpp:653-736).  psi is the node's cubic-Mach
// state, updated in place; nb holds the neighbours' pre-collision macros.
template <int ORDER, bool HYBRID, bool CUBIC>
__host__ __device__ __forceinline__ void collide_rr64(const K64& k, double f[19], double psi[19],
                                                       const double nb[6][4]) {
    double rho, ux, uy, uz, eq[19];
    if constexpr (CUBIC) moments64_psi(k, f, psi, rho, ux, uy, uz);
    else moments64(k, f, rho, ux, uy, uz);
    feq64<ORDER>(k, rho, ux, uy, uz, eq);
    double S[6] = {0., 0., 0., 0., 0., 0.};
    // the reference always evaluates psi for HRR (Node.cpp:709); it only enters the
    // result with the cubic correction, so it is skipped otherwise
    if constexpr (HYBRID || CUBIC) strain_psi64<CUBIC>(k, nb, S, psi);
    rr64_relax<ORDER, HYBRID, CUBIC>(k, f, psi, S, eq, ux, uy, uz);
}

// HRR without the cubic-Mach terms with a given strain rate S: interface
// nodes of a two-level step collide with the strain rate their partners hand
// them (a grid-coupling node keeps the strain rate it received, Node.cpp:
// 1795-1798).
template <int ORDER>
__host__ __device__ __forceinline__ void collide_hrr64_strain(const K64& k, double f[19], const double S[6]) {
    double rho, ux, uy, uz, eq[19];
    moments64(k, f, rho, ux, uy, uz);
    feq64<ORDER>(k, rho, ux, uy, uz, eq);
    rr64_relax<ORDER, true, false>(k, f, nullptr, S, eq, ux, uy, uz);
}

// RR or HRR with the cubic-Mach terms at an interface node of a two-level
// step: its partner hands it psi (Node.cpp:2069) and, for HRR, the strain
// rate S before it collides, and a grid-coupling node's own strain-rate
// function returns at once (Node.cpp:1790-1793), so the handed psi enters its
// density and velocity (Node.cpp:11-30, 85-118), Pi1 and the result.
template <int ORDER, bool HYBRID>
__host__ __device__ __forceinline__ void collide_rr64_handed(const K64& k, double f[19], const double psi[19],
                                                              const double S[6]) {
    double rho, ux, uy, uz, eq[19];
    moments64_psi(k, f, psi, rho, ux, uy, uz);
    feq64<ORDER>(k, rho, ux, uy, uz, eq);
    rr64_relax<ORDER, HYBRID, true>(k, f, psi, S, eq, ux, uy, uz);
}

// raw moments (Node.cpp:2693-2762)
__host__ __device__ __forceinline__ void rm_moments64(const K64& k, const double f[19], double m[19]) {
    // coefficients of M_abc per link: row r, link q -> -1, 0, 1
    auto coef = [](int r, int q) constexpr -> int {
        const int x = EX[q], y = EY[q], z = EZ[q];
        const int X = x * x, Y = y * y, Z = z * z;
        switch (r) {
            case 0: return 1;
            case 1: return x;
            case 2: return y;
            case 3: return z;
            case 5: return X - Z;
            case 6: return Y - Z;
            case 7: return x * y;
            case 8: return x * z;
            case 9: return y * z;
            case 10: return X * z;
            case 11: return x * Z;
            case 12: return Y * z;
            case 13: return y * Z;
            case 14: return X * y;
            case 15: return x * Y;
            case 16: return X * Z;
            case 17: return X * Y;
            case 18: return Y * Z;
            default: return 0;
        }
    };
    constexpr int PW[19] = {0, 1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 3, 4, 4, 4};
    static_for<0, 19>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        double s = 0.;
        if constexpr (r == 4) {
            static_for<0, 6>([&](auto qc) { s += f[decltype(qc)::value]; });
            double t = 0.;
            static_for<6, 18>([&](auto qc) { t += f[decltype(qc)::value]; });
            s += 2. * t;
        } else {
            static_for<0, 19>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int a = coef(r, q);
                if constexpr (a == 1) s += f[q];
                else if constexpr (a == -1) s -= f[q];
            });
        }
        if constexpr (r == 0) m[r] = s;
        else m[r] = s * k.rm_cp[PW[r]];
    });
}

// Coreixas' extended raw-moment equilibrium (Node.cpp:2650-2690), per link class
__host__ __device__ __forceinline__ void rm_feq64(const K64& k, double rho, double ux, double uy, double uz,
                                                   double out[19]) {
    const double x1 = k.rm_ix[1], x2 = k.rm_ix[2], x3 = k.rm_ix[3], x4 = k.rm_ix[4];
    const double U[3] = {ux, uy, uz};
    const double S[3] = {ux * ux, uy * uy, uz * uz};
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int n = (EX[q] != 0) + (EY[q] != 0) + (EZ[q] != 0);
        double t;
        if constexpr (n == 0) {
            t = 1 - (S[0] + S[1] + S[2]) * x2 + 3 * (S[0] * S[1] + S[0] * S[2] + S[1] * S[2]) * x4;
        } else if constexpr (n == 1) {
            constexpr int a = EX[q] ? 0 : (EY[q] ? 1 : 2);
            constexpr int b = (a == 0) ? 1 : 0, cdim = (a == 2) ? 1 : 2;
            constexpr int s = ecomp(q, a);
            const double ua = s * U[a];
            t = 1 + 3 * ua * x1 + 3 * (S[a] - S[b] - S[cdim]) * x2 - 9 * ua * (S[b] + S[cdim]) * x3 -
                9 * S[a] * (S[b] + S[cdim]) * x4;
        } else {
            constexpr int a = EX[q] ? 0 : 1;
            constexpr int b = EZ[q] ? 2 : 1;
            constexpr int sa = ecomp(q, a), sb = ecomp(q, b);
            const double ua = sa * U[a], ub = sb * U[b];
            t = 1 + 3 * (ua + ub) * x1 + 3 * (S[a] + S[b]) * x2 + 9 * ua * ub * x2 + 9 * (S[a] * ub + ua * S[b]) * x3 +
                9 * S[a] * S[b] * x4;
        }
        out[q] = rho * weight(q) * t;
    });
}

// inverse raw-moment transform (Node.cpp:2765-2795)
__host__ __device__ __forceinline__ void rm_inverse64(const K64& k, const double m[19], double f[19]) {
    const double x1 = k.rm_ix[1], x2 = k.rm_ix[2], x3 = k.rm_ix[3], x4 = k.rm_ix[4];
    const double o6 = 1. / 6.;
    const double Pxx = o6 * (m[4] + 2. * m[5] - m[6]), Pyy = o6 * (m[4] - m[5] + 2. * m[6]),
                 Pzz = o6 * (m[4] - m[5] - m[6]);
    f[0] = 0.5 * m[1] * x1 + Pxx * x2 - 0.5 * (m[11] + m[15]) * x3 - 0.5 * (m[16] + m[17]) * x4;
    f[2] = -0.5 * m[1] * x1 + Pxx * x2 + 0.5 * (m[11] + m[15]) * x3 - 0.5 * (m[16] + m[17]) * x4;
    f[1] = 0.5 * m[2] * x1 + Pyy * x2 - 0.5 * (m[13] + m[14]) * x3 - 0.5 * (m[17] + m[18]) * x4;
    f[3] = -0.5 * m[2] * x1 + Pyy * x2 + 0.5 * (m[13] + m[14]) * x3 - 0.5 * (m[17] + m[18]) * x4;
    f[4] = 0.5 * m[3] * x1 + Pzz * x2 - 0.5 * (m[10] + m[12]) * x3 - 0.5 * (m[16] + m[18]) * x4;
    f[5] = -0.5 * m[3] * x1 + Pzz * x2 + 0.5 * (m[10] + m[12]) * x3 - 0.5 * (m[16] + m[18]) * x4;
    // diagonal link (sa, sb) in plane (a, b): M_ab, M_aab (sign of b), M_abb (sign of a), M_aabb
    auto diag = [&](int sa, int sb, int mab, int maab, int mabb, int m4) -> double {
        return 0.25 * (double)(sa * sb) * m[mab] * x2 + 0.25 * (sb * m[maab] + sa * m[mabb]) * x3 + 0.25 * m[m4] * x4;
    };
    f[6] = diag(1, 1, 8, 10, 11, 16);
    f[7] = diag(1, -1, 8, 10, 11, 16);
    f[10] = diag(-1, 1, 8, 10, 11, 16);
    f[11] = diag(-1, -1, 8, 10, 11, 16);
    f[8] = diag(1, 1, 9, 12, 13, 18);
    f[9] = diag(1, -1, 9, 12, 13, 18);
    f[12] = diag(-1, 1, 9, 12, 13, 18);
    f[13] = diag(-1, -1, 9, 12, 13, 18);
    f[16] = diag(1, 1, 7, 14, 15, 17);
    f[15] = diag(1, -1, 7, 14, 15, 17);
    f[14] = diag(-1, 1, 7, 14, 15, 17);
    f[17] = diag(-1, -1, 7, 14, 15, 17);
    f[18] = m[0] - m[4] * x2 + (m[16] + m[17] + m[18]) * x4;
}

// MRT raw moments with the relaxation applied (Node.cpp:2565-2605 without discarding the result)
__host__ __device__ __forceinline__ void collide_mrt_rm64(const K64& k, double f[19]) {
    double m[19], eq[19], meq[19];
    rm_moments64(k, f, m);
    rm_feq64(k, m[0], m[1] / m[0], m[2] / m[0], m[3] / m[0], eq);
    rm_moments64(k, eq, meq);
    static_for<0, 19>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if constexpr (r >= 5 && r <= 9) m[r] += k.omega * (meq[r] - m[r]);
        else m[r] += 1. * (meq[r] - m[r]);
    });
    rm_inverse64(k, m, f);
}

// Gram-Schmidt moments (Node.cpp:2934-3004) in lattice units and row powers of c
__host__ __device__ constexpr int gs_coef(int r, int q) {
    const int x = EX[q], y = EY[q], z = EZ[q];
    const int e2 = x * x + y * y + z * z;
    const int axis = e2 == 1 ? -2 : 1;
    switch (r) {
        case 0: return 1;
        case 1: return e2 == 2 ? 1 : (e2 == 0 ? -1 : 0);
        case 2: return e2 == 1 ? -2 : 1;
        case 3: return x;
        case 4: return axis * x;
        case 5: return y;
        case 6: return axis * y;
        case 7: return z;
        case 8: return axis * z;
        case 9: return 3 * x * x - e2;
        case 10: return e2 == 1 ? -(3 * x * x - e2) : (e2 == 2 ? 3 * x * x - e2 : 0);
        case 11: return y * y - z * z;
        case 12: return e2 == 1 ? -(y * y - z * z) : (e2 == 2 ? y * y - z * z : 0);
        case 13: return x * y;
        case 14: return y * z;
        case 15: return x * z;
        case 16: return x * (y * y - z * z);
        case 17: return y * (z * z - x * x);
        default: return z * (x * x - y * y);
    }
}
constexpr int GS_PW[19] = {0, 2, 4, 1, 3, 1, 3, 1, 3, 2, 4, 2, 4, 2, 2, 2, 3, 3, 3};

// Gram-Schmidt moments of a population set (calcAllMoments_GSbasis,
// Node.cpp:2934-3004: each row summed over q in order, then times its power of c)
__host__ __device__ __forceinline__ void gs_moments64(const K64& k, const double f[19], double m[19]) {
    static_for<0, 19>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        double s = 0.;
        static_for<0, 19>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int a = gs_coef(r, q);
            if constexpr (a == 1) s += f[q];
            else if constexpr (a == -1) s -= f[q];
            else if constexpr (a == 2) s += 2. * f[q];
            else if constexpr (a == -2) s -= 2. * f[q];
        });
        if constexpr (r == 0) m[r] = s;
        else m[r] = s * k.gs_cp[GS_PW[r]];
    });
}

// Populations of a set of Gram-Schmidt moments
// (transformMomentsToDistributions_GSbasis, Node.cpp:2870-2931: each
// population summed over the rows in order, starting from the first term)
__host__ __device__ __forceinline__ void gs_inverse64(const K64& k, const double m[19], double f[19]) {
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        double s = 0.;
        bool first = true;
        static_for<0, 19>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            if constexpr (gs_coef(r, q) != 0) {
                const double t = k.gs_pre[r][q] * m[r];
                s = first ? t : s + t;
                first = false;
            }
        });
        f[q] = s;
    });
}

__host__ __device__ __forceinline__ void collide_mrt_gs64(const K64& k, double f[19]) {
    double m[19];
    gs_moments64(k, f, m);
    // equilibrium moments (Node.cpp:2838-2867) and rates (Node.cpp:2815-2825)
    const double ir = 1. / m[0], jx = m[3], jy = m[5], jz = m[7];
    const double xx = jx * jx, yy = jy * jy, zz = jz * jz;
    double meq[19] = {0.};
    meq[1] = (xx + yy + zz) * ir;
    meq[9] = (2. * xx - yy - zz) * ir;
    meq[11] = (yy - zz) * ir;
    meq[13] = jx * jy * ir;
    meq[14] = jy * jz * ir;
    meq[15] = jx * jz * ir;
    double dm[19];
    static_for<0, 19>([&](auto rc) {
        constexpr int r = decltype(rc)::value;
        if constexpr (r == 0 || r == 3 || r == 5 || r == 7) dm[r] = 0.;
        else if constexpr (r == 1 || r == 9 || r == 11 || r == 13 || r == 14 || r == 15) dm[r] = k.omega * (meq[r] - m[r]);
        else dm[r] = 1. * (meq[r] - m[r]);
    });
    double df[19];
    gs_inverse64(k, dm, df);
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[q] += df[q];
    });
}

}  // namespace acp
