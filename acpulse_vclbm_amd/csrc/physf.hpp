// Float32 node physics of the shifted tiers, in lattice units (rho0 = 1,
// dx = dt = 1, cs^2 = 1/3), on population deviations d_q = f_q - w_q.
//
// The collision operators are the reference's (Node.cpp:511-736, 2565-2835)
// rewritten for deviations: density enters as rho - 1 = sum_q d_q and every
// equilibrium is produced as feq_q - w_q = w_q (drho + rho P_q(u)), so that the
// small acoustic perturbation keeps float32 precision instead of being added to
// and subtracted from O(1) weights.
#pragma once

#include "lattice.hpp"
#include "static_for.hpp"

namespace acp {

__host__ __device__ constexpr float wf(int q) { return q < 6 ? 1.f / 18.f : (q < 18 ? 1.f / 36.f : 1.f / 3.f); }

// d_q -> (rho - 1, u)
__device__ __forceinline__ void moments_dev(const float d[19], float& dr, float& ux, float& uy, float& uz) {
    float r = 0.f, jx = 0.f, jy = 0.f, jz = 0.f;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        r += d[q];
        if constexpr (EX[q] > 0) jx += d[q];
        if constexpr (EX[q] < 0) jx -= d[q];
        if constexpr (EY[q] > 0) jy += d[q];
        if constexpr (EY[q] < 0) jy -= d[q];
        if constexpr (EZ[q] > 0) jz += d[q];
        if constexpr (EZ[q] < 0) jz -= d[q];
    });
    dr = r;
    const float ir = 1.f / (1.f + r);
    ux = jx * ir;
    uy = jy * ir;
    uz = jz * ir;
}

// The fourth-order terms of the reference (Node.cpp:1032-1045 for the
// equilibrium, 1343-1357 for the off-equilibrium coefficients, and the
// Hermite values of each link) are written in SI units as, e.g.,
// H_xxyy + (1/6) H_zz: a
// fourth-order and a second-order Hermite polynomial of the molecular velocity
// added with a factor 1/6 that carries no units.  In lattice units (velocities
// divided by c = sqrt(3) c_s) the same expression reads H_xxyy + k6 H_zz with
// k6 = 1 / (6 c^2), c in m/s; the velocity polynomials A, B, C and the
// off-equilibrium coefficients mix orders the same way.  k6 is that factor.

// feq_q - w_q for the order-N Gauss-Hermite equilibrium (Node.cpp:739-1132).
template <int ORDER>
__device__ __forceinline__ void feq_dev(float dr, float ux, float uy, float uz, float out[19], float k6) {
    const float rho = 1.f + dr;
    const float xx = ux * ux, yy = uy * uy, zz = uz * uz;
    const float u2 = xx + yy + zz;
    const float XXY = uy * xx, XYY = ux * yy, XXZ = uz * xx, XZZ = ux * zz, YYZ = uz * yy, YZZ = uy * zz;
    const float a3[6] = {XXY + YZZ, XXZ + YYZ, XYY + XZZ, YZZ - XXY, YYZ - XXZ, XZZ - XYY};
    float A = 0.f, B = 0.f, C = 0.f;
    if constexpr (ORDER >= 4) {
        A = uy * XXY + k6 * zz;
        B = uz * XXZ + k6 * yy + (2.f / 7.f) * A;
        C = uy * YZZ + k6 * xx + (2.f / 7.f) * A + (2.f / 5.f) * B;
    }
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const float cu = (float)EX[q] * ux + (float)EY[q] * uy + (float)EZ[q] * uz;
        float p = 3.f * cu + 4.5f * cu * cu - 1.5f * u2;
        if constexpr (ORDER >= 3) {
            // orthogonal third-order Hermite values of link q (lattice units)
            constexpr float hx = EX[q], hy = EY[q], hz = EZ[q];
            constexpr float hxx = hx * hx - 1.f / 3.f, hyy = hy * hy - 1.f / 3.f, hzz = hz * hz - 1.f / 3.f;
            constexpr float H[6] = {hxx * hy + hzz * hy, hxx * hz + hyy * hz, hyy * hx + hzz * hx,
                                    hzz * hy - hxx * hy, hyy * hz - hxx * hz, hzz * hx - hyy * hx};
            const float s3 = 3.f * (H[0] * a3[0] + H[1] * a3[1] + H[2] * a3[2]) + H[3] * a3[3] + H[4] * a3[4] +
                             H[5] * a3[5];
            p += 4.5f * s3;
        }
        if constexpr (ORDER >= 4) {
            constexpr float hx = EX[q], hy = EY[q], hz = EZ[q];
            constexpr float hxx = hx * hx - 1.f / 3.f, hyy = hy * hy - 1.f / 3.f, hzz = hz * hz - 1.f / 3.f;
            const float a4 = hxx * hyy + k6 * hzz;
            const float b4 = hxx * hzz + k6 * hyy + (2.f / 7.f) * a4;
            const float c4 = hyy * hzz + k6 * hxx + (2.f / 7.f) * a4 + (2.f / 5.f) * b4;
            const float s4 = (8.f / 7.f) * a4 * A + (56.f / 45.f) * b4 * B + (40.f / 27.f) * c4 * C;
            p += 20.25f * s4;
        }
        out[q] = wf(q) * (dr + rho * p);
    });
}

// Pi1 = sum_q H2_q fneq_q, {xx, yy, zz, xy, xz, yz} (Node.cpp:2493-2558)
__device__ __forceinline__ void pi1_dev(const float fneq[19], float pi[6]) {
    float xx = 0.f, yy = 0.f, zz = 0.f, xy = 0.f, xz = 0.f, yz = 0.f, tr = 0.f;
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const float d = fneq[q];
        tr += d;
        if constexpr (EX[q] != 0) xx += d;
        if constexpr (EY[q] != 0) yy += d;
        if constexpr (EZ[q] != 0) zz += d;
        if constexpr (EX[q] * EY[q] > 0) xy += d;
        if constexpr (EX[q] * EY[q] < 0) xy -= d;
        if constexpr (EX[q] * EZ[q] > 0) xz += d;
        if constexpr (EX[q] * EZ[q] < 0) xz -= d;
        if constexpr (EY[q] * EZ[q] > 0) yz += d;
        if constexpr (EY[q] * EZ[q] < 0) yz -= d;
    });
    const float t3 = tr * (1.f / 3.f);
    pi[0] = xx - t3;
    pi[1] = yy - t3;
    pi[2] = zz - t3;
    pi[3] = xy;
    pi[4] = xz;
    pi[5] = yz;
}

// recursive regularised off-equilibrium from second-order coefficients
// (Node.cpp:1136-1427, 1430-1747), lattice units
template <int ORDER>
__device__ __forceinline__ void offeq_dev(const float a2[6], float ux, float uy, float uz, float out[19], float k6) {
    const float XX = a2[0], YY = a2[1], ZZ = a2[2], XY = a2[3], XZ = a2[4], YZ = a2[5];
    float b3[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if constexpr (ORDER >= 3) {
        const float tXXY = 2.f * ux * XY + uy * XX, tXYY = 2.f * uy * XY + ux * YY, tXXZ = 2.f * ux * XZ + uz * XX;
        const float tXZZ = 2.f * uz * XZ + ux * ZZ, tYYZ = 2.f * uy * YZ + uz * YY, tYZZ = 2.f * uz * YZ + uy * ZZ;
        b3[0] = tXXY + tYZZ; b3[1] = tXXZ + tYYZ; b3[2] = tXYY + tXZZ;
        b3[3] = tYZZ - tXXY; b3[4] = tYYZ - tXXZ; b3[5] = tXZZ - tXYY;
    }
    float A = 0.f, B = 0.f, C = 0.f;
    if constexpr (ORDER >= 4) {
        const float fXXYY = uy * uy * XX + ux * ux * YY + 4.f * ux * uy * XY;
        const float fXXZZ = uz * uz * XX + ux * ux * ZZ + 4.f * ux * uz * XZ;
        const float fYYZZ = uy * uy * ZZ + uz * uz * YY + 4.f * uy * uz * YZ;
        A = fXXYY + k6 * ZZ;
        B = fXXZZ + k6 * YY + (2.f / 7.f) * A;
        C = fYYZZ + k6 * XX + (2.f / 7.f) * A + (2.f / 5.f) * B;
    }
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr float hx = EX[q], hy = EY[q], hz = EZ[q];
        constexpr float hxx = hx * hx - 1.f / 3.f, hyy = hy * hy - 1.f / 3.f, hzz = hz * hz - 1.f / 3.f;
        float s2 = hxx * XX + hyy * YY + hzz * ZZ + 2.f * (hx * hy * XY + hx * hz * XZ + hy * hz * YZ);
        float x = 4.5f * s2;
        if constexpr (ORDER >= 3) {
            constexpr float H[6] = {hxx * hy + hzz * hy, hxx * hz + hyy * hz, hyy * hx + hzz * hx,
                                    hzz * hy - hxx * hy, hyy * hz - hxx * hz, hzz * hx - hyy * hx};
            x += 4.5f * (3.f * (H[0] * b3[0] + H[1] * b3[1] + H[2] * b3[2]) + H[3] * b3[3] + H[4] * b3[4] +
                         H[5] * b3[5]);
        }
        if constexpr (ORDER >= 4) {
            const float a4 = hxx * hyy + k6 * hzz;
            const float b4 = hxx * hzz + k6 * hyy + (2.f / 7.f) * a4;
            const float c4 = hyy * hzz + k6 * hxx + (2.f / 7.f) * a4 + (2.f / 5.f) * b4;
            x += 20.25f * ((8.f / 7.f) * a4 * A + (56.f / 45.f) * b4 * B + (40.f / 27.f) * c4 * C);
        }
        out[q] = wf(q) * x;
    });
}

// Coreixas raw-moment equilibrium minus the weights (Node.cpp:2650-2690), lattice units
__device__ __forceinline__ void rm_feq_dev(float dr, float ux, float uy, float uz, float out[19]) {
    const float rho = 1.f + dr;
    const float U[3] = {ux, uy, uz};
    const float S[3] = {ux * ux, uy * uy, uz * uz};
    static_for<0, 19>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int n = (EX[q] != 0) + (EY[q] != 0) + (EZ[q] != 0);
        float p;
        if constexpr (n == 0) {
            p = -(S[0] + S[1] + S[2]) + 3.f * (S[0] * S[1] + S[0] * S[2] + S[1] * S[2]);
        } else if constexpr (n == 1) {
            constexpr int a = EX[q] ? 0 : (EY[q] ? 1 : 2);
            constexpr int b = (a == 0) ? 1 : 0, cdim = (a == 2) ? 1 : 2;
            const float ua = (float)ecomp(q, a) * U[a];
            p = 3.f * ua + 3.f * (S[a] - S[b] - S[cdim]) - 9.f * ua * (S[b] + S[cdim]) - 9.f * S[a] * (S[b] + S[cdim]);
        } else {
            constexpr int a = EX[q] ? 0 : 1;
            constexpr int b = EZ[q] ? 2 : 1;
            const float ua = (float)ecomp(q, a) * U[a], ub = (float)ecomp(q, b) * U[b];
            p = 3.f * (ua + ub) + 3.f * (S[a] + S[b]) + 9.f * ua * ub + 9.f * (S[a] * ub + ua * S[b]) + 9.f * S[a] * S[b];
        }
        out[q] = wf(q) * (dr + rho * p);
    });
}

// Gram-Schmidt row coefficients (Node.cpp:2934-3004), lattice units
__host__ __device__ constexpr int gsf_coef(int r, int q) {
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
// sum_q w_q M_r[q]^2 of each row as a float (exact small rationals)
__host__ __device__ constexpr float gsf_norm(int r) {
    float s = 0.f;
    for (int q = 0; q < 19; ++q) s += wf(q) * (float)(gsf_coef(r, q) * gsf_coef(r, q));
    return s;
}

}  // namespace acp
