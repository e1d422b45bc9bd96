// Host side of the engine: case lifetime, constants, the LODI schedule and the
// C ABI of include/acp.h.
#include "engine.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

namespace acp {

static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}
void clear_error() { g_err.clear(); }
int hip_fail(hipError_t e, const char* where) {
    g_err = std::string(where) + ": " + hipGetErrorString(e);
    return ACP_E_HIP;
}

#define HIPCHK(call)                                        \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return acp::hip_fail(e_, #call); \
    } while (0)

static long gcdl(long a, long b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) {
        long t = a % b;
        a = b;
        b = t;
    }
    return a;
}

void k64_init(K64& k, const acp_case_desc& d) {
    std::memset(&k, 0, sizeof(k));
    const double cs = d.cs;
    k.cs = cs;
    k.c = cs * std::sqrt(3.);                                            // Control.cpp:52
    k.dt = d.spacing / k.c;                                              // Control.cpp:53
    k.omega = (std::pow(cs, 2.) * (k.dt / std::pow(2., 0))) /           // Lattice.cpp:1404
              (d.nu + 0.5 * std::pow(cs, 2.) * (k.dt / std::pow(2., 0)));
    k.cs2 = std::pow(cs, 2.);
    k.rcs2 = 1. / k.cs2;
    k.c2 = 1. / 2. * (1. / std::pow(cs, 4.));
    k.c3 = 1. / 6. * (1. / std::pow(cs, 6.));
    k.c4 = 1. / 4. * (1. / std::pow(cs, 8.));
    k.hA = k.c * k.c - k.cs2;
    k.hB = 0. * 0. - k.cs2;
    k.pA = std::pow(k.c, 2.) - k.cs2;
    k.pB = std::pow(0., 2.) - k.cs2;
    k.cc = k.c * k.c;
    const double o6 = 1. / 6., t7 = 2. / 7., t5 = 2. / 5.;
    for (int q = 0; q < 19; ++q) {
        const double x = EX[q] == 0 ? 0. : (EX[q] > 0 ? k.c : -k.c);
        const double y = EY[q] == 0 ? 0. : (EY[q] > 0 ? k.c : -k.c);
        const double z = EZ[q] == 0 ? 0. : (EZ[q] > 0 ? k.c : -k.c);
        const double xx = x * x - k.cs2, yy = y * y - k.cs2, zz = z * z - k.cs2;   // Node.cpp:1016-1018
        const double XXY = xx * y, XYY = yy * x, XXZ = xx * z, XZZ = zz * x, YYZ = yy * z, YZZ = zz * y;
        k.h3[0][q] = XXY + YZZ;
        k.h3[1][q] = XXZ + YYZ;
        k.h3[2][q] = XYY + XZZ;
        k.h3[3][q] = YZZ - XXY;
        k.h3[4][q] = YYZ - XXZ;
        k.h3[5][q] = XZZ - XYY;
        const double a = xx * yy + o6 * zz;                                          // Node.cpp:1040-1052
        const double b = xx * zz + o6 * yy + t7 * a;
        const double c = yy * zz + o6 * xx + t7 * a + t5 * b;
        k.h4[0][q] = a;
        k.h4[1][q] = b;
        k.h4[2][q] = c;
    }
    k.half_dt = 0.5 * k.dt;
    k.rec_cs4_strain = 1. / (k.cs2 * k.cs2);
    k.inv_pre = -((k.dt * k.cs2) / (k.omega * d.nu));
    k.nu = d.nu;
    k.sigma = d.sigma;
    k.dist = d.spacing / std::pow(2., 0);
    k.lodi_cs2 = std::pow(cs, 2.);
    k.dom_len[0] = d.nx * d.spacing;
    k.dom_len[1] = d.ny * d.spacing;
    k.dom_len[2] = d.nz * d.spacing;
    k.ma = d.mach;
    k.ma2 = d.mach * d.mach;
    k.k1 = d.k1;
    k.rho0 = d.rho0;
    k.ls = d.spacing * std::pow(2., -0);
    k.dt_l = k.dt / std::pow(2., 0);
    k.rm_cp[0] = 1.0;
    k.rm_cp[1] = k.c;
    k.rm_cp[2] = std::pow(k.c, 2.0);
    k.rm_cp[3] = std::pow(k.c, 3.0);
    k.rm_cp[4] = std::pow(k.c, 4.0);
    const double ix = 1. / k.c;
    k.rm_ix[0] = 1.0;
    k.rm_ix[1] = ix;
    k.rm_ix[2] = std::pow(ix, 2.);
    k.rm_ix[3] = std::pow(ix, 3.);
    k.rm_ix[4] = std::pow(ix, 4.);
    for (int p = 0; p < 5; ++p) k.gs_cp[p] = p == 0 ? 1.0 : (p == 1 ? k.c : std::pow(k.c, (double)p));
    // Inverse Gram-Schmidt transform: the basis is orthogonal under the lattice
    // weights, so M^-1[q][r] = w_q M_r[q] / sum_q' w_q' M_r[q']^2 (each entry is
    // +-1/den), times the row's power of 1/c.
    const long wd[19] = {18, 18, 18, 18, 18, 18, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 36, 3};
    for (int r = 0; r < 19; ++r) {
        long nn = 0, nd = 1;
        for (int q = 0; q < 19; ++q) {
            const long m = gs_coef(r, q);
            nn = nn * wd[q] + m * m * nd;
            nd = nd * wd[q];
            const long g = gcdl(nn, nd);
            if (g) {
                nn /= g;
                nd /= g;
            }
        }
        for (int q = 0; q < 19; ++q) {
            const long m = gs_coef(r, q);
            if (m == 0) {
                k.gs_pre[r][q] = 0.;
                continue;
            }
            long en = m * nd, ed = wd[q] * nn;
            const long g = gcdl(en, ed);
            en /= g;
            ed /= g;
            if (ed < 0) {
                ed = -ed;
                en = -en;
            }
            const double v = (double)en / (double)ed;
            k.gs_pre[r][q] = r == 0 ? v : v * k.rm_ix[GS_PW[r]];
        }
    }
}

void kshift_init(KShift& k, const acp_case_desc& d) {
    K64 k64;
    k64_init(k64, d);
    k.omega = (float)k64.omega;
    k.s = d.precision == ACP_PREC_F16_SHIFTED ? (float)d.f16_scale : 1.0f;
    k.inv_s = 1.0f / k.s;
    k.sigma = (float)d.sigma;
    k.hrr_pre = (float)(-(2.0 / 3.0) / k64.omega);
    k.k1 = (float)d.k1;
    k.ma = (float)d.mach;
    k.ma2 = (float)(d.mach * d.mach);
    k.inv_len[0] = 1.0f / (float)d.nx;
    k.inv_len[1] = 1.0f / (float)d.ny;
    k.inv_len[2] = 1.0f / (float)d.nz;
    k.rho0 = d.rho0;
    k.c_si = k64.c;
}

LodiSchedule build_lodi_schedule(int nx, int ny, int nz) {
    const int im = nx - 1, jm = ny - 1, km = nz - 1;
    auto lin = [&](int i, int j, int k) { return (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k); };
    std::vector<LodiNode> seq;
    seq.reserve(2 * ((size_t)ny * nz + (size_t)nx * ny + (size_t)nx * nz));
    // Normal, second-normal and tangent neighbours, inward normal direction and
    // corner flag as Lattice::setConnectivity assigns them (Lattice.cpp:500-1392).
    auto add = [&](int i, int j, int k) {
        LodiNode n;
        n.node = lin(i, j, k);
        for (int t = 0; t < 4; ++t) n.t[t] = -1;
        const int bi = i == 0 ? -1 : (i == im ? 1 : 0);
        const int bj = j == 0 ? -1 : (j == jm ? 1 : 0);
        const int bk = k == 0 ? -1 : (k == km ? 1 : 0);
        const int nb = (bi != 0) + (bj != 0) + (bk != 0);
        if (nb == 1) {
            n.corner = 0;
            if (bk) {  // z faces
                n.dir = bk < 0 ? 4 : 5;
                n.n1 = lin(i, j, k - bk);
                n.n2 = lin(i, j, k - 2 * bk);
                n.t[0] = lin(i + 1, j, k); n.t[1] = lin(i - 1, j, k); n.t[2] = lin(i, j + 1, k); n.t[3] = lin(i, j - 1, k);
            } else if (bj) {  // y faces
                n.dir = bj < 0 ? 1 : 3;
                n.n1 = lin(i, j - bj, k);
                n.n2 = lin(i, j - 2 * bj, k);
                n.t[0] = lin(i + 1, j, k); n.t[1] = lin(i - 1, j, k); n.t[2] = lin(i, j, k + 1); n.t[3] = lin(i, j, k - 1);
            } else {  // x faces
                n.dir = bi < 0 ? 0 : 2;
                n.n1 = lin(i - bi, j, k);
                n.n2 = lin(i - 2 * bi, j, k);
                n.t[0] = lin(i, j + 1, k); n.t[1] = lin(i, j - 1, k); n.t[2] = lin(i, j, k + 1); n.t[3] = lin(i, j, k - 1);
            }
        } else {
            n.corner = 1;
            if (bi == 0) {  // edges along x: normal along the inward yz diagonal
                n.dir = bj < 0 ? 1 : 3;
                n.n1 = lin(i, j - bj, k - bk);
                n.n2 = lin(i, j - 2 * bj, k - 2 * bk);
            } else if (bi < 0) {  // inlet edges and corners
                n.dir = 0;
                if (nb == 3) {
                    n.n1 = lin(1, j, k);
                    n.n2 = lin(2, j, k);
                } else {
                    n.n1 = lin(1, j - bj, k - bk);
                    n.n2 = lin(2, j - 2 * bj, k - 2 * bk);
                }
            } else {  // outlet edges and corners
                n.dir = bk ? (bk < 0 ? 4 : 5) : (bj < 0 ? 1 : 3);
                n.n1 = lin(im - 1, j - bj, k - bk);
                n.n2 = lin(im - 2, j - 2 * bj, k - 2 * bk);
            }
        }
        seq.push_back(n);
    };
    for (int j = 0; j <= jm; ++j)
        for (int k = 0; k <= km; ++k) {
            add(0, j, k);
            add(im, j, k);
        }
    for (int i = 1; i <= im - 1; ++i)
        for (int j = 0; j <= jm; ++j) {
            add(i, j, 0);
            add(i, j, km);
        }
    for (int i = 1; i <= im - 1; ++i)
        for (int k = 1; k <= km - 1; ++k) {
            add(i, 0, k);
            add(i, jm, k);
        }
    // phases: a node that reads the B entries of a boundary node handled before it
    // goes to a later phase; one that reads a node handled after it forces that
    // node into a later phase.
    const size_t n = seq.size();
    std::unordered_map<int64_t, int32_t> pos;
    pos.reserve(n * 2);
    for (size_t s = 0; s < n; ++s) pos[seq[s].node] = (int32_t)s;
    std::vector<int32_t> phase(n, 1), lb(n, 1);
    int32_t maxph = 1;
    for (size_t s = 0; s < n; ++s) {
        int32_t ph = lb[s];
        const int64_t rd[2] = {seq[s].n1, seq[s].n2};
        for (int r = 0; r < 2; ++r) {
            auto it = pos.find(rd[r]);
            if (it == pos.end() || it->second == (int32_t)s) continue;
            if ((size_t)it->second < s) ph = std::max(ph, phase[it->second] + 1);
        }
        phase[s] = ph;
        maxph = std::max(maxph, ph);
        for (int r = 0; r < 2; ++r) {
            auto it = pos.find(rd[r]);
            if (it != pos.end() && (size_t)it->second > s) lb[it->second] = std::max(lb[it->second], ph + 1);
        }
    }
    LodiSchedule out;
    out.nodes.reserve(n);
    out.phase_start.push_back(0);
    for (int32_t ph = 1; ph <= maxph; ++ph) {
        for (size_t s = 0; s < n; ++s)
            if (phase[s] == ph) out.nodes.push_back(seq[s]);
        out.phase_start.push_back((int64_t)out.nodes.size());
    }
    return out;
}

}  // namespace acp

using acp::fail;

extern "C" const char* acp_last_error(void) { return acp::g_err.c_str(); }

namespace {

bool shifted(const acp_case* c) { return c->precision != ACP_PREC_F64; }
size_t elem_bytes(int prec) { return prec == ACP_PREC_F64 ? 8 : (prec == ACP_PREC_F32_SHIFTED ? 4 : 2); }

template <typename T>
int dalloc(acp_case* c, T** p, size_t count) {
    const size_t bytes = count * sizeof(T);
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(ACP_E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                                   hipGetErrorString(e));
    }
    c->device_bytes += (int64_t)bytes;
    return ACP_OK;
}

void release(acp_case* c) {
    void* ptrs[] = {c->dk, c->A, c->B, c->psi, c->M, c->dks, c->SA, c->SB, c->SM, c->lodi, c->probe_nodes,
                    c->probe_buf, c->partial};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
}

// Upload host float64 populations [19][N] (SI units) into both arrays.
int upload_f64(acp_case* c, const double* f) {
    const int64_t N = c->g.N;
    if (!shifted(c)) {
        for (int q = 0; q < 19; ++q) {
            HIPCHK(hipMemcpyAsync(c->A + (size_t)q * N, f + (size_t)q * N, sizeof(double) * N, hipMemcpyHostToDevice,
                                  c->stream));
            HIPCHK(hipMemcpyAsync(c->B + (size_t)q * N, c->A + (size_t)q * N, sizeof(double) * N,
                                  hipMemcpyDeviceToDevice, c->stream));
        }
        if (c->psi) HIPCHK(hipMemsetAsync(c->psi, 0, sizeof(double) * 19 * N, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return ACP_OK;
    }
    // staged conversion, one q-plane segment at a time
    const int64_t chunk = std::min<int64_t>(N, (int64_t)1 << 24);
    double* stage = nullptr;
    HIPCHK(hipMalloc((void**)&stage, sizeof(double) * chunk));
    const size_t eb = elem_bytes(c->precision);
    int rc = ACP_OK;
    for (int q = 0; q < 19 && rc == ACP_OK; ++q)
        for (int64_t l0 = 0; l0 < N && rc == ACP_OK; l0 += chunk) {
            const int64_t n = std::min(chunk, N - l0);
            hipError_t e = hipMemcpyAsync(stage, f + (size_t)q * N + l0, sizeof(double) * n, hipMemcpyHostToDevice,
                                          c->stream);
            if (e == hipSuccess)
                e = acp::launch_to_store(c->precision, c->dks, stage, (char*)c->SA + ((size_t)q * N + l0) * eb,
                                         (char*)c->SB + ((size_t)q * N + l0) * eb, n, q, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = acp::hip_fail(e, "population upload");
        }
    (void)hipFree(stage);
    return rc;
}

int check_case(const acp_case* c) {
    if (!c) return fail(ACP_E_ARG, "null case handle");
    return ACP_OK;
}

// host-side initial condition: rho(i,j) with u = 0, populations feq_N, uploaded in chunks
template <typename RhoFn>
int init_rho_field(acp_case* c, RhoFn rho_of) {
    const acp::Grid& g = c->g;
    const int64_t N = g.N;
    std::vector<double> f((size_t)19 * N);
    acp::K64 k = c->hk;
    for (int kz = 0; kz < g.nz; ++kz)
        for (int j = 0; j < g.ny; ++j)
            for (int i = 0; i < g.nx; ++i) {
                const int64_t l = (int64_t)i + (int64_t)g.nx * ((int64_t)j + (int64_t)g.ny * kz);
                const double rho = rho_of(i, j, kz);
                double eq[19];
                if (c->d.order == 2) acp::feq64<2>(k, rho, 0., 0., 0., eq);
                else if (c->d.order == 3) acp::feq64<3>(k, rho, 0., 0., 0., eq);
                else acp::feq64<4>(k, rho, 0., 0., 0., eq);
                for (int q = 0; q < 19; ++q) f[(size_t)q * N + l] = eq[q];
            }
    int rc = upload_f64(c, f.data());
    if (rc == ACP_OK) {
        c->initialized = true;
        c->steps = 0;
        c->probe_count = 0;
    }
    return rc;
}

}  // namespace

extern "C" int acp_case_create(const acp_case_desc* desc, void* stream, acp_case** out) {
    if (!desc || !out) return fail(ACP_E_ARG, "acp_case_create: null argument");
    *out = nullptr;
    const acp_case_desc& d = *desc;
    if (d.max_level != 0)
        return fail(ACP_E_UNSUPPORTED, "MaxLevel > 0: two-level refinement is not part of this engine");
    if (d.nx < 3 || d.ny < 3 || d.nz < 3) return fail(ACP_E_ARG, "grid needs at least 3 nodes per direction");
    if (d.order < 2 || d.order > 4) return fail(ACP_E_CONFIG, "Wrong order of equilibrium!");
    if (d.op < ACP_OP_SRT || d.op > ACP_OP_HRR) return fail(ACP_E_ARG, "unknown collision operator");
    if (d.precision < ACP_PREC_F64 || d.precision > ACP_PREC_F16_SHIFTED) return fail(ACP_E_ARG, "unknown precision");
    if (d.boundary != ACP_BC_LODI && d.boundary != ACP_BC_PERIODIC) return fail(ACP_E_ARG, "unknown boundary");
    if (d.precision != ACP_PREC_F64 && d.cubic)
        return fail(ACP_E_UNSUPPORTED, "the cubic-Mach correction is available in the float64 tier only");
    if (!(d.cs > 0) || !(d.spacing > 0) || !(d.nu > 0) || !(d.rho0 > 0))
        return fail(ACP_E_ARG, "SpeedOfSound, Spacing, KinematicViscosity and Density must be positive");
    if (d.precision == ACP_PREC_F16_SHIFTED && !(d.f16_scale > 0)) return fail(ACP_E_ARG, "FP16 scale must be positive");

    acp_case* c = new acp_case();
    c->d = d;
    c->g = acp::Grid{d.nx, d.ny, d.nz, (int64_t)d.nx * d.ny * d.nz};
    c->precision = d.precision;
    c->periodic = d.boundary == ACP_BC_PERIODIC;
    switch (d.op) {
        case ACP_OP_SRT: c->bulk_op = acp::OP_SRT; break;
        case ACP_OP_RR: c->bulk_op = acp::OP_RR; break;
        case ACP_OP_HRR: c->bulk_op = acp::OP_HRR; break;
        case ACP_OP_MRT_GS: c->bulk_op = d.mrt_mode == ACP_MRT_RELAXED ? acp::OP_MRT_GS : acp::OP_STREAM; break;
        default: c->bulk_op = d.mrt_mode == ACP_MRT_RELAXED ? acp::OP_MRT_RM : acp::OP_STREAM; break;
    }
    c->need_macro = d.op == ACP_OP_HRR || (d.op == ACP_OP_RR && d.cubic);
    c->need_psi = d.cubic && (d.op == ACP_OP_RR || d.op == ACP_OP_HRR);
    acp::k64_init(c->hk, d);
    acp::kshift_init(c->hks, d);
    int rc = ACP_OK;
    if (stream) {
        c->stream = (hipStream_t)stream;
    } else {
        hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete c;
            return acp::hip_fail(e, "hipStreamCreate");
        }
        c->own_stream = true;
    }
    const int64_t N = c->g.N;
    auto bail = [&](int code) {
        release(c);
        delete c;
        return code;
    };
    if (!shifted(c)) {
        if ((rc = dalloc(c, &c->dk, 1)) || (rc = dalloc(c, &c->A, (size_t)19 * N)) ||
            (rc = dalloc(c, &c->B, (size_t)19 * N)))
            return bail(rc);
        if (c->need_psi && (rc = dalloc(c, &c->psi, (size_t)19 * N))) return bail(rc);
        if (c->need_macro && (rc = dalloc(c, &c->M, (size_t)4 * N))) return bail(rc);
        hipError_t e = hipMemcpy(c->dk, &c->hk, sizeof(acp::K64), hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(acp::hip_fail(e, "constants upload"));
    } else {
        const size_t eb = elem_bytes(c->precision);
        char *a = nullptr, *b = nullptr;
        if ((rc = dalloc(c, &c->dks, 1)) || (rc = dalloc(c, &a, (size_t)19 * N * eb)) ||
            (rc = dalloc(c, &b, (size_t)19 * N * eb)))
            return bail(rc);
        c->SA = a;
        c->SB = b;
        if (d.op == ACP_OP_HRR && (rc = dalloc(c, &c->SM, (size_t)4 * N))) return bail(rc);
        hipError_t e = hipMemcpy(c->dks, &c->hks, sizeof(acp::KShift), hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(acp::hip_fail(e, "constants upload"));
        // the float64 constants are used by the conversions of the reference initial condition
        if ((rc = dalloc(c, &c->dk, 1))) return bail(rc);
        e = hipMemcpy(c->dk, &c->hk, sizeof(acp::K64), hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(acp::hip_fail(e, "constants upload"));
    }
    if (!c->periodic) {
        acp::LodiSchedule sch = acp::build_lodi_schedule(d.nx, d.ny, d.nz);
        c->phase_start = sch.phase_start;
        if ((rc = dalloc(c, &c->lodi, sch.nodes.size()))) return bail(rc);
        hipError_t e = hipMemcpy(c->lodi, sch.nodes.data(), sizeof(acp::LodiNode) * sch.nodes.size(),
                                 hipMemcpyHostToDevice);
        if (e != hipSuccess) return bail(acp::hip_fail(e, "LODI schedule upload"));
    }
    c->partial_blocks = (int)std::min<int64_t>(1024, (N + 255) / 256);
    if ((rc = dalloc(c, &c->partial, (size_t)c->partial_blocks))) return bail(rc);
    *out = c;
    acp::clear_error();
    return ACP_OK;
}

extern "C" void acp_case_destroy(acp_case* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    delete c;
}

extern "C" int acp_case_info_get(const acp_case* c, acp_case_info* out) {
    if (check_case(c) || !out) return fail(ACP_E_ARG, "acp_case_info_get: null argument");
    std::memset(out, 0, sizeof(*out));
    out->nodes = c->g.N;
    out->device_bytes = c->device_bytes;
    const int64_t s = (int64_t)elem_bytes(c->precision);
    int64_t b = 2 * 19 * s;
    if (c->need_macro) b += 19 * s + 2 * 4 * (c->precision == ACP_PREC_F64 ? 8 : 4);
    if (c->need_psi) b += 2 * 19 * 8;
    out->bytes_per_node_step = b;
    out->lodi_nodes = c->phase_start.empty() ? 0 : c->phase_start.back();
    out->lodi_launches = c->phase_start.empty() ? 0 : (int32_t)c->phase_start.size() - 1;
    out->steps_done = c->steps;
    out->omega = c->hk.omega;
    out->dt = c->hk.dt;
    return ACP_OK;
}

extern "C" int acp_case_init_reference(acp_case* c) {
    if (int rc = check_case(c)) return rc;
    // Lattice.cpp:1415-1433 with node coordinates of createLattice (Lattice.cpp:33-35)
    const double Rc = 0.06, A = 0.01, xc = -8. * Rc;
    const double dens = c->d.rho0, sp = c->d.spacing, lx = c->d.domain[0], ly = c->d.domain[1];
    return init_rho_field(c, [&](int i, int j, int) {
        const double x = i * (sp / std::pow(2., 0)), y = j * (sp / std::pow(2., 0));
        const double xn = x - lx / 2., yn = y - ly / 2.;
        const double xr = xn - xc;
        return dens * (1 + A * std::exp(-(xr * xr + yn * yn) / (2 * Rc * Rc)));
    });
}

extern "C" int acp_case_init_pulse(acp_case* c, double amp, double x0, double y0, double b) {
    if (int rc = check_case(c)) return rc;
    if (!(b > 0)) return fail(ACP_E_ARG, "pulse half-width must be positive");
    const double ln2 = 0.69314718055994530942, rho0 = c->d.rho0;
    return init_rho_field(c, [&](int i, int j, int) {
        const double dx = i - x0, dy = j - y0;
        return rho0 * (1. + amp * std::exp(-ln2 * (dx * dx + dy * dy) / (b * b)));
    });
}

extern "C" int acp_case_init_random(acp_case* c, uint64_t seed, double rho_amp, double u_amp) {
    if (int rc = check_case(c)) return rc;
    hipError_t e;
    if (!shifted(c)) {
        e = acp::launch_init_random64(c->d.order, c->dk, c->A, c->B, c->g, seed, c->d.rho0, rho_amp, u_amp,
                                      c->stream);
        if (e == hipSuccess && c->psi) e = hipMemsetAsync(c->psi, 0, sizeof(double) * 19 * c->g.N, c->stream);
    } else {
        e = acp::launch_init_random_shift(c->precision, c->d.order, c->dks, c->SA, c->SB, c->g, seed, rho_amp, u_amp,
                                          c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return acp::hip_fail(e, "acp_case_init_random");
    c->initialized = true;
    c->steps = 0;
    c->probe_count = 0;
    return ACP_OK;
}

extern "C" int acp_case_set_populations(acp_case* c, const double* f) {
    if (int rc = check_case(c)) return rc;
    if (!f) return fail(ACP_E_ARG, "null population array");
    int rc = upload_f64(c, f);
    if (rc == ACP_OK) {
        c->initialized = true;
        c->steps = 0;
        c->probe_count = 0;
    }
    return rc;
}

extern "C" int acp_case_step(acp_case* c, int32_t n) {
    if (int rc = check_case(c)) return rc;
    if (!c->initialized) return fail(ACP_E_STATE, "acp_case_step before initialisation");
    if (n < 0) return fail(ACP_E_ARG, "negative step count");
    if (c->nprobe > 0 && c->probe_count + n > c->probe_cap)
        return fail(ACP_E_STATE, "probe buffer full: read the probes first");
    const int64_t N = c->g.N;
    const int nph = c->phase_start.empty() ? 0 : (int)c->phase_start.size() - 1;
    for (int32_t s = 0; s < n; ++s) {
        hipError_t e = hipSuccess;
        if (!shifted(c)) {
            if (c->need_macro) e = acp::launch_macro64(c->dk, c->A, c->M, c->g, c->periodic, c->stream);
            if (e == hipSuccess)
                e = acp::launch_bulk64(c->bulk_op, c->d.order, c->d.cubic != 0, c->periodic, c->dk, c->A, c->B,
                                       c->psi, c->M, c->g, c->stream);
            for (int p = 0; p < nph && e == hipSuccess; ++p)
                e = acp::launch_lodi64(c->d.order, c->d.cubic != 0, c->d.transverse != 0, c->dk, c->A, c->B, c->psi,
                                       c->lodi + c->phase_start[p], c->phase_start[p + 1] - c->phase_start[p], N,
                                       c->stream);
            if (e == hipSuccess && c->nprobe > 0)
                e = acp::launch_probe64(c->B, N, c->probe_nodes, c->nprobe,
                                        c->probe_buf + (size_t)c->probe_count * c->nprobe, c->stream);
            std::swap(c->A, c->B);
        } else {
            if (c->SM) e = acp::launch_macro_shift(c->precision, c->dks, c->SA, c->SM, c->g, c->stream);
            if (e == hipSuccess)
                e = acp::launch_bulk_shift(c->precision, c->bulk_op, c->d.order, c->periodic, c->dks, c->SA, c->SB,
                                           c->SM, c->g, c->stream);
            for (int p = 0; p < nph && e == hipSuccess; ++p)
                e = acp::launch_lodi_shift(c->precision, c->d.order, c->d.transverse != 0, c->dks, c->SA, c->SB,
                                           c->lodi + c->phase_start[p], c->phase_start[p + 1] - c->phase_start[p], N,
                                           c->stream);
            if (e == hipSuccess && c->nprobe > 0)
                e = acp::launch_probe_shift(c->precision, c->dks, c->SB, N, c->probe_nodes, c->nprobe,
                                            c->probe_buf + (size_t)c->probe_count * c->nprobe, c->stream);
            std::swap(c->SA, c->SB);
        }
        if (e != hipSuccess) return acp::hip_fail(e, "acp_case_step launch");
        if (c->nprobe > 0) ++c->probe_count;
        ++c->steps;
    }
    return ACP_OK;
}

extern "C" int acp_case_sync(acp_case* c) {
    if (int rc = check_case(c)) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return ACP_OK;
}

extern "C" int acp_case_get_populations(acp_case* c, double* f) {
    if (int rc = check_case(c)) return rc;
    if (!f) return fail(ACP_E_ARG, "null population array");
    const int64_t N = c->g.N;
    if (!shifted(c)) {
        HIPCHK(hipMemcpyAsync(f, c->A, sizeof(double) * 19 * N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return ACP_OK;
    }
    const int64_t chunk = std::min<int64_t>(N, (int64_t)1 << 24);
    double* stage = nullptr;
    HIPCHK(hipMalloc((void**)&stage, sizeof(double) * chunk));
    const size_t eb = elem_bytes(c->precision);
    int rc = ACP_OK;
    for (int q = 0; q < 19 && rc == ACP_OK; ++q)
        for (int64_t l0 = 0; l0 < N && rc == ACP_OK; l0 += chunk) {
            const int64_t n = std::min(chunk, N - l0);
            hipError_t e = acp::launch_from_store(c->precision, c->dks, (const char*)c->SA + ((size_t)q * N + l0) * eb,
                                                  stage, n, q, c->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(f + (size_t)q * N + l0, stage, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = acp::hip_fail(e, "population download");
        }
    (void)hipFree(stage);
    return rc;
}

extern "C" int acp_case_get_macros(acp_case* c, double* rho, double* u) {
    if (int rc = check_case(c)) return rc;
    if (!rho || !u) return fail(ACP_E_ARG, "null output array");
    const int64_t N = c->g.N;
    double* buf = nullptr;
    HIPCHK(hipMalloc((void**)&buf, sizeof(double) * 4 * N));
    hipError_t e = shifted(c) ? acp::launch_macros_out_shift(c->precision, c->dks, c->SA, c->g, buf, buf + N, c->stream)
                              : acp::launch_macros_out64(c->dk, c->A, c->g, buf, buf + N, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rho, buf, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(u, buf + N, sizeof(double) * 3 * N, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(buf);
    if (e != hipSuccess) return acp::hip_fail(e, "acp_case_get_macros");
    return ACP_OK;
}

extern "C" int acp_case_mass(acp_case* c, double* mass) {
    if (int rc = check_case(c)) return rc;
    if (!mass) return fail(ACP_E_ARG, "null output");
    hipError_t e = shifted(c) ? acp::launch_mass_shift(c->precision, c->dks, c->SA, c->g.N, c->partial,
                                                       c->partial_blocks, c->stream)
                              : acp::launch_mass64(c->A, c->g.N, c->partial, c->partial_blocks, c->stream);
    if (e != hipSuccess) return acp::hip_fail(e, "acp_case_mass");
    std::vector<double> h((size_t)c->partial_blocks);
    HIPCHK(hipMemcpyAsync(h.data(), c->partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double s = 0.;
    for (double v : h) s += v;
    // the shifted tiers reduce the deviations only; add the rest-state mass rho0 * N
    if (shifted(c)) s += c->d.rho0 * (double)c->g.N;
    *mass = s;
    return ACP_OK;
}

extern "C" int acp_case_set_probes(acp_case* c, const int64_t* nodes, int32_t nprobe, int32_t capacity) {
    if (int rc = check_case(c)) return rc;
    if (nprobe < 0 || capacity < 0 || (nprobe > 0 && !nodes)) return fail(ACP_E_ARG, "bad probe arguments");
    for (int32_t p = 0; p < nprobe; ++p)
        if (nodes[p] < 0 || nodes[p] >= c->g.N) return fail(ACP_E_ARG, "probe node out of range");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->probe_nodes) (void)hipFree(c->probe_nodes);
    if (c->probe_buf) (void)hipFree(c->probe_buf);
    c->probe_nodes = nullptr;
    c->probe_buf = nullptr;
    c->nprobe = 0;
    c->probe_cap = 0;
    c->probe_count = 0;
    if (nprobe == 0) return ACP_OK;
    HIPCHK(hipMalloc((void**)&c->probe_nodes, sizeof(int64_t) * nprobe));
    HIPCHK(hipMalloc((void**)&c->probe_buf, sizeof(double) * (size_t)nprobe * std::max(capacity, 1)));
    HIPCHK(hipMemcpy(c->probe_nodes, nodes, sizeof(int64_t) * nprobe, hipMemcpyHostToDevice));
    c->nprobe = nprobe;
    c->probe_cap = capacity;
    return ACP_OK;
}

extern "C" int acp_case_read_probes(acp_case* c, double* out, int32_t* nsteps) {
    if (int rc = check_case(c)) return rc;
    if (!nsteps) return fail(ACP_E_ARG, "null step counter");
    *nsteps = c->probe_count;
    if (c->nprobe > 0 && c->probe_count > 0) {
        if (!out) return fail(ACP_E_ARG, "null probe output");
        HIPCHK(hipMemcpyAsync(out, c->probe_buf, sizeof(double) * (size_t)c->probe_count * c->nprobe,
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->probe_count = 0;
    return ACP_OK;
}

extern "C" int64_t acp_lodi_schedule(int32_t nx, int32_t ny, int32_t nz, int64_t* nodes, int32_t* phases,
                                     int64_t cap) {
    if (nx < 3 || ny < 3 || nz < 3) return 0;
    acp::LodiSchedule s = acp::build_lodi_schedule(nx, ny, nz);
    const int64_t n = (int64_t)s.nodes.size();
    int64_t t = 0;
    for (size_t p = 0; p + 1 < s.phase_start.size(); ++p)
        for (int64_t i = s.phase_start[p]; i < s.phase_start[p + 1]; ++i, ++t)
            if (t < cap) {
                if (nodes) nodes[t] = s.nodes[i].node;
                if (phases) phases[t] = (int32_t)p + 1;
            }
    return n;
}

namespace {
int download_array(acp_case* c, bool current, double* f) {
    const int64_t N = c->g.N;
    if (!shifted(c)) {
        HIPCHK(hipMemcpyAsync(f, current ? c->A : c->B, sizeof(double) * 19 * N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return ACP_OK;
    }
    const int64_t chunk = std::min<int64_t>(N, (int64_t)1 << 24);
    double* stage = nullptr;
    HIPCHK(hipMalloc((void**)&stage, sizeof(double) * chunk));
    const size_t eb = elem_bytes(c->precision);
    const char* src = (const char*)(current ? c->SA : c->SB);
    int rc = ACP_OK;
    for (int q = 0; q < 19 && rc == ACP_OK; ++q)
        for (int64_t l0 = 0; l0 < N && rc == ACP_OK; l0 += chunk) {
            const int64_t n = std::min(chunk, N - l0);
            hipError_t e = acp::launch_from_store(c->precision, c->dks, src + ((size_t)q * N + l0) * eb, stage, n, q,
                                                  c->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(f + (size_t)q * N + l0, stage, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = acp::hip_fail(e, "population download");
        }
    (void)hipFree(stage);
    return rc;
}
}  // namespace

extern "C" int acp_case_write_vtk(acp_case* c, const char* stem) {
    if (int rc = check_case(c)) return rc;
    if (!stem) return fail(ACP_E_ARG, "null file name");
    const int64_t N = c->g.N;
    std::vector<double> cur((size_t)19 * N), prev;
    if (int rc = download_array(c, true, cur.data())) return rc;
    if (c->d.nonreflecting) {
        prev.resize((size_t)19 * N);
        if (int rc = download_array(c, false, prev.data())) return rc;
    }
    std::vector<double> psi;
    if (c->psi) {
        psi.resize((size_t)19 * N);
        HIPCHK(hipMemcpy(psi.data(), c->psi, sizeof(double) * 19 * N, hipMemcpyDeviceToHost));
    }
    return acp::write_vtk_host(c->d, c->hk, cur.data(), prev.empty() ? nullptr : prev.data(),
                               psi.empty() ? nullptr : psi.data(), stem);
}
