// Host side of the engine: case lifetime, constants, the LODI schedule and the
// C ABI of include/acp.h.
#include "engine.hpp"

#include <algorithm>
#include <cstddef>
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
    // Node.cpp:1016-1052: xi xi - cs^2 is hA for a non-zero component and hB
    // for a zero one; the third-order values are sums and differences of
    // (xi xi - cs^2) xi_b, the fourth-order ones depend on which components
    // are non-zero (h3_of, h4_pattern)
    const double PA = k.hA * k.c, PB = k.hB * k.c;
    k.h3m[0] = PB + PB;
    k.h3m[1] = PA + PB;
    k.h3m[2] = PA + PA;
    k.h3m[3] = PA - PB;
    const double o6 = 1. / 6., t7 = 2. / 7., t5 = 2. / 5.;
    const double e7 = 8. / 7., e45 = 56. / 45., e27 = 40. / 27.;
    for (int p = 0; p < 8; ++p) {
        const double xx = (p & 1) ? k.hA : k.hB, yy = (p & 2) ? k.hA : k.hB, zz = (p & 4) ? k.hA : k.hB;
        const double a = xx * yy + o6 * zz;
        const double b = xx * zz + o6 * yy + t7 * a;
        const double c = yy * zz + o6 * xx + t7 * a + t5 * b;
        k.h4e[p][0] = e7 * a;
        k.h4e[p][1] = e45 * b;
        k.h4e[p][2] = e27 * c;
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
    k.k6 = (float)(1. / (6. * k64.c * k64.c));
}

LodiSchedule build_lodi_schedule(int nx, int ny, int nz) {
    const int im = nx - 1, jm = ny - 1, km = nz - 1;
    auto lin = [&](int i, int j, int k) { return (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k); };
    // Position of a boundary node in the pass order below (x faces by (j, k), then
    // z faces by (i, j), then y faces by (i, k)); -1 for an interior node.
    const int64_t z0 = 2 * (int64_t)ny * nz, y0 = z0 + 2 * (int64_t)(nx - 2) * ny;
    auto pos3 = [&](int i, int j, int k) -> int64_t {
        if (i == 0 || i == im) return 2 * ((int64_t)j * nz + k) + (i == im);
        if (k == 0 || k == km) return z0 + 2 * ((int64_t)(i - 1) * ny + j) + (k == km);
        if (j == 0 || j == jm) return y0 + 2 * ((int64_t)(i - 1) * (nz - 2) + (k - 1)) + (j == jm);
        return -1;
    };
    auto neighbour = [&](int64_t& node, int64_t& pos, int i, int j, int k) {
        node = lin(i, j, k);
        pos = pos3(i, j, k);
    };
    std::vector<int64_t> rd1, rd2;  // pass-order positions of the normal neighbours
    const size_t nbnd = 2 * ((size_t)ny * nz + (size_t)nx * ny + (size_t)nx * nz);
    rd1.reserve(nbnd);
    rd2.reserve(nbnd);
    // Normal, second-normal and tangent neighbours, inward normal direction and
    // corner flag as Lattice::setConnectivity assigns them (Lattice.cpp:500-1392).
    auto record = [&](int i, int j, int k, LodiNode& n, int64_t& p1, int64_t& p2) {
        p1 = p2 = -1;
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
                neighbour(n.n1, p1, i, j, k - bk);
                neighbour(n.n2, p2, i, j, k - 2 * bk);
                n.t[0] = lin(i + 1, j, k); n.t[1] = lin(i - 1, j, k); n.t[2] = lin(i, j + 1, k); n.t[3] = lin(i, j - 1, k);
            } else if (bj) {  // y faces
                n.dir = bj < 0 ? 1 : 3;
                neighbour(n.n1, p1, i, j - bj, k);
                neighbour(n.n2, p2, i, j - 2 * bj, k);
                n.t[0] = lin(i + 1, j, k); n.t[1] = lin(i - 1, j, k); n.t[2] = lin(i, j, k + 1); n.t[3] = lin(i, j, k - 1);
            } else {  // x faces
                n.dir = bi < 0 ? 0 : 2;
                neighbour(n.n1, p1, i - bi, j, k);
                neighbour(n.n2, p2, i - 2 * bi, j, k);
                n.t[0] = lin(i, j + 1, k); n.t[1] = lin(i, j - 1, k); n.t[2] = lin(i, j, k + 1); n.t[3] = lin(i, j, k - 1);
            }
        } else {
            n.corner = 1;
            if (bi == 0) {  // edges along x: normal along the inward yz diagonal
                n.dir = bj < 0 ? 1 : 3;
                neighbour(n.n1, p1, i, j - bj, k - bk);
                neighbour(n.n2, p2, i, j - 2 * bj, k - 2 * bk);
            } else if (bi < 0) {  // inlet edges and corners
                n.dir = 0;
                if (nb == 3) {
                    neighbour(n.n1, p1, 1, j, k);
                    neighbour(n.n2, p2, 2, j, k);
                } else {
                    neighbour(n.n1, p1, 1, j - bj, k - bk);
                    neighbour(n.n2, p2, 2, j - 2 * bj, k - 2 * bk);
                }
            } else {  // outlet edges and corners
                n.dir = bk ? (bk < 0 ? 4 : 5) : (bj < 0 ? 1 : 3);
                neighbour(n.n1, p1, im - 1, j - bj, k - bk);
                neighbour(n.n2, p2, im - 2, j - 2 * bj, k - 2 * bk);
            }
        }
    };
    auto add = [&](int i, int j, int k) {
        LodiNode n;
        int64_t p1, p2;
        record(i, j, k, n, p1, p2);
        rd1.push_back(p1);
        rd2.push_back(p2);
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
    const size_t n = rd1.size();
    std::vector<int32_t> phase(n, 1), lb(n, 1);
    int32_t maxph = 1;
    for (size_t s = 0; s < n; ++s) {
        int32_t ph = lb[s];
        const int64_t rd[2] = {rd1[s], rd2[s]};
        for (int r = 0; r < 2; ++r)
            if (rd[r] >= 0 && rd[r] < (int64_t)s) ph = std::max(ph, phase[rd[r]] + 1);
        phase[s] = ph;
        maxph = std::max(maxph, ph);
        for (int r = 0; r < 2; ++r)
            if (rd[r] > (int64_t)s) lb[rd[r]] = std::max(lb[rd[r]], ph + 1);
    }
    // Emit each phase in linear node order (i fastest): nodes of one phase are
    // independent, and this order makes consecutive threads of the LODI kernel
    // touch consecutive nodes.  Only boundary nodes are visited.
    auto visit = [&](auto&& f) {
        for (int k = 0; k <= km; ++k)
            for (int j = 0; j <= jm; ++j) {
                if (k == 0 || k == km || j == 0 || j == jm) {
                    for (int i = 0; i <= im; ++i) f(i, j, k);
                } else {
                    f(0, j, k);
                    f(im, j, k);
                }
            }
    };
    std::vector<int64_t> cursor((size_t)maxph + 2, 0);
    visit([&](int i, int j, int k) { ++cursor[phase[pos3(i, j, k)] + 1]; });
    for (int32_t ph = 1; ph <= maxph; ++ph) cursor[ph + 1] += cursor[ph];
    LodiSchedule out;
    out.phase_start.assign(cursor.begin() + 1, cursor.end());
    out.nodes.resize(n);
    visit([&](int i, int j, int k) {
        int64_t p1, p2;
        record(i, j, k, out.nodes[cursor[phase[pos3(i, j, k)]]++], p1, p2);
    });
    return out;
}

LodiSchedule build_lodi_schedule_x(int nx, int ny, int nz, int faces) {
    const int im = nx - 1;
    auto lin = [&](int i, int j, int k) { return (int64_t)i + (int64_t)nx * ((int64_t)j + (int64_t)ny * k); };
    // x-face records as for the reference's x faces (Lattice.cpp:826-908),
    // tangent neighbours wrapped in y and z.  With nx >= 4 the normal
    // neighbours are interior nodes, so no record reads another's B entries
    // and the faces form one phase, emitted in linear node order.
    LodiSchedule out;
    out.nodes.reserve((size_t)((faces & 1) + ((faces >> 1) & 1)) * ny * nz);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int side = 0; side < 2; ++side) {
                if (!(faces & (1 << side))) continue;
                const int i = side ? im : 0, s = side ? -1 : 1;  // s: inward x direction
                LodiNode n;
                n.node = lin(i, j, k);
                n.n1 = lin(i + s, j, k);
                n.n2 = lin(i + 2 * s, j, k);
                n.t[0] = lin(i, (j + 1) % ny, k);
                n.t[1] = lin(i, (j + ny - 1) % ny, k);
                n.t[2] = lin(i, j, (k + 1) % nz);
                n.t[3] = lin(i, j, (k + nz - 1) % nz);
                n.dir = side ? 2 : 0;
                n.corner = 0;
                out.nodes.push_back(n);
            }
    out.phase_start = {0, (int64_t)out.nodes.size()};
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
    const int64_t rounded = acp::round_alloc((int64_t)bytes);
    if (c->arena) {
        if (c->arena_used + rounded > c->arena_size) {
            *p = nullptr;
            return fail(ACP_E_OOM, "case region too small: needs more than " + std::to_string(c->arena_size) +
                                       " bytes");
        }
        *p = reinterpret_cast<T*>(c->arena + c->arena_used);
        c->arena_used += rounded;
    } else {
        hipError_t e = hipMalloc((void**)p, bytes);
        if (e != hipSuccess) {
            *p = nullptr;
            (void)hipGetLastError();
            return fail(ACP_E_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed: " +
                                       hipGetErrorString(e));
        }
    }
    c->device_bytes += rounded;
    return ACP_OK;
}

// Give back a buffer obtained from dalloc (nothing to do inside a region).
template <typename T>
void dfree(acp_case* c, T*& p) {
    if (p && !c->arena) (void)hipFree(p);
    p = nullptr;
}

void drop_graphs(acp_case* c) {
    for (auto& g : c->gexec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

void release(acp_case* c) {
    drop_graphs(c);
    if (!c->arena)
        for (void* p : c->extra)
            if (p) (void)hipFree(p);
    c->extra.clear();
    void* ptrs[] = {c->dk, c->A, c->B, c->psi, c->M, c->dks, c->SA, c->SB, c->SM, c->lodi, c->probe_nodes,
                    c->probe_buf, c->probe_slot, c->partial, c->stage};
    if (!c->arena)
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
    if (c->mfma_ct) (void)hipFree(c->mfma_ct);
    if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
}

// Host-to-device copy ordered on the case's own stream (never the null stream,
// so that it neither waits for nor disturbs work other threads issue on other
// streams), complete on return.
hipError_t upload(acp_case* c, void* dst, const void* src, size_t bytes) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e;
}

// After a new initial state: step and probe counters back to zero.
int restart(acp_case* c) {
    c->initialized = true;
    c->steps = 0;
    c->probe_count = 0;
    if (c->probe_slot) {
        HIPCHK(hipMemsetAsync(c->probe_slot, 0, sizeof(int32_t), c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return ACP_OK;
}

// Upload host float64 populations [19][N] (SI units) into both arrays: one
// q-plane segment at a time through the staging buffer, placed in the block
// layout (and converted for the shifted tiers) on the device.
int upload_f64(acp_case* c, const double* f) {
    const int64_t N = c->g.N;
    const int64_t chunk = c->stage_elems;
    double* stage = c->stage;
    void* A = shifted(c) ? c->SA : (void*)c->A;
    void* B = shifted(c) ? c->SB : (void*)c->B;
    int rc = ACP_OK;
    for (int q = 0; q < 19 && rc == ACP_OK; ++q)
        for (int64_t l0 = 0; l0 < N && rc == ACP_OK; l0 += chunk) {
            const int64_t n = std::min(chunk, N - l0);
            hipError_t e = hipMemcpyAsync(stage, f + (size_t)q * N + l0, sizeof(double) * n, hipMemcpyHostToDevice,
                                          c->stream);
            if (e == hipSuccess)
                e = acp::launch_put_plane(c->precision, c->dks, stage, A, B, c->g.L, q, l0, n, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = acp::hip_fail(e, "population upload");
        }
    if (rc == ACP_OK && c->psi) {
        HIPCHK(hipMemsetAsync(c->psi, 0, sizeof(double) * 19 * N, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return rc;
}

// Download one population array of the case (A when `current`, else B) into
// host float64 [19][N] (SI units).
int download_array(acp_case* c, bool current, double* f) {
    const int64_t N = c->g.N;
    const int64_t chunk = c->stage_elems;
    double* stage = c->stage;
    const void* src = shifted(c) ? (current ? c->SA : c->SB) : (const void*)(current ? c->A : c->B);
    int rc = ACP_OK;
    for (int q = 0; q < 19 && rc == ACP_OK; ++q)
        for (int64_t l0 = 0; l0 < N && rc == ACP_OK; l0 += chunk) {
            const int64_t n = std::min(chunk, N - l0);
            hipError_t e = acp::launch_get_plane(c->precision, c->dks, src, stage, c->g.L, q, l0, n, c->stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(f + (size_t)q * N + l0, stage, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = acp::hip_fail(e, "population download");
        }
    return rc;
}

// A case that a pack steps belongs to the pack until acp_pack_remove: its
// host-side arrays and counters are stale until then.
int check_case(const acp_case* c) {
    if (!c) return fail(ACP_E_ARG, "null case handle");
    if (c->pack) return fail(ACP_E_STATE, "the case is stepped by a pack: acp_pack_remove it first");
    return ACP_OK;
}

// Initial condition rho(i, j) (the same on every k) with u = 0, populations
// feq_N: the host evaluates the density one band of rows at a time into the
// staging buffer; the device evaluates the equilibria and fills both arrays.
template <typename RhoFn>
int init_rho_field(acp_case* c, RhoFn rho_of) {
    const acp::Grid& g = c->g;
    const int rows = (int)std::min<int64_t>(g.ny, c->stage_elems / g.nx);
    if (rows < 1) return fail(ACP_E_ARG, "grid rows longer than the staging buffer");
    std::vector<double> band((size_t)rows * g.nx);
    void* A = shifted(c) ? c->SA : (void*)c->A;
    void* B = shifted(c) ? c->SB : (void*)c->B;
    for (int j0 = 0; j0 < g.ny; j0 += rows) {
        const int nr = std::min(rows, g.ny - j0);
        for (int j = 0; j < nr; ++j)
            for (int i = 0; i < g.nx; ++i) band[(size_t)j * g.nx + i] = rho_of(i, j0 + j, 0);
        // the upload waits on the stream, so the previous band's kernel is done with the stage
        HIPCHK(upload(c, c->stage, band.data(), sizeof(double) * (size_t)nr * g.nx));
        HIPCHK(acp::launch_init_rows(c->precision, c->d.order, c->dk, c->dks, c->stage, j0, nr, g, A, B, c->stream));
    }
    if (c->psi) HIPCHK(hipMemsetAsync(c->psi, 0, sizeof(double) * 19 * g.N, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return restart(c);
}

}  // namespace

namespace {

// float64 staging buffer of the host transfers: 1 Mi elements (8 MiB)
constexpr int64_t kStageElems = (int64_t)1 << 20;

// Nodes per population-layout block by tier (plane.hpp).  Measured with the
// collide + push kernels on the 512^3 periodic case: blocks of 8192 nodes run
// 9-12 % faster than whole planes in float64 and 3-7 % in float32; float16
// ran faster with whole planes.
int64_t layout_block_nodes(int precision) { return precision == ACP_PREC_F16_SHIFTED ? 0 : 8192; }

// The LODI x faces of a boundary kind: bit 0 the face i = 0, bit 1 i = nx - 1.
int x_lodi_faces(int32_t boundary) {
    return boundary == ACP_BC_LODI_X ? 3 : (boundary == ACP_BC_OPEN_LODI_X ? 2 : 0);
}

int validate_desc(const acp_case_desc& d) {
    if (d.max_level != 0)
        return fail(ACP_E_UNSUPPORTED, "MaxLevel > 0: two-level refinement is not part of this engine");
    if (d.nx < 3 || d.ny < 3 || d.nz < 3) return fail(ACP_E_ARG, "grid needs at least 3 nodes per direction");
    if (d.order < 2 || d.order > 4) return fail(ACP_E_CONFIG, "Wrong order of equilibrium!");
    if (d.op < ACP_OP_SRT || d.op > ACP_OP_HRR) return fail(ACP_E_ARG, "unknown collision operator");
    if (d.precision < ACP_PREC_F64 || d.precision > ACP_PREC_F16_SHIFTED) return fail(ACP_E_ARG, "unknown precision");
    if (d.boundary < ACP_BC_LODI || d.boundary > ACP_BC_OPEN_LODI_X) return fail(ACP_E_ARG, "unknown boundary");
    if ((d.boundary == ACP_BC_LODI_X || d.boundary == ACP_BC_OPEN_LODI_X) && d.nx < 4)
        return fail(ACP_E_ARG, "LODI x faces need at least 4 nodes in x");
    if (d.precision != ACP_PREC_F64 && d.cubic)
        return fail(ACP_E_UNSUPPORTED, "the cubic-Mach correction is available in the float64 tier only");
    if (!(d.cs > 0) || !(d.spacing > 0) || !(d.nu > 0) || !(d.rho0 > 0))
        return fail(ACP_E_ARG, "SpeedOfSound, Spacing, KinematicViscosity and Density must be positive");
    if (d.precision == ACP_PREC_F16_SHIFTED && !(d.f16_scale > 0)) return fail(ACP_E_ARG, "FP16 scale must be positive");
    return ACP_OK;
}

// Device bytes a case of `d` holds: the same buffers, in the same 256-byte
// units, as acp_case_create allocates, plus nprobe probes of `cap` steps.
int64_t footprint_bytes(const acp_case_desc& d, int32_t nprobe, int32_t cap) {
    using acp::round_alloc;
    const int64_t N = (int64_t)d.nx * d.ny * d.nz;
    const int64_t E = acp::make_layout(N, layout_block_nodes(d.precision)).elems;
    const bool need_macro = d.op == ACP_OP_HRR || (d.op == ACP_OP_RR && d.cubic);
    const bool need_psi = d.cubic && (d.op == ACP_OP_RR || d.op == ACP_OP_HRR);
    int64_t b = round_alloc(sizeof(acp::K64));
    if (d.precision == ACP_PREC_F64) {
        b += 2 * round_alloc(E * 8);
        if (need_psi) b += round_alloc(19 * N * 8);
        if (need_macro) b += round_alloc(4 * N * 8);
    } else {
        const int64_t eb = d.precision == ACP_PREC_F32_SHIFTED ? 4 : 2;
        b += round_alloc(sizeof(acp::KShift)) + 2 * round_alloc(E * eb);
        if (d.op == ACP_OP_HRR) b += round_alloc(4 * N * 4);
    }
    b += round_alloc(std::min<int64_t>(N, kStageElems) * 8);
    if (d.boundary == ACP_BC_LODI)
        b += round_alloc((N - (int64_t)(d.nx - 2) * (d.ny - 2) * (d.nz - 2)) * (int64_t)sizeof(acp::LodiNode));
    else if (x_lodi_faces(d.boundary))
        b += round_alloc((int64_t)(x_lodi_faces(d.boundary) == 3 ? 2 : 1) * d.ny * d.nz *
                         (int64_t)sizeof(acp::LodiNode));
    b += round_alloc(std::min<int64_t>(1024, (N + 255) / 256) * 8);
    if (nprobe > 0)
        b += round_alloc((int64_t)nprobe * 8) + round_alloc((int64_t)nprobe * std::max(cap, 1) * 8) + round_alloc(4);
    return b;
}

int create_case(const acp_case_desc* desc, void* stream, void* region, int64_t region_bytes, acp_case** out);

}  // namespace

extern "C" int acp_case_footprint(const acp_case_desc* desc, int32_t nprobe, int32_t probe_capacity,
                                  int64_t* bytes) {
    if (!desc || !bytes || nprobe < 0 || probe_capacity < 0) return fail(ACP_E_ARG, "acp_case_footprint: bad argument");
    if (int rc = validate_desc(*desc)) return rc;
    *bytes = footprint_bytes(*desc, nprobe, probe_capacity);
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_case_create(const acp_case_desc* desc, void* stream, acp_case** out) {
    return create_case(desc, stream, nullptr, 0, out);
}

extern "C" int acp_case_create_in(const acp_case_desc* desc, void* stream, void* region, int64_t region_bytes,
                                  acp_case** out) {
    if (!region || region_bytes <= 0 || ((uintptr_t)region % 256) != 0)
        return fail(ACP_E_ARG, "acp_case_create_in: region must be non-null, 256-byte aligned and non-empty");
    return create_case(desc, stream, region, region_bytes, out);
}

namespace {

int create_case(const acp_case_desc* desc, void* stream, void* region, int64_t region_bytes, acp_case** out) {
    if (!desc || !out) return fail(ACP_E_ARG, "acp_case_create: null argument");
    *out = nullptr;
    const acp_case_desc& d = *desc;
    if (int rc = validate_desc(d)) return rc;

    acp_case* c = new acp_case();
    c->arena = (char*)region;
    c->arena_size = region_bytes;
    c->d = d;
    c->g = acp::Grid{d.nx, d.ny, d.nz, (int64_t)d.nx * d.ny * d.nz,
                     d.boundary == ACP_BC_OPEN_X ? 3 : (d.boundary == ACP_BC_OPEN_LODI_X ? 1 : 0)};
    c->g.L = acp::make_layout(c->g.N, layout_block_nodes(d.precision));
    c->g.xlodi = x_lodi_faces(d.boundary);
    c->precision = d.precision;
    c->periodic = d.boundary != ACP_BC_LODI;
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
        if ((rc = dalloc(c, &c->dk, 1)) || (rc = dalloc(c, &c->A, (size_t)c->g.L.elems)) ||
            (rc = dalloc(c, &c->B, (size_t)c->g.L.elems)))
            return bail(rc);
        if (c->need_psi && (rc = dalloc(c, &c->psi, (size_t)19 * N))) return bail(rc);
        if (c->need_macro && (rc = dalloc(c, &c->M, (size_t)4 * N))) return bail(rc);
        hipError_t e = upload(c, c->dk, &c->hk, sizeof(acp::K64));
        if (e != hipSuccess) return bail(acp::hip_fail(e, "constants upload"));
    } else {
        const size_t eb = elem_bytes(c->precision);
        char *a = nullptr, *b = nullptr;
        if ((rc = dalloc(c, &c->dks, 1)) || (rc = dalloc(c, &a, (size_t)c->g.L.elems * eb)) ||
            (rc = dalloc(c, &b, (size_t)c->g.L.elems * eb)))
            return bail(rc);
        c->SA = a;
        c->SB = b;
        if (d.op == ACP_OP_HRR && (rc = dalloc(c, &c->SM, (size_t)4 * N))) return bail(rc);
        hipError_t e = upload(c, c->dks, &c->hks, sizeof(acp::KShift));
        if (e != hipSuccess) return bail(acp::hip_fail(e, "constants upload"));
        // the float64 constants are used by the conversions of the reference initial condition
        if ((rc = dalloc(c, &c->dk, 1))) return bail(rc);
        e = upload(c, c->dk, &c->hk, sizeof(acp::K64));
        if (e != hipSuccess) return bail(acp::hip_fail(e, "constants upload"));
    }
    if (!c->periodic || c->g.xlodi) {
        acp::LodiSchedule sch = c->periodic ? acp::build_lodi_schedule_x(d.nx, d.ny, d.nz, c->g.xlodi)
                                            : acp::build_lodi_schedule(d.nx, d.ny, d.nz);
        c->phase_start = sch.phase_start;
        if ((rc = dalloc(c, &c->lodi, sch.nodes.size()))) return bail(rc);
        hipError_t e = upload(c, c->lodi, sch.nodes.data(), sizeof(acp::LodiNode) * sch.nodes.size());
        if (e != hipSuccess) return bail(acp::hip_fail(e, "LODI schedule upload"));
    }
    c->partial_blocks = (int)std::min<int64_t>(1024, (N + 255) / 256);
    if ((rc = dalloc(c, &c->partial, (size_t)c->partial_blocks))) return bail(rc);
    c->stage_elems = std::min<int64_t>(N, kStageElems);
    if ((rc = dalloc(c, &c->stage, (size_t)c->stage_elems))) return bail(rc);
    // the padding of the last layout block is never written: zero it once, so
    // that whole-array reductions (the checksum) see fixed values there
    if (c->g.L.elems > 19 * N) {
        const size_t bytes = (size_t)c->g.L.elems * (shifted(c) ? elem_bytes(c->precision) : sizeof(double));
        hipError_t e = hipMemsetAsync(shifted(c) ? c->SA : (void*)c->A, 0, bytes, c->stream);
        if (e == hipSuccess) e = hipMemsetAsync(shifted(c) ? c->SB : (void*)c->B, 0, bytes, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return bail(acp::hip_fail(e, "population arrays"));
    }
    c->probe_mark = c->arena_used;
    c->arr0 = shifted(c) ? c->SA : (const void*)c->A;
    *out = c;
    acp::clear_error();
    return ACP_OK;
}

}  // namespace

extern "C" void acp_case_destroy(acp_case* c) {
    if (!c) return;
    acp::pack_detach(c);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    release(c);
    delete c;
}

extern "C" int acp_case_info_get(const acp_case* c, acp_case_info* out) {
    if (!c || !out) return fail(ACP_E_ARG, "acp_case_info_get: null argument");
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

namespace acp {
// Lattice::initialization (Lattice.cpp:1397-1493) on one level of a lattice:
// the reference's hard-coded pulse (Lattice.cpp:1415-1433) at the level's node
// coordinates (createLattice, Lattice.cpp:33-35): x = (i + i0) * (spacing /
// 2^level), y = j * (spacing / 2^level), with the level-0 spacing, density
// and domain size of d0; i0 is the level's first node in the level's own
// index space of the whole domain.
int init_reference_level(acp_case* c, int level, int i0, const acp_case_desc& d0) {
    if (int rc = check_case(c)) return rc;
    const double Rc = 0.06, A = 0.01, xc = -8. * Rc;
    const double dens = d0.rho0, sp = d0.spacing, lx = d0.domain[0], ly = d0.domain[1];
    return init_rho_field(c, [&](int i, int j, int) {
        const double x = (i + i0) * (sp / std::pow(2., level)), y = j * (sp / std::pow(2., level));
        const double xn = x - lx / 2., yn = y - ly / 2.;
        const double xr = xn - xc;
        return dens * (1 + A * std::exp(-(xr * xr + yn * yn) / (2 * Rc * Rc)));
    });
}
}  // namespace acp

extern "C" int acp_case_init_reference(acp_case* c) {
    if (int rc = check_case(c)) return rc;
    return acp::init_reference_level(c, 0, 0, c->d);
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

extern "C" int acp_case_init_density(acp_case* c, const double* rho) {
    if (int rc = check_case(c)) return rc;
    if (!rho) return fail(ACP_E_ARG, "null density field");
    void* A = shifted(c) ? c->SA : (void*)c->A;
    void* B = shifted(c) ? c->SB : (void*)c->B;
    for (int64_t l0 = 0; l0 < c->g.N; l0 += c->stage_elems) {
        const int64_t n = std::min(c->stage_elems, c->g.N - l0);
        // the upload waits on the stream, so the previous range's kernel is done with the stage
        HIPCHK(upload(c, c->stage, rho + l0, sizeof(double) * n));
        HIPCHK(acp::launch_init_nodes(c->precision, c->d.order, c->dk, c->dks, c->stage, l0, n, c->g, A, B,
                                      c->stream));
    }
    if (c->psi) HIPCHK(hipMemsetAsync(c->psi, 0, sizeof(double) * 19 * c->g.N, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return restart(c);
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
    return restart(c);
}

extern "C" int acp_case_set_populations(acp_case* c, const double* f) {
    if (int rc = check_case(c)) return rc;
    if (!f) return fail(ACP_E_ARG, "null population array");
    int rc = upload_f64(c, f);
    if (rc == ACP_OK) rc = restart(c);
    return rc;
}

namespace {

// Launch n steps on the case stream, one kernel sequence per step.
int step_eager(acp_case* c, int32_t n) {
    const int nph = c->phase_start.empty() ? 0 : (int)c->phase_start.size() - 1;
    for (int32_t s = 0; s < n; ++s) {
        hipError_t e = hipSuccess;
        if (!shifted(c)) {
            if (c->need_macro) e = acp::launch_macro64(c->dk, c->A, c->M, c->g, c->periodic, c->stream);
            if (e == hipSuccess && c->g.variant == acp::BULK_MFMA)
                e = acp::launch_bulk64_mfma(c->dk, c->A, c->B, c->mfma_ct, c->g, c->stream);
            else if (e == hipSuccess)
                e = acp::launch_bulk64(c->bulk_op, c->d.order, c->d.cubic != 0, c->periodic, c->dk, c->A, c->B,
                                       c->psi, c->M, c->g, c->stream);
            for (int x = 0; x < c->nfix && e == hipSuccess; ++x)
                e = acp::launch_fix64(c->bulk_op, c->d.order, c->d.cubic != 0, c->dk, c->A, c->B, c->psi, c->g,
                                      c->fix_col[x], c->fix_S[x], c->stream);
            if (e == hipSuccess && c->fix_list_n > 0)
                e = acp::launch_fix64_list(c->bulk_op, c->d.order, c->d.cubic != 0, c->dk, c->A, c->B, c->psi, c->g,
                                           c->fix_list, c->fix_list_n, c->fix_list_P, c->stream);
            for (int p = 0; p < nph && e == hipSuccess; ++p)
                e = acp::launch_lodi64(c->d.order, c->d.cubic != 0, c->d.transverse != 0, c->dk, c->A, c->B, c->psi,
                                       c->lodi + c->phase_start[p], c->phase_start[p + 1] - c->phase_start[p], c->g.L,
                                       c->stream);
            if (e == hipSuccess && c->nprobe > 0)
                e = acp::launch_probe64(c->B, c->g.L, c->probe_nodes, c->nprobe, c->probe_buf, c->probe_slot,
                                        c->probe_cap, c->stream);
            std::swap(c->A, c->B);
        } else {
            if (c->SM) e = acp::launch_macro_shift(c->precision, c->dks, c->SA, c->SM, c->g, c->stream);
            if (e == hipSuccess)
                e = acp::launch_bulk_shift(c->precision, c->bulk_op, c->d.order, c->periodic, c->dks, c->SA, c->SB,
                                           c->SM, c->g, c->stream);
            for (int x = 0; x < c->nfix && e == hipSuccess; ++x)
                e = acp::launch_fix_hrr_shift(c->precision, c->d.order, c->dks, c->SA, c->SB, c->g, c->fix_col[x],
                                              c->fix_S[x], c->hk.dt, c->stream);
            for (int p = 0; p < nph && e == hipSuccess; ++p)
                e = acp::launch_lodi_shift(c->precision, c->d.order, c->d.transverse != 0, c->dks, c->SA, c->SB,
                                           c->lodi + c->phase_start[p], c->phase_start[p + 1] - c->phase_start[p],
                                           c->g.L, c->stream);
            if (e == hipSuccess && c->nprobe > 0)
                e = acp::launch_probe_shift(c->precision, c->dks, c->SB, c->g.L, c->probe_nodes, c->nprobe,
                                            c->probe_buf, c->probe_slot, c->probe_cap, c->stream);
            std::swap(c->SA, c->SB);
        }
        if (e != hipSuccess) return acp::hip_fail(e, "acp_case_step launch");
        if (c->nprobe > 0) ++c->probe_count;
        ++c->steps;
    }
    return ACP_OK;
}

int parity(const acp_case* c) { return (shifted(c) ? c->SA : (const void*)c->A) == c->arr0 ? 0 : 1; }

// One chunk of graph_chunk steps (an even count, so the arrays end where they
// started) as a captured graph; captured on first use for each parity.
int step_graph_chunk(acp_case* c) {
    const int p = parity(c);
    if (!c->gexec[p]) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        const int32_t steps0 = c->steps, count0 = c->probe_count;
        const int rc = step_eager(c, c->graph_chunk);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        c->steps = steps0;
        c->probe_count = count0;
        if (rc != ACP_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&c->gexec[p], g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            c->gexec[p] = nullptr;
            return acp::hip_fail(e, "step graph capture");
        }
    }
    HIPCHK(hipGraphLaunch(c->gexec[p], c->stream));
    c->steps += c->graph_chunk;
    if (c->nprobe > 0) c->probe_count += c->graph_chunk;
    return ACP_OK;
}

}  // namespace

namespace acp {
int case_step_launch(acp_case* c, int32_t n) { return step_eager(c, n); }

int case_alloc_extra(acp_case* c, int64_t bytes, void** p) {
    char* q = nullptr;
    if (int rc = dalloc(c, &q, (size_t)bytes)) return rc;
    c->extra.push_back(q);
    // probes set later go after this buffer in a placed case
    c->probe_mark = c->arena_used;
    *p = q;
    return ACP_OK;
}
}  // namespace acp

extern "C" int acp_case_step(acp_case* c, int32_t n) {
    if (int rc = check_case(c)) return rc;
    if (!c->initialized) return fail(ACP_E_STATE, "acp_case_step before initialisation");
    if (n < 0) return fail(ACP_E_ARG, "negative step count");
    if (c->nprobe > 0 && c->probe_count + n > c->probe_cap)
        return fail(ACP_E_STATE, "probe buffer full: read the probes first");
    while (c->graph_chunk > 0 && n >= c->graph_chunk) {
        if (int rc = step_graph_chunk(c)) return rc;
        n -= c->graph_chunk;
    }
    return step_eager(c, n);
}

extern "C" int acp_case_enable_graphs(acp_case* c, int32_t chunk) {
    if (int rc = check_case(c)) return rc;
    if (chunk < 0 || chunk == 1 || (chunk % 2) != 0)
        return fail(ACP_E_ARG, "graph chunk must be 0 (off) or an even number of steps");
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graphs(c);
    c->graph_chunk = chunk;
    return ACP_OK;
}

extern "C" int acp_case_set_bulk_variant(acp_case* c, int32_t variant) {
    if (int rc = check_case(c)) return rc;
    if (variant < acp::BULK_AUTO || variant > acp::BULK_RELATIVE_NT)
        return fail(ACP_E_ARG, "unknown bulk-kernel variant");
    if (variant == acp::BULK_MFMA) {
        if (shifted(c) || c->bulk_op != acp::OP_MRT_RM || !c->periodic || c->g.xopen || c->g.xlodi ||
            c->g.nx % 256 != 0)
            return fail(ACP_E_UNSUPPORTED, "the matrix-core form needs float64 MRT-RM (relaxed) on a periodic grid "
                                           "with nx a multiple of 256");
        if (!c->mfma_ct) {
            double h[acp::kMfmaRows * acp::kMfmaCols];
            acp::mrt_rm_collision_matrix(c->hk, h);
            HIPCHK(hipMalloc((void**)&c->mfma_ct, sizeof(h)));
            HIPCHK(upload(c, c->mfma_ct, h, sizeof(h)));
        }
    }
    const int eb = shifted(c) ? (int)elem_bytes(c->precision) : 8;
    if ((variant == acp::BULK_NARROW_NT || variant == acp::BULK_NARROW_PLAIN) &&
        !acp::narrow_planes(c->g.L, eb, shifted(c) ? 4 : 8))
        return fail(ACP_E_ARG, "the arrays of this case need 64-bit offsets");
    HIPCHK(hipStreamSynchronize(c->stream));
    drop_graphs(c);  // captured graphs hold the old kernels
    c->g.variant = variant;
    ++c->graph_epoch;
    return ACP_OK;
}

extern "C" int acp_case_checksum(acp_case* c, uint64_t* out) {
    if (int rc = check_case(c)) return rc;
    if (!out) return fail(ACP_E_ARG, "null checksum output");
    const void* a = shifted(c) ? c->SA : (const void*)c->A;
    const int eb = shifted(c) ? (int)elem_bytes(c->precision) : 8;
    unsigned long long* part = reinterpret_cast<unsigned long long*>(c->partial);
    HIPCHK(acp::launch_checksum(a, eb, c->g.L.elems, part, c->partial_blocks, c->stream));
    std::vector<unsigned long long> h(c->partial_blocks);
    HIPCHK(hipMemcpyAsync(h.data(), part, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    uint64_t sum = 0;
    for (unsigned long long v : h) sum += (uint64_t)v;
    *out = sum;
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
    return download_array(c, true, f);
}

extern "C" int acp_case_get_previous_populations(acp_case* c, double* f) {
    if (int rc = check_case(c)) return rc;
    if (!f) return fail(ACP_E_ARG, "null population array");
    return download_array(c, false, f);
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

extern "C" int acp_case_get_psi(acp_case* c, double* psi) {
    if (int rc = check_case(c)) return rc;
    if (!psi) return fail(ACP_E_ARG, "null output array");
    const size_t n = (size_t)19 * c->g.N;
    if (!c->psi) {
        std::fill(psi, psi + n, 0.);
        return ACP_OK;
    }
    HIPCHK(hipMemcpyAsync(psi, c->psi, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ACP_OK;
}

extern "C" int acp_case_mass(acp_case* c, double* mass) {
    if (int rc = check_case(c)) return rc;
    if (!mass) return fail(ACP_E_ARG, "null output");
    hipError_t e = shifted(c) ? acp::launch_mass_shift(c->precision, c->dks, c->SA, c->g.L, c->partial,
                                                       c->partial_blocks, c->stream)
                              : acp::launch_mass64(c->A, c->g.L, c->partial, c->partial_blocks, c->stream);
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
    drop_graphs(c);  // captured graphs hold the old buffers
    ++c->graph_epoch;
    dfree(c, c->probe_nodes);
    dfree(c, c->probe_buf);
    dfree(c, c->probe_slot);
    c->device_bytes -= c->probe_bytes;
    c->probe_bytes = 0;
    if (c->arena) c->arena_used = c->probe_mark;
    c->nprobe = 0;
    c->probe_cap = 0;
    c->probe_count = 0;
    if (nprobe == 0) return ACP_OK;
    const int64_t before = c->device_bytes;
    int rc = dalloc(c, &c->probe_nodes, (size_t)nprobe);
    if (rc == ACP_OK) rc = dalloc(c, &c->probe_buf, (size_t)nprobe * std::max(capacity, 1));
    if (rc == ACP_OK) rc = dalloc(c, &c->probe_slot, 1);
    c->probe_bytes = c->device_bytes - before;
    if (rc != ACP_OK) {
        dfree(c, c->probe_nodes);
        dfree(c, c->probe_buf);
        dfree(c, c->probe_slot);
        c->device_bytes = before;
        c->probe_bytes = 0;
        if (c->arena) c->arena_used = c->probe_mark;
        return rc;
    }
    HIPCHK(hipMemcpyAsync(c->probe_nodes, nodes, sizeof(int64_t) * nprobe, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemsetAsync(c->probe_slot, 0, sizeof(int32_t), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
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
    }
    if (c->probe_slot) HIPCHK(hipMemsetAsync(c->probe_slot, 0, sizeof(int32_t), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
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

static_assert(sizeof(acp_lodi_record) == sizeof(acp::LodiNode) &&
                  offsetof(acp_lodi_record, corner) == offsetof(acp::LodiNode, corner) &&
                  offsetof(acp_lodi_record, dir) == offsetof(acp::LodiNode, dir),
              "acp_lodi_record mirrors the kernels' record");

extern "C" int64_t acp_lodi_launches(int32_t nx, int32_t ny, int32_t nz, acp_lodi_record* records, int64_t cap,
                                     int64_t* launch_start, int32_t launch_cap, int32_t* nlaunch) {
    if (nx < 3 || ny < 3 || nz < 3) return 0;
    acp::LodiSchedule s = acp::build_lodi_schedule(nx, ny, nz);
    const int64_t n = (int64_t)s.nodes.size();
    if (records && cap > 0) std::memcpy(records, s.nodes.data(), sizeof(acp_lodi_record) * (size_t)std::min(n, cap));
    const int32_t nl = (int32_t)s.phase_start.size() - 1;
    for (int32_t p = 0; p <= nl && p < launch_cap; ++p)
        if (launch_start) launch_start[p] = s.phase_start[p];
    if (nlaunch) *nlaunch = nl;
    return n;
}

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
        HIPCHK(hipMemcpyAsync(psi.data(), c->psi, sizeof(double) * 19 * N, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return acp::write_vtk_host(c->d, c->hk, cur.data(), prev.empty() ? nullptr : prev.data(),
                               psi.empty() ? nullptr : psi.data(), stem);
}
