// VTK output with the byte layout of Lattice::writeResultsVTK
// (Lattice.cpp:2212-2243 header, 4094-4349 single level): legacy ASCII
// UNSTRUCTURED_GRID of VOXEL cells with POINT_DATA Pressure, Density, Omega and
// Velocity, numbers as C++ streams print them with scientific | showpoint and
// precision 8 (printf "%.8e"), and the POINTS / CELLS counts patched into the
// blank fields reserved after their keywords.
//
// With Non-Reflecting yes the reference prints boundary nodes with i > 0 from
// the other population array (Lattice.cpp:4213-4217, 4257-4261, 4324-4330);
// `fprev` is that array.  The boundary nodes are the LODI nodes of the
// boundary kind: every face node with LODI on all faces, the LODI x faces'
// nodes of a grid wrapped in y and z, none of a periodic grid.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "engine.hpp"

namespace {

struct Out {
    std::string s;
    void str(const char* t) { s += t; }
    void num(double v) {
        char b[64];
        std::snprintf(b, sizeof(b), "%.8e", v);
        s += b;
    }
    void inum(long long v) { s += std::to_string(v); }
};

}  // namespace

namespace acp {

namespace {

// Whether node (i, j, k) of a level is a LODI boundary node (of its boundary kind).
bool lodi_node(const acp_case_desc& d, int i, int j, int kz) {
    if (d.boundary == ACP_BC_LODI)
        return i == 0 || j == 0 || kz == 0 || i == d.nx - 1 || j == d.ny - 1 || kz == d.nz - 1;
    if (d.boundary == ACP_BC_LODI_X) return i == 0 || i == d.nx - 1;
    return d.boundary == ACP_BC_OPEN_LODI_X && i == d.nx - 1;
}

// Density and velocity of every node of a level as calcDensity / calcVelocity
// print them; LODI nodes with i > 0 from the other array under Non-Reflecting.
void level_macros(const acp_case_desc& d, const K64& k, const double* fcur, const double* fprev, const double* psi,
                  std::vector<double>& rho, std::vector<double>& ux, std::vector<double>& uy, std::vector<double>& uz) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int64_t N = (int64_t)nx * ny * nz;
    rho.resize((size_t)N), ux.resize((size_t)N), uy.resize((size_t)N), uz.resize((size_t)N);
    for (int kz = 0; kz < nz; ++kz)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t l = i + (int64_t)nx * (j + (int64_t)ny * kz);
                const double* F = (d.nonreflecting && i > 0 && lodi_node(d, i, j, kz)) ? fprev : fcur;
                double f[19], ps[19];
                for (int q = 0; q < 19; ++q) f[q] = F[(size_t)q * N + l];
                if (d.cubic) {
                    for (int q = 0; q < 19; ++q) ps[q] = psi ? psi[(size_t)q * N + l] : 0.;
                    moments64_psi(k, f, ps, rho[l], ux[l], uy[l], uz[l]);
                } else {
                    moments64(k, f, rho[l], ux[l], uy[l], uz[l]);
                }
            }
}

int write_file(const std::string& text, const char* stem) {
    std::string fname = std::string(stem) + ".vtk";
    std::ofstream fh(fname, std::ios::binary);
    if (!fh) return fail(ACP_E_IO, "An error occured while writing paraview file " + fname);
    fh.write(text.data(), (std::streamsize)text.size());
    if (!fh) return fail(ACP_E_IO, "short write " + fname);
    return ACP_OK;
}

void voxel(Out& o, long long n0, long long sx, long long sxy) {
    const long long v[8] = {n0, n0 + 1, n0 + sx, n0 + sx + 1, n0 + sxy, n0 + sxy + 1, n0 + sxy + sx, n0 + sxy + sx + 1};
    o.str("8 ");
    for (int a = 0; a < 8; ++a) {
        o.inum(v[a]);
        o.str(a == 7 ? "\n" : " ");
    }
}

}  // namespace

// Two-level output (the multi-level branch of Lattice::writeResultsVTK,
// Lattice.cpp:2243-4092, for this engine's geometry): the reference writes
// the fine regions' nodes first and then the coarse nodes of the unrefined
// core, interface nodes on both levels, VOXEL cells of each level between its
// printed nodes, and POINT_DATA Omega, NodeID, Pressure, Density, Velocity in
// that point order.  Here the fine level covers coarse x in [x0, x1]: its
// nodes come first (k, j, i ascending, x = (2 x0 + i) spacing / 2), then the
// coarse nodes outside the fine region's interior, columns 0..x0 and x1..nx-1
// (none beyond x1 when x1 = nx wraps a periodic level).  NodeID: 0 fluid, 1
// LODI boundary, 2 grid-coupling (the fine level's coupled faces).
int write_vtk2_host(const acp_case_desc& cd, const K64& ck, const double* ccur, const double* cprev,
                    const acp_case_desc& fd, const K64& fk, const double* fcur, const double* fprev, int x0, int x1,
                    const char* stem, const double* cpsi, const double* fpsi) {
    const int nx = cd.nx, ny = cd.ny, nz = cd.nz, fnx = fd.nx, fny = fd.ny, fnz = fd.nz;
    const int64_t FN = (int64_t)fnx * fny * fnz;
    std::vector<int> cols;
    for (int i = 0; i <= x0; ++i) cols.push_back(i);
    for (int i = x1; i < nx; ++i) cols.push_back(i);
    const long long ncol = (long long)cols.size();
    const long long CP = ncol * ny * nz, P = FN + CP;
    const double sp0 = cd.spacing / std::pow(2., 0), sp1 = cd.spacing / std::pow(2., 1);
    Out o;
    o.s.reserve((size_t)P * 160 + 1024);
    o.str("# vtk DataFile Version 2.0\n");
    o.str("LBM results for grid refinement benchmark in channel flow\n");
    o.str("ASCII\n");
    o.str("DATASET UNSTRUCTURED_GRID\n");
    o.str("POINTS ");
    const size_t pnt_pos = o.s.size();
    o.str((std::string(14, ' ') + " float\n").c_str());
    auto point = [&](double x, double y, double z) {
        o.num(x);
        o.str(" ");
        o.num(y);
        o.str(" ");
        o.num(z);
        o.str("\n");
    };
    for (int kz = 0; kz < fnz; ++kz)
        for (int j = 0; j < fny; ++j)
            for (int i = 0; i < fnx; ++i) point((2 * x0 + i) * sp1, j * sp1, kz * sp1);
    for (int kz = 0; kz < nz; ++kz)
        for (int j = 0; j < ny; ++j)
            for (int i : cols) point(i * sp0, j * sp0, kz * sp0);
    o.str("\n");
    o.str("CELLS ");
    const size_t cell_pos = o.s.size();
    o.str((std::string(9, ' ') + " " + std::string(9, ' ') + "\n").c_str());
    long long cells = 0;
    for (int kz = 0; kz < fnz - 1; ++kz)
        for (int j = 0; j < fny - 1; ++j)
            for (int i = 0; i < fnx - 1; ++i, ++cells)
                voxel(o, i + (long long)fnx * (j + (long long)fny * kz), fnx, (long long)fnx * fny);
    for (int kz = 0; kz < nz - 1; ++kz)
        for (int j = 0; j < ny - 1; ++j)
            for (long long c = 0; c + 1 < ncol; ++c) {
                if (cols[c + 1] != cols[c] + 1) continue;  // no cell across the fine region
                voxel(o, FN + c + ncol * (j + (long long)ny * kz), ncol, ncol * ny);
                ++cells;
            }
    o.str("\n");
    o.str("CELL_TYPES ");
    o.inum(cells);
    o.str("\n");
    for (long long c = 0; c < cells; ++c) o.str("11\n");
    o.str("\n");
    o.str("POINT_DATA ");
    o.inum(P);
    o.str("\n");
    std::vector<double> fr, fu, fv, fw, cr, cu, cv, cw;
    level_macros(fd, fk, fcur, fprev, fpsi, fr, fu, fv, fw);
    level_macros(cd, ck, ccur, cprev, cpsi, cr, cu, cv, cw);
    // every printed point of both levels in the POINTS order: fn(level, l)
    auto each = [&](auto fn) {
        for (int64_t l = 0; l < FN; ++l) fn(1, l, (int)(l % fnx));
        for (int kz = 0; kz < nz; ++kz)
            for (int j = 0; j < ny; ++j)
                for (int i : cols) fn(0, i + (int64_t)nx * (j + (int64_t)ny * kz), i);
    };
    o.str("SCALARS Omega float 1\n");
    o.str("LOOKUP_TABLE default\n");
    each([&](int lv, int64_t, int) {
        o.num(lv ? fk.omega : ck.omega);
        o.str("\n");
    });
    o.str("SCALARS NodeID float 1\n");
    o.str("LOOKUP_TABLE default\n");
    each([&](int lv, int64_t l, int i) {
        const acp_case_desc& d = lv ? fd : cd;
        const int j = (int)((l / d.nx) % d.ny), kz = (int)(l / ((int64_t)d.nx * d.ny));
        int id = lodi_node(d, i, j, kz) ? 1 : 0;
        if (lv && !id && (i == 0 || (fd.boundary == ACP_BC_OPEN_X && i == fnx - 1))) id = 2;
        o.inum(id);
        o.str("\n");
    });
    const double cs2 = std::pow(cd.cs, 2.);
    o.str("SCALARS Pressure float 1\n");
    o.str("LOOKUP_TABLE default\n");
    each([&](int lv, int64_t l, int) {
        o.num((lv ? fr : cr)[l] * cs2);
        o.str("\n");
    });
    o.str("SCALARS Density float 1\n");
    o.str("LOOKUP_TABLE default\n");
    each([&](int lv, int64_t l, int) {
        o.num((lv ? fr : cr)[l]);
        o.str("\n");
    });
    o.str("\nVECTORS Velocity float \n");
    each([&](int lv, int64_t l, int) {
        o.num((lv ? fu : cu)[l]);
        o.str(" ");
        o.num((lv ? fv : cv)[l]);
        o.str(" ");
        o.num((lv ? fw : cw)[l]);
        o.str(" \n");
    });
    const std::string pc = std::to_string(P);
    o.s.replace(pnt_pos, pc.size(), pc);
    const std::string cc = std::to_string(cells) + " " + std::to_string(9 * cells);
    o.s.replace(cell_pos, cc.size(), cc);
    return write_file(o.s, stem);
}

// Two-level output of the reference's own geometry (a fine shell around a
// coarse core; the multi-level branch of Lattice::writeResultsVTK,
// Lattice.cpp:2243-4092, as written).  Points in seven parts: on level 1
// (indices over the domain, refLayer = 2 ly, refLayerX = 2 lx) the bottom
// slab k <= refLayer, the top slab k >= kmax - refLayer, the left slab
// j <= refLayer and the right slab j >= jmax - refLayer (both for refLayer <= k
// <= kmax - refLayer), the inlet block i <= refLayerX and the outlet block
// i >= imax - refLayer (both for refLayer <= j, k <= max - refLayer), then on
// level 0 the core [lx, imax - ly] x [ly, jmax - ly] x [ly, kmax - ly]; the
// planes where two parts meet are printed by both.  Cells: VOXELs between
// each part's points, with the part's own node strides, except that the core
// part's z stride is the level-1 plane size (the reference computes it once
// for level 1 and does not reset it for level 0), which its cells keep.
// NodeID: 1 for level-1 boundary nodes, 2 for grid-coupling nodes (level-1
// nodes on the hole box's surface; the core part holds no level-0
// grid-coupling node), 0 otherwise.
int write_vtk_shell_host(const acp_case_desc& cd, const K64& ck, const double* ccur, const double* cprev,
                         const double* cpsi, const acp_case_desc& fd, const K64& fk, const double* fcur,
                         const double* fprev, const double* fpsi, int ly, int lx, const char* stem) {
    const int imax = fd.nx - 1, jmax = fd.ny - 1, kmax = fd.nz - 1, R = 2 * ly, RX = 2 * lx;
    const int imax0 = cd.nx - 1, jmax0 = cd.ny - 1, kmax0 = cd.nz - 1;
    struct Part {
        int level, i0, i1, j0, j1, k0, k1;
    };
    const Part parts[7] = {{1, 0, imax, 0, jmax, 0, R},
                           {1, 0, imax, 0, jmax, kmax - R, kmax},
                           {1, 0, imax, 0, R, R, kmax - R},
                           {1, 0, imax, jmax - R, jmax, R, kmax - R},
                           {1, 0, RX, R, jmax - R, R, kmax - R},
                           {1, imax - R, imax, R, jmax - R, R, kmax - R},
                           {0, lx, imax0 - ly, ly, jmax0 - ly, ly, kmax0 - ly}};
    long long cnt[7], P = 0;
    for (int p = 0; p < 7; ++p) {
        const Part& q = parts[p];
        cnt[p] = (long long)(q.i1 - q.i0 + 1) * (q.j1 - q.j0 + 1) * (q.k1 - q.k0 + 1);
        P += cnt[p];
    }
    auto each = [&](auto fn) {
        for (const Part& q : parts)
            for (int k = q.k0; k <= q.k1; ++k)
                for (int j = q.j0; j <= q.j1; ++j)
                    for (int i = q.i0; i <= q.i1; ++i) fn(q.level, i, j, k);
    };
    Out o;
    o.s.reserve((size_t)P * 160 + 1024);
    o.str("# vtk DataFile Version 2.0\n");
    o.str("LBM results for grid refinement benchmark in channel flow\n");
    o.str("ASCII\n");
    o.str("DATASET UNSTRUCTURED_GRID\n");
    o.str("POINTS ");
    const size_t pnt_pos = o.s.size();
    o.str((std::string(14, ' ') + " float\n").c_str());
    const double sp[2] = {cd.spacing / std::pow(2., 0), cd.spacing / std::pow(2., 1)};
    each([&](int lv, int i, int j, int k) {
        o.num(i * sp[lv]);
        o.str(" ");
        o.num(j * sp[lv]);
        o.str(" ");
        o.num(k * sp[lv]);
        o.str("\n");
    });
    o.str("\n");
    o.str("CELLS ");
    const size_t cell_pos = o.s.size();
    o.str((std::string(9, ' ') + " " + std::string(9, ' ') + "\n").c_str());
    long long cells = 0;
    // level-1 node counts and strides (nodeCntrX, nodeCntrXY, nodeCntrRef, nodeCntrRefX)
    const long long X = imax + 1, Y = jmax + 1, XY = X * Y, Rn = R + 1, RXn = RX + 1, Ymid = Y - 2 * (Rn - 1);
    long long base = 0;
    auto cell = [&](long long n0, long long sx, long long sz) {
        voxel(o, n0, sx, sz);
        ++cells;
    };
    for (int k = 0; k <= R - 1; ++k)  // bottom
        for (int j = 0; j <= jmax - 1; ++j)
            for (int i = 0; i <= imax - 1; ++i) cell(i + j * X + k * XY, X, XY);
    base = cnt[0];
    for (int k = 0; k < R; ++k)  // top
        for (int j = 0; j <= jmax - 1; ++j)
            for (int i = 0; i <= imax - 1; ++i) cell(i + j * X + k * XY + base, X, XY);
    base += cnt[1];
    for (int k = 0; k <= kmax - 1 - 2 * R; ++k)  // left
        for (int j = 0; j <= R - 1; ++j)
            for (int i = 0; i <= imax - 1; ++i) cell(i + j * X + k * Rn * X + base, X, Rn * X);
    base += cnt[2];
    for (int k = 0; k <= kmax - 1 - 2 * R; ++k)  // right
        for (int j = 0; j < R; ++j)
            for (int i = 0; i <= imax - 1; ++i) cell(i + j * X + k * Rn * X + base, X, Rn * X);
    base += cnt[3];
    for (int k = 0; k <= kmax - 1 - 2 * R; ++k)  // inlet
        for (int j = 0; j <= jmax - 1 - 2 * R; ++j)
            for (int i = 0; i <= RX - 1; ++i) cell(i + j * RXn + k * RXn * Ymid + base, RXn, RXn * Ymid);
    base += cnt[4];
    for (int k = 0; k <= kmax - 1 - 2 * R; ++k)  // outlet
        for (int j = 0; j <= jmax - 1 - 2 * R; ++j)
            for (int i = 0; i < R; ++i) cell(i + j * Rn + k * Rn * Ymid + base, Rn, Rn * Ymid);
    base += cnt[5];
    // the core (level 0): rows of nodeCntrX - (refLayerX + refLayer) points, z stride XY of level 1
    const long long W = (imax0 + 1) - (lx + ly);
    for (int k = 0; k <= kmax0 - 1 - 2 * ly; ++k)
        for (int j = 0; j <= jmax0 - 1 - 2 * ly; ++j)
            for (int i = 0; i <= imax0 - 1 - ly - lx; ++i) cell(i + j * W + k * XY + base, W, XY);
    o.str("\n");
    o.str("CELL_TYPES ");
    o.inum(cells);
    o.str("\n");
    for (long long c = 0; c < cells; ++c) o.str("11\n");
    o.str("\n");
    o.str("POINT_DATA ");
    o.inum(P);
    o.str("\n");
    std::vector<double> fr, fu, fv, fw, cr, cu, cv, cw;
    level_macros(fd, fk, fcur, fprev, fpsi, fr, fu, fv, fw);
    level_macros(cd, ck, ccur, cprev, cpsi, cr, cu, cv, cw);
    auto idx = [&](int lv, int i, int j, int k) {
        const acp_case_desc& d = lv ? fd : cd;
        return (int64_t)i + (int64_t)d.nx * (j + (int64_t)d.ny * k);
    };
    o.str("SCALARS Omega float 1\n");
    o.str("LOOKUP_TABLE default\n");
    each([&](int lv, int, int, int) {
        o.num(lv ? fk.omega : ck.omega);
        o.str("\n");
    });
    o.str("SCALARS NodeID float 1\n");
    o.str("LOOKUP_TABLE default\n");
    each([&](int lv, int i, int j, int k) {
        int id = 0;
        if (lv) {
            if (lodi_node(fd, i, j, k)) id = 1;
            else {
                const bool in = i >= RX && i <= imax - R && j >= R && j <= jmax - R && k >= R && k <= kmax - R;
                const bool inner = i > RX && i < imax - R && j > R && j < jmax - R && k > R && k < kmax - R;
                if (in && !inner) id = 2;
            }
        }
        o.inum(id);
        o.str("\n");
    });
    const double cs2 = std::pow(cd.cs, 2.);
    o.str("SCALARS Pressure float 1\n");
    o.str("LOOKUP_TABLE default\n");
    each([&](int lv, int i, int j, int k) {
        o.num((lv ? fr : cr)[idx(lv, i, j, k)] * cs2);
        o.str("\n");
    });
    o.str("SCALARS Density float 1\n");
    o.str("LOOKUP_TABLE default\n");
    each([&](int lv, int i, int j, int k) {
        o.num((lv ? fr : cr)[idx(lv, i, j, k)]);
        o.str("\n");
    });
    o.str("\nVECTORS Velocity float \n");
    each([&](int lv, int i, int j, int k) {
        const int64_t l = idx(lv, i, j, k);
        o.num((lv ? fu : cu)[l]);
        o.str(" ");
        o.num((lv ? fv : cv)[l]);
        o.str(" ");
        o.num((lv ? fw : cw)[l]);
        o.str(" \n");
    });
    const std::string pc = std::to_string(P);
    o.s.replace(pnt_pos, pc.size(), pc);
    const std::string cc = std::to_string(cells) + " " + std::to_string(9 * cells);
    o.s.replace(cell_pos, cc.size(), cc);
    return write_file(o.s, stem);
}

int write_vtk_host(const acp_case_desc& d, const K64& k, const double* fcur, const double* fprev, const double* psi,
                   const char* stem) {
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int64_t N = (int64_t)nx * ny * nz;
    const double sp = d.spacing / std::pow(2., 0);
    Out o;
    o.s.reserve((size_t)N * 120 + 1024);
    o.str("# vtk DataFile Version 2.0\n");
    o.str("LBM results for grid refinement benchmark in channel flow\n");
    o.str("ASCII\n");
    o.str("DATASET UNSTRUCTURED_GRID\n");
    o.str("POINTS ");
    const size_t pnt_pos = o.s.size();
    o.str((std::string(14, ' ') + " float\n").c_str());
    for (int kz = 0; kz < nz; ++kz)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                o.num(i * sp);
                o.str(" ");
                o.num(j * sp);
                o.str(" ");
                o.num(kz * sp);
                o.str("\n");
            }
    o.str("\n");
    o.str("CELLS ");
    const size_t cell_pos = o.s.size();
    o.str((std::string(9, ' ') + " " + std::string(9, ' ') + "\n").c_str());
    long long cells = 0;
    const long long sx = nx, sxy = (long long)nx * ny;
    for (int kz = 0; kz < nz - 1; ++kz)
        for (int j = 0; j < ny - 1; ++j)
            for (int i = 0; i < nx - 1; ++i) {
                const long long n0 = i + j * sx + kz * sxy;
                const long long v[8] = {n0, n0 + 1, n0 + sx, n0 + sx + 1, n0 + sxy, n0 + sxy + 1, n0 + sxy + sx,
                                        n0 + sxy + sx + 1};
                o.str("8 ");
                for (int a = 0; a < 8; ++a) {
                    o.inum(v[a]);
                    o.str(a == 7 ? "\n" : " ");
                }
                ++cells;
            }
    o.str("\n");
    o.str("CELL_TYPES ");
    o.inum(cells);
    o.str("\n");
    for (long long c = 0; c < cells; ++c) o.str("11\n");
    o.str("\n");
    o.str("POINT_DATA ");
    o.inum(N);
    o.str("\n");

    // density and velocity as calcDensity / calcVelocity print them
    std::vector<double> rho, ux, uy, uz;
    level_macros(d, k, fcur, fprev, psi, rho, ux, uy, uz);
    const double cs2 = std::pow(d.cs, 2.);
    o.str("SCALARS Pressure float 1\n");
    o.str("LOOKUP_TABLE default\n");
    for (int64_t l = 0; l < N; ++l) {
        o.num(rho[l] * cs2);
        o.str("\n");
    }
    o.str("SCALARS Density float 1\n");
    o.str("LOOKUP_TABLE default\n");
    for (int64_t l = 0; l < N; ++l) {
        o.num(rho[l]);
        o.str("\n");
    }
    o.str("SCALARS Omega float 1\n");
    o.str("LOOKUP_TABLE default\n");
    for (int64_t l = 0; l < N; ++l) {
        o.num(k.omega);
        o.str("\n");
    }
    o.str("\nVECTORS Velocity float \n");
    for (int64_t l = 0; l < N; ++l) {
        o.num(ux[l]);
        o.str(" ");
        o.num(uy[l]);
        o.str(" ");
        o.num(uz[l]);
        o.str(" \n");
    }
    // overwrite the reserved blanks with the counts, as the reference's seekp +
    // write does (Lattice.cpp:4344-4347)
    const std::string pc = std::to_string(N);
    o.s.replace(pnt_pos, pc.size(), pc);
    const std::string cc = std::to_string(cells) + " " + std::to_string(9 * cells);
    o.s.replace(cell_pos, cc.size(), cc);
    return write_file(o.s, stem);
}

}  // namespace acp

extern "C" int acp_write_vtk_host(const acp_case_desc* d, const double* fcur, const double* fprev, const double* psi,
                                  const char* stem) {
    if (!d || !fcur || !stem) return acp::fail(ACP_E_ARG, "acp_write_vtk_host: null argument");
    if (d->nonreflecting && !fprev) return acp::fail(ACP_E_ARG, "Non-Reflecting output needs the previous array");
    acp::K64 k;
    acp::k64_init(k, *d);
    return acp::write_vtk_host(*d, k, fcur, fprev, psi, stem);
}

extern "C" int acp_write_vtk2_host(const acp_case_desc* coarse, const double* ccur, const double* cprev, int32_t x0,
                                   int32_t x1, const double* fcur, const double* fprev, const char* stem) {
    if (!coarse || !ccur || !fcur || !stem) return acp::fail(ACP_E_ARG, "acp_write_vtk2_host: null argument");
    if (coarse->nonreflecting && (!cprev || !fprev))
        return acp::fail(ACP_E_ARG, "Non-Reflecting output needs the previous arrays");
    acp_case_desc fd;
    if (int rc = acp_refined_fine_desc(coarse, x0, x1, &fd)) return rc;
    acp::K64 ck, fk;
    acp::k64_init(ck, *coarse);
    acp::k64_init(fk, fd);
    return acp::write_vtk2_host(*coarse, ck, ccur, cprev, fd, fk, fcur, fprev, x0, x1, stem);
}

extern "C" int acp_write_vtk_shell_host(const acp_case_desc* coarse, int32_t layer_yz, int32_t layer_x,
                                        const double* ccur, const double* cprev, const double* cpsi,
                                        const double* fcur, const double* fprev, const double* fpsi,
                                        const char* stem) {
    if (!coarse || !ccur || !fcur || !stem) return acp::fail(ACP_E_ARG, "acp_write_vtk_shell_host: null argument");
    if (coarse->nonreflecting && (!cprev || !fprev))
        return acp::fail(ACP_E_ARG, "Non-Reflecting output needs the previous arrays");
    if (layer_yz < 2 || layer_x < 2 || coarse->nx - 1 - layer_yz - layer_x < 2 || coarse->ny - 1 - 2 * layer_yz < 2 ||
        coarse->nz - 1 - 2 * layer_yz < 2)
        return acp::fail(ACP_E_ARG, "shell geometry: the refinement layers leave no coarse core");
    acp_case_desc fd = *coarse;
    fd.nx = 2 * coarse->nx - 1;
    fd.ny = 2 * coarse->ny - 1;
    fd.nz = 2 * coarse->nz - 1;
    fd.spacing = coarse->spacing / 2.;
    acp::K64 ck, fk;
    acp::k64_init(ck, *coarse);
    acp::k64_init(fk, fd);
    return acp::write_vtk_shell_host(*coarse, ck, ccur, cprev, cpsi, fd, fk, fcur, fprev, fpsi, layer_yz, layer_x,
                                     stem);
}
