// Single-level VTK output with the byte layout of Lattice::writeResultsVTK
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
    std::vector<double> rho((size_t)N), ux((size_t)N), uy((size_t)N), uz((size_t)N);
    for (int kz = 0; kz < nz; ++kz)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const int64_t l = i + (int64_t)nx * (j + (int64_t)ny * kz);
                bool bnd;
                if (d.boundary == ACP_BC_LODI)
                    bnd = i == 0 || j == 0 || kz == 0 || i == nx - 1 || j == ny - 1 || kz == nz - 1;
                else if (d.boundary == ACP_BC_LODI_X)
                    bnd = i == 0 || i == nx - 1;
                else
                    bnd = d.boundary == ACP_BC_OPEN_LODI_X && i == nx - 1;
                const double* F = (d.nonreflecting && i > 0 && bnd) ? fprev : fcur;
                double f[19], ps[19];
                for (int q = 0; q < 19; ++q) f[q] = F[(size_t)q * N + l];
                if (d.cubic) {
                    for (int q = 0; q < 19; ++q) ps[q] = psi ? psi[(size_t)q * N + l] : 0.;
                    moments64_psi(k, f, ps, rho[l], ux[l], uy[l], uz[l]);
                } else {
                    moments64(k, f, rho[l], ux[l], uy[l], uz[l]);
                }
            }
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
    std::string fname = std::string(stem) + ".vtk";
    std::ofstream fh(fname, std::ios::binary);
    if (!fh) return fail(ACP_E_IO, "An error occured while writing paraview file " + fname);
    fh.write(o.s.data(), (std::streamsize)o.s.size());
    if (!fh) return fail(ACP_E_IO, "short write " + fname);
    return ACP_OK;
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
