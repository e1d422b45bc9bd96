// Test harness around the reference solver's own uniform-grid code.
//
// This file is test infrastructure: it is linked against objects compiled
// from the unmodified reference sources (Control.cpp, Lattice.cpp, Node.cpp)
// and produces known-answer data for the tests.  It never ships and never
// runs on the GPU host; only the data it writes (tests/golden/) does.
//
// Modes
//   run  <bc.txt> <nsteps> <outdir> [--ic <file>] [--probe i,j,k]...
//        Lattice(Control&) -> initialization() -> nsteps x solve(alt), the
//        loop of main.cpp:109-157 without the VTK output.  --ic replaces the
//        initial populations of every array (as initialization() does with
//        its hard-coded pulse, Lattice.cpp:1457-1475).  Writes
//          state.bin   current populations, f64 [q][k][j][i]
//          probe.bin   rho = sum_q f_q at each probe node after every step
//          time.txt    wall seconds spent inside solve()
//   geom <bc.txt> <outdir>
//        Dumps the connectivity built by Lattice::setConnectivity
//        (Lattice.cpp:237-1394) through the public Node getters.
//   kat  <bc.txt> <outdir> <nsamples> <seed>
//        Per-function known-answer vectors from standalone FluidNode objects.
//   partner <bc.txt> <outdir> <nsamples> <seed>
//        Known answers of the strain-rate partner functions of the two-level
//        step (Node.cpp:1892-2490) on standalone node clusters.
//   vtk  <bc.txt> <nsteps> <outfile-stem> [--ic <file>]
//        Runs nsteps and writes the reference's own VTK file.
//
// The reference reads never-assigned members of heap objects (neighbour
// pointers of boundary nodes, isCorner_) and so relies on zero-filled
// memory; the global operator new below makes that deterministic.

#include <cstdio>
#include <cstdlib>
#include <cstdint>
#include <cstring>
#include <cmath>
#include <new>
#include <memory>
#include <chrono>
#include <iostream>
#include <fstream>
#include <sstream>
#include <iomanip>
#include <string>
#include <vector>
#include <map>
#include <unordered_map>
#include <algorithm>
#include <array>
#include <omp.h>

// Reach Lattice::mesh_ for IC injection; access specifiers do not change the
// object layout, and every standard header is already included above.
#define private public
#define protected public
#include "Lattice.h"
#undef private
#undef protected

void* operator new(std::size_t n) {
    void* p = std::calloc(1, n ? n : 1);
    if (!p) throw std::bad_alloc();
    return p;
}
void* operator new[](std::size_t n) { return operator new(n); }
void operator delete(void* p) noexcept { std::free(p); }
void operator delete[](void* p) noexcept { std::free(p); }
void operator delete(void* p, std::size_t) noexcept { std::free(p); }
void operator delete[](void* p, std::size_t) noexcept { std::free(p); }

static void die(const std::string& m) { std::cerr << "ref_driver: " << m << "\n"; std::exit(2); }

static void write_f64(const std::string& path, const std::vector<double>& v) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) die("cannot write " + path);
    if (!v.empty() && std::fwrite(v.data(), sizeof(double), v.size(), fp) != v.size()) die("short write " + path);
    std::fclose(fp);
}
static void write_i64(const std::string& path, const std::vector<int64_t>& v) {
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) die("cannot write " + path);
    if (!v.empty() && std::fwrite(v.data(), sizeof(int64_t), v.size(), fp) != v.size()) die("short write " + path);
    std::fclose(fp);
}
static std::vector<double> read_f64(const std::string& path, size_t n) {
    std::vector<double> v(n);
    FILE* fp = std::fopen(path.c_str(), "rb");
    if (!fp) die("cannot read " + path);
    if (std::fread(v.data(), sizeof(double), n, fp) != n) die("short read " + path);
    std::fclose(fp);
    return v;
}

struct Dims { int nx, ny, nz; };
static Dims dims_of(Control& bc) { return {bc.getNumberOfNodesX(), bc.getNumberOfNodesY(), bc.getNumberOfNodesZ()}; }

static void inject_ic(Lattice& lat, const Dims& d, const std::vector<double>& f) {
    const size_t N = (size_t)d.nx * d.ny * d.nz;
    for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i) {
                Node* n = lat.mesh_[0][i][j][k];
                const size_t l = (size_t)i + (size_t)d.nx * ((size_t)j + (size_t)d.ny * k);
                for (int q = 0; q < 19; ++q) {
                    double v = f[(size_t)q * N + l];
                    n->setDistribution(true, q, v);
                    n->setDistribution(false, q, v);
                    n->setDistributionPreCol(q, v);
                    n->setDistributionFictitious(q, v);
                }
            }
}

static std::vector<double> dump_state(Lattice& lat, const Dims& d, bool arr) {
    const size_t N = (size_t)d.nx * d.ny * d.nz;
    std::vector<double> f((size_t)19 * N);
    int level = 0;
    for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i)
                for (int q = 0; q < 19; ++q) {
                    int ii = i, jj = j, kk = k, qq = q;
                    f[(size_t)q * N + (size_t)i + (size_t)d.nx * ((size_t)j + (size_t)d.ny * k)] =
                        lat.getDistribution(arr, qq, level, ii, jj, kk);
                }
    return f;
}

static int mode_run(int argc, char** argv, bool vtk) {
    if (argc < 5) die("usage: run <bc.txt> <nsteps> <outdir|stem> [--ic f] [--probe i,j,k]");
    std::string bcfile = argv[2];
    int nsteps = std::atoi(argv[3]);
    std::string out = argv[4];
    std::string icfile;
    std::vector<std::array<int, 3>> probes;
    for (int a = 5; a < argc; ++a) {
        std::string s = argv[a];
        if (s == "--ic" && a + 1 < argc) icfile = argv[++a];
        else if (s == "--probe" && a + 1 < argc) {
            std::array<int, 3> p{};
            if (std::sscanf(argv[++a], "%d,%d,%d", &p[0], &p[1], &p[2]) != 3) die("bad --probe");
            probes.push_back(p);
        } else die("unknown option " + s);
    }
    Control bc(bcfile);
    if (bc.getLevelMax() != 0) die("only MaxLevel 0 is runnable (GridCoupling.cpp is absent)");
    Dims d = dims_of(bc);
    Lattice lat(bc);
    lat.initialization(bc);
    if (!icfile.empty()) inject_ic(lat, d, read_f64(icfile, (size_t)19 * d.nx * d.ny * d.nz));

    bool alt = true;  // main.cpp:55
    bool alt1 = true;
    std::vector<double> probe;
    double secs = 0.0;
    int level = 0;
    for (int t = 0; t < nsteps; ++t) {
        auto t0 = std::chrono::steady_clock::now();
        lat.solve(alt);
        auto t1 = std::chrono::steady_clock::now();
        secs += std::chrono::duration<double>(t1 - t0).count();
        bool cur = !alt;
        for (auto& p : probes) {
            double r = 0.0;
            for (int q = 0; q < 19; ++q) {
                int qq = q, ii = p[0], jj = p[1], kk = p[2];
                r += lat.getDistribution(cur, qq, level, ii, jj, kk);
            }
            probe.push_back(r);
        }
        alt = !alt;  // main.cpp:155
    }
    if (vtk) {
        bool cur = alt;  // the array the last step wrote
        lat.writeResultsVTK(out, cur, alt1);
        return 0;
    }
    write_f64(out + "/state.bin", dump_state(lat, d, alt));
    write_f64(out + "/probe.bin", probe);
    std::ofstream tf(out + "/time.txt");
    tf << std::setprecision(9) << secs << " " << d.nx << " " << d.ny << " " << d.nz << " " << nsteps << " "
       << bc.getNumberOfThreads() << "\n";
    return 0;
}

// Connectivity table, one row of 27 int64 per node (linear index l = i + nx*(j + ny*k)):
//   [0..17]  neighbors_[q] target index or -1
//   [18]     normalNeighbor_  [19] normalNeighborSnd_
//   [20,21]  tangentNeighbor1_[0,1]   [22,23] tangentNeighbor2_[0,1]
//   [24]     normalDir_  [25] isCorner_  [26] 1 for boundary node (ID_ == 1)
static int mode_geom(int argc, char** argv) {
    if (argc < 4) die("usage: geom <bc.txt> <outdir>");
    Control bc(argv[2]);
    Dims d = dims_of(bc);
    Lattice lat(bc);
    std::unordered_map<Node*, int64_t> idx;
    for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i)
                idx[lat.mesh_[0][i][j][k]] = (int64_t)i + (int64_t)d.nx * ((int64_t)j + (int64_t)d.ny * k);
    auto id = [&](Node* p) -> int64_t {
        if (!p) return -1;
        auto it = idx.find(p);
        if (it == idx.end()) die("dangling pointer");
        return it->second;
    };
    std::vector<int64_t> rows;
    for (int k = 0; k < d.nz; ++k)
        for (int j = 0; j < d.ny; ++j)
            for (int i = 0; i < d.nx; ++i) {
                Node* n = lat.mesh_[0][i][j][k];
                for (int q = 0; q < 18; ++q) rows.push_back(id(n->neighbors_[q]));
                rows.push_back(id(n->normalNeighbor_));
                rows.push_back(id(n->normalNeighborSnd_));
                rows.push_back(id(n->tangentNeighbor1_[0]));
                rows.push_back(id(n->tangentNeighbor1_[1]));
                rows.push_back(id(n->tangentNeighbor2_[0]));
                rows.push_back(id(n->tangentNeighbor2_[1]));
                rows.push_back(n->normalDir_);
                rows.push_back(n->isCorner_ ? 1 : 0);
                rows.push_back(n->getNodeID());
            }
    write_i64(std::string(argv[3]) + "/geom.bin", rows);
    return 0;
}

// splitmix64 -> uniform in [-1, 1)
struct Rng {
    uint64_t s;
    double next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return (double)(z >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    }
};

static int mode_kat(int argc, char** argv) {
    if (argc < 6) die("usage: kat <bc.txt> <outdir> <nsamples> <seed>");
    Control bc(argv[2]);
    std::string out = argv[3];
    int ns = std::atoi(argv[4]);
    Rng rng{(uint64_t)std::strtoull(argv[5], nullptr, 10)};
    const double rho0 = bc.getDensity(), cs = bc.getSpeedOfSound();
    double dt = bc.getTimeStep();
    double om = bc.getCollisionFrequency();
    double* w = bc.getWeightFactors();

    std::vector<double> in_rho, in_u, in_f, in_nb;  // inputs
    std::vector<double> o_feq[3], o_pi1, o_rr[3], o_srt, o_rrc, o_hrr, o_strain, o_psi;
    std::vector<double> o_rm_mom, o_rm_eq, o_rm_inv, o_gs_mom, o_gs_eqm, o_gs_inv, o_hrrc;
    const int order = bc.getOrderOfEquilibrium();

    for (int s = 0; s < ns; ++s) {
        double rho = rho0 * (1.0 + 0.01 * rng.next());
        double u[3] = {0.05 * cs * rng.next(), 0.05 * cs * rng.next(), 0.05 * cs * rng.next()};
        in_rho.push_back(rho);
        for (double x : u) in_u.push_back(x);

        FluidNode n(0, 0, 0, 1, 1, 1, 0);
        n.setTimeStep(dt);
        n.setCollisionFrequency(om);
        double feq[3][19];
        n.calcEquilibrium2_GHbasis(feq[0], bc, rho, u);
        n.calcEquilibrium3_GHbasis(feq[1], bc, rho, u);
        n.calcEquilibrium4_GHbasis(feq[2], bc, rho, u);
        for (int o = 0; o < 3; ++o)
            for (int q = 0; q < 19; ++q) o_feq[o].push_back(feq[o][q]);

        // populations: order-3 equilibrium plus a non-equilibrium perturbation
        double f[19];
        for (int q = 0; q < 19; ++q) {
            f[q] = feq[1][q] + rho0 * w[q] * 0.01 * rng.next();
            in_f.push_back(f[q]);
        }
        bool alt = true;
        for (int q = 0; q < 19; ++q) { int qq = q; n.setDistribution(true, qq, f[q]); n.setDistribution(false, qq, f[q]); }
        for (int q = 0; q < 19; ++q) { int qq = q; n.setCubicMachCorrection(qq, 0.0); }

        // Pi1 and RR off-equilibria from the order-`order` equilibrium of f's own moments
        double dens = n.calcDensity(bc, alt);
        double velo[3];
        n.calcVelocity(bc, alt, dens, velo);
        double eq[19];
        if (order == 2) n.calcEquilibrium2_GHbasis(eq, bc, dens, velo);
        else if (order == 3) n.calcEquilibrium3_GHbasis(eq, bc, dens, velo);
        else n.calcEquilibrium4_GHbasis(eq, bc, dens, velo);
        n.calcPi1Tensor(alt, eq, bc, velo);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) o_pi1.push_back(n.getPi1Tensor(a, b));
        double rr[19];
        n.calcRRoffEquilibrium2_GHbasis(alt, rr, eq, bc, velo);
        for (double x : rr) o_rr[0].push_back(x);
        n.calcRRoffEquilibrium3_GHbasis(alt, rr, eq, bc, velo);
        for (double x : rr) o_rr[1].push_back(x);
        n.calcRRoffEquilibrium4_GHbasis(alt, rr, eq, bc, velo);
        for (double x : rr) o_rr[2].push_back(x);

        // collideSRT / collideRR in place
        n.collideSRT(bc, alt);
        for (int q = 0; q < 19; ++q) { int qq = q; o_srt.push_back(n.getDistribution(alt, qq)); }
        for (int q = 0; q < 19; ++q) { int qq = q; n.setDistribution(true, qq, f[q]); }
        n.collideRR(bc, alt);
        for (int q = 0; q < 19; ++q) { int qq = q; o_rrc.push_back(n.getDistribution(alt, qq)); }

        // collideHRR with six axis neighbours holding random pre-collision populations
        FluidNode nb[6] = {FluidNode(0, 0, 0, 2, 1, 1, 0), FluidNode(0, 0, 0, 1, 2, 1, 0), FluidNode(0, 0, 0, 0, 1, 1, 0),
                           FluidNode(0, 0, 0, 1, 0, 1, 0), FluidNode(0, 0, 0, 1, 1, 2, 0), FluidNode(0, 0, 0, 1, 1, 0, 0)};
        for (int m = 0; m < 6; ++m) {
            double rn = rho0 * (1.0 + 0.01 * rng.next());
            double un[3] = {0.05 * cs * rng.next(), 0.05 * cs * rng.next(), 0.05 * cs * rng.next()};
            double fe[19];
            n.calcEquilibrium3_GHbasis(fe, bc, rn, un);
            for (int q = 0; q < 19; ++q) {
                double v = fe[q] + rho0 * w[q] * 0.01 * rng.next();
                in_nb.push_back(v);
                int qq = q;
                nb[m].setDistributionPreCol(qq, v);
                nb[m].setCubicMachCorrectionPreCol(qq, 0.0);
            }
            int dm = m;
            Node* p = &nb[m];
            n.setNeighbor(dm, p);
        }
        for (int q = 0; q < 19; ++q) { int qq = q; n.setDistribution(true, qq, f[q]); n.setCubicMachCorrection(qq, 0.0); }
        n.collideHRR(bc, alt);
        for (int q = 0; q < 19; ++q) { int qq = q; o_hrr.push_back(n.getDistribution(alt, qq)); }
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) o_strain.push_back(n.getStrainRateTensorFD(a, b));
        for (int q = 0; q < 19; ++q) { int qq = q; o_psi.push_back(n.getCubicMachCorrection(qq)); }

        // MRT building blocks (their composition is discarded by the reference collide)
        double mom[19], feqrm[19], inv[19];
        n.calcAllMoments_RMbasis(bc, mom, f);
        for (double x : mom) o_rm_mom.push_back(x);
        Vector3D vv(u[0], u[1], u[2]);
        n.calcEquilibrium4_RMbasis(feqrm, rho, &vv, bc);
        for (double x : feqrm) o_rm_eq.push_back(x);
        n.transformMomentsToDistributions_RMbasis(inv, mom, bc);
        for (double x : inv) o_rm_inv.push_back(x);
        for (int q = 0; q < 19; ++q) { int qq = q; n.setDistribution(true, qq, f[q]); }
        n.calcAllMoments_GSbasis(bc, mom, alt);
        for (double x : mom) o_gs_mom.push_back(x);
        double eqm[19] = {0};
        n.calcEquilibriumMoments2_GSbasis(eqm, mom[0], mom[3], mom[5], mom[7]);
        for (double x : eqm) o_gs_eqm.push_back(x);
        n.transformMomentsToDistributions_GSbasis(inv, mom, bc);
        for (double x : inv) o_gs_inv.push_back(x);
    }
    write_f64(out + "/in_rho.bin", in_rho);
    write_f64(out + "/in_u.bin", in_u);
    write_f64(out + "/in_f.bin", in_f);
    write_f64(out + "/in_nb.bin", in_nb);
    for (int o = 0; o < 3; ++o) write_f64(out + "/feq" + std::to_string(o + 2) + ".bin", o_feq[o]);
    write_f64(out + "/pi1.bin", o_pi1);
    for (int o = 0; o < 3; ++o) write_f64(out + "/rroffeq" + std::to_string(o + 2) + ".bin", o_rr[o]);
    write_f64(out + "/collide_srt.bin", o_srt);
    write_f64(out + "/collide_rr.bin", o_rrc);
    write_f64(out + "/collide_hrr.bin", o_hrr);
    write_f64(out + "/hrr_strain.bin", o_strain);
    write_f64(out + "/hrr_psi.bin", o_psi);
    write_f64(out + "/rm_moments.bin", o_rm_mom);
    write_f64(out + "/rm_feq4.bin", o_rm_eq);
    write_f64(out + "/rm_inverse.bin", o_rm_inv);
    write_f64(out + "/gs_moments.bin", o_gs_mom);
    write_f64(out + "/gs_eqmoments.bin", o_gs_eqm);
    write_f64(out + "/gs_inverse.bin", o_gs_inv);
    return 0;
}

// Known answers of the strain-rate partner functions of the two-level step
// (Node.cpp:1892-2490) on standalone node clusters (heap nodes, whose unset
// neighbour pointers the zero-filling operator new leaves NULL): a centre node, its six
// axis neighbours, their partners on the other level and, for the direct
// coupling, the second neighbours along each axis.  Every node holds random
// populations in both arrays and in its pre-collision copy, so that the
// array each function reads is pinned.  The level-0 flag is true and the
// level-1 flag false.
static int mode_partner(int argc, char** argv) {
    if (argc < 6) die("usage: partner <bc.txt> <outdir> <nsamples> <seed>");
    Control bc(argv[2]);
    std::string out = argv[3];
    int ns = std::atoi(argv[4]);
    Rng rng{(uint64_t)std::strtoull(argv[5], nullptr, 10)};
    const double rho0 = bc.getDensity(), cs = bc.getSpeedOfSound();
    double* w = bc.getWeightFactors();
    bool alt0 = true, alt1 = false;
    std::vector<double> i_now, i_other, i_pre, i_fine, i_fine_other, i_mask;
    std::vector<double> o_full_s, o_full_psi, o_half_s, o_half_psi, o_fine_s, o_fine_psi;
    std::vector<double> o_dc0_s, o_dc0_psi, o_dc0p_s, o_dc0p_psi, o_dc1_s, o_dc1_psi;
    auto pops = [&](double* f) {
        double rn = rho0 * (1.0 + 0.01 * rng.next());
        double un[3] = {0.05 * cs * rng.next(), 0.05 * cs * rng.next(), 0.05 * cs * rng.next()};
        FluidNode t(0, 0, 0, 0, 0, 0, 0);
        t.calcEquilibrium3_GHbasis(f, bc, rn, un);
        for (int q = 0; q < 19; ++q) f[q] += rho0 * w[q] * 0.01 * rng.next();
    };
    auto put = [&](Node& n, bool which, const double* f) {
        for (int q = 0; q < 19; ++q) { int qq = q; n.setDistribution(which, qq, f[q]); }
    };
    auto put_pre = [&](Node& n, const double* f) {
        for (int q = 0; q < 19; ++q) { int qq = q; double v = f[q]; n.setDistributionPreCol(qq, v); }
    };
    auto strain = [&](Node& n, int which, std::vector<double>& s) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                s.push_back(which == 0 ? n.getStrainRateTensorFD(a, b)
                                       : (which == 1 ? n.getStrainRateTensorFDhalf(a, b) : n.getStrainRateTensorFDfull(a, b)));
    };
    auto psi = [&](Node& n, int which, std::vector<double>& s) {
        for (int q = 0; q < 19; ++q) {
            int qq = q;
            s.push_back(which == 0 ? n.getCubicMachCorrection(qq)
                                   : (which == 1 ? n.getCubicMachCorrectionHalf(qq) : n.getCubicMachCorrectionFull(qq)));
        }
    };
    for (int s = 0; s < ns; ++s) {
        // per axis neighbour m: its populations in the level-0 array (now), the
        // other array, its pre-collision copy, and the populations of the node
        // on the other level it leads to (level-1 array, and the other array)
        double now[6][19], other[6][19], pre[6][19], fine[6][19], fine_other[6][19];
        int mask[6];
        for (int m = 0; m < 6; ++m) {
            pops(now[m]); pops(other[m]); pops(pre[m]); pops(fine[m]); pops(fine_other[m]);
            mask[m] = rng.next() > 0.0 ? 1 : 0;
            for (int q = 0; q < 19; ++q) {
                i_now.push_back(now[m][q]); i_other.push_back(other[m][q]); i_pre.push_back(pre[m][q]);
                i_fine.push_back(fine[m][q]); i_fine_other.push_back(fine_other[m][q]);
            }
            i_mask.push_back(mask[m]);
        }
        // coarse partner: neighbours on level 0, grid-coupling ones (mask) lead
        // to their fine partner (Node.cpp:2114-2240, 2244-2380)
        {
            std::unique_ptr<FluidNode> cu(new FluidNode(0, 0, 0, 5, 5, 5, 0));
            FluidNode& c = *cu;
            std::vector<FluidNode> nb, fp;
            nb.reserve(6); fp.reserve(6);
            for (int m = 0; m < 6; ++m) { nb.emplace_back(0, 0, 0, 5, 5, 5, 0); fp.emplace_back(0, 0, 0, 10, 10, 10, 1); }
            for (int m = 0; m < 6; ++m) {
                put(nb[m], alt0, now[m]); put(nb[m], !alt0, other[m]); put_pre(nb[m], pre[m]);
                put(fp[m], alt1, fine[m]); put(fp[m], !alt1, fine_other[m]);
                if (mask[m]) { nb[m].tag_ = "gridcoupling"; Node* pp = &fp[m]; nb[m].setPartner(pp); }
                int dm = m; Node* np = &nb[m]; c.setNeighbor(dm, np);
            }
            c.calcStrainRateTensorFDandCubicMachCorrectionCoarsePartners(alt0, alt1, bc);
            strain(c, 2, o_full_s); psi(c, 2, o_full_psi);
            c.calcStrainRateTensorFDandCubicMachCorrectionHalfCoarsePartners(alt0, alt1, bc);
            strain(c, 1, o_half_s); psi(c, 1, o_half_psi);
        }
        // fine partner: a level-1 node and its six fine neighbours (Node.cpp:2383-2490)
        {
            std::unique_ptr<FluidNode> cu(new FluidNode(0, 0, 0, 10, 10, 10, 1));
            FluidNode& c = *cu;
            std::vector<FluidNode> nb;
            nb.reserve(6);
            for (int m = 0; m < 6; ++m) nb.emplace_back(0, 0, 0, 10, 10, 10, 1);
            for (int m = 0; m < 6; ++m) {
                put(nb[m], alt1, fine[m]); put(nb[m], !alt1, fine_other[m]);
                int dm = m; Node* np = &nb[m]; c.setNeighbor(dm, np);
            }
            c.calcStrainRateTensorFDandCubicMachCorrectionFinePartners(alt1, bc);
            strain(c, 0, o_fine_s); psi(c, 0, o_fine_psi);
        }
        // direct coupling, level 0: a missing neighbour (mask) is replaced by
        // the partner's second fine neighbour along the axis (Node.cpp:1929-1953)
        {
            std::unique_ptr<FluidNode> cu(new FluidNode(0, 0, 0, 5, 5, 5, 0)), pu(new FluidNode(0, 0, 0, 10, 10, 10, 1));
            FluidNode &c = *cu, &p = *pu;
            std::vector<FluidNode> nb, a1, a2;
            nb.reserve(6); a1.reserve(6); a2.reserve(6);
            for (int m = 0; m < 6; ++m) {
                nb.emplace_back(0, 0, 0, 5, 5, 5, 0); a1.emplace_back(0, 0, 0, 10, 10, 10, 1); a2.emplace_back(0, 0, 0, 10, 10, 10, 1);
            }
            Node* pp = &p;
            c.setPartner(pp);
            for (int m = 0; m < 6; ++m) {
                int dm = m;
                put(nb[m], alt0, now[m]); put(nb[m], !alt0, other[m]); put_pre(nb[m], pre[m]);
                put(a1[m], alt1, other[m]); put(a1[m], !alt1, pre[m]);
                put(a2[m], alt1, fine[m]); put(a2[m], !alt1, fine_other[m]);
                Node* x1 = &a1[m]; p.setNeighbor(dm, x1);
                Node* x2 = &a2[m]; a1[m].setNeighbor(dm, x2);
                if (!mask[m]) { Node* np = &nb[m]; c.setNeighbor(dm, np); }
            }
            c.calcStrainRateTensorFDandCubicMachCorrectionDirectCoupling(alt0, alt1, bc);
            strain(c, 0, o_dc0_s); psi(c, 0, o_dc0_psi);
            strain(p, 0, o_dc0p_s); psi(p, 0, o_dc0p_psi);
        }
        // direct coupling, level 1: a missing neighbour (mask) is replaced by the
        // coarse partner's neighbour, averaged over its pre-collision and level-0
        // states; an existing one by the second fine neighbour (Node.cpp:1976-2033)
        {
            std::unique_ptr<FluidNode> cu(new FluidNode(0, 0, 0, 10, 10, 10, 1)), qu(new FluidNode(0, 0, 0, 5, 5, 5, 0));
            FluidNode &c = *cu, &q = *qu;
            std::vector<FluidNode> cn, a1, a2;
            cn.reserve(6); a1.reserve(6); a2.reserve(6);
            for (int m = 0; m < 6; ++m) {
                cn.emplace_back(0, 0, 0, 5, 5, 5, 0); a1.emplace_back(0, 0, 0, 10, 10, 10, 1); a2.emplace_back(0, 0, 0, 10, 10, 10, 1);
            }
            Node* qp = &q;
            c.setPartner(qp);
            for (int m = 0; m < 6; ++m) {
                int dm = m;
                put(cn[m], alt0, now[m]); put(cn[m], !alt0, other[m]); put_pre(cn[m], pre[m]);
                Node* cp = &cn[m]; q.setNeighbor(dm, cp);
                put(a1[m], alt1, other[m]); put(a1[m], !alt1, pre[m]);
                put(a2[m], alt1, fine[m]); put(a2[m], !alt1, fine_other[m]);
                Node* x2 = &a2[m]; a1[m].setNeighbor(dm, x2);
                if (!mask[m]) { Node* x1 = &a1[m]; c.setNeighbor(dm, x1); }
            }
            c.calcStrainRateTensorFDandCubicMachCorrectionDirectCoupling(alt0, alt1, bc);
            strain(c, 0, o_dc1_s); psi(c, 0, o_dc1_psi);
        }
    }
    write_f64(out + "/in_now.bin", i_now);
    write_f64(out + "/in_other.bin", i_other);
    write_f64(out + "/in_pre.bin", i_pre);
    write_f64(out + "/in_fine.bin", i_fine);
    write_f64(out + "/in_fine_other.bin", i_fine_other);
    write_f64(out + "/in_mask.bin", i_mask);
    write_f64(out + "/full_s.bin", o_full_s); write_f64(out + "/full_psi.bin", o_full_psi);
    write_f64(out + "/half_s.bin", o_half_s); write_f64(out + "/half_psi.bin", o_half_psi);
    write_f64(out + "/fine_s.bin", o_fine_s); write_f64(out + "/fine_psi.bin", o_fine_psi);
    write_f64(out + "/dc0_s.bin", o_dc0_s); write_f64(out + "/dc0_psi.bin", o_dc0_psi);
    write_f64(out + "/dc0p_s.bin", o_dc0p_s); write_f64(out + "/dc0p_psi.bin", o_dc0p_psi);
    write_f64(out + "/dc1_s.bin", o_dc1_s); write_f64(out + "/dc1_psi.bin", o_dc1_psi);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) die("usage: ref_driver run|geom|kat|partner|vtk ...");
    std::string m = argv[1];
    if (m == "run") return mode_run(argc, argv, false);
    if (m == "vtk") return mode_run(argc, argv, true);
    if (m == "geom") return mode_geom(argc, argv);
    if (m == "kat") return mode_kat(argc, argv);
    if (m == "partner") return mode_partner(argc, argv);
    die("unknown mode " + m);
    return 2;
}
