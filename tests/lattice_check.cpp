// A program written against the reference's Control / Lattice / Node
// interface (Control.h, Lattice.h:102-120, Node.h:38-99), compiled against
// include/acp_lattice.hpp and linked to libacpulse.so.  It calls every public
// stepping member of the reference's Lattice and prints what the GPU suite
// compares with the engine's own entry points and with the oracle.
//
//   lattice_check <case file> uniform   createLattice, setConnectivity,
//        initialization, one step through the timeStepping member of the case
//        file's operator, then solve(flag) steps; prints populations of
//        probe nodes of both arrays and NodeView fields of selected nodes
//   lattice_check <case file> nested    the same for a MaxLevel 1 lattice:
//        one step through the nestedTimeStepping member of the operator,
//        then solve(flagL0, flagL1); prints both levels' populations
//   lattice_check <case file> mismatch  calls a stepping member of another
//        operator (the lattice must refuse it and exit with status 1)
//   lattice_check <case file> create    only constructs the lattice (a
//        geometry the engine does not build must be refused with a reason)
#include <cstdio>
#include <cstdlib>
#include <string>

#include "acp_lattice.hpp"

using namespace acp;

static void print_node(Lattice& lat, int level, int i, int j, int k) {
    NodeView n = lat.node(level, i, j, k);
    std::printf("node %d %d %d %d tag %s id %d dir %d corner %d x %.17g y %.17g z %.17g omega %.17g dt %.17g\n",
                level, i, j, k, n.getTag().c_str(), n.getNodeID(), n.getNormalDir(), (int)n.getIsCorner(), n.getX(),
                n.getY(), n.getZ(), n.getCollisionFrequency(), n.getTimeStep());
    std::printf("macro %d %d %d %d rho %.17g u %.17g %.17g %.17g\n", level, i, j, k, n.getNodeDensity(),
                n.getNodeVelocityComponent(0), n.getNodeVelocityComponent(1), n.getNodeVelocityComponent(2));
    std::printf("push %d %d %d %d", level, i, j, k);
    for (int q = 0; q < 18; ++q) {
        NodeView m = n.getNeighbor(q);
        if (m.valid()) std::printf(" %d:%d,%d,%d", q, m.getIndexI(), m.getIndexJ(), m.getIndexK());
        else std::printf(" %d:-", q);
    }
    std::printf("\n");
    NodeView a = n.getNormalNeighbor(), b = n.getNormalNeighborSnd();
    std::printf("lodi %d %d %d %d", level, i, j, k);
    for (NodeView m : {a, b, n.getTangentNeighbor1(0), n.getTangentNeighbor1(1), n.getTangentNeighbor2(0),
                       n.getTangentNeighbor2(1)}) {
        if (m.valid()) std::printf(" %d,%d,%d", m.getIndexI(), m.getIndexJ(), m.getIndexK());
        else std::printf(" -");
    }
    std::printf("\n");
    NodeView p = n.getPartner();
    if (p.valid()) std::printf("partner %d %d %d %d -> %d %d %d %d\n", level, i, j, k, p.getLevel(), p.getIndexI(),
                               p.getIndexJ(), p.getIndexK());
    std::printf("pop %d %d %d %d", level, i, j, k);
    for (int q = 0; q < 19; ++q) std::printf(" %.17g", n.getDistribution(true, q));
    std::printf("\n");
    std::printf("popprev %d %d %d %d", level, i, j, k);
    for (int q = 0; q < 19; ++q) std::printf(" %.17g", n.getDistributionPreCol(q));
    std::printf("\n");
    std::printf("psi %d %d %d %d", level, i, j, k);
    for (int q = 0; q < 19; ++q) std::printf(" %.17g", n.getCubicMachCorrection(q));
    std::printf("\n");
}

static void print_get(Lattice& lat, bool flag, int level, int i, int j, int k) {
    std::printf("get %d %d %d %d %d", (int)flag, level, i, j, k);
    for (int q = 0; q < 19; ++q) std::printf(" %.17g", lat.getDistribution(flag, q, level, i, j, k));
    std::printf("\n");
}

static void step_named(Lattice& lat, Control& bc, bool& alt) {
    int level = 0;
    const std::string m = bc.getCollisionModel();
    if (m == "SRT") lat.timeSteppingSRT(level, alt);
    else if (m == "MRT-GS") lat.timeSteppingMRT_GSbasis(level, alt);
    else if (m == "MRT-RM") lat.timeSteppingMRT_RMbasis(level, alt);
    else if (m == "RR") lat.timeSteppingRR(level, alt);
    else lat.timeSteppingHRR(level, alt);
}

static void nested_named(Lattice& lat, Control& bc, bool& alt0, bool& alt1) {
    const std::string m = bc.getCollisionModel();
    if (m == "SRT") lat.nestedTimeSteppingSRT(alt1, alt0, bc);
    else if (m == "MRT-GS") lat.nestedTimeSteppingMRT_GSbasis(alt1, alt0, bc);
    else if (m == "MRT-RM") lat.nestedTimeSteppingMRT_RMbasis(alt1, alt0, bc);
    else if (m == "RR") lat.nestedTimeSteppingRR(alt1, alt0, bc);
    else lat.nestedTimeSteppingHRR(alt1, alt0, bc);
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: lattice_check <case file> uniform|nested|mismatch|create\n");
        return 2;
    }
    const std::string mode = argv[2];
    Control bc(argv[1]);
    Lattice lat(bc);
    if (mode == "create") {
        std::printf("created %d levels\n", lat.levels());
        return 0;
    }
    lat.createLattice();
    lat.setConnectivity();
    lat.initialization(bc);
    const int steps = bc.getTimeStepMax();
    const int nx = bc.getNumberOfNodesX(), ny = bc.getNumberOfNodesY(), nz = bc.getNumberOfNodesZ();
    if (mode == "mismatch") {
        bool alt = true;
        int level = 0;
        if (bc.getCollisionModel() == std::string("RR")) lat.timeSteppingSRT(level, alt);
        else lat.timeSteppingRR(level, alt);
        std::printf("stepped\n");
        return 0;
    }
    bool alt = true;
    if (mode == "uniform") {
        step_named(lat, bc, alt);
        alt = !alt;
        for (int t = 1; t < steps; ++t) {
            lat.solve(alt);
            alt = !alt;
        }
        std::printf("steps %d levels %d\n", steps, lat.levels());
        const int pts[][3] = {{0, 0, 0}, {nx - 1, ny / 2, nz / 2}, {nx / 2, 0, nz / 2}, {nx / 2, ny / 2, 0},
                              {nx / 2, ny / 2, nz / 2}, {1, 1, 1}, {0, ny - 1, nz / 2}};
        for (auto& p : pts) print_node(lat, 0, p[0], p[1], p[2]);
        print_get(lat, alt, 0, nx / 2, ny / 2, nz / 2);
        print_get(lat, !alt, 0, nx / 2, ny / 2, nz / 2);
        return 0;
    }
    if (mode == "nested") {
        bool alt1 = true;
        nested_named(lat, bc, alt, alt1);
        alt = !alt;
        for (int t = 1; t < steps; ++t) {
            lat.solve(alt, alt1);
            alt = !alt;
        }
        std::printf("steps %d levels %d x0 %d x1 %d\n", steps, lat.levels(), lat.refinedX0(), lat.refinedX1());
        const int x0 = lat.refinedX0(), x1 = lat.refinedX1();
        const int pts0[][3] = {{0, ny / 2, nz / 2}, {x0, ny / 2, nz / 2}, {(x0 + x1) / 2, ny / 2, nz / 2}, {nx - 1, 1, 1}};
        for (auto& p : pts0) print_node(lat, 0, p[0], p[1], p[2]);
        const int fx = 2 * (x1 - x0);
        const int pts1[][3] = {{0, ny, nz}, {1, ny, nz}, {fx / 2 + 1, ny + 1, nz}, {fx, 3, 1}};
        for (auto& p : pts1) print_node(lat, 1, p[0], p[1], p[2]);
        print_get(lat, alt, 0, (x0 + x1) / 2, ny / 2, nz / 2);
        print_get(lat, alt1, 1, 2 * x0 + 3, ny, nz);
        if (x0 == 0 && x1 == 0) {
            // the shell geometry: nodes of the core box's surface (level 0) and of
            // the hole's faces, edges and corners (level 1)
            const int lx = bc.getRefinementLayerX(), ly = bc.getRefinementLayer();
            const int c0[][3] = {{lx - 1, ny / 2, nz / 2}, {lx - 1, ly - 1, ly - 1}, {nx - ly, ny / 2, nz - ly},
                                 {lx, ny / 2, nz / 2}, {nx / 2, ny / 2, nz / 2}};
            for (auto& p : c0) print_node(lat, 0, p[0], p[1], p[2]);
            const int f1[][3] = {{2 * lx, ny - 1, nz - 1}, {2 * lx, 2 * ly, 2 * ly}, {2 * lx, ny, nz + 1},
                                 {2 * lx + 1, 2 * ly, nz}, {2 * (nx - 1 - ly), 2 * ly + 1, 2 * ly + 1},
                                 {2 * lx - 2, ny - 1, nz - 1}, {1, 1, 1}};
            for (auto& p : f1) print_node(lat, 1, p[0], p[1], p[2]);
        }
        return 0;
    }
    std::fprintf(stderr, "unknown mode %s\n", mode.c_str());
    return 2;
}
