// acp_run -- command-line driver with the behaviour of the reference's main()
// (main.cpp:37-271): reads the case file, prints the same parameter summary
// (the two-level one for MaxLevel 1), writes <Save-as>_<t>.vtk at t = 0 and
// every WriteInterval steps, and prints the same progress and time estimates.
// A MaxLevel 1 case file takes this engine's refined geometry (Boundary
// LODI-X and Refined-Region x0 x1, acp_lattice.hpp).  The case file path
// is the first argument (the reference hard-codes it, main.cpp:44); with no
// argument ./BoundaryConditions.txt is read.  Written against the Control /
// Lattice interface of acp_lattice.hpp, as the reference's main() is against
// its own classes.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include "acp_lattice.hpp"

using namespace std::chrono;
using acp::Control;
using acp::Lattice;

static std::ostream& display(std::ostream& os, nanoseconds ns) {
    typedef duration<int, std::ratio<86400>> days;
    const char fill = os.fill();
    os.fill('0');
    auto d = duration_cast<days>(ns);
    ns -= d;
    auto h = duration_cast<hours>(ns);
    ns -= h;
    auto m = duration_cast<minutes>(ns);
    ns -= m;
    auto s = duration_cast<seconds>(ns);
    os << std::setw(2) << d.count() << "d:" << std::setw(2) << h.count() << "h:" << std::setw(2) << m.count()
       << "m:" << std::setw(2) << s.count() << 's' << std::endl
       << std::endl;
    os.fill(fill);
    return os;
}

// main.cpp:158-263: the two-level summary and loop, solve(flagL0, flagL1)
static int runTwoLevel(Control& bc, Lattice& lat, steady_clock::time_point beginTotal) {
    int time = 0;
    const int initTime = 0;
    const int writeInterval = bc.getWriteInterval();
    const int maxTimeStep = bc.getTimeStepMax();
    bc.setTime_(static_cast<double>(time));
    bool alternatingL0 = true, negAlternatingL0 = false, alternatingL1 = true;
    std::cout << "all parameters are being displayed in SI units" << std::endl << std::endl;
    std::cout << "result file(s): " << bc.getCaseName() << "_t.vtk" << std::endl << std::endl;
    std::cout << "dimensionless collision frequency on coarse grid = " << bc.getCollisionFrequency() << std::endl;
    std::cout << "  dimensionless collision frequency on fine grid = " << 1. / ((2. / bc.getCollisionFrequency()) - 0.5)
              << std::endl;
    std::cout << "                   time step size on coarse grid = " << bc.getTimeStep() << std::endl;
    std::cout << "                     time step size on fine grid = " << bc.getTimeStep() / 2. << std::endl;
    std::cout << "                          spacing on coarse grid = " << bc.getSpacing() << std::endl;
    std::cout << "                            spacing on fine grid = " << bc.getSpacing() / 2. << std::endl;
    std::cout << "                                 REYNOLDS number = " << bc.getReynoldsNumber() << std::endl;
    std::cout << "                             kinematic viscosity = " << bc.getKinematicViscosity() << std::endl;
    std::cout << "                                         density = " << bc.getDensity() << std::endl;
    std::cout << "                                  speed of sound = " << bc.getSpeedOfSound() << std::endl;
    std::cout << "                                     MACH number = " << bc.getMachNumber() << std::endl;
    std::cout << "                            order of equilibrium = " << bc.getOrderOfEquilibrium() << std::endl;
    std::cout << "                                 collision model = " << bc.getCollisionModel() << std::endl;
    if (bc.getCollisionModel() == "HRR")
        std::cout << "                         hybridization parameter = " << bc.getHybridParameter() << std::endl;
    std::cout << "                                cubic correction = " << bc.getCubicMachCorrection() << std::endl;
    std::cout << "              non-reflecting boundary conditions = " << bc.getReflectionHandling() << std::endl;
    if (bc.getReflectionHandling() == "yes") {
        std::cout << "                                transverse terms = " << bc.getTransverseInclusion() << std::endl;
        std::cout << "                            K1 relaxation factor = " << bc.getLODIK1relaxation() << std::endl;
    }
    std::cout << "                               coupling approach = " << bc.getCouplingProcedure() << std::endl;
    std::cout << "          fine links to be replaced at interface = " << bc.getCouplingReplacementC2F() << std::endl;
    std::cout << "        coarse links to be replaced at interface = " << bc.getCouplingReplacementF2C() << std::endl;
    std::cout << "                     hanging node reconstruction = " << bc.getHangingNodeReconstruction() << std::endl;
    std::cout << "                                   f2c filtering = " << bc.getF2Cfilter() << std::endl;
    std::cout << "                                   scaling style = " << bc.getScalingStyle() << std::endl;
    std::cout << "                                   domain size x = " << bc.getChannelSizeX_() << std::endl;
    std::cout << "                                   domain size y = " << bc.getChannelSizeY_() << std::endl;
    std::cout << "                                   domain size z = " << bc.getChannelSizeZ_() << std::endl;
    std::cout << "                            number of time steps = " << bc.getTimeStepMax() << std::endl;
    std::cout << "                                 write intervall = " << bc.getWriteInterval() << std::endl;
    std::cout << "refinement layers at Y and Z boundaries and back = " << bc.getRefinementLayer() << std::endl;
    std::cout << "                                         overlap = " << bc.getOverlap() << std::endl;
    std::cout << "                    refinement layers at X front = " << bc.getRefinementLayerX() << std::endl;
    std::cout << "                   refined region (coarse x0 x1) = " << lat.refinedX0() << " " << lat.refinedX1()
              << std::endl
              << std::endl;
    lat.initialization(bc);
    const steady_clock::time_point beginCalc = steady_clock::now();
    const time_t now = system_clock::to_time_t(system_clock::now());
    std::cout << "current date and time is: " << std::ctime(&now) << std::endl;
    auto fileName = [&](int t) {
        std::ostringstream ts;
        ts << t;
        return bc.getCaseName() + "_" + ts.str();
    };
    while (time < maxTimeStep) {
        if (time == 0) lat.writeResultsVTK(fileName(0), alternatingL0, alternatingL1);
        lat.solve(alternatingL0, alternatingL1);
        ++time;
        bc.setTime_(static_cast<double>(time));
        if (time - initTime == 1) {
            acp_case_sync(lat.handle(0));  // both levels run on one stream
            const nanoseconds calcTime =
                duration_cast<nanoseconds>((steady_clock::now() - beginCalc) * (maxTimeStep - initTime));
            std::cout << "estimated total computation time: ";
            display(std::cout, calcTime);
        } else if (writeInterval != 0 && (time % writeInterval) == 0) {
            lat.writeResultsVTK(fileName(time), negAlternatingL0, alternatingL1);
            std::cout << "solved " << std::fixed << std::setprecision(2)
                      << (static_cast<double>(time) - static_cast<double>(initTime)) /
                             (maxTimeStep - static_cast<double>(initTime)) * 100
                      << "% of the case" << std::endl;
            if (time - initTime == writeInterval) {
                const nanoseconds calcTime = duration_cast<nanoseconds>(
                    (steady_clock::now() - beginCalc) * ((maxTimeStep - initTime) / writeInterval));
                std::cout << std::endl << "corrected computation time estimation (including time for file creation): ";
                display(std::cout, calcTime);
            }
        }
        alternatingL0 = !alternatingL0;
        negAlternatingL0 = !negAlternatingL0;
    }
    acp_case_sync(lat.handle(0));  // both levels run on one stream
    std::cout << std::endl << "given maximum number of time steps reached" << std::endl << std::endl;
    std::cout << "total computation time: ";
    display(std::cout, duration_cast<nanoseconds>(steady_clock::now() - beginTotal));
    return 0;
}

int main(int argc, char** argv) {
    const steady_clock::time_point beginTotal = steady_clock::now();
    const std::string inputFile = argc > 1 ? argv[1] : "BoundaryConditions.txt";
    Control boundaryConditions(inputFile);
    Lattice simulationCase(boundaryConditions);
    const bool twoLevel = simulationCase.levels() == 2;
    int time = 0;
    const int initTime = 0;
    const int writeInterval = boundaryConditions.getWriteInterval();
    const int maxTimeStep = boundaryConditions.getTimeStepMax();
    boundaryConditions.setTime_(static_cast<double>(time));
    bool alternatingL0 = true;  // which of the two population arrays is current
    bool negAlternatingL0 = !alternatingL0;
    bool alternatingL1 = true;

    std::cout << "LB acoustic pulse solver" << std::endl;
    std::cout << "initializing from uniform zero-velocity equilibrium distribution" << std::endl << std::endl;
    if (twoLevel) return runTwoLevel(boundaryConditions, simulationCase, beginTotal);
    std::cout << "all parameters are being displayed in SI units" << std::endl << std::endl;
    std::cout << "result file(s): " << boundaryConditions.getCaseName() << "_t.vtk" << std::endl << std::endl;
    std::cout << "        dimensionless collision frequency = " << boundaryConditions.getCollisionFrequency() << std::endl;
    std::cout << "                           time step size = " << boundaryConditions.getTimeStep() << std::endl;
    std::cout << "                                  spacing = " << boundaryConditions.getSpacing() << std::endl;
    std::cout << "                          REYNOLDS number = " << boundaryConditions.getReynoldsNumber() << std::endl;
    std::cout << "                      kinematic viscosity = " << boundaryConditions.getKinematicViscosity() << std::endl;
    std::cout << "                                  density = " << boundaryConditions.getDensity() << std::endl;
    std::cout << "                           speed of sound = " << boundaryConditions.getSpeedOfSound() << std::endl;
    std::cout << "                              MACH number = " << boundaryConditions.getMachNumber() << std::endl;
    std::cout << "                     order of equilibrium = " << boundaryConditions.getOrderOfEquilibrium() << std::endl;
    std::cout << "                          collision model = " << boundaryConditions.getCollisionModel() << std::endl;
    if (boundaryConditions.getCollisionModel() == "HRR")
        std::cout << "                 hybridization parameter = " << boundaryConditions.getHybridParameter() << std::endl;
    std::cout << "       non-reflecting boundary conditions = " << boundaryConditions.getReflectionHandling() << std::endl;
    if (boundaryConditions.getReflectionHandling() == "yes") {
        std::cout << "                        transverse terms = " << boundaryConditions.getTransverseInclusion() << std::endl;
        std::cout << "                    K1 relaxation factor = " << boundaryConditions.getLODIK1relaxation() << std::endl;
    }
    std::cout << "                         cubic correction = " << boundaryConditions.getCubicMachCorrection() << std::endl;
    std::cout << "                            domain size x = " << boundaryConditions.getChannelSizeX_() << std::endl;
    std::cout << "                            domain size y = " << boundaryConditions.getChannelSizeY_() << std::endl;
    std::cout << "                            domain size z = " << boundaryConditions.getChannelSizeZ_() << std::endl;
    std::cout << "                     number of time steps = " << boundaryConditions.getTimeStepMax() << std::endl;
    std::cout << "                          write intervall = " << boundaryConditions.getWriteInterval() << std::endl
              << std::endl;

    simulationCase.initialization(boundaryConditions);
    const steady_clock::time_point beginCalc = steady_clock::now();
    const time_t now = system_clock::to_time_t(system_clock::now());
    std::cout << "current date and time is: " << std::ctime(&now) << std::endl;

    auto fileName = [&](int t) {
        std::ostringstream ts;
        ts << t;
        return boundaryConditions.getCaseName() + "_" + ts.str();
    };
    while (time < maxTimeStep) {
        if (time == 0) simulationCase.writeResultsVTK(fileName(0), alternatingL0, alternatingL1);
        simulationCase.solve(alternatingL0);
        ++time;
        boundaryConditions.setTime_(static_cast<double>(time));
        if (time - initTime == 1) {
            // the launches are asynchronous: the estimate waits for the first step
            acp_case_sync(simulationCase.handle());
            const nanoseconds calcTime =
                duration_cast<nanoseconds>((steady_clock::now() - beginCalc) * (maxTimeStep - initTime));
            std::cout << "estimated total computation time: ";
            display(std::cout, calcTime);
        } else if (writeInterval != 0 && (time % writeInterval) == 0) {
            simulationCase.writeResultsVTK(fileName(time), negAlternatingL0, alternatingL1);
            std::cout << "solved " << std::fixed << std::setprecision(2)
                      << (static_cast<double>(time) - static_cast<double>(initTime)) /
                             (maxTimeStep - static_cast<double>(initTime)) * 100
                      << "% of the case" << std::endl;
            if (time - initTime == writeInterval) {
                const nanoseconds calcTime = duration_cast<nanoseconds>(
                    (steady_clock::now() - beginCalc) * ((maxTimeStep - initTime) / writeInterval));
                std::cout << std::endl << "corrected computation time estimation (including time for file creation): ";
                display(std::cout, calcTime);
            }
        }
        alternatingL0 = !alternatingL0;
        negAlternatingL0 = !negAlternatingL0;
    }
    acp_case_sync(simulationCase.handle());
    std::cout << std::endl << "given maximum number of time steps reached" << std::endl << std::endl;
    std::cout << "total computation time: ";
    display(std::cout, duration_cast<nanoseconds>(steady_clock::now() - beginTotal));
    return 0;
}
