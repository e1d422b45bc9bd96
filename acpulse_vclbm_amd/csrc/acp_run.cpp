// acp_run -- command-line driver with the behaviour of the reference's main()
// (main.cpp:37-157, single level): reads the case file, prints the same
// parameter summary, writes <Save-as>_<t>.vtk at t = 0 and every WriteInterval
// steps, and prints the same progress and time estimates.  The case file path
// is the first argument (the reference hard-codes it, main.cpp:44); with no
// argument ./BoundaryConditions.txt is read.  Uses only the C ABI of acp.h.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include "acp.h"

using namespace std::chrono;

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

static void die(const char* what) {
    std::cerr << what << acp_last_error() << std::endl;
    std::exit(1);
}

int main(int argc, char** argv) {
    const steady_clock::time_point beginTotal = steady_clock::now();
    const std::string input = argc > 1 ? argv[1] : "BoundaryConditions.txt";
    acp_case_desc d;
    if (acp_control_read(input.c_str(), &d) != ACP_OK) die("");
    if (d.max_level != 0) {
        std::cerr << "MaxLevel " << d.max_level << ": two-level refinement is not available in this engine" << std::endl;
        return 1;
    }
    acp_case* c = nullptr;
    if (acp_case_create(&d, nullptr, &c) != ACP_OK) die("case creation failed: ");
    acp_case_info info;
    acp_case_info_get(c, &info);
    const double cs = d.cs, dt = info.dt;
    const double omega = (std::pow(cs, 2) * dt) / (d.nu + 0.5 * std::pow(cs, 2) * dt);  // Control.cpp:54

    std::cout << "LB acoustic pulse solver" << std::endl;
    std::cout << "initializing from uniform zero-velocity equilibrium distribution" << std::endl << std::endl;
    std::cout << "all parameters are being displayed in SI units" << std::endl << std::endl;
    std::cout << "result file(s): " << d.save_as << "_t.vtk" << std::endl << std::endl;
    std::cout << "        dimensionless collision frequency = " << omega << std::endl;
    std::cout << "                           time step size = " << dt << std::endl;
    std::cout << "                                  spacing = " << d.spacing << std::endl;
    std::cout << "                          REYNOLDS number = " << 0. << std::endl;
    std::cout << "                      kinematic viscosity = " << d.nu << std::endl;
    std::cout << "                                  density = " << d.rho0 << std::endl;
    std::cout << "                           speed of sound = " << d.cs << std::endl;
    std::cout << "                              MACH number = " << d.mach << std::endl;
    std::cout << "                     order of equilibrium = " << d.order << std::endl;
    std::cout << "                          collision model = " << d.collision_model << std::endl;
    if (std::string(d.collision_model) == "HRR")
        std::cout << "                 hybridization parameter = " << d.sigma << std::endl;
    std::cout << "       non-reflecting boundary conditions = " << (d.nonreflecting ? "yes" : "no") << std::endl;
    if (d.nonreflecting) {
        std::cout << "                        transverse terms = " << (d.transverse ? "yes" : "no") << std::endl;
        std::cout << "                    K1 relaxation factor = " << d.k1 << std::endl;
    }
    std::cout << "                         cubic correction = " << (d.cubic ? "yes" : "no") << std::endl;
    std::cout << "                            domain size x = " << d.domain[0] << std::endl;
    std::cout << "                            domain size y = " << d.domain[1] << std::endl;
    std::cout << "                            domain size z = " << d.domain[2] << std::endl;
    std::cout << "                     number of time steps = " << d.timesteps << std::endl;
    std::cout << "                          write intervall = " << d.write_interval << std::endl << std::endl;

    if (acp_case_init_reference(c) != ACP_OK) die("initialization failed: ");
    const steady_clock::time_point beginCalc = steady_clock::now();
    const time_t now = system_clock::to_time_t(system_clock::now());
    std::cout << "current date and time is: " << std::ctime(&now) << std::endl;

    const int maxTimeStep = d.timesteps, writeInterval = d.write_interval, initTime = 0;
    int time = 0;
    auto write = [&](int t) {
        std::ostringstream name;
        name << d.save_as << "_" << t;
        if (acp_case_write_vtk(c, name.str().c_str()) != ACP_OK) std::cerr << acp_last_error() << "\n";
    };
    while (time < maxTimeStep) {
        if (time == 0) write(0);
        if (acp_case_step(c, 1) != ACP_OK) die("step failed: ");
        ++time;
        if (time - initTime == 1) {
            acp_case_sync(c);
            const nanoseconds calcTime = duration_cast<nanoseconds>((steady_clock::now() - beginCalc) * (maxTimeStep - initTime));
            std::cout << "estimated total computation time: ";
            display(std::cout, calcTime);
        } else if (writeInterval != 0 && (time % writeInterval) == 0) {
            write(time);
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
    }
    acp_case_sync(c);
    acp_case_destroy(c);
    std::cout << std::endl << "given maximum number of time steps reached" << std::endl << std::endl;
    std::cout << "total computation time: ";
    display(std::cout, duration_cast<nanoseconds>(steady_clock::now() - beginTotal));
    return 0;
}
