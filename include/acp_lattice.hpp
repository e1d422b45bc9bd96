// C++ classes with the public interface of the reference's Control and Lattice
// (Control.h:17-75, Lattice.h:23-120), implemented over the C ABI of acp.h, so
// that a program written against the reference -- its main() (main.cpp:37-157)
// in particular -- runs on this engine by including this header, using
// namespace acp and linking libacpulse.so.  Header only.
//
// Kept from the reference:
//   * Control(filename) reads the case file with first-match keyword
//     semantics and, like the reference, prints the message and exits when the
//     file or a mandatory keyword is missing (Control.cpp:8-273); its getters
//     return the same values (node counts, time step, collision frequency,
//     lattice tables in the reference's direction order).
//   * Lattice(Control&) holds one uniform level, initialization() writes the
//     reference's pulse into both arrays, solve(flag) advances one step reading
//     the array the flag names and writing the other (the caller flips the
//     flag, main.cpp:155-156), writeResultsVTK(name, flag, flagL1) writes
//     <name>.vtk with the reference's layout, getDistribution() returns one
//     population of either array.
// Different:
//   * The two arrays live on the device and the case tracks which one holds
//     the newest state.  solve(flag) and writeResultsVTK(flag, ...) must name
//     it (after initialisation both arrays are equal and either flag is
//     accepted); a flag naming the other array, which the reference would
//     silently compute with, prints a message and exits.
//   * getDistribution with the flag of the previous step returns that step's
//     input state; the reference collides in place, so its array holds the
//     post-collision values of interior nodes there (boundary nodes agree).
//   * MaxLevel > 0 exits: the reference's fine-shell refinement is not built;
//     two-level cases run through acp_refined_* (acp.h).
//   * The private stepping members (timeStepping*, nestedTimeStepping*,
//     createLattice, setConnectivity) are not part of this interface.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "acp.h"

namespace acp {

namespace detail {
[[noreturn]] inline void die(const std::string& what) {
    std::cerr << what << std::flush;
    std::exit(1);
}
inline void check(int rc, const char* where) {
    if (rc != ACP_OK) die(std::string(where) + ": " + acp_last_error() + "\n");
}
}  // namespace detail

class Control {
public:
    Control() { acp_case_desc_default(&d_); }

    explicit Control(std::string filename) : file_(filename) {
        if (acp_control_read(filename.c_str(), &d_) != ACP_OK) detail::die(acp_last_error());
        c2f_ = readString(filename, "Coarse-to-Fine");
        f2c_ = readString(filename, "Fine-to-Coarse");
        hanging_ = readString(filename, "Hanging-Node-Reconstruction");
        scaling_ = readString(filename, "Scaling-Style");
        f2c_filter_ = readString(filename, "F2C-Filter");
        coupling_ = readString(filename, "Coupling-Approach");
        reflecting_ = readString(filename, "Non-Reflecting");
        transverse_ = readString(filename, "Transverse");
        cubic_ = readString(filename, "Cubic-Mach-Correction");
        ref_layer_ = static_cast<int>(readValue(filename, "RefinementLayerYZ"));
        ref_layer_x_ = static_cast<int>(readValue(filename, "RefinementLayerX"));
        overlap_ = readValue(filename, "Overlap");
        tables();
    }

    // Control.cpp:135-273
    double readValue(std::string filename, std::string keyword) {
        double v = 0.;
        if (acp_control_read_value(filename.c_str(), keyword.c_str(), &v) != ACP_OK) detail::die(acp_last_error());
        return v;
    }
    std::string readString(std::string filename, std::string keyword) {
        std::vector<char> b(4096);
        if (acp_control_read_string(filename.c_str(), keyword.c_str(), b.data(), (int64_t)b.size()) != ACP_OK)
            detail::die(acp_last_error());
        return b.data();
    }
    std::string readFileName(std::string filename, std::string keyword) {
        std::vector<char> b(4096);
        if (acp_control_read_filename(filename.c_str(), keyword.c_str(), b.data(), (int64_t)b.size()) != ACP_OK)
            detail::die(acp_last_error());
        return b.data();
    }
    void readVector(std::string filename, std::string keyword, double* vector) {
        if (acp_control_read_vector(filename.c_str(), keyword.c_str(), vector) != ACP_OK)
            detail::die(acp_last_error());
    }

    int getNumberOfNodesX() { return d_.nx; }
    int getNumberOfNodesY() { return d_.ny; }
    int getNumberOfNodesZ() { return d_.nz; }
    int getLevelMax() { return d_.max_level; }
    int getTimeStepMax() { return d_.timesteps; }
    int getWriteInterval() { return d_.write_interval; }
    int getNumberOfThreads() { return d_.threads; }
    int getRefinementLayer() { return ref_layer_; }
    int getRefinementLayerX() { return ref_layer_x_; }
    int getOrderOfEquilibrium() { return d_.order; }
    void setTime_(double time) { time_ = time; }
    double getSpacing() { return d_.spacing; }
    double getHybridParameter() { return d_.sigma; }
    double getDensity() { return d_.rho0; }
    double getMachNumber() { return d_.mach; }
    // Control.cpp:51-55
    double getMolecularVelocity() { return d_.cs * std::sqrt(3.); }
    double getReynoldsNumber() { return 0.; }
    double getSpeedOfSound() { return d_.cs; }
    double getKinematicViscosity() { return d_.nu; }
    double getTimeStep() { return d_.spacing / getMolecularVelocity(); }
    double getTime_() { return time_; }
    double getChannelSizeX_() { return d_.domain[0]; }
    double getChannelSizeY_() { return d_.domain[1]; }
    double getChannelSizeZ_() { return d_.domain[2]; }
    double getXsi(int& dir, int& comp) { return xsi_[dir][comp]; }
    double getXsi(unsigned short& dir, unsigned short& comp) { return xsi_[dir][comp]; }
    double* getXsiX() { return xsi_x_; }
    double* getXsiY() { return xsi_y_; }
    double* getXsiZ() { return xsi_z_; }
    double* getXsi(int& dir) { return xsi_[dir]; }
    double* getXsi(unsigned short& dir) { return xsi_[dir]; }
    double* getWeightFactors() { return weights_; }
    double getLODIK1relaxation() { return d_.k1; }
    double getLODIK2relaxation() { return d_.k2; }
    virtual double getCollisionFrequency() {
        const double dt = getTimeStep();
        return (std::pow(d_.cs, 2) * dt) / (d_.nu + 0.5 * std::pow(d_.cs, 2) * dt);
    }
    virtual double getOverlap() { return overlap_; }
    std::string getCaseName() { return d_.save_as; }
    std::string getCollisionModel() { return d_.collision_model; }
    std::string getCouplingReplacementC2F() { return c2f_; }
    std::string getCouplingReplacementF2C() { return f2c_; }
    std::string getHangingNodeReconstruction() { return hanging_; }
    std::string getScalingStyle() { return scaling_; }
    std::string getReflectionHandling() { return reflecting_; }
    std::string getTransverseInclusion() { return transverse_; }
    std::string getCubicMachCorrection() { return cubic_; }
    std::string getF2Cfilter() { return f2c_filter_; }
    std::string getCouplingProcedure() { return coupling_; }
    virtual ~Control() = default;

    // the engine's description of the case (includes the options this engine adds)
    const acp_case_desc& desc() const { return d_; }
    acp_case_desc& desc() { return d_; }

private:
    // D3Q19 in the reference's order (Control.cpp:58-131): +x +y -x -y +z -z,
    // then (+x,+z) (+x,-z) (+y,+z) (+y,-z) (-x,+z) (-x,-z) (-y,+z) (-y,-z)
    // (-x,+y) (+x,-y) (+x,+y) (-x,-y), rest; components scaled by c = sqrt(3) c_s.
    void tables() {
        static const int e[19][3] = {{1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0}, {0, 0, 1},  {0, 0, -1}, {1, 0, 1},
                                     {1, 0, -1}, {0, 1, 1},  {0, 1, -1}, {-1, 0, 1}, {-1, 0, -1}, {0, -1, 1}, {0, -1, -1},
                                     {-1, 1, 0}, {1, -1, 0}, {1, 1, 0},  {-1, -1, 0}, {0, 0, 0}};
        const double c = getMolecularVelocity();
        for (int q = 0; q < 19; ++q) {
            const double v[3] = {e[q][0] * c, e[q][1] * c, e[q][2] * c};
            xsi_x_[q] = e[q][0] ? v[0] : 0.;
            xsi_y_[q] = e[q][1] ? v[1] : 0.;
            xsi_z_[q] = e[q][2] ? v[2] : 0.;
            xsi_[q][0] = xsi_x_[q];
            xsi_[q][1] = xsi_y_[q];
            xsi_[q][2] = xsi_z_[q];
            weights_[q] = q < 6 ? 1. / 18. : (q < 18 ? 1. / 36. : 1. / 3.);
        }
    }

    acp_case_desc d_{};
    std::string file_;
    std::string c2f_, f2c_, hanging_, scaling_, f2c_filter_, coupling_, reflecting_, transverse_, cubic_;
    int ref_layer_ = 0, ref_layer_x_ = 0;
    double overlap_ = 0., time_ = 0.;
    double xsi_[19][3] = {}, xsi_x_[19] = {}, xsi_y_[19] = {}, xsi_z_[19] = {}, weights_[19] = {};
};

class Lattice {
public:
    explicit Lattice(Control& bc) : bc_(bc) {
        if (bc.getLevelMax() != 0)
            detail::die("MaxLevel " + std::to_string(bc.getLevelMax()) +
                        ": the fine-shell refinement is not available; use acp_refined_* for two-level cases\n");
        detail::check(acp_case_create(&bc.desc(), nullptr, &c_), "Lattice");
        N_ = (int64_t)bc.getNumberOfNodesX() * bc.getNumberOfNodesY() * bc.getNumberOfNodesZ();
    }
    Lattice(const Lattice&) = delete;
    Lattice& operator=(const Lattice&) = delete;
    virtual ~Lattice() { acp_case_destroy(c_); }

    // Lattice.cpp:1397-1493: both arrays hold the reference's initial field
    virtual void initialization(Control&) {
        detail::check(acp_case_init_reference(c_), "Lattice::initialization");
        both_ = true;
        cached_ = false;
    }

    // Lattice.cpp:2158-2183 for one level
    virtual void solve(bool& alternating) {
        if (both_) cur_ = alternating;
        if (alternating != cur_) detail::die("Lattice::solve: the flag names the array of the previous step\n");
        detail::check(acp_case_step(c_, 1), "Lattice::solve");
        cur_ = !cur_;
        both_ = false;
        cached_ = false;
    }
    virtual void solve(bool&, bool&) {
        detail::die("Lattice::solve(bool&, bool&): two-level cases run through acp_refined_*\n");
    }

    // Lattice.cpp:2212-2243, 4094-4349: <filename>.vtk from the newest state
    virtual void writeResultsVTK(std::string filename, bool& alternating, bool& /*alternatingL1*/) {
        if (!both_ && alternating != cur_)
            detail::die("Lattice::writeResultsVTK: the flag names the array of the previous step\n");
        if (acp_case_write_vtk(c_, filename.c_str()) != ACP_OK)
            std::cerr << "An error occured while writing paraview file\n";
    }

    // Lattice.h:119: population dir of node (i, j, k) on level 0 from the
    // array the flag names; a direction >= 19 gives 0 as the reference's NULL.
    virtual double getDistribution(bool& alternating, int& dir, int& level, int& i, int& j, int& k) {
        if (dir >= 19) return 0.;
        if (level != 0) detail::die("Lattice::getDistribution: single level only\n");
        if (!cached_) {
            cur_f_.resize((size_t)19 * N_);
            prev_f_.resize((size_t)19 * N_);
            detail::check(acp_case_get_populations(c_, cur_f_.data()), "Lattice::getDistribution");
            detail::check(acp_case_get_previous_populations(c_, prev_f_.data()), "Lattice::getDistribution");
            cached_ = true;
        }
        const int64_t l = i + (int64_t)bc_.getNumberOfNodesX() * (j + (int64_t)bc_.getNumberOfNodesY() * k);
        const std::vector<double>& f = (both_ || alternating == cur_) ? cur_f_ : prev_f_;
        return f[(size_t)dir * N_ + l];
    }

    // Lattice::readValue (Lattice.h:120): a number from a keyword file
    virtual double readValue(std::string filename, std::string keyword) { return bc_.readValue(filename, keyword); }

    // the case handle, for the engine's own entry points (probes, graphs, ...)
    acp_case* handle() { return c_; }

private:
    Control& bc_;
    acp_case* c_ = nullptr;
    int64_t N_ = 0;
    bool cur_ = true;   // the flag naming the array with the newest state
    bool both_ = true;  // both arrays hold the same state (after initialisation)
    bool cached_ = false;
    std::vector<double> cur_f_, prev_f_;
};

}  // namespace acp
