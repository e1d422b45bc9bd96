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
//   * The public stepping members of Lattice.h:102-112 are all here:
//     createLattice() (re)creates the device state of every level,
//     setConnectivity() is kept for programs that call it (the connectivity
//     is derived when the device state is created), timeStepping<Op>(level,
//     flag) is one uniform step of level 0, nestedTimeStepping<Op>(flagL1,
//     flagL0, control) one coupled two-level step, solve(flagL0, flagL1)
//     dispatches to it on the collision model as Lattice.cpp:2185-2209 does.
//     A lattice steps with the operator its case file names: a stepping
//     member of another operator prints a message and exits.
//   * MaxLevel 1 builds a two-level lattice (acp_refined_*, acp.h).  The
//     reference's keys alone select its own geometry, a fine shell around a
//     coarse core (Lattice.cpp:6-213: RefinementLayerYZ, RefinementLayerX,
//     Overlap 1 for the standard coupling; acp_refined_create_shell); the
//     direct coupling (Overlap 0) is not available on it and exits with that
//     reason.  The engine key `Boundary LODI-X` or `Boundary PERIODIC` with
//     the fine region `Refined-Region x0 x1` in coarse node indices (default
//     for LODI-X: the right half, x0 = nx / 2, x1 = nx - 1) selects this
//     engine's x band instead.
//   * NodeView (node(level, i, j, k)) gives the getters of the reference's
//     Node (Node.h:38-99) on a host copy of the newest state: populations of
//     both arrays, density and velocity, cubic-Mach terms, node kind, LODI
//     neighbours and the push links; it replaces the Node objects, which do
//     not exist on the host here.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <unordered_map>
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
        // optional key of this engine: the fine region of a refined geometry
        std::vector<char> b(256);
        if (acp_control_read_filename(filename.c_str(), "Refined-Region", b.data(), (int64_t)b.size()) == ACP_OK) {
            int x0 = 0, x1 = 0;
            if (std::sscanf(b.data(), "%d %d", &x0, &x1) != 2)
                detail::die("Refined-Region needs two coarse node indices x0 x1\n");
            region_[0] = x0;
            region_[1] = x1;
        }
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
    int getRefinementLayer() const { return ref_layer_; }
    int getRefinementLayerX() const { return ref_layer_x_; }
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
    // Refined-Region of the case file ({-1, -1} when absent)
    const int* refinedRegion() const { return region_; }

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
    int region_[2] = {-1, -1};
    double overlap_ = 0., time_ = 0.;
    double xsi_[19][3] = {}, xsi_x_[19] = {}, xsi_y_[19] = {}, xsi_z_[19] = {}, weights_[19] = {};
};

class Lattice;

// Host view of one node of a Lattice with the getters of the reference's Node
// (Node.h:38-99), read from a host copy of the lattice's newest state (taken
// on first use after each step).  An invalid view stands for the reference's
// NULL node pointer (a neighbour a boundary node does not push to).
class NodeView {
public:
    NodeView() = default;
    bool valid() const { return lat_ != nullptr; }

    int getIndexI() const { return i_; }
    int getIndexJ() const { return j_; }
    int getIndexK() const { return k_; }
    int getLevel() const { return level_; }
    inline int getNormalDir();
    inline int getNodeID();
    inline bool getIsCorner();
    inline double getX() const;
    inline double getY() const;
    inline double getZ() const;
    // density and velocity of the node's newest state (Node.cpp:11-30, 85-118)
    inline double getNodeDensity();
    inline double getNodeVelocityComponent(int dir);
    inline double getCollisionFrequency() const;
    double getTurbulentViscosity() const { return 0.; }
    // population dir of the array the flag names; a direction >= 19 gives 0
    // as the reference's NULL
    inline double getDistribution(bool distr, int dir);
    // the state the last step read (the reference's pre-collision copy,
    // Lattice.cpp:1504-1522, taken from the same array before the step)
    inline double getDistributionPreCol(int dir);
    inline double getCubicMachCorrection(int dir);
    inline double getTimeStep() const;
    inline std::string getTag();
    inline NodeView getNeighbor(int dir);
    inline NodeView getNormalNeighbor();
    inline NodeView getNormalNeighborSnd();
    inline NodeView getTangentNeighbor1(int dir);
    inline NodeView getTangentNeighbor2(int dir);
    inline NodeView getPartner();

private:
    friend class Lattice;
    NodeView(Lattice* lat, int level, int i, int j, int k) : lat_(lat), level_(level), i_(i), j_(j), k_(k) {}
    inline int64_t lin() const;
    inline const acp_lodi_record* lodi();
    inline NodeView at_node(int64_t l) const;
    Lattice* lat_ = nullptr;
    int level_ = 0, i_ = 0, j_ = 0, k_ = 0;
};

class Lattice {
public:
    explicit Lattice(Control& bc) : bc_(bc) {
        createLattice();
        setConnectivity();
    }
    Lattice(const Lattice&) = delete;
    Lattice& operator=(const Lattice&) = delete;
    virtual ~Lattice() { release(); }

    // Lattice.cpp:6-234: the device state of every level (the constructor
    // calls it; calling it again discards the state, as the reference's new
    // nodes start empty, and initialization() must run before the next step)
    virtual void createLattice() {
        release();
        const acp_case_desc& d = bc_.desc();
        if (d.max_level == 0) {
            detail::check(acp_case_create(&d, nullptr, &lv_[0]), "Lattice::createLattice");
        } else if (d.max_level == 1) {
            acp_case_desc cd = d;
            cd.max_level = 0;
            shell_ = d.boundary == ACP_BC_LODI;
            if (shell_) {
                // the reference's geometry (Lattice.cpp:6-213)
                if (bc_.getOverlap() != 1.)
                    detail::die("Lattice::createLattice: MaxLevel 1: Overlap " + std::to_string(bc_.getOverlap()) +
                                ": the fine shell couples with the standard coupling (Overlap 1); the direct "
                                "coupling (Overlap 0) is available on this engine's x band (Boundary LODI-X)\n");
                x0_ = x1_ = 0;
                detail::check(acp_refined_create_shell(&cd, bc_.getRefinementLayer(), bc_.getRefinementLayerX(),
                                                       nullptr, &r_),
                              "Lattice::createLattice");
            } else {
                std::string why;
                if (!refined_geometry(x0_, x1_, why))
                    detail::die("Lattice::createLattice: MaxLevel 1: " + why + "\n");
                detail::check(acp_refined_create(&cd, x0_, x1_, nullptr, &r_), "Lattice::createLattice");
            }
            detail::check(acp_refined_level(r_, 0, &lv_[0]), "Lattice::createLattice");
            detail::check(acp_refined_level(r_, 1, &lv_[1]), "Lattice::createLattice");
            apply_coupling_keys();
        } else {
            detail::die("Lattice::createLattice: MaxLevel " + std::to_string(d.max_level) +
                        ": at most one refined level\n");
        }
        for (int l = 0; l < levels(); ++l) {
            acp_case_info info{};
            detail::check(acp_case_info_get(lv_[l], &info), "Lattice::createLattice");
            N_[l] = info.nodes;
            omega_[l] = info.omega;
            dt_[l] = info.dt;
            cur_[l] = true;
            both_[l] = true;
        }
        initialized_ = false;
        invalidate();
    }
    // The case file's coupling keys (BoundaryConditions.txt:52-90):
    // Coupling-Approach STD / DC1 / DC2 with Coarse-to-Fine, Fine-to-Coarse
    // and Hanging-Node-Reconstruction, F2C-Filter NOFI / LAG / TOU and Scaling-Style
    // SRTSTYLE / MRTSTYLE.  A value the engine does not provide ends the program with the
    // key and the value, as the reference ends on a malformed case file.
    void apply_coupling_keys() {
        const std::string filt = bc_.getF2Cfilter();
        if (filt != "NOFI" && filt != "LAG" && filt != "TOU")
            detail::die("Lattice::createLattice: F2C-Filter " + filt + " (NOFI, LAG or TOU)\n");
        detail::check(acp_refined_set_filter(r_, filt == "LAG" ? ACP_F2C_LAG : (filt == "TOU" ? ACP_F2C_TOU : ACP_F2C_NOFI)),
                      "Lattice::createLattice");
        const std::string sc = bc_.getScalingStyle();
        if (sc != "SRTSTYLE" && sc != "MRTSTYLE")
            detail::die("Lattice::createLattice: Scaling-Style " + sc + " (SRTSTYLE or MRTSTYLE)\n");
        detail::check(acp_refined_set_scaling(r_, sc == "MRTSTYLE" ? ACP_SCALING_MRT : ACP_SCALING_SRT),
                      "Lattice::createLattice");
        const std::string cp = bc_.getCouplingProcedure();
        if (cp != "STD" && cp != "DC1" && cp != "DC2")
            detail::die("Lattice::createLattice: Coupling-Approach " + cp + " (STD, DC1 or DC2)\n");
        auto links = [](const std::string& key, const std::string& v, const char* m, const char* mp, const char* all) {
            if (v == m) return (int32_t)ACP_LINKS_MISSING;
            if (v == mp) return (int32_t)ACP_LINKS_MISSING_COMMON;
            if (v == all) return (int32_t)ACP_LINKS_ALL;
            detail::die("Lattice::createLattice: " + key + " " + v + " (" + m + ", " + mp + " or " + all + ")\n");
            return (int32_t)-1;
        };
        const int32_t c2f = links("Coarse-to-Fine", bc_.getCouplingReplacementC2F(), "Mf", "MfP", "fAll");
        const int32_t f2c = links("Fine-to-Coarse", bc_.getCouplingReplacementF2C(), "Mc", "McP", "cAll");
        const int32_t hang = links("Hanging-Node-Reconstruction", bc_.getHangingNodeReconstruction(), "hMf", "hMfP", "hAll");
        const std::string where = "Lattice::createLattice: Coupling-Approach " + cp;
        detail::check(acp_refined_set_coupling(r_, cp == "STD" ? ACP_COUPLING_STD : ACP_COUPLING_DC, c2f, f2c, hang),
                      where.c_str());
    }

    // Lattice.cpp:237-1394: the neighbour, LODI and partner relations are a
    // function of the grid and its boundary kind; the engine derives them
    // when the device state is created (the LODI schedule on the host, the
    // push links in the kernels), so nothing remains to do here
    virtual void setConnectivity() {
        if (!lv_[0]) detail::die("Lattice::setConnectivity before createLattice\n");
    }

    // Lattice.cpp:1397-1493: both arrays of every level hold the reference's
    // initial field
    virtual void initialization(Control&) {
        if (r_) detail::check(acp_refined_init_reference(r_), "Lattice::initialization");
        else detail::check(acp_case_init_reference(lv_[0]), "Lattice::initialization");
        for (int l = 0; l < 2; ++l) both_[l] = true;
        l1_flag_set_ = false;
        initialized_ = true;
        invalidate();
    }

    // Lattice.cpp:2158-2183: one step of level 0 with the case file's operator
    virtual void solve(bool& alternating) {
        int level = 0;
        const std::string m = bc_.getCollisionModel();
        if (m == "SRT") timeSteppingSRT(level, alternating);
        else if (m == "MRT-GS") timeSteppingMRT_GSbasis(level, alternating);
        else if (m == "MRT-RM") timeSteppingMRT_RMbasis(level, alternating);
        else if (m == "RR") timeSteppingRR(level, alternating);
        else if (m == "HRR") timeSteppingHRR(level, alternating);
    }
    // Lattice.cpp:2185-2209: one coupled two-level step
    virtual void solve(bool& alternatingL0, bool& alternatingL1) {
        const std::string m = bc_.getCollisionModel();
        if (m == "SRT") nestedTimeSteppingSRT(alternatingL1, alternatingL0, bc_);
        else if (m == "MRT-GS") nestedTimeSteppingMRT_GSbasis(alternatingL1, alternatingL0, bc_);
        else if (m == "MRT-RM") nestedTimeSteppingMRT_RMbasis(alternatingL1, alternatingL0, bc_);
        else if (m == "RR") nestedTimeSteppingRR(alternatingL1, alternatingL0, bc_);
        else if (m == "HRR") nestedTimeSteppingHRR(alternatingL1, alternatingL0, bc_);
    }

    // Lattice.cpp:1496-2155: one uniform step of `level` reading the array
    // `alternating` names and writing the other
    virtual void timeSteppingSRT(int& level, bool& alternating) { uniform(ACP_OP_SRT, "timeSteppingSRT", level, alternating); }
    virtual void timeSteppingMRT_GSbasis(int& level, bool& alternating) {
        uniform(ACP_OP_MRT_GS, "timeSteppingMRT_GSbasis", level, alternating);
    }
    virtual void timeSteppingMRT_RMbasis(int& level, bool& alternating) {
        uniform(ACP_OP_MRT_RM, "timeSteppingMRT_RMbasis", level, alternating);
    }
    virtual void timeSteppingRR(int& level, bool& alternating) { uniform(ACP_OP_RR, "timeSteppingRR", level, alternating); }
    virtual void timeSteppingHRR(int& level, bool& alternating) { uniform(ACP_OP_HRR, "timeSteppingHRR", level, alternating); }

    // Lattice.cpp:4354-15051: one coarse step and two fine steps with the
    // coupling between them
    virtual void nestedTimeSteppingSRT(bool& alternatingL1, bool& alternatingL0, Control&) {
        nested(ACP_OP_SRT, "nestedTimeSteppingSRT", alternatingL1, alternatingL0);
    }
    virtual void nestedTimeSteppingMRT_GSbasis(bool& alternatingL1, bool& alternatingL0, Control&) {
        nested(ACP_OP_MRT_GS, "nestedTimeSteppingMRT_GSbasis", alternatingL1, alternatingL0);
    }
    virtual void nestedTimeSteppingMRT_RMbasis(bool& alternatingL1, bool& alternatingL0, Control&) {
        nested(ACP_OP_MRT_RM, "nestedTimeSteppingMRT_RMbasis", alternatingL1, alternatingL0);
    }
    virtual void nestedTimeSteppingRR(bool& alternatingL1, bool& alternatingL0, Control&) {
        nested(ACP_OP_RR, "nestedTimeSteppingRR", alternatingL1, alternatingL0);
    }
    virtual void nestedTimeSteppingHRR(bool& alternatingL1, bool& alternatingL0, Control&) {
        nested(ACP_OP_HRR, "nestedTimeSteppingHRR", alternatingL1, alternatingL0);
    }

    // Lattice.cpp:2212-4349: <filename>.vtk from the newest state of every
    // level (the two-level layout of acp_refined_write_vtk)
    virtual void writeResultsVTK(std::string filename, bool& alternating, bool& /*alternatingL1*/) {
        if (!both_[0] && alternating != cur_[0])
            detail::die("Lattice::writeResultsVTK: the flag names the array of the previous step\n");
        const int rc = r_ ? acp_refined_write_vtk(r_, filename.c_str()) : acp_case_write_vtk(lv_[0], filename.c_str());
        if (rc != ACP_OK) std::cerr << "An error occured while writing paraview file\n";
    }

    // Lattice.h:119: population dir of node (i, j, k) of `level` from the
    // array the flag names; a direction >= 19 gives 0 as the reference's NULL.
    // Level 1 takes the reference's fine indices over the whole domain; nodes
    // outside the fine region give 0.
    virtual double getDistribution(bool& alternating, int& dir, int& level, int& i, int& j, int& k) {
        if (dir >= 19) return 0.;
        if (level < 0 || level >= levels()) detail::die("Lattice::getDistribution: no level " + std::to_string(level) + "\n");
        const int il = level == 1 ? i - 2 * x0_ : i;
        if (il < 0 || il >= nx(level)) return 0.;
        return host(level).f[(alternating == cur_[level] || both_[level]) ? 0 : 1][(size_t)dir * N_[level] + lin(level, il, j, k)];
    }

    // Lattice::readValue (Lattice.h:120): a number from a keyword file
    virtual double readValue(std::string filename, std::string keyword) { return bc_.readValue(filename, keyword); }

    // The reference's Node objects as host views (no counterpart: the
    // reference hands out Node* through its mesh).  Level 1 takes the fine
    // region's own indices (0 <= i < 2 (x1 - x0) + 1).
    NodeView node(int level, int i, int j, int k) {
        if (level < 0 || level >= levels() || i < 0 || j < 0 || k < 0 || i >= nx(level) || j >= ny(level) || k >= nz(level))
            return NodeView();
        return NodeView(this, level, i, j, k);
    }

    // the case handle of level 0 (and of level 1 of a two-level lattice), for
    // the engine's own entry points (probes, graphs, ...), and the two-level handle
    acp_case* handle(int level = 0) { return lv_[level]; }
    acp_refined* refinedHandle() { return r_; }
    int levels() const { return r_ ? 2 : 1; }
    int refinedX0() const { return x0_; }
    int refinedX1() const { return x1_; }

private:
    friend class NodeView;
    struct Host {
        bool have = false, have_macros = false, have_psi = false;
        std::vector<double> f[2], rho, u, psi;
        std::unordered_map<int64_t, acp_lodi_record> lodi;
        bool have_lodi = false;
    };

    void release() {
        if (r_) acp_refined_destroy(r_);
        else if (lv_[0]) acp_case_destroy(lv_[0]);
        r_ = nullptr;
        lv_[0] = lv_[1] = nullptr;
    }
    void invalidate() {
        for (Host& h : host_) {
            h.have = h.have_macros = h.have_psi = false;
        }
    }
    int nx(int level) const {
        return level == 0 ? bc_.desc().nx : (shell_ ? 2 * bc_.desc().nx - 1 : 2 * (x1_ - x0_) + 1);
    }
    int ny(int level) const { return level == 0 ? bc_.desc().ny : 2 * bc_.desc().ny - (shell_ ? 1 : 0); }
    int nz(int level) const { return level == 0 ? bc_.desc().nz : 2 * bc_.desc().nz - (shell_ ? 1 : 0); }
    // the shell geometry's grid-coupling surfaces (createLattice): the core
    // box on level 0, the hole box on level 1
    bool shell_surface(int level, int i, int j, int k) const {
        const acp_case_desc& d = bc_.desc();
        const int ly = bc_.getRefinementLayer(), lx = bc_.getRefinementLayerX(), m = level == 0 ? 1 : 2;
        const int o = level == 0 ? 1 : 0;  // Overlap 1 on level 0
        const int b[6] = {m * lx - o, m * (d.nx - 1 - ly) + o, m * ly - o, m * (d.ny - 1 - ly) + o, m * ly - o,
                          m * (d.nz - 1 - ly) + o};
        const bool in = i >= b[0] && i <= b[1] && j >= b[2] && j <= b[3] && k >= b[4] && k <= b[5];
        const bool inner = i > b[0] && i < b[1] && j > b[2] && j < b[3] && k > b[4] && k < b[5];
        return in && !inner;
    }
    int64_t lin(int level, int i, int j, int k) const {
        return i + (int64_t)nx(level) * (j + (int64_t)ny(level) * k);
    }
    // the boundary kind of a level: the case file's for level 0; the fine
    // level's faces are open (coupled), or LODI at x = nx - 1 when the region
    // reaches the high-x face of a LODI-X coarse level
    int boundary(int level) const {
        const int b = bc_.desc().boundary;
        if (level == 0 || shell_) return b;
        return (b == ACP_BC_LODI_X && x1_ == bc_.desc().nx - 1) ? ACP_BC_OPEN_LODI_X : ACP_BC_OPEN_X;
    }

    bool refined_geometry(int& x0, int& x1, std::string& why) const {
        const acp_case_desc& d = bc_.desc();
        const int* rr = bc_.refinedRegion();
        if (d.boundary != ACP_BC_LODI_X && d.boundary != ACP_BC_PERIODIC) {
            why = "refinement needs Boundary LODI-X or PERIODIC";
            return false;
        }
        if (rr[0] >= 0) {
            x0 = rr[0];
            x1 = rr[1];
        } else if (d.boundary == ACP_BC_LODI_X) {
            x0 = d.nx / 2;
            x1 = d.nx - 1;
        } else {
            why = "a periodic grid needs Refined-Region x0 x1";
            return false;
        }
        return true;
    }

    void check_op(int op, const char* name) const {
        if (bc_.desc().op != op)
            detail::die(std::string("Lattice::") + name + ": this lattice collides with the case file's Collision-Model " +
                        std::string(bc_.desc().collision_model) + " (one operator per case)\n");
    }
    void uniform(int op, const char* name, int& level, bool& alternating) {
        check_op(op, name);
        if (r_ || level != 0)
            detail::die(std::string("Lattice::") + name + ": level " + std::to_string(level) +
                        (r_ ? " of a two-level lattice steps through solve(bool&, bool&)\n" : ": single-level lattice\n"));
        if (both_[0]) cur_[0] = alternating;
        if (alternating != cur_[0]) detail::die("Lattice::solve: the flag names the array of the previous step\n");
        detail::check(acp_case_step(lv_[0], 1), name);
        cur_[0] = !cur_[0];
        both_[0] = false;
        invalidate();
    }
    void nested(int op, const char* name, bool& alternatingL1, bool& alternatingL0) {
        check_op(op, name);
        if (!r_) detail::die(std::string("Lattice::") + name + ": MaxLevel 0 lattice\n");
        if (both_[0]) cur_[0] = alternatingL0;
        if (alternatingL0 != cur_[0]) detail::die("Lattice::solve: the level-0 flag names the array of the previous step\n");
        // the fine level takes two steps per coarse step and ends each coarse
        // step on the array it started from: its flag stays (main.cpp:262-263)
        if (!l1_flag_set_) {
            cur_[1] = alternatingL1;
            l1_flag_set_ = true;
        }
        if (alternatingL1 != cur_[1]) detail::die("Lattice::solve: the level-1 flag changed between steps\n");
        detail::check(acp_refined_step(r_, 1), name);
        cur_[0] = !cur_[0];
        both_[0] = false;
        both_[1] = false;
        invalidate();
    }

    Host& host(int level) {
        Host& h = host_[level];
        if (!h.have) {
            for (int a = 0; a < 2; ++a) h.f[a].resize((size_t)19 * N_[level]);
            detail::check(acp_case_get_populations(lv_[level], h.f[0].data()), "Lattice: host copy");
            detail::check(acp_case_get_previous_populations(lv_[level], h.f[1].data()), "Lattice: host copy");
            h.have = true;
        }
        return h;
    }
    Host& macros(int level) {
        Host& h = host_[level];
        if (!h.have_macros) {
            h.rho.resize((size_t)N_[level]);
            h.u.resize((size_t)3 * N_[level]);
            detail::check(acp_case_get_macros(lv_[level], h.rho.data(), h.u.data()), "Lattice: macros");
            h.have_macros = true;
        }
        return h;
    }
    Host& psi(int level) {
        Host& h = host_[level];
        if (!h.have_psi) {
            h.psi.resize((size_t)19 * N_[level]);
            detail::check(acp_case_get_psi(lv_[level], h.psi.data()), "Lattice: psi");
            h.have_psi = true;
        }
        return h;
    }
    // LODI records of a level's boundary nodes, as Lattice::setConnectivity
    // assigns the normal, second-normal and tangent neighbours
    // (Lattice.cpp:500-1392); the faces of a grid wrapped in y and z have no
    // edges or corners and wrap their tangents
    const acp_lodi_record* lodi(int level, int64_t l) {
        Host& h = host_[level];
        if (!h.have_lodi) {
            const int b = boundary(level), X = nx(level), Y = ny(level), Z = nz(level);
            if (b == ACP_BC_LODI) {
                const int64_t n = acp_lodi_launches(X, Y, Z, nullptr, 0, nullptr, 0, nullptr);
                std::vector<acp_lodi_record> rec((size_t)n);
                acp_lodi_launches(X, Y, Z, rec.data(), n, nullptr, 0, nullptr);
                for (const acp_lodi_record& r : rec) h.lodi[r.node] = r;
            } else if (b == ACP_BC_LODI_X || b == ACP_BC_OPEN_LODI_X) {
                for (int k = 0; k < Z; ++k)
                    for (int j = 0; j < Y; ++j)
                        for (int side = (b == ACP_BC_LODI_X ? 0 : 1); side < 2; ++side) {
                            const int i = side ? X - 1 : 0, s = side ? -1 : 1;
                            acp_lodi_record r{};
                            r.node = lin(level, i, j, k);
                            r.n1 = lin(level, i + s, j, k);
                            r.n2 = lin(level, i + 2 * s, j, k);
                            r.t[0] = lin(level, i, (j + 1) % Y, k);
                            r.t[1] = lin(level, i, (j + Y - 1) % Y, k);
                            r.t[2] = lin(level, i, j, (k + 1) % Z);
                            r.t[3] = lin(level, i, j, (k + Z - 1) % Z);
                            r.dir = side ? 2 : 0;
                            r.corner = 0;
                            h.lodi[r.node] = r;
                        }
            }
            h.have_lodi = true;
        }
        auto it = h.lodi.find(l);
        return it == h.lodi.end() ? nullptr : &it->second;
    }

    Control& bc_;
    acp_case* lv_[2] = {nullptr, nullptr};
    acp_refined* r_ = nullptr;
    int x0_ = 0, x1_ = 0;
    bool shell_ = false;
    int64_t N_[2] = {0, 0};
    double omega_[2] = {0., 0.}, dt_[2] = {0., 0.};
    bool cur_[2] = {true, true};   // the flag naming the array with the newest state, per level
    bool both_[2] = {true, true};  // both arrays hold the same state (after initialisation)
    bool l1_flag_set_ = false;
    bool initialized_ = false;
    Host host_[2];
};

// ---------------------------------------------------------------- NodeView

namespace detail {
// D3Q19 directions in the reference's order (Control.cpp:58-116)
constexpr int kE[19][3] = {{1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0}, {0, 0, 1},  {0, 0, -1}, {1, 0, 1},
                           {1, 0, -1}, {0, 1, 1},  {0, 1, -1}, {-1, 0, 1}, {-1, 0, -1}, {0, -1, 1}, {0, -1, -1},
                           {-1, 1, 0}, {1, -1, 0}, {1, 1, 0},  {-1, -1, 0}, {0, 0, 0}};
}  // namespace detail

int64_t NodeView::lin() const { return lat_->lin(level_, i_, j_, k_); }

const acp_lodi_record* NodeView::lodi() { return lat_ ? lat_->lodi(level_, lin()) : nullptr; }

NodeView NodeView::at_node(int64_t l) const {
    if (l < 0) return NodeView();
    const int X = lat_->nx(level_), Y = lat_->ny(level_);
    return NodeView(lat_, level_, (int)(l % X), (int)((l / X) % Y), (int)(l / ((int64_t)X * Y)));
}

int NodeView::getNormalDir() {
    const acp_lodi_record* r = lodi();
    return r ? r->dir : 0;  // the reference never assigns it elsewhere (zero-filled heap)
}
bool NodeView::getIsCorner() {
    const acp_lodi_record* r = lodi();
    return r ? r->corner != 0 : false;
}
int NodeView::getNodeID() {
    const std::string t = getTag();
    return t == "fluid" ? 0 : (t == "boundary" ? 1 : 2);
}
// FluidNode / BoundaryNode / GridCoupling tags (FluidNode.h:11-16,
// BoundaryNode.h:11-16, GridCoupling.h:12-17): boundary nodes are the LODI
// nodes; on the fine level of a two-level lattice the coupled x faces are
// grid-coupling nodes
std::string NodeView::getTag() {
    if (lat_->shell_) {
        // createLattice with MaxLevel 1: level 0 has no boundary nodes
        if (level_ == 1 && lodi()) return "boundary";
        return lat_->shell_surface(level_, i_, j_, k_) ? "gridcoupling" : "fluid";
    }
    if (lodi()) return "boundary";
    if (level_ == 1) {
        const int b = lat_->boundary(1), X = lat_->nx(1);
        if (i_ == 0 || (b == ACP_BC_OPEN_X && i_ == X - 1)) return "gridcoupling";
    }
    return "fluid";
}
double NodeView::getX() const { return (i_ + (level_ == 1 ? 2 * lat_->x0_ : 0)) * (lat_->bc_.getSpacing() / std::pow(2., level_)); }
double NodeView::getY() const { return j_ * (lat_->bc_.getSpacing() / std::pow(2., level_)); }
double NodeView::getZ() const { return k_ * (lat_->bc_.getSpacing() / std::pow(2., level_)); }
double NodeView::getNodeDensity() { return lat_->macros(level_).rho[(size_t)lin()]; }
double NodeView::getNodeVelocityComponent(int dir) {
    return lat_->macros(level_).u[(size_t)dir * lat_->N_[level_] + lin()];
}
double NodeView::getCollisionFrequency() const { return lat_->omega_[level_]; }
double NodeView::getTimeStep() const { return lat_->dt_[level_]; }
double NodeView::getDistribution(bool distr, int dir) {
    if (dir >= 19) return 0.;
    const int a = (distr == lat_->cur_[level_] || lat_->both_[level_]) ? 0 : 1;
    return lat_->host(level_).f[a][(size_t)dir * lat_->N_[level_] + lin()];
}
double NodeView::getDistributionPreCol(int dir) {
    if (dir >= 19) return 0.;
    return lat_->host(level_).f[lat_->both_[level_] ? 0 : 1][(size_t)dir * lat_->N_[level_] + lin()];
}
double NodeView::getCubicMachCorrection(int dir) {
    if (dir >= 19) return 0.;
    return lat_->psi(level_).psi[(size_t)dir * lat_->N_[level_] + lin()];
}
// Node::getNeighbor (Node.h:66): the node link dir pushes to, or none, by
// the push rule of Lattice::setConnectivity (Lattice.cpp:244-1392): a LODI
// boundary node pushes only the links that point inward in every dimension
// in which it lies on the boundary; a LODI x face of a grid wrapped in y and
// z its inward x links; links leaving an open face go nowhere; periodic
// directions wrap.
NodeView NodeView::getNeighbor(int dir) {
    if (!lat_ || dir < 0 || dir >= 18) return NodeView();
    const int b = lat_->boundary(level_), X = lat_->nx(level_), Y = lat_->ny(level_), Z = lat_->nz(level_);
    const int* e = detail::kE[dir];
    if (b == ACP_BC_LODI) {
        const int bi = i_ == 0 ? 1 : (i_ == X - 1 ? -1 : 0);
        const int bj = j_ == 0 ? 1 : (j_ == Y - 1 ? -1 : 0);
        const int bk = k_ == 0 ? 1 : (k_ == Z - 1 ? -1 : 0);
        if ((bi && e[0] != bi) || (bj && e[1] != bj) || (bk && e[2] != bk)) return NodeView();
        return NodeView(lat_, level_, i_ + e[0], j_ + e[1], k_ + e[2]);
    }
    const bool lodi_x = lodi() != nullptr;
    if (lodi_x && e[0] != (i_ == 0 ? 1 : -1)) return NodeView();
    int i = i_ + e[0];
    if (b == ACP_BC_PERIODIC) i = (i + X) % X;
    else if (i < 0 || i >= X) return NodeView();
    return NodeView(lat_, level_, i, (j_ + e[1] + Y) % Y, (k_ + e[2] + Z) % Z);
}
NodeView NodeView::getNormalNeighbor() {
    const acp_lodi_record* r = lodi();
    return r ? at_node(r->n1) : NodeView();
}
NodeView NodeView::getNormalNeighborSnd() {
    const acp_lodi_record* r = lodi();
    return r ? at_node(r->n2) : NodeView();
}
NodeView NodeView::getTangentNeighbor1(int dir) {
    const acp_lodi_record* r = lodi();
    return (r && (dir == 0 || dir == 1)) ? at_node(r->t[dir]) : NodeView();
}
NodeView NodeView::getTangentNeighbor2(int dir) {
    const acp_lodi_record* r = lodi();
    return (r && (dir == 0 || dir == 1)) ? at_node(r->t[2 + dir]) : NodeView();
}
// Node::getPartner (Node.h:71): the coincident node of the other level of a
// two-level lattice (coarse (I, J, K) with x0 <= I <= x1 and fine
// (2 (I - x0), 2 J, 2 K); on the shell geometry every coarse node and fine
// (2 I, 2 J, 2 K))
NodeView NodeView::getPartner() {
    if (!lat_ || !lat_->r_) return NodeView();
    if (lat_->shell_) {  // the shell geometry: coarse (I, J, K) and fine (2 I, 2 J, 2 K)
        if (level_ == 0) return NodeView(lat_, 1, 2 * i_, 2 * j_, 2 * k_);
        if ((i_ & 1) || (j_ & 1) || (k_ & 1)) return NodeView();
        return NodeView(lat_, 0, i_ / 2, j_ / 2, k_ / 2);
    }
    const int x0 = lat_->x0_, x1 = lat_->x1_;
    if (level_ == 0) {
        if (i_ < x0 || i_ > x1) return NodeView();
        return NodeView(lat_, 1, 2 * (i_ - x0), 2 * j_, 2 * k_);
    }
    if ((i_ & 1) || (j_ & 1) || (k_ & 1)) return NodeView();
    return NodeView(lat_, 0, x0 + i_ / 2, j_ / 2, k_ / 2);
}

}  // namespace acp
