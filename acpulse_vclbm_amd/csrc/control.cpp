// Case-file reader with the semantics of the reference's Control class
// (Control.cpp:8-273): every key is looked up by scanning the file from the top
// and taking the first line whose first whitespace-separated token equals the
// key; numbers are read with stream extraction (so "1.17621;" and
// "1.49e-5 //!air" parse), strings are the next token, the Save-as path is the
// rest of the line.  A missing mandatory key is an error.
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "acp.h"
#include "engine.hpp"

namespace {

struct Reader {
    std::string path;
    std::string err;

    bool line_for(const std::string& key, std::string& line, std::istringstream& rest) {
        std::ifstream fh(path);
        if (!fh) {
            err = "Error reading boundary conditions file " + path;
            return false;
        }
        std::string s;
        while (std::getline(fh, s)) {
            std::istringstream sin(s);
            std::string op;
            sin >> op;
            if (op == key) {
                line = s;
                rest.str(s);
                rest.clear();
                std::string skip;
                rest >> skip;
                return true;
            }
        }
        return false;
    }
    bool value(const std::string& key, double& out, bool required = true) {
        std::string line;
        std::istringstream rest;
        if (!line_for(key, line, rest)) {
            if (required && err.empty()) err = "No Value for keyword \"" + key + "\" found in " + path;
            return false;
        }
        double v = 0.;
        rest >> v;
        if (rest.fail()) v = 0.;
        out = v;
        return true;
    }
    bool str(const std::string& key, std::string& out, bool required = true) {
        std::string line;
        std::istringstream rest;
        if (!line_for(key, line, rest)) {
            if (required && err.empty()) err = "No Value for keyword \"" + key + "\" found in " + path;
            return false;
        }
        rest >> out;
        return true;
    }
    bool rest_of_line(const std::string& key, std::string& out) {
        std::string line;
        std::istringstream rest;
        if (!line_for(key, line, rest)) {
            if (err.empty()) err = "No Value for keyword \"" + key + "\" found in " + path;
            return false;
        }
        std::string t = line.substr(key.size());
        const size_t p = t.find_first_not_of(' ');
        out = p == std::string::npos ? std::string() : t.substr(p);
        return true;
    }
};

void copy_str(char* dst, size_t cap, const std::string& s) {
    std::strncpy(dst, s.c_str(), cap - 1);
    dst[cap - 1] = 0;
}

}  // namespace

extern "C" void acp_case_desc_default(acp_case_desc* d) {
    std::memset(d, 0, sizeof(*d));
    // the reference's shipped BoundaryConditions.txt
    d->timesteps = 100;
    d->write_interval = 10;
    d->threads = 16;
    d->rho0 = 1.17621;
    d->nu = 1.49e-5;
    d->cs = 347.3;
    d->mach = 0.0;
    d->order = 3;
    d->op = ACP_OP_HRR;
    d->sigma = 0.98;
    d->cubic = 0;
    d->nonreflecting = 1;
    d->transverse = 0;
    d->k1 = 1.0;
    d->k2 = 1.0;
    d->domain[0] = 3.;
    d->domain[1] = 3.;
    d->domain[2] = 0.4;
    d->spacing = 2.0e-2;
    d->max_level = 0;
    d->boundary = ACP_BC_LODI;
    d->mrt_mode = ACP_MRT_AS_REFERENCE;
    d->precision = ACP_PREC_F64;
    d->f16_scale = 1000.0;
    const double eps = 1e-8;
    d->nx = (int32_t)std::ceil(d->domain[0] / d->spacing + eps);
    d->ny = (int32_t)std::ceil(d->domain[1] / d->spacing + eps);
    d->nz = (int32_t)std::ceil(d->domain[2] / d->spacing + eps);
    copy_str(d->save_as, sizeof(d->save_as), "/path-to-results-dir/results-filename");
    copy_str(d->collision_model, sizeof(d->collision_model), "HRR");
}

extern "C" int acp_control_read(const char* path, acp_case_desc* out) {
    if (!path || !out) return acp::fail(ACP_E_ARG, "acp_control_read: null argument");
    acp_case_desc d;
    acp_case_desc_default(&d);
    Reader r{path, ""};
    double v;
    std::string s;
    // mandatory keys, in the order of Control.h's member declarations
    if (!r.value("Max-Time-Step", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.timesteps = (int32_t)v;
    if (!r.value("WriteInterval", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.write_interval = (int32_t)v;
    if (!r.value("Density", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.rho0 = v;
    if (!r.value("Threads", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.threads = (int32_t)v;
    if (!r.value("KinematicViscosity", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.nu = v;
    if (!r.value("SpeedOfSound", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.cs = v;
    if (!r.rest_of_line("Save-as", s)) return acp::fail(ACP_E_CONFIG, r.err);
    copy_str(d.save_as, sizeof(d.save_as), s);
    if (!r.str("Collision-Model", s)) return acp::fail(ACP_E_CONFIG, r.err);
    copy_str(d.collision_model, sizeof(d.collision_model), s);
    if (s == "SRT") d.op = ACP_OP_SRT;
    else if (s == "MRT-GS") d.op = ACP_OP_MRT_GS;
    else if (s == "MRT-RM") d.op = ACP_OP_MRT_RM;
    else if (s == "RR") d.op = ACP_OP_RR;
    else if (s == "HRR") d.op = ACP_OP_HRR;
    else return acp::fail(ACP_E_CONFIG, "unknown Collision-Model \"" + s + "\"");
    if (!r.value("Order-Of-Equilibrium", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.order = (int32_t)v;
    if (d.order < 2 || d.order > 4) return acp::fail(ACP_E_CONFIG, "Wrong order of equilibrium!");
    if (!r.str("Cubic-Mach-Correction", s)) return acp::fail(ACP_E_CONFIG, r.err);
    d.cubic = s == "yes";
    const char* strings[] = {"Coarse-to-Fine", "Fine-to-Coarse", "Hanging-Node-Reconstruction", "Scaling-Style",
                             "F2C-Filter", "Coupling-Approach"};
    for (const char* key : strings)
        if (!r.str(key, s)) return acp::fail(ACP_E_CONFIG, r.err);
    if (!r.value("K1-Relaxation", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.k1 = v;
    if (!r.value("K2-Relaxation", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.k2 = v;
    if (!r.value("Hybrid-Parameter", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.sigma = v;
    if (!r.str("Non-Reflecting", s)) return acp::fail(ACP_E_CONFIG, r.err);
    d.nonreflecting = s == "yes";
    if (!r.str("Transverse", s)) return acp::fail(ACP_E_CONFIG, r.err);
    d.transverse = s == "yes";
    const char* dom[] = {"Domain-Size-X", "Domain-Size-Y", "Domain-Size-Z"};
    for (int a = 0; a < 3; ++a) {
        if (!r.value(dom[a], v)) return acp::fail(ACP_E_CONFIG, r.err);
        d.domain[a] = v;
    }
    if (!r.value("Spacing", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.spacing = v;
    for (const char* key : {"RefinementLayerYZ", "RefinementLayerX"})
        if (!r.value(key, v)) return acp::fail(ACP_E_CONFIG, r.err);
    if (!r.value("MaxLevel", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.max_level = (int32_t)v;
    if (!r.value("MACH-Number", v)) return acp::fail(ACP_E_CONFIG, r.err);
    d.mach = v;
    if (!r.value("Overlap", v)) return acp::fail(ACP_E_CONFIG, r.err);
    // Control.cpp:43-45
    const double eps = 1e-8;
    d.nx = (int32_t)std::ceil(d.domain[0] / d.spacing + eps);
    d.ny = (int32_t)std::ceil(d.domain[1] / d.spacing + eps);
    d.nz = (int32_t)std::ceil(d.domain[2] / d.spacing + eps);
    // optional keys of this engine (absent from the reference's key set, so a
    // reference case file selects the reference behaviour)
    if (r.str("Precision", s, false)) {
        if (s == "F64") d.precision = ACP_PREC_F64;
        else if (s == "F32-SHIFTED") d.precision = ACP_PREC_F32_SHIFTED;
        else if (s == "F16-SHIFTED") d.precision = ACP_PREC_F16_SHIFTED;
        else return acp::fail(ACP_E_CONFIG, "unknown Precision \"" + s + "\"");
    }
    if (r.str("Boundary", s, false)) {
        if (s == "LODI") d.boundary = ACP_BC_LODI;
        else if (s == "PERIODIC") d.boundary = ACP_BC_PERIODIC;
        else if (s == "LODI-X") d.boundary = ACP_BC_LODI_X;
        else return acp::fail(ACP_E_CONFIG, "unknown Boundary \"" + s + "\"");
    }
    if (r.str("MRT-Mode", s, false)) {
        if (s == "REFERENCE") d.mrt_mode = ACP_MRT_AS_REFERENCE;
        else if (s == "RELAXED") d.mrt_mode = ACP_MRT_RELAXED;
        else return acp::fail(ACP_E_CONFIG, "unknown MRT-Mode \"" + s + "\"");
    }
    if (r.value("FP16-Scale", v, false)) d.f16_scale = v;
    *out = d;
    acp::clear_error();
    return ACP_OK;
}

// Control::readValue / readString / readFileName / readVector
// (Control.cpp:135-273) for any keyword: the first line whose first token is
// the keyword; a missing file or keyword is ACP_E_CONFIG (the reference prints
// the same message and exits).
extern "C" int acp_control_read_value(const char* path, const char* keyword, double* out) {
    if (!path || !keyword || !out) return acp::fail(ACP_E_ARG, "acp_control_read_value: null argument");
    Reader r{path, {}};
    if (!r.value(keyword, *out)) return acp::fail(ACP_E_CONFIG, r.err);
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_control_read_string(const char* path, const char* keyword, char* out, int64_t cap) {
    if (!path || !keyword || !out || cap <= 0) return acp::fail(ACP_E_ARG, "acp_control_read_string: bad argument");
    Reader r{path, {}};
    std::string s;
    if (!r.str(keyword, s)) return acp::fail(ACP_E_CONFIG, r.err);
    if ((int64_t)s.size() >= cap) return acp::fail(ACP_E_ARG, "acp_control_read_string: output buffer too small");
    copy_str(out, (size_t)cap, s);
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_control_read_filename(const char* path, const char* keyword, char* out, int64_t cap) {
    if (!path || !keyword || !out || cap <= 0) return acp::fail(ACP_E_ARG, "acp_control_read_filename: bad argument");
    Reader r{path, {}};
    std::string s;
    if (!r.rest_of_line(keyword, s)) return acp::fail(ACP_E_CONFIG, r.err);
    if ((int64_t)s.size() >= cap) return acp::fail(ACP_E_ARG, "acp_control_read_filename: output buffer too small");
    copy_str(out, (size_t)cap, s);
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_control_read_vector(const char* path, const char* keyword, double* out3) {
    if (!path || !keyword || !out3) return acp::fail(ACP_E_ARG, "acp_control_read_vector: null argument");
    Reader r{path, {}};
    std::string line;
    std::istringstream rest;
    if (!r.line_for(keyword, line, rest)) {
        if (r.err.empty()) r.err = std::string("No Value for keyword \"") + keyword + "\" found in " + path;
        return acp::fail(ACP_E_CONFIG, r.err);
    }
    double x = 0., y = 0., z = 0.;
    rest >> x >> y >> z;
    out3[0] = x;
    out3[1] = y;
    out3[2] = z;
    acp::clear_error();
    return ACP_OK;
}
