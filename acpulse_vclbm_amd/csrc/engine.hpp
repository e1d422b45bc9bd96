// Internal declarations shared by the host-side translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "acp.h"
#include "k64.hpp"
#include "kshift.hpp"

namespace acp {

int fail(int code, const std::string& msg);

// Every device buffer of a case is accounted (and placed inside a caller's
// region) in multiples of 256 bytes.
inline int64_t round_alloc(int64_t bytes) { return (bytes + 255) / 256 * 256; }
void clear_error();
int hip_fail(hipError_t e, const char* where);

// Constants of one case, derived from the case description exactly as the
// reference derives them (Control.cpp:39-131, Lattice.cpp:1404, Node.cpp).
void k64_init(K64& k, const acp_case_desc& d);
void kshift_init(KShift& k, const acp_case_desc& d);

// Boundary-node schedule of the LODI pass: nodes in the reference's pass order
// (x faces, z faces, y faces; Lattice.cpp:1582-1624) grouped into phases so that
// a node reading another boundary node's B entries sees them exactly as the
// reference's sequential loops do.
// LODI work of one step: records [phase_start[p], phase_start[p+1]) form
// launch p, the nodes of dependency phase p + 1 of the reference's sequential
// pass order.
struct LodiSchedule {
    std::vector<LodiNode> nodes;
    std::vector<int64_t> phase_start;  // [nphase + 1]
};
LodiSchedule build_lodi_schedule(int nx, int ny, int nz);
// The LODI x faces of a grid wrapped in y and z (faces: bit 0 the face i = 0,
// bit 1 the face i = nx - 1; nx >= 4).
LodiSchedule build_lodi_schedule_x(int nx, int ny, int nz, int faces);

// The reference's initial pulse on one level of a lattice (acp_case_init_reference
// for level 0; refined cases for their fine level).
int init_reference_level(acp_case* c, int level, int i0, const acp_case_desc& d0);

// The nodes of x column `col` collide with a strain rate handed over by the
// other level of a two-level step (S: float64 planes [6][ny nz], SI units;
// with the cubic-Mach terms followed by the handed psi [19][ny nz], which
// replaces the nodes' psi state) and push as the bulk kernel does; launched
// right after the bulk kernel (refine.hip, kshift.hip).  op: OP_HRR, or OP_RR
// with the cubic-Mach terms.
hipError_t launch_fix64(int op, int order, bool cubic, const K64* k, const double* A, double* B, double* psi,
                        const Grid& g, int col, const double* S, hipStream_t s);
// The same for a list of nodes (none on a face of the grid), P planes
// [nv][n] over the list.
hipError_t launch_fix64_list(int op, int order, bool cubic, const K64* k, const double* A, double* B, double* psi,
                             const Grid& g, const int64_t* list, int64_t n, const double* P, hipStream_t s);
hipError_t launch_fix_hrr_shift(int prec, int order, const KShift* k, const void* A, void* B, const Grid& g, int col,
                                const double* S, double dt, hipStream_t s);

// Device memory a two-level case adds to one of its levels (counted in the
// level's footprint, placed in its region, freed with it).
int case_alloc_extra(acp_case* c, int64_t bytes, void** p);

// Launch n plain steps of a case on its stream (no graphs, no checks); the
// current array is c->A (c->SA) afterwards and the one read, c->B (c->SB).
int case_step_launch(acp_case* c, int32_t n);

// VTK output of one level from host float64 arrays (current array, the other
// array for the Non-Reflecting boundary quirk, cubic-Mach psi or null).
int write_vtk2_host(const acp_case_desc& cd, const K64& ck, const double* ccur, const double* cprev,
                    const acp_case_desc& fd, const K64& fk, const double* fcur, const double* fprev, int x0, int x1,
                    const char* stem, const double* cpsi = nullptr, const double* fpsi = nullptr);
// the same for the shell geometry (RefinementLayerYZ ly, RefinementLayerX lx)
int write_vtk_shell_host(const acp_case_desc& cd, const K64& ck, const double* ccur, const double* cprev,
                         const double* cpsi, const acp_case_desc& fd, const K64& fk, const double* fcur,
                         const double* fprev, const double* fpsi, int ly, int lx, const char* stem);
int write_vtk_host(const acp_case_desc& d, const K64& k, const double* fcur, const double* fprev, const double* psi,
                   const char* stem);

// Packed stepping (pack.cpp): whether a case may join a pack of a tier and
// order, and the removal of a member case that is being destroyed.
int pack_eligible(const acp_case* c, int32_t prec, int32_t order);
void pack_detach(acp_case* c);

}  // namespace acp

struct acp_case {
    acp_case_desc d;
    acp::Grid g;
    int precision;
    int bulk_op;
    bool periodic, need_macro, need_psi;
    // float64 tier
    acp::K64 hk;
    acp::K64* dk = nullptr;
    double *A = nullptr, *B = nullptr, *psi = nullptr, *M = nullptr;
    // shifted tiers
    acp::KShift hks;
    acp::KShift* dks = nullptr;
    void *SA = nullptr, *SB = nullptr;
    float* SM = nullptr;
    // LODI
    acp::LodiNode* lodi = nullptr;
    std::vector<int64_t> phase_start;
    // probes and reductions
    int64_t* probe_nodes = nullptr;
    double* probe_buf = nullptr;
    int32_t* probe_slot = nullptr;  // device row counter of the probe series
    int32_t nprobe = 0, probe_cap = 0, probe_count = 0;
    int64_t probe_bytes = 0;
    // staging buffer of the host transfers
    double* stage = nullptr;
    int64_t stage_elems = 0;
    // caller-owned device region the case lives in (acp_case_create_in), or null
    char* arena = nullptr;
    int64_t arena_size = 0, arena_used = 0, probe_mark = 0;
    // captured step chunks: one executable graph per ping-pong parity
    int32_t graph_chunk = 0;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    const void* arr0 = nullptr;  // the array that is current at parity 0
    double* partial = nullptr;
    int partial_blocks = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool initialized = false;
    int32_t steps = 0;
    // bumped whenever something a captured graph holds changes (probe buffers,
    // kernel form): graphs that other objects captured over this case are
    // stale once it differs from the value they recorded
    uint64_t graph_epoch = 0;
    // x columns whose nodes collide with a strain rate handed over by the other
    // level of a two-level HRR case (fix_S: float64 planes [6][ny nz], SI; with
    // the cubic-Mach terms, RR or HRR, followed by psi [19][ny nz])
    int32_t nfix = 0;
    int32_t fix_col[2] = {0, 0};
    const double* fix_S[2] = {nullptr, nullptr};
    // or a list of such nodes (the shell geometry of a two-level case, float64):
    // fix_list [fix_list_n] node indices off the grid's faces, fix_list_P
    // planes [nv][fix_list_n] over the list
    const int64_t* fix_list = nullptr;
    int64_t fix_list_n = 0;
    const double* fix_list_P = nullptr;
    std::vector<void*> extra;  // buffers of case_alloc_extra (freed with the case)
    // the pack that steps the case (acp_pack_add .. acp_pack_remove): its
    // arrays, step and probe counters are the pack's until then
    acp_pack* pack = nullptr;
    // collision matrix of the matrix-core MRT-RM form (BULK_MFMA), outside any region
    double* mfma_ct = nullptr;
    int64_t device_bytes = 0;
};
