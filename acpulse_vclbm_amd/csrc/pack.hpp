// Packed step launches: several one-level cases of one storage tier and
// equilibrium order advanced by the same kernel launches.
//
// A pack keeps a table of its member cases in device memory.  Every kernel of
// a packed step is a persistent grid that walks a work list of the table: the
// bulk collide-and-push launch of one operator walks the 256-node chunks of
// every member with that operator, a LODI launch the 256-record chunks of one
// phase of every member, and so on.  The step's last kernel swaps each
// member's two arrays and counts its steps down; a member whose count is zero
// is skipped by every kernel.  The host rewrites the table (in stream order,
// between step chunks) only when members join or leave, so one captured graph
// of a chunk of steps serves every membership with the same operators.
#pragma once

#include <cstdint>

#include "k64.hpp"

namespace acp {

constexpr int kPackSlots = 32;      // members of one pack
constexpr int kPackPhases = 8;      // LODI launches per step of one member
constexpr int kPackChunk = 256;     // nodes or LODI records per work item (one workgroup)
constexpr int kPackBulkGrid = 2048; // workgroups of a persistent bulk or macro launch
constexpr int kPackLodiGrid = 1024; // workgroups of a persistent LODI launch

struct PackSlot {
    const void* k;  // K64 (float64 tier) or KShift (shifted tiers) of the case
    void* A;        // populations at time t (read)
    void* B;        // populations at t + 1 (written); swapped with A after every step
    void* M;        // macro planes of the HRR pre-pass, or null
    const LodiNode* lodi;
    int64_t phase_start[kPackPhases + 1];  // LODI launch p: records [phase_start[p], phase_start[p + 1])
    const int64_t* probe_nodes;
    double* probe_buf;
    int32_t* probe_slot;
    int32_t nprobe, probe_cap;
    PopLayout L;
    int32_t nx, ny, nz, op;
    int32_t steps_left;  // steps still to take; 0: skipped
    int32_t transverse;
};

// Work list of one launch: members slot[0..n) in order, member m owning the
// work items [prefix[m], prefix[m + 1]).
struct PackList {
    int32_t n;
    int32_t slot[kPackSlots];
    int32_t prefix[kPackSlots + 1];
};

struct PackTable {
    PackList bulk[6];              // one per operator code (OP_SRT .. OP_STREAM); HRR's also drives the pre-pass
    PackList lodi[kPackPhases];
    int32_t nslot;                 // slots [0, nslot) may hold members
    PackSlot slot[kPackSlots];
};

// Persistent launches over the table (k64_bulk.hip, k64_aux.hip, kshift.hip).
hipError_t launch_bulk64_packed(int op, int order, const PackTable* t, hipStream_t s);
hipError_t launch_macro64_packed(const PackTable* t, hipStream_t s);
hipError_t launch_lodi64_packed(int order, int phase, const PackTable* t, hipStream_t s);
hipError_t launch_probe64_packed(const PackTable* t, hipStream_t s);
hipError_t launch_bulk_shift_packed(int prec, int op, int order, const PackTable* t, hipStream_t s);
hipError_t launch_macro_shift_packed(int prec, const PackTable* t, hipStream_t s);
hipError_t launch_lodi_shift_packed(int prec, int order, int phase, const PackTable* t, hipStream_t s);
hipError_t launch_probe_shift_packed(int prec, const PackTable* t, hipStream_t s);
// the step's last kernel: swap the arrays of every member with steps left and count one down
hipError_t launch_pack_advance(PackTable* t, hipStream_t s);

// Locate work item w of list pl, starting the search at member m (work items
// are visited in increasing order by each workgroup).
__device__ __forceinline__ int pack_member(const PackList& pl, int w, int m) {
    while (pl.prefix[m + 1] <= w) ++m;
    return m;
}

}  // namespace acp
