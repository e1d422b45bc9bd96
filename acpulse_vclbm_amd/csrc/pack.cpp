// Packed stepping of several resident one-level cases (acp_pack_*).
//
// The reference steps one case per process (main.cpp:37-271); this component
// has no counterpart there.  A pack keeps a member table in device memory
// (pack.hpp) and advances all members with shared persistent launches: per
// step the HRR pre-pass, one bulk launch per operator present, one launch per
// LODI phase, one probe launch and the bookkeeping launch that swaps each
// member's arrays and counts its steps down.  The host keeps an exact mirror
// of what the device table holds after the launches it has issued (each
// member's current array and steps left), and rewrites the table in stream
// order only when members join or leave.  Because the launches read the table
// from device memory, a captured graph of a chunk of steps stays valid across
// membership changes; graphs are cached per set of launches (operators
// present, LODI phases, probes).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "acp.h"
#include "engine.hpp"
#include "pack.hpp"

#define HIPCHK_RET(call)                                    \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return acp::hip_fail(e_, #call); \
    } while (0)

namespace {

using acp::fail;

constexpr int kRing = 8;      // pinned staging copies of the table
constexpr int kEvents = 64;   // events marking the ends of launched chunks

struct Member {
    acp_case* c = nullptr;
    int32_t total = 0;     // steps given at acp_pack_add
    int32_t left = 0;      // steps not yet launched
    void* cur = nullptr;   // array holding time t after the launched steps
    void* nxt = nullptr;
    int64_t last_event = -1;  // sequence number of the chunk that launched its last step
};

}  // namespace

struct acp_pack {
    int32_t prec = 0, order = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    acp::PackTable* dtab = nullptr;
    acp::PackTable* ring[kRing] = {};
    hipEvent_t ring_ev[kRing] = {};
    bool ring_used[kRing] = {};
    int ring_next = 0;
    hipEvent_t ev[kEvents] = {};
    int64_t ev_seq = 0;  // chunks launched
    Member slot[acp::kPackSlots];
    bool dirty = true;
    int32_t graph_chunk = 0;
    std::map<uint32_t, hipGraphExec_t> graphs;  // by launch-set key
};

namespace {

int hipchk(hipError_t e, const char* where) { return e == hipSuccess ? ACP_OK : acp::hip_fail(e, where); }

bool is_shifted(const acp_pack* p) { return p->prec != ACP_PREC_F64; }
int elem_bytes(int prec) { return prec == ACP_PREC_F64 ? 8 : (prec == ACP_PREC_F32_SHIFTED ? 4 : 2); }

void* case_cur(const acp_case* c) { return c->precision == ACP_PREC_F64 ? (void*)c->A : c->SA; }
void* case_nxt(const acp_case* c) { return c->precision == ACP_PREC_F64 ? (void*)c->B : c->SB; }

int64_t chunks(int64_t n) { return (n + acp::kPackChunk - 1) / acp::kPackChunk; }

// Launch set of the next chunk: bits 0-5 the bulk operators present, bits 8-11
// the LODI launches per step, bit 12 probes.  Members with no step left are
// skipped by every kernel and count for nothing here.
uint32_t launch_key(const acp_pack* p) {
    uint32_t ops = 0, nph = 0, probes = 0;
    for (const Member& m : p->slot) {
        if (!m.c || m.left <= 0) continue;
        ops |= 1u << m.c->bulk_op;
        if (!m.c->phase_start.empty()) nph = std::max<uint32_t>(nph, (uint32_t)m.c->phase_start.size() - 1);
        if (m.c->nprobe > 0) probes = 1;
    }
    return ops | (nph << 8) | (probes << 12);
}

// Write the mirror of the member table into the next pinned copy and queue
// its upload on the pack's stream (after every chunk launched so far, so the
// device sees exactly the mirrored state).
int upload_table(acp_pack* p) {
    const int r = p->ring_next;
    if (p->ring_used[r]) HIPCHK_RET(hipEventSynchronize(p->ring_ev[r]));
    acp::PackTable* t = p->ring[r];
    std::memset(t, 0, sizeof(*t));
    int32_t nslot = 0;
    for (int s = 0; s < acp::kPackSlots; ++s) {
        const Member& m = p->slot[s];
        if (!m.c) continue;
        nslot = s + 1;
        const acp_case* c = m.c;
        acp::PackSlot& e = t->slot[s];
        e.k = is_shifted(p) ? (const void*)c->dks : (const void*)c->dk;
        e.A = m.cur;
        e.B = m.nxt;
        e.M = is_shifted(p) ? (void*)c->SM : (void*)c->M;
        e.lodi = c->lodi;
        const int nph = c->phase_start.empty() ? 0 : (int)c->phase_start.size() - 1;
        for (int q = 0; q <= acp::kPackPhases && nph > 0; ++q) e.phase_start[q] = c->phase_start[std::min(q, nph)];
        e.probe_nodes = c->probe_nodes;
        e.probe_buf = c->probe_buf;
        e.probe_slot = c->probe_slot;
        e.nprobe = c->nprobe;
        e.probe_cap = c->probe_cap;
        e.L = c->g.L;
        e.nx = c->g.nx;
        e.ny = c->g.ny;
        e.nz = c->g.nz;
        e.op = c->bulk_op;
        e.steps_left = m.left;
        e.transverse = c->d.transverse != 0;
        if (m.left <= 0) continue;
        acp::PackList& bl = t->bulk[c->bulk_op];
        bl.slot[bl.n] = s;
        bl.prefix[bl.n + 1] = bl.prefix[bl.n] + (int32_t)chunks(c->g.N);
        ++bl.n;
        for (int q = 0; q < nph; ++q) {
            const int64_t cnt = c->phase_start[q + 1] - c->phase_start[q];
            if (cnt <= 0) continue;
            acp::PackList& ll = t->lodi[q];
            ll.slot[ll.n] = s;
            ll.prefix[ll.n + 1] = ll.prefix[ll.n] + (int32_t)chunks(cnt);
            ++ll.n;
        }
    }
    t->nslot = nslot;
    HIPCHK_RET(hipMemcpyAsync(p->dtab, t, sizeof(*t), hipMemcpyHostToDevice, p->stream));
    HIPCHK_RET(hipEventRecord(p->ring_ev[r], p->stream));
    p->ring_used[r] = true;
    p->ring_next = (r + 1) % kRing;
    p->dirty = false;
    return ACP_OK;
}

// One packed step's launches for the launch set `key`.
hipError_t launch_step(acp_pack* p, uint32_t key) {
    const int nph = (int)((key >> 8) & 0xf);
    hipError_t e = hipSuccess;
    if (key & (1u << acp::OP_HRR))
        e = is_shifted(p) ? acp::launch_macro_shift_packed(p->prec, p->dtab, p->stream)
                          : acp::launch_macro64_packed(p->dtab, p->stream);
    for (int op = 0; op < 6 && e == hipSuccess; ++op) {
        if (!(key & (1u << op))) continue;
        e = is_shifted(p) ? acp::launch_bulk_shift_packed(p->prec, op, p->order, p->dtab, p->stream)
                          : acp::launch_bulk64_packed(op, p->order, p->dtab, p->stream);
    }
    for (int q = 0; q < nph && e == hipSuccess; ++q)
        e = is_shifted(p) ? acp::launch_lodi_shift_packed(p->prec, p->order, q, p->dtab, p->stream)
                          : acp::launch_lodi64_packed(p->order, q, p->dtab, p->stream);
    if (e == hipSuccess && (key & (1u << 12)))
        e = is_shifted(p) ? acp::launch_probe_shift_packed(p->prec, p->dtab, p->stream)
                          : acp::launch_probe64_packed(p->dtab, p->stream);
    if (e == hipSuccess) e = acp::launch_pack_advance(p->dtab, p->stream);
    return e;
}

// Launch n steps (graph replay when n is the graph chunk), then advance the
// host mirror of every member.
int launch_chunk(acp_pack* p, int32_t n) {
    if (p->dirty)
        if (int rc = upload_table(p)) return rc;
    const uint32_t key = launch_key(p);
    if (p->graph_chunk > 0 && n == p->graph_chunk) {
        hipGraphExec_t& gx = p->graphs[key];
        if (!gx) {
            hipGraph_t g = nullptr;
            HIPCHK_RET(hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed));
            hipError_t e = hipSuccess;
            for (int32_t s = 0; s < n && e == hipSuccess; ++s) e = launch_step(p, key);
            const hipError_t e2 = hipStreamEndCapture(p->stream, &g);
            if (e == hipSuccess) e = e2;
            if (e == hipSuccess) e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                gx = nullptr;
                p->graphs.erase(key);
                return acp::hip_fail(e, "packed step graph capture");
            }
        }
        HIPCHK_RET(hipGraphLaunch(gx, p->stream));
    } else {
        for (int32_t s = 0; s < n; ++s) HIPCHK_RET(launch_step(p, key));
    }
    const int64_t seq = p->ev_seq++;
    HIPCHK_RET(hipEventRecord(p->ev[seq % kEvents], p->stream));
    for (Member& m : p->slot) {
        if (!m.c || m.left <= 0) continue;
        const int32_t k = std::min(n, m.left);
        m.left -= k;
        if (k & 1) std::swap(m.cur, m.nxt);
        if (m.left == 0) m.last_event = seq;
    }
    return ACP_OK;
}

Member* find(acp_pack* p, const acp_case* c) {
    for (Member& m : p->slot)
        if (m.c == c) return &m;
    return nullptr;
}

void destroy(acp_pack* p) {
    for (auto& kv : p->graphs)
        if (kv.second) (void)hipGraphExecDestroy(kv.second);
    for (int r = 0; r < kRing; ++r) {
        if (p->ring[r]) (void)hipHostFree(p->ring[r]);
        if (p->ring_ev[r]) (void)hipEventDestroy(p->ring_ev[r]);
    }
    for (hipEvent_t e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->dtab) (void)hipFree(p->dtab);
    if (p->own_stream && p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
}

}  // namespace

namespace acp {
// Whether case c may join a pack of tier `prec` and order `order` (message in
// the thread's error text when not).
int pack_eligible(const acp_case* c, int32_t prec, int32_t order) {
    if (!c) return fail(ACP_E_ARG, "null case handle");
    if (c->precision != prec || c->d.order != order)
        return fail(ACP_E_ARG, "acp_pack_add: the case's tier and equilibrium order differ from the pack's");
    if (c->d.boundary != ACP_BC_LODI)
        return fail(ACP_E_UNSUPPORTED, "acp_pack_add: packs hold cases with LODI on every face (ACP_BC_LODI)");
    if (c->d.cubic) return fail(ACP_E_UNSUPPORTED, "acp_pack_add: no cubic-Mach correction in packs");
    if (c->nfix > 0) return fail(ACP_E_UNSUPPORTED, "acp_pack_add: a level of a two-level case cannot join a pack");
    if ((int64_t)c->phase_start.size() - 1 > kPackPhases)
        return fail(ACP_E_UNSUPPORTED, "acp_pack_add: more LODI launches per step than a pack holds");
    const int eb = elem_bytes(prec), pb = prec == ACP_PREC_F64 ? 8 : 4;
    if (!narrow_planes(c->g.L, eb, pb) || c->g.N >= ((int64_t)1 << 31) ||
        chunks(c->g.N) * kPackSlots >= ((int64_t)1 << 31))
        return fail(ACP_E_UNSUPPORTED, "acp_pack_add: case too large for a pack");
    return ACP_OK;
}

// A member case destroyed while in its pack: the pack finishes its queued
// work and forgets the member.
void pack_detach(acp_case* c) {
    acp_pack* p = c->pack;
    if (!p) return;
    (void)hipStreamSynchronize(p->stream);
    if (Member* m = find(p, c)) *m = Member();
    c->pack = nullptr;
    p->dirty = true;
}
}  // namespace acp

extern "C" int acp_pack_create(int32_t precision, int32_t order, void* stream, acp_pack** out) {
    if (!out) return fail(ACP_E_ARG, "acp_pack_create: null output");
    *out = nullptr;
    if (precision < ACP_PREC_F64 || precision > ACP_PREC_F16_SHIFTED || order < 2 || order > 4)
        return fail(ACP_E_ARG, "acp_pack_create: tier ACP_PREC_* and equilibrium order 2, 3 or 4");
    acp_pack* p = new acp_pack();
    p->prec = precision;
    p->order = order;
    hipError_t e = hipSuccess;
    if (stream) {
        p->stream = (hipStream_t)stream;
    } else {
        e = hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking);
        p->own_stream = e == hipSuccess;
    }
    if (e == hipSuccess) e = hipMalloc((void**)&p->dtab, sizeof(acp::PackTable));
    for (int r = 0; r < kRing && e == hipSuccess; ++r) {
        e = hipHostMalloc((void**)&p->ring[r], sizeof(acp::PackTable), hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->ring_ev[r], hipEventDisableTiming);
    }
    for (int i = 0; i < kEvents && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&p->ev[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        destroy(p);
        return acp::hip_fail(e, "acp_pack_create");
    }
    *out = p;
    acp::clear_error();
    return ACP_OK;
}

extern "C" void acp_pack_destroy(acp_pack* p) {
    if (!p) return;
    (void)hipStreamSynchronize(p->stream);
    for (Member& m : p->slot)
        if (m.c) (void)acp_pack_remove(p, m.c);
    destroy(p);
}

extern "C" int acp_pack_add(acp_pack* p, acp_case* c, int32_t steps) {
    if (!p) return fail(ACP_E_ARG, "acp_pack_add: null pack");
    if (int rc = acp::pack_eligible(c, p->prec, p->order)) return rc;
    if (!c->initialized) return fail(ACP_E_STATE, "acp_pack_add: case not initialised");
    if (c->pack) return fail(ACP_E_STATE, "acp_pack_add: case already in a pack");
    if (steps < 0) return fail(ACP_E_ARG, "acp_pack_add: negative step count");
    if (c->nprobe > 0 && c->probe_count + steps > c->probe_cap)
        return fail(ACP_E_STATE, "acp_pack_add: probe buffer too small for the steps: read the probes first");
    Member* m = find(p, nullptr);
    if (!m) return fail(ACP_E_STATE, "acp_pack_add: pack full");
    // the case's own work (initial condition, probe set-up) completes before
    // the pack's stream touches its arrays
    if (int rc = hipchk(hipStreamSynchronize(c->stream), "acp_pack_add")) return rc;
    m->c = c;
    m->total = steps;
    m->left = steps;
    m->cur = case_cur(c);
    m->nxt = case_nxt(c);
    m->last_event = -1;
    c->pack = p;
    p->dirty = true;
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_pack_step(acp_pack* p, int32_t n) {
    if (!p) return fail(ACP_E_ARG, "acp_pack_step: null pack");
    if (n < 0) return fail(ACP_E_ARG, "acp_pack_step: negative step count");
    while (n > 0) {
        bool any = false;
        for (const Member& m : p->slot) any = any || (m.c && m.left > 0);
        if (!any) break;
        const int32_t k = p->graph_chunk > 0 && n >= p->graph_chunk ? p->graph_chunk : n;
        if (int rc = launch_chunk(p, k)) return rc;
        n -= k;
    }
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_pack_steps_left(const acp_pack* p, const acp_case* c, int32_t* left) {
    if (!p || !c || !left) return fail(ACP_E_ARG, "acp_pack_steps_left: null argument");
    for (const Member& m : p->slot)
        if (m.c == c) {
            *left = m.left;
            return ACP_OK;
        }
    return fail(ACP_E_ARG, "acp_pack_steps_left: case not in this pack");
}

extern "C" int acp_pack_remove(acp_pack* p, acp_case* c) {
    if (!p || !c) return fail(ACP_E_ARG, "acp_pack_remove: null argument");
    Member* m = find(p, c);
    if (!m) return fail(ACP_E_ARG, "acp_pack_remove: case not in this pack");
    // every launch that touched the member has finished (the event of a later
    // chunk, should the ring have reused it, is later still)
    const int64_t seq = m->last_event >= 0 ? m->last_event : p->ev_seq - 1;
    if (seq >= 0) {
        if (int rc = hipchk(hipEventSynchronize(p->ev[seq % kEvents]), "acp_pack_remove")) return rc;
    }
    const int32_t done = m->total - m->left;
    if (c->precision == ACP_PREC_F64) {
        c->A = (double*)m->cur;
        c->B = (double*)m->nxt;
    } else {
        c->SA = m->cur;
        c->SB = m->nxt;
    }
    c->steps += done;
    if (c->nprobe > 0) c->probe_count += done;
    c->pack = nullptr;
    *m = Member();
    p->dirty = true;
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_pack_enable_graphs(acp_pack* p, int32_t chunk) {
    if (!p) return fail(ACP_E_ARG, "acp_pack_enable_graphs: null pack");
    if (chunk < 0) return fail(ACP_E_ARG, "acp_pack_enable_graphs: chunk must be 0 (off) or positive");
    p->graph_chunk = chunk;
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_pack_members(const acp_pack* p, int32_t* members) {
    if (!p || !members) return fail(ACP_E_ARG, "acp_pack_members: null argument");
    int32_t n = 0;
    for (const Member& m : p->slot) n += m.c != nullptr;
    *members = n;
    return ACP_OK;
}
