// Serving loop: many independent acoustic-pulse cases resident on one GPU.
//
// The reference runs one case per process (main.cpp:37-271); this component
// has no counterpart there.  Requests carry a case description, an initial
// condition, a step count and an arrival time.  The server
//
//   * admits requests in arrival order once they have arrived, a stream slot
//     is free and their HBM footprint (acp_case_footprint) fits a free range of
//     one device arena allocated up front, so that creating and retiring cases
//     never calls hipMalloc/hipFree (which would wait on every stream);
//   * creates and initialises admitted cases on an admission thread;
//   * drives every resident case on its own non-blocking stream from a driver
//     thread, keeping a bounded number of step chunks (captured HIP graphs)
//     queued per case, so that the cases' kernels overlap on the device; or,
//     with packing on, steps the eligible small one-level cases of one tier
//     and order together (acp_pack_*): one pack per tier and order, on a
//     stream of its own, with the same bounded number of chunks queued per
//     pack;
//   * on completion reads the probe series, mass and state checksum, releases
//     the case's arena range and slot, and posts a result with its times.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "acp.h"
#include "engine.hpp"
#include "pack.hpp"

namespace {

// First-fit allocator of [0, size) with coalescing free ranges.
class RangeAllocator {
  public:
    void reset(int64_t size) {
        free_.clear();
        if (size > 0) free_[0] = size;
    }
    int64_t alloc(int64_t bytes) {
        for (auto it = free_.begin(); it != free_.end(); ++it)
            if (it->second >= bytes) {
                const int64_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > bytes) free_[off + bytes] = len - bytes;
                return off;
            }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<int64_t, int64_t> free_;
};

struct Job {
    acp_request req;
    int64_t bytes = 0;   // footprint
    int64_t off = -1;    // arena offset
    int slot = -1;
    acp_case* c = nullptr;      // the case, or the coarse level of r
    acp_refined* r = nullptr;   // two-level case
    int32_t launched = 0;
    std::deque<hipEvent_t> inflight;
    double t_submit = 0, t_admit = 0, t_start = 0;
    int status = ACP_OK;
    std::string error;
    int64_t seq = 0;  // submission number
};

// The eligible one-level cases of one tier and order, stepped together.
struct PackRun {
    int32_t prec = 0, order = 0;
    acp_pack* p = nullptr;
    hipStream_t stream = nullptr;
    std::vector<Job> members;
    // launched chunks: completion event and the members whose last step it holds
    std::deque<std::pair<hipEvent_t, std::vector<int64_t>>> inflight;
};

}  // namespace

struct acp_server {
    acp_server_desc cfg{};
    int device = 0;
    char* arena = nullptr;
    RangeAllocator ranges;
    std::vector<hipStream_t> streams;
    std::vector<int> free_slots;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_admit, cv_drive, cv_done;
    std::deque<Job> queue;          // submitted, sorted by arrival
    std::deque<Job> ready;          // created and initialised, waiting for the driver
    std::vector<acp_result> done;   // results not yet polled
    bool stopping = false;
    int admit_running = 0;  // admission threads still running
    int64_t submitted = 0, completed = 0, bytes_in_use = 0, peak_bytes = 0;
    int32_t in_admission = 0, running = 0, peak_active = 0;

    std::vector<std::thread> admit_threads;
    std::thread drive_thread;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

namespace {

void post_result(acp_server* s, Job& j, double t_done, double mass, uint64_t checksum) {
    acp_result r;
    std::memset(&r, 0, sizeof(r));
    r.tag = j.req.tag;
    r.status = j.status;
    r.steps = j.launched;
    r.device_bytes = j.bytes;
    r.t_arrival = j.req.arrival;
    r.t_submit = j.t_submit;
    r.t_admit = j.t_admit;
    r.t_start = j.t_start;
    r.t_done = t_done;
    r.mass = mass;
    r.checksum = checksum;
    std::strncpy(r.error, j.error.c_str(), sizeof(r.error) - 1);
    std::lock_guard<std::mutex> lk(s->mu);
    if (j.off >= 0) {
        s->ranges.release(j.off, j.bytes);
        s->bytes_in_use -= j.bytes;
    }
    if (j.slot >= 0) s->free_slots.push_back(j.slot);
    s->done.push_back(r);
    ++s->completed;
    s->cv_done.notify_all();
    s->cv_admit.notify_all();
}

int init_case(acp_case* c, const acp_request& q);

bool two_level(const acp_request& q) { return q.refine[1] > q.refine[0]; }

void destroy_job_case(Job& j) {
    if (j.r) acp_refined_destroy(j.r);
    else if (j.c) acp_case_destroy(j.c);
    j.r = nullptr;
    j.c = nullptr;
}

// Create and initialise the job's case inside its arena range.
int create_job_case(acp_server* s, Job& j) {
    const acp_request& q = j.req;
    int rc;
    if (two_level(q)) {
        if (q.init_kind != ACP_INIT_PULSE)
            return acp::fail(ACP_E_UNSUPPORTED, "two-level requests start from a pulse (ACP_INIT_PULSE)");
        rc = acp_refined_create_in(&q.desc, q.refine[0], q.refine[1], s->streams[j.slot], s->arena + j.off, j.bytes,
                                   &j.r);
        if (rc == ACP_OK) rc = acp_refined_level(j.r, 0, &j.c);
        if (rc == ACP_OK) rc = acp_refined_set_filter(j.r, q.f2c_filter);
        if (rc == ACP_OK) {
            const int32_t links = q.coupling == ACP_COUPLING_DC ? ACP_LINKS_MISSING : ACP_LINKS_ALL;
            rc = acp_refined_set_coupling(j.r, q.coupling, links, links, links);
        }
        if (rc == ACP_OK) rc = acp_refined_init_pulse(j.r, q.amp, q.x0, q.y0, q.halfwidth);
    } else {
        rc = acp_case_create_in(&q.desc, s->streams[j.slot], s->arena + j.off, j.bytes, &j.c);
        if (rc == ACP_OK) rc = init_case(j.c, q);
    }
    if (rc == ACP_OK && q.probe[0] >= 0) {
        const acp_case_desc& d = q.desc;
        const int64_t node = (int64_t)q.probe[0] + (int64_t)d.nx * ((int64_t)q.probe[1] + (int64_t)d.ny * q.probe[2]);
        rc = acp_case_set_probes(j.c, &node, 1, std::max(q.steps, 1));
    }
    if (rc == ACP_OK && s->cfg.graph_chunk > 0)
        rc = j.r ? acp_refined_enable_graphs(j.r, s->cfg.graph_chunk) : acp_case_enable_graphs(j.c, s->cfg.graph_chunk);
    return rc;
}

int init_case(acp_case* c, const acp_request& q) {
    switch (q.init_kind) {
        case ACP_INIT_REFERENCE: return acp_case_init_reference(c);
        case ACP_INIT_PULSE: return acp_case_init_pulse(c, q.amp, q.x0, q.y0, q.halfwidth);
        case ACP_INIT_RANDOM: return acp_case_init_random(c, q.seed, q.rho_amp, q.u_amp);
        default: return acp::fail(ACP_E_ARG, "unknown initial-condition kind");
    }
}

// Admission threads: take arrived requests in order while a slot and an arena
// range are free; create and initialise them (several at once, one per thread:
// building a large case's LODI schedule takes a noticeable fraction of a
// second of host time); hand them to the driver.
void admit_loop(acp_server* s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        if (s->queue.empty()) {
            if (s->stopping) break;
            s->cv_admit.wait(lk);
            continue;
        }
        Job& head = s->queue.front();
        const double t = s->now();
        if (head.req.arrival > t) {
            s->cv_admit.wait_for(lk, std::chrono::duration<double>(std::min(head.req.arrival - t, 0.05)));
            continue;
        }
        if (s->free_slots.empty()) {
            s->cv_admit.wait_for(lk, std::chrono::milliseconds(50));
            continue;
        }
        const int64_t off = s->ranges.alloc(head.bytes);
        if (off < 0) {  // head of line waits for memory
            s->cv_admit.wait_for(lk, std::chrono::milliseconds(50));
            continue;
        }
        Job j = std::move(head);
        s->queue.pop_front();
        j.off = off;
        j.slot = s->free_slots.back();
        s->free_slots.pop_back();
        s->bytes_in_use += j.bytes;
        s->peak_bytes = std::max(s->peak_bytes, s->bytes_in_use);
        ++s->in_admission;
        s->peak_active = std::max(s->peak_active, s->in_admission + s->running);
        j.t_admit = s->now();
        lk.unlock();

        const int rc = create_job_case(s, j);
        if (rc != ACP_OK) {
            j.status = rc;
            j.error = acp_last_error();
            destroy_job_case(j);
            lk.lock();
            --s->in_admission;
            lk.unlock();
            post_result(s, j, s->now(), 0., 0);
            lk.lock();
            continue;
        }
        lk.lock();
        --s->in_admission;
        ++s->running;
        s->ready.push_back(std::move(j));
        s->cv_drive.notify_all();
    }
    --s->admit_running;
    s->cv_drive.notify_all();
}

// Finish a case: probe series, mass, checksum, then release it.
void finish_job(acp_server* s, Job& j) {
    double mass = 0.;
    uint64_t cs = 0;
    int rc = j.status;
    if (rc == ACP_OK) rc = acp_case_sync(j.c);
    if (rc == ACP_OK && j.req.probe[0] >= 0) {
        int32_t n = 0;
        std::vector<double> tmp;
        double* out = j.req.probe_out;
        if (!out) {
            tmp.resize(std::max(j.req.steps, 1));
            out = tmp.data();
        }
        rc = acp_case_read_probes(j.c, out, &n);
    }
    if (rc == ACP_OK) rc = acp_case_mass(j.c, &mass);
    if (rc == ACP_OK) rc = acp_case_checksum(j.c, &cs);
    if (rc == ACP_OK && j.r) {
        acp_case* fine = nullptr;
        uint64_t fcs = 0;
        rc = acp_refined_level(j.r, 1, &fine);
        if (rc == ACP_OK) rc = acp_case_checksum(fine, &fcs);
        cs += fcs;
    }
    if (rc != ACP_OK && j.status == ACP_OK) {
        j.status = rc;
        j.error = acp_last_error();
    }
    const double t_done = s->now();
    destroy_job_case(j);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        --s->running;
    }
    post_result(s, j, t_done, mass, cs);
}

// Driver thread: keep up to `chunks_ahead` step chunks queued per resident
// case; retire chunks by event; finish cases whose steps are all done.
void drive_loop(acp_server* s) {
    (void)hipSetDevice(s->device);
    std::vector<Job> active;
    std::vector<PackRun> packs;
    std::vector<hipEvent_t> pool;
    const int32_t chunk = s->cfg.graph_chunk > 0 ? s->cfg.graph_chunk : std::max(s->cfg.launch_steps, 1);
    const int32_t ahead = std::max(s->cfg.chunks_ahead, 1);
    auto take_event = [&](hipEvent_t* ev) -> int {
        if (!pool.empty()) {
            *ev = pool.back();
            pool.pop_back();
            return ACP_OK;
        }
        if (hipEventCreateWithFlags(ev, hipEventDisableTiming) != hipSuccess) {
            *ev = nullptr;
            return acp::fail(ACP_E_HIP, "hipEventCreate");
        }
        return ACP_OK;
    };
    while (true) {
        std::vector<Job> arrived;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            while (!s->ready.empty()) {
                arrived.push_back(std::move(s->ready.front()));
                s->ready.pop_front();
            }
            bool packed = false;
            for (const PackRun& pr : packs) packed = packed || !pr.members.empty();
            if (active.empty() && arrived.empty() && !packed) {
                if (s->admit_running == 0) break;
                s->cv_drive.wait_for(lk, std::chrono::milliseconds(20));
                continue;
            }
        }
        bool progressed = false;
        // arrived cases join the pack of their tier and order when packing is
        // on and they are eligible; the others run on their own streams
        for (Job& j : arrived) {
            const int64_t nodes = (int64_t)j.req.desc.nx * j.req.desc.ny * j.req.desc.nz;
            if (s->cfg.pack_max_nodes > 0 && nodes <= s->cfg.pack_max_nodes && !j.r && j.status == ACP_OK &&
                j.req.steps > 0 &&
                acp::pack_eligible(j.c, j.req.desc.precision, j.req.desc.order) == ACP_OK) {
                PackRun* pr = nullptr;
                for (PackRun& x : packs)
                    if (x.prec == j.req.desc.precision && x.order == j.req.desc.order) pr = &x;
                if (!pr) {
                    PackRun x;
                    x.prec = j.req.desc.precision;
                    x.order = j.req.desc.order;
                    int rc = hipStreamCreateWithFlags(&x.stream, hipStreamNonBlocking) == hipSuccess
                                 ? acp_pack_create(x.prec, x.order, x.stream, &x.p)
                                 : acp::fail(ACP_E_HIP, "pack stream");
                    if (rc == ACP_OK && s->cfg.graph_chunk > 0) rc = acp_pack_enable_graphs(x.p, s->cfg.graph_chunk);
                    if (rc == ACP_OK) {
                        packs.push_back(std::move(x));
                        pr = &packs.back();
                    } else {
                        if (x.p) acp_pack_destroy(x.p);
                        if (x.stream) (void)hipStreamDestroy(x.stream);
                    }
                }
                if (pr && (int)pr->members.size() < acp::kPackSlots && acp_pack_add(pr->p, j.c, j.req.steps) == ACP_OK) {
                    pr->members.push_back(std::move(j));
                    progressed = true;
                    continue;
                }
            }
            active.push_back(std::move(j));
        }
        for (PackRun& pr : packs) {
            // retire finished chunks; finish the members whose last step they held
            while (!pr.inflight.empty()) {
                const hipError_t e = hipEventQuery(pr.inflight.front().first);
                if (e == hipErrorNotReady) break;
                for (Job& j : pr.members)
                    if (e != hipSuccess && j.status == ACP_OK) {
                        j.status = ACP_E_HIP;
                        j.error = hipGetErrorString(e);
                    }
                const std::vector<int64_t> fin = std::move(pr.inflight.front().second);
                pool.push_back(pr.inflight.front().first);
                pr.inflight.pop_front();
                for (int64_t seq : fin)
                    for (size_t i = 0; i < pr.members.size(); ++i)
                        if (pr.members[i].seq == seq) {
                            Job& j = pr.members[i];
                            const int rc = acp_pack_remove(pr.p, j.c);
                            if (rc != ACP_OK && j.status == ACP_OK) {
                                j.status = rc;
                                j.error = acp_last_error();
                            }
                            finish_job(s, j);
                            pr.members.erase(pr.members.begin() + i);
                            break;
                        }
                progressed = true;
            }
            // launch chunks while members have steps left
            while ((int32_t)pr.inflight.size() < ahead) {
                std::vector<int64_t> fin;
                bool any = false;
                for (Job& j : pr.members) {
                    if (j.launched >= j.req.steps) continue;
                    any = true;
                    if (j.launched == 0) j.t_start = s->now();
                    if (j.launched + chunk >= j.req.steps) fin.push_back(j.seq);
                }
                if (!any) break;
                hipEvent_t ev = nullptr;
                int rc = acp_pack_step(pr.p, chunk);
                if (rc == ACP_OK) rc = take_event(&ev);
                if (rc == ACP_OK && hipEventRecord(ev, pr.stream) != hipSuccess)
                    rc = acp::fail(ACP_E_HIP, "hipEventRecord");
                if (rc != ACP_OK) {
                    // the pack stops: every member finishes with the error
                    if (ev) pool.push_back(ev);
                    (void)hipStreamSynchronize(pr.stream);
                    const std::string msg = acp_last_error();
                    for (Job& j : pr.members) {
                        (void)acp_pack_remove(pr.p, j.c);
                        j.status = rc;
                        j.error = msg;
                        finish_job(s, j);
                    }
                    pr.members.clear();
                    for (auto& x : pr.inflight) pool.push_back(x.first);
                    pr.inflight.clear();
                    break;
                }
                for (Job& j : pr.members) j.launched = std::min(j.req.steps, j.launched + chunk);
                pr.inflight.emplace_back(ev, std::move(fin));
                progressed = true;
            }
        }
        for (size_t i = 0; i < active.size();) {
            Job& j = active[i];
            while (!j.inflight.empty()) {
                const hipError_t e = hipEventQuery(j.inflight.front());
                if (e == hipErrorNotReady) break;
                if (e != hipSuccess && j.status == ACP_OK) {
                    j.status = ACP_E_HIP;
                    j.error = hipGetErrorString(e);
                }
                pool.push_back(j.inflight.front());
                j.inflight.pop_front();
                progressed = true;
            }
            while (j.status == ACP_OK && j.launched < j.req.steps && (int32_t)j.inflight.size() < ahead) {
                if (j.launched == 0) j.t_start = s->now();
                const int32_t k = std::min(chunk, j.req.steps - j.launched);
                int rc = j.r ? acp_refined_step(j.r, k) : acp_case_step(j.c, k);
                hipEvent_t ev = nullptr;
                if (rc == ACP_OK) {
                    if (!pool.empty()) {
                        ev = pool.back();
                        pool.pop_back();
                    } else if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
                        ev = nullptr;
                        rc = acp::fail(ACP_E_HIP, "hipEventCreate");
                    }
                }
                if (rc == ACP_OK && hipEventRecord(ev, s->streams[j.slot]) != hipSuccess)
                    rc = acp::fail(ACP_E_HIP, "hipEventRecord");
                if (rc != ACP_OK) {
                    if (ev) pool.push_back(ev);
                    j.status = rc;
                    j.error = acp_last_error();
                    break;
                }
                j.inflight.push_back(ev);
                j.launched += k;
                progressed = true;
            }
            const bool all_launched = j.launched >= j.req.steps || j.status != ACP_OK;
            if (all_launched && j.inflight.empty()) {
                finish_job(s, j);
                active.erase(active.begin() + i);
                progressed = true;
            } else {
                ++i;
            }
        }
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(50));
    }
    for (PackRun& pr : packs) {
        acp_pack_destroy(pr.p);
        (void)hipStreamDestroy(pr.stream);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
}

}  // namespace

extern "C" int acp_server_create(const acp_server_desc* desc, acp_server** out) {
    if (!desc || !out) return acp::fail(ACP_E_ARG, "acp_server_create: null argument");
    *out = nullptr;
    if (desc->max_active < 1 || desc->hbm_budget_bytes < 256 || desc->chunks_ahead < 0 || desc->pack_max_nodes < 0 ||
        (desc->graph_chunk != 0 && (desc->graph_chunk < 2 || desc->graph_chunk % 2 != 0)))
        return acp::fail(ACP_E_ARG, "acp_server_create: need max_active >= 1, a budget, an even graph chunk or 0");
    acp_server* s = new acp_server();
    s->cfg = *desc;
    s->cfg.hbm_budget_bytes = desc->hbm_budget_bytes / 256 * 256;
    (void)hipGetDevice(&s->device);
    hipError_t e = hipMalloc((void**)&s->arena, (size_t)s->cfg.hbm_budget_bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        return acp::fail(ACP_E_OOM, "acp_server_create: arena of " + std::to_string(desc->hbm_budget_bytes) +
                                        " bytes: " + hipGetErrorString(e));
    }
    s->ranges.reset(s->cfg.hbm_budget_bytes);
    for (int i = 0; i < desc->max_active; ++i) {
        hipStream_t st = nullptr;
        e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (hipStream_t x : s->streams) (void)hipStreamDestroy(x);
            (void)hipFree(s->arena);
            delete s;
            return acp::hip_fail(e, "acp_server_create: stream");
        }
        s->streams.push_back(st);
        s->free_slots.push_back(desc->max_active - 1 - i);
    }
    s->t0 = std::chrono::steady_clock::now();
    const int na = desc->admit_threads > 0 ? desc->admit_threads : 4;
    s->admit_running = na;
    for (int i = 0; i < na; ++i) s->admit_threads.emplace_back(admit_loop, s);
    s->drive_thread = std::thread(drive_loop, s);
    *out = s;
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_server_submit(acp_server* s, const acp_request* r) {
    if (!s || !r) return acp::fail(ACP_E_ARG, "acp_server_submit: null argument");
    if (r->steps < 0) return acp::fail(ACP_E_ARG, "acp_server_submit: negative step count");
    const acp_case_desc& d = r->desc;
    if (r->probe[0] >= 0 && (r->probe[0] >= d.nx || r->probe[1] < 0 || r->probe[1] >= d.ny || r->probe[2] < 0 ||
                             r->probe[2] >= d.nz))
        return acp::fail(ACP_E_ARG, "acp_server_submit: probe outside the grid");
    int64_t bytes = 0;
    const int32_t np = r->probe[0] >= 0 ? 1 : 0, cap = std::max(r->steps, 1);
    if (two_level(*r)) {
        if (r->init_kind != ACP_INIT_PULSE)
            return acp::fail(ACP_E_UNSUPPORTED, "acp_server_submit: two-level requests start from a pulse");
        if (r->f2c_filter != ACP_F2C_NOFI && r->f2c_filter != ACP_F2C_LAG && r->f2c_filter != ACP_F2C_TOU)
            return acp::fail(ACP_E_UNSUPPORTED, "acp_server_submit: fine-to-coarse filter NOFI, LAG or TOU");
        if (r->coupling != ACP_COUPLING_STD && r->coupling != ACP_COUPLING_DC)
            return acp::fail(ACP_E_ARG, "acp_server_submit: coupling ACP_COUPLING_STD or ACP_COUPLING_DC");
        if (int rc = acp_refined_footprint(&d, r->refine[0], r->refine[1], np, cap, &bytes)) return rc;
    } else if (int rc = acp_case_footprint(&d, np, cap, &bytes)) {
        return rc;
    }
    if (bytes > s->cfg.hbm_budget_bytes)
        return acp::fail(ACP_E_OOM, "acp_server_submit: case needs " + std::to_string(bytes) +
                                        " bytes, more than the server's HBM budget");
    Job j;
    j.req = *r;
    j.bytes = bytes;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stopping) return acp::fail(ACP_E_STATE, "acp_server_submit: server is shutting down");
    j.t_submit = s->now();
    j.seq = s->submitted;
    auto pos = std::upper_bound(s->queue.begin(), s->queue.end(), r->arrival,
                                [](double a, const Job& x) { return a < x.req.arrival; });
    s->queue.insert(pos, std::move(j));
    ++s->submitted;
    s->cv_admit.notify_all();
    acp::clear_error();
    return ACP_OK;
}

extern "C" int acp_server_poll(acp_server* s, acp_result* out, int32_t max, int32_t wait_ms, int32_t* n) {
    if (!s || !n || (max > 0 && !out)) return acp::fail(ACP_E_ARG, "acp_server_poll: null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->done.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->done.empty(); });
    const int32_t k = std::min<int32_t>(max, (int32_t)s->done.size());
    for (int32_t i = 0; i < k; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + k);
    *n = k;
    return ACP_OK;
}

extern "C" int acp_server_get_stats(acp_server* s, acp_server_stats* st) {
    if (!s || !st) return acp::fail(ACP_E_ARG, "acp_server_get_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->queued = (int64_t)s->queue.size();
    st->active = s->in_admission + s->running;
    st->peak_active = s->peak_active;
    st->bytes_in_use = s->bytes_in_use;
    st->peak_bytes = s->peak_bytes;
    st->budget_bytes = s->cfg.hbm_budget_bytes;
    st->clock = s->now();
    return ACP_OK;
}

extern "C" double acp_server_clock(acp_server* s) { return s ? s->now() : 0.; }

extern "C" void acp_server_destroy(acp_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
        s->cv_admit.notify_all();
    }
    for (std::thread& t : s->admit_threads)
        if (t.joinable()) t.join();
    if (s->drive_thread.joinable()) s->drive_thread.join();
    for (hipStream_t st : s->streams) (void)hipStreamDestroy(st);
    if (s->arena) (void)hipFree(s->arena);
    delete s;
}
