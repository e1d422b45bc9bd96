"""Serving loop: many acoustic-pulse cases resident on one GPU (acp_server_*).

The reference runs one case per process (main.cpp:37-271).  Here a native
server (csrc/server.cpp) keeps cases resident, admits them by HBM footprint
and steps them concurrently on separate streams in captured-graph chunks.
This module builds request mixes, drives a server and reports cases/hour and
latency percentiles (arrival to result).

Request mix (one draw per request, in this order, from numpy's PCG64 with the
given seed): inter-arrival time (exponential, unit mean, scaled by 1/rate),
nx and ny uniform in [128, 1024], a refinement flag with probability 1/2 and a
coupling scheme (overlap or direct), the collision operator uniform over
BGK/MRT-GS/MRT-RM/RR/HRR, the pulse amplitude log-uniform in [1e-4, 1e-2], the
half-width uniform in [4, 20] nodes, the step count uniform in [1000, 8000]
(times `step_scale`), and the tier (float64 or float32-shifted, 1/2 each).
Grids are nx x ny x 3, tau = 0.5 + 3e-4, N_eq = 3, the probe at
(3/4 nx, ny/2, 1).

  * One-level requests: LODI on all faces, the pulse at the centre.
  * Refined requests (acp_refined_*): the refined acoustic-pulse geometry, a
    coarse grid with LODI x faces, periodic in y and z, and a fine level
    (ratio 2) over x in [nx/2, nx - 1], which ends on a LODI face at the
    high-x end; the pulse at (nx/4, ny/2) in the coarse half, the probe read
    on the coarse level inside the fine region; F2C filter NOFI; the
    standard coupling for requests that draw "overlap", the direct coupling
    (missing links: Mf, Mc, hMf) for those that draw "direct".
"""
import ctypes
import time

import numpy as np

from . import _capi
from ._capi import INIT, Request, Result, ServerDesc, ServerStats, check, lib

OPS = ["SRT", "MRT-GS", "MRT-RM", "RR", "HRR"]
TAU = 0.5 + 3e-4


def make_requests(n=256, seed=1, rate=None, step_scale=1.0, nxy=(128, 1024), steps=(1000, 8000), precision=None):
    """The serving workload as a list of dicts (arrival times in seconds when
    `rate` is given in cases/s, else all zero)."""
    rng = np.random.default_rng(seed)
    out = []
    t = 0.0
    for i in range(n):
        gap = rng.exponential(1.0)
        nx = int(rng.integers(nxy[0], nxy[1] + 1))
        ny = int(rng.integers(nxy[0], nxy[1] + 1))
        refined = bool(rng.random() < 0.5)
        scheme = "overlap" if rng.random() < 0.5 else "direct"
        op = OPS[int(rng.integers(0, 5))]
        amp = float(10.0 ** rng.uniform(-4.0, -2.0))
        hw = float(rng.uniform(4.0, 20.0))
        nsteps = int(rng.integers(steps[0], steps[1] + 1))
        prec = "F64" if rng.random() < 0.5 else "F32-SHIFTED"
        if rate:
            t += gap / rate
        nsteps = max(2, int(round(nsteps * step_scale)))
        out.append({"tag": i, "arrival": t if rate else 0.0, "nx": nx, "ny": ny, "nz": 3, "op": op, "amp": amp,
                    "halfwidth": hw, "steps": nsteps, "precision": precision or prec, "refined": refined,
                    "scheme": scheme})
    return out


def two_level(r):
    """Whether the request runs as a refined case."""
    return bool(r.get("refined"))


def fine_region(r):
    return (r["nx"] // 2, r["nx"] - 1)


def request_desc(r):
    from . import make_desc
    return make_desc(r["nx"], r["ny"], r["nz"], op=r["op"], order=3,
                     boundary="LODI-X" if two_level(r) else "LODI", mrt_mode="RELAXED",
                     precision=r["precision"], tau=TAU, sigma=0.98)


def pulse_centre(r):
    return (r["nx"] / 4.0 if two_level(r) else r["nx"] / 2.0, r["ny"] / 2.0)


def probe_point(r):
    return (3 * r["nx"] // 4, r["ny"] // 2, r["nz"] // 2)


def footprint(r):
    """Device bytes the server reserves for request `r` (one probe of `steps` densities)."""
    from . import footprint as case_footprint, refined_footprint
    cap = max(int(r["steps"]), 1)
    if two_level(r):
        return refined_footprint(request_desc(r), *fine_region(r), 1, cap)
    return case_footprint(request_desc(r), 1, cap)


def mix_counts(reqs):
    """How the refinement flags of a request list were served."""
    return {"requests": len(reqs), "refined_flag": sum(bool(r["refined"]) for r in reqs),
            "two_level": sum(two_level(r) for r in reqs),
            "direct_coupling": sum(two_level(r) and r["scheme"] == "direct" for r in reqs)}


class Server:
    """acp_server_* handle: submit requests, poll results, read statistics."""

    def __init__(self, hbm_budget_bytes, max_active=8, graph_chunk=16, launch_steps=16, chunks_ahead=4,
                 admit_threads=4, pack_max_nodes=0):
        L = lib()
        d = ServerDesc(int(hbm_budget_bytes), int(max_active), int(graph_chunk), int(launch_steps),
                       int(chunks_ahead), int(admit_threads), int(pack_max_nodes))
        h = ctypes.c_void_p()
        check(L.acp_server_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self._probe_bufs = {}

    def submit(self, r, arrival=None):
        q = Request()
        q.desc = request_desc(r)
        q.steps = int(r["steps"])
        q.init_kind = INIT["PULSE"]
        cx, cy = pulse_centre(r)
        q.amp, q.x0, q.y0, q.halfwidth = r["amp"], cx, cy, r["halfwidth"]
        if two_level(r):
            q.refine[0], q.refine[1] = fine_region(r)
            q.f2c_filter = _capi.F2C_FILTER["NOFI"]
            q.coupling = 1 if r.get("scheme") == "direct" else 0
        pi, pj, pk = probe_point(r)
        q.probe[0], q.probe[1], q.probe[2] = pi, pj, pk
        tag = int(r["tag"])
        # the native server writes the probe series into this buffer when the
        # case finishes, so it must stay alive and unshared until then
        if tag in self._probe_bufs:
            raise ValueError(f"request tag {tag} is already in flight")
        buf = np.zeros(max(q.steps, 1))
        q.probe_out = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        q.arrival = float(r["arrival"] if arrival is None else arrival)
        q.tag = tag
        check(lib().acp_server_submit(self.h, ctypes.byref(q)))
        self._probe_bufs[tag] = buf

    def poll(self, max_results=64, wait_ms=0):
        out = (Result * max_results)()
        n = ctypes.c_int32()
        check(lib().acp_server_poll(self.h, out, max_results, int(wait_ms), ctypes.byref(n)))
        res = []
        for k in range(n.value):
            x = out[k]
            d = {f: getattr(x, f) for f, _ in Result._fields_}
            d["error"] = x.error.decode(errors="replace")
            d["probe"] = self._probe_bufs.pop(int(x.tag), None)
            res.append(d)
        return res

    def stats(self):
        st = ServerStats()
        check(lib().acp_server_get_stats(self.h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in ServerStats._fields_}

    def clock(self):
        return lib().acp_server_clock(self.h)

    def close(self):
        if self.h:
            lib().acp_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def run_mix(reqs, hbm_budget_bytes, max_active=8, graph_chunk=16, launch_steps=16, chunks_ahead=4,
            admit_threads=4, timeout_s=3600.0, pack_max_nodes=0, on_result=None):
    """Offer `reqs` (arrival times relative to the start) to a fresh server and
    collect every result; eligible one-level cases of at most `pack_max_nodes`
    nodes share packed step launches; `on_result(result)` is called as each
    result arrives.  Returns (results sorted by tag, summary dict)."""
    with Server(hbm_budget_bytes, max_active, graph_chunk, launch_steps, chunks_ahead, admit_threads,
                pack_max_nodes) as srv:
        for r in reqs:
            srv.submit(r)
        results = []
        t_end = time.monotonic() + timeout_s
        while len(results) < len(reqs):
            got = srv.poll(64, wait_ms=200)
            if on_result:
                for x in got:
                    on_result(x)
            results += got
            if time.monotonic() > t_end:
                raise TimeoutError(f"{len(results)} of {len(reqs)} cases finished within {timeout_s} s")
        stats = srv.stats()
    results.sort(key=lambda d: d["tag"])
    return results, summarize(results, stats)


def summarize(results, stats=None):
    ok = [r for r in results if r["status"] == 0]
    lat = np.array([r["t_done"] - r["t_arrival"] for r in ok]) if ok else np.zeros(1)
    wait = np.array([r["t_admit"] - r["t_arrival"] for r in ok]) if ok else np.zeros(1)
    t0 = min(r["t_arrival"] for r in results) if results else 0.0
    t1 = max(r["t_done"] for r in results) if results else 0.0
    span = max(t1 - t0, 1e-9)
    node_steps = sum(r["steps"] * (r["device_bytes"] > 0) for r in ok)
    out = {
        "cases": len(results), "failed": len(results) - len(ok), "makespan_s": span,
        "cases_per_hour": len(ok) / span * 3600.0,
        "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
        "latency_mean_s": float(lat.mean()), "queue_wait_p50_s": float(np.percentile(wait, 50)),
        "queue_wait_p99_s": float(np.percentile(wait, 99)), "steps_total": int(node_steps),
    }
    if stats:
        out.update({"peak_active": stats["peak_active"], "peak_hbm_bytes": stats["peak_bytes"],
                    "budget_bytes": stats["budget_bytes"]})
    return out


def run_alone(r, graph_chunk=0):
    """One request on its own case (no server): final checksum, mass and probe series."""
    from . import Case, Refined
    cx, cy = pulse_centre(r)
    if two_level(r):
        with Refined(request_desc(r), *fine_region(r)) as rc:
            rc.set_filter("NOFI")
            if r.get("scheme") == "direct":
                rc.set_coupling("DC1", "Mf", "Mc", "hMf")
            rc.init_pulse(r["amp"], cx, cy, r["halfwidth"])
            c = rc.level(0)
            c.set_probes([probe_point(r)], capacity=max(r["steps"], 1))
            if graph_chunk:
                rc.enable_graphs(graph_chunk)
            rc.step(r["steps"])
            c.sync()
            probe = c.read_probes()[:, 0].copy()
            return {"checksum": rc.checksum(), "mass": c.mass(), "probe": probe}
    with Case(request_desc(r)) as c:
        c.init_pulse(r["amp"], cx, cy, r["halfwidth"])
        c.set_probes([probe_point(r)], capacity=max(r["steps"], 1))
        if graph_chunk:
            c.enable_graphs(graph_chunk)
        c.step(r["steps"])
        c.sync()
        probe = c.read_probes()[:, 0].copy()
        return {"checksum": c.checksum(), "mass": c.mass(), "probe": probe}


def main():
    """Run the seed-1 request mix on one GPU and print its summary as JSON.

    The full mix as specified is ``--requests 256 --step-scale 1``.
    ``--sequential-f64`` runs every request in float64, one at a time with one
    launch sequence per step (the sequential baseline); ``--rate-per-hour``
    offers Poisson arrivals (otherwise every request arrives at once, which
    measures capacity).  ``--progress`` prints a line per finished case."""
    import argparse
    import json

    ap = argparse.ArgumentParser(description="run the serving request mix on one GPU")
    ap.add_argument("--requests", type=int, default=48)
    ap.add_argument("--step-scale", type=float, default=0.25)
    ap.add_argument("--rate-per-hour", type=float, default=0.0, help="Poisson arrivals per hour; 0: all at once")
    ap.add_argument("--max-active", type=int, default=16)
    ap.add_argument("--graph-chunk", type=int, default=16)
    ap.add_argument("--budget-gib", type=float, default=64.0)
    ap.add_argument("--pack-max-nodes", type=int, default=0,
                    help="one-level cases of at most this many nodes share packed step launches")
    ap.add_argument("--sequential-f64", action="store_true", help="float64, one case at a time")
    ap.add_argument("--all-f64", action="store_true", help="every request in float64")
    ap.add_argument("--progress", action="store_true")
    a = ap.parse_args()
    rate = a.rate_per_hour / 3600.0 if a.rate_per_hour > 0 else None
    reqs = make_requests(a.requests, seed=1, rate=rate, step_scale=a.step_scale)
    if a.sequential_f64 or a.all_f64:
        reqs = [dict(r, precision="F64") for r in reqs]
    t0 = time.monotonic()
    done = [0]

    def progress(x):
        done[0] += 1
        print(f"{done[0]}/{len(reqs)} tag {x['tag']} status {x['status']} steps {x['steps']} "
              f"latency {x['t_done'] - x['t_arrival']:.2f} s wall {time.monotonic() - t0:.1f} s", flush=True)

    kw = dict(on_result=progress if a.progress else None, pack_max_nodes=a.pack_max_nodes)
    budget = int(a.budget_gib * (1 << 30))
    if a.sequential_f64:
        _, s = run_mix(reqs, budget, max_active=1, graph_chunk=0, launch_steps=1, chunks_ahead=64, admit_threads=1,
                       **kw)
    else:
        _, s = run_mix(reqs, budget, max_active=a.max_active, graph_chunk=a.graph_chunk, **kw)
    s.update({"requests": len(reqs), "step_scale": a.step_scale, "rate_per_hour": a.rate_per_hour,
              "sequential_f64": a.sequential_f64, "all_f64": a.all_f64, "pack_max_nodes": a.pack_max_nodes,
              "max_active": 1 if a.sequential_f64 else a.max_active, "mix": mix_counts(reqs),
              "node_updates": int(sum(_node_updates(r) for r in reqs))})
    print(json.dumps(s), flush=True)


def _node_updates(r):
    """Node updates of a request: steps x nodes, a two-level case counting its
    fine level twice per coarse step."""
    n = r["nx"] * r["ny"] * r["nz"]
    if two_level(r):
        x0, x1 = fine_region(r)
        n += 2 * (2 * (x1 - x0) + 1) * (2 * r["ny"]) * (2 * r["nz"])
    return n * r["steps"]


if __name__ == "__main__":
    main()
