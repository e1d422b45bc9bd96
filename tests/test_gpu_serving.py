"""Captured-graph stepping and the serving loop on the GPU.

Every comparison here is exact: a case gives bit-identical populations (via
the order-independent checksum and, for small cases, the full arrays), probe
series and mass whether it is stepped with plain launches, with graph replays,
alone, or inside the server among other cases.
"""
import numpy as np
import pytest

import acpulse_vclbm_amd as acp
from acpulse_vclbm_amd import _capi, serve

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("spec", [
    dict(op="HRR", precision="F64", transverse=True, cubic=True),
    dict(op="SRT", precision="F32-SHIFTED"),
    dict(op="MRT-RM", precision="F16-SHIFTED", boundary="PERIODIC", mrt_mode="RELAXED"),
], ids=["hrr-f64-cubic", "srt-f32", "mrtrm-f16-periodic"])
def test_footprint_is_what_a_case_allocates(gpu, spec):
    d = acp.make_desc(40, 36, 5, tau=0.6, **spec)
    with acp.Case(d) as c:
        assert c.info().device_bytes == acp.footprint(d)
        c.set_probes([(3, 4, 2), (20, 18, 2)], capacity=77)
        assert c.info().device_bytes == acp.footprint(d, 2, 77)


@pytest.mark.parametrize("spec", [
    dict(op="HRR", precision="F64", transverse=True),
    dict(op="RR", precision="F64", cubic=True),
    dict(op="SRT", precision="F32-SHIFTED"),
    dict(op="MRT-GS", precision="F64", boundary="PERIODIC", mrt_mode="RELAXED"),
], ids=["hrr-f64", "rr-cubic-f64", "srt-f32", "mrtgs-periodic"])
def test_graph_replays_match_plain_steps(gpu, spec):
    """Chunk of 4 steps: every replay (both ping-pong parities, odd remainders in
    between, eager work and synchronisation between replays) must leave the case
    in exactly the state plain stepping gives."""
    d = acp.make_desc(30, 26, 3, tau=0.53, **spec)
    with acp.Case(d) as a, acp.Case(d) as b:
        for c in (a, b):
            c.init_pulse(1e-3, 15.0, 13.0, 4.0)
            c.set_probes([(22, 13, 1), (5, 5, 1)], capacity=200)
        b.enable_graphs(4)
        plan = [4, 4, 1, 4, 8, 3, 4, 12, 1, 6]
        for n in plan:
            a.step(n)
            b.step(n)
            a.sync()
            b.sync()
            assert a.checksum() == b.checksum()
            assert a.mass() == b.mass()
        assert np.array_equal(a.populations(), b.populations())
        pa, pb = a.read_probes(), b.read_probes()
        assert pa.shape == (sum(plan), 2)
        assert np.array_equal(pa, pb)


@pytest.mark.parametrize("spec", [dict(op="RR", precision="F64"), dict(op="SRT", precision="F32-SHIFTED"),
                                  dict(op="MRT-RM", precision="F64", mrt_mode="RELAXED", lag=True),
                                  dict(op="HRR", precision="F64", duct=True),
                                  dict(op="RR", precision="F32-SHIFTED", duct=True)],
                         ids=["rr-f64", "srt-f32", "mrtrm-f64-lag", "hrr-f64-duct", "rr-f32-duct"])
def test_refined_graph_replays_match_plain_steps(gpu, spec):
    """Refined cases replayed from captured graphs of 4 coarse steps (both
    levels' parities, odd remainders, synchronisation between replays) against
    plain refined steps: both levels bit-identical, probe series equal; and the
    refined footprint is what the two levels allocate.  Periodic coarse grids
    with the fine level over [16, 32], and ducts (LODI x faces) with the fine
    level over [16, 31], ending on a LODI face."""
    spec = dict(spec)
    lag = spec.pop("lag", False)
    duct = spec.pop("duct", False)
    d = acp.make_desc(32, 12, 3, tau=0.56, boundary="LODI-X" if duct else "PERIODIC", **spec)
    x1 = 31 if duct else 32
    with acp.Refined(d, 16, x1) as a, acp.Refined(d, 16, x1) as b:
        for r in (a, b):
            if lag:
                r.set_filter("LAG")
            r.init_pulse(1e-3, 8.0, 6.0, 3.0)
            r.level(0).set_probes([(24, 6, 1)], capacity=100)
        assert (a.level(0).info().device_bytes + a.level(1).info().device_bytes ==
                acp.refined_footprint(d, 16, x1, 1, 100))
        b.enable_graphs(4)
        plan = [4, 1, 4, 8, 3, 12]
        for n in plan:
            a.step(n)
            b.step(n)
            a.level(0).sync()
            b.level(0).sync()
            assert a.checksum() == b.checksum()
        for lv in (0, 1):
            assert np.array_equal(a.level(lv).populations(), b.level(lv).populations())
        assert np.array_equal(a.level(0).read_probes(), b.level(0).read_probes())


def test_refined_graphs_follow_new_probes(gpu):
    """Probes set on a level after graph replays: the refined case captures
    its graphs anew, so the next replays record into the new probe buffer
    (different node, different capacity) exactly as plain steps do."""
    d = acp.make_desc(32, 12, 3, tau=0.56, boundary="LODI-X", op="RR")
    with acp.Refined(d, 16, 31) as a, acp.Refined(d, 16, 31) as b:
        for r in (a, b):
            r.init_pulse(1e-3, 8.0, 6.0, 3.0)
            r.level(0).set_probes([(24, 6, 1)], capacity=16)
        b.enable_graphs(4)
        for r in (a, b):
            r.step(8)
        assert np.array_equal(a.level(0).read_probes(), b.level(0).read_probes())
        for r in (a, b):
            r.level(0).set_probes([(20, 3, 1), (12, 6, 1)], capacity=40)
            r.step(12)
        pa, pb = a.level(0).read_probes(), b.level(0).read_probes()
        assert pa.shape == (12, 2)
        assert np.array_equal(pa, pb)
        assert a.checksum() == b.checksum()


def test_graph_chunk_argument_checks(gpu):
    d = acp.make_desc(10, 10, 3, tau=0.6)
    with acp.Case(d) as c:
        for bad in (-2, 1, 3):
            with pytest.raises(_capi.AcpError):
                c.enable_graphs(bad)
        c.enable_graphs(0)


def _small_mix(n, seed):
    reqs = serve.make_requests(n, seed=seed, nxy=(24, 72), steps=(20, 90))
    for i, r in enumerate(reqs):
        r["arrival"] = 0.002 * i
    return reqs


def test_server_results_equal_alone(gpu):
    """The mix holds one-level and two-level requests (seed 5: refined requests
    among them); each result equals the same request run alone."""
    reqs = _small_mix(12, seed=5)
    assert 0 < sum(serve.two_level(r) for r in reqs) < len(reqs)
    alone = [serve.run_alone(r) for r in reqs]
    results, summary = serve.run_mix(reqs, hbm_budget_bytes=1 << 30, max_active=4, graph_chunk=8, chunks_ahead=2,
                                     timeout_s=120)
    assert summary["failed"] == 0, [(r["tag"], r["status"], r["error"]) for r in results if r["status"]]
    assert len(results) == len(reqs)
    for r, a, q in zip(results, alone, reqs):
        assert r["status"] == 0, r["error"]
        assert r["steps"] == q["steps"]
        assert r["checksum"] == a["checksum"], q
        assert r["mass"] == a["mass"], q
        assert np.array_equal(r["probe"], a["probe"]), q
        assert r["t_arrival"] <= r["t_admit"] + 1e-9 <= r["t_done"] + 2e-9
    assert summary["peak_active"] <= 4
    assert summary["latency_p99_s"] >= summary["latency_p50_s"] > 0


def test_server_admits_by_footprint(gpu):
    reqs = _small_mix(8, seed=9)
    fps = [serve.footprint(r) for r in reqs]
    budget = (2 * max(fps) + 255) // 256 * 256  # room for at most two of the largest at once
    alone = [serve.run_alone(r) for r in reqs]
    results, summary = serve.run_mix(reqs, hbm_budget_bytes=budget, max_active=8, graph_chunk=0, launch_steps=5,
                                     chunks_ahead=3, timeout_s=120)
    assert summary["failed"] == 0, [(r["tag"], r["status"], r["error"]) for r in results if r["status"]]
    assert summary["peak_hbm_bytes"] <= budget
    for r, a in zip(results, alone):
        assert r["checksum"] == a["checksum"]
        assert np.array_equal(r["probe"], a["probe"])
    # a case larger than the whole budget is refused at submission
    with serve.Server(budget, max_active=2) as srv:
        big = dict(reqs[0], nx=400, ny=400, tag=99)
        with pytest.raises(_capi.AcpError) as e:
            srv.submit(big)
        assert e.value.code == _capi.ACP_E_OOM


def _mix64(x):
    x = x ^ (x >> np.uint64(30))
    x = x * np.uint64(0xbf58476d1ce4e5b9)
    x = x ^ (x >> np.uint64(27))
    x = x * np.uint64(0x94d049bb133111eb)
    return x ^ (x >> np.uint64(31))


def layout_checksum64(f):
    """acp_case_checksum of a float64 state [19][N] computed on the host: the
    engine lays float64 populations out in blocks of 8192 nodes (element
    (q, l) at (l >> 13) 19 8192 + 8192 q + (l & 8191), the last block padded
    with zeros; a grid of one block holds whole planes) and sums
    mix64(bits ^ index * 0x9e3779b97f4a7c15) over every element mod 2^64."""
    N = f.shape[1]
    if N <= 8192:
        flat = f.reshape(-1)
    else:
        nb = (N + 8191) // 8192
        pad = np.zeros((19, nb * 8192))
        pad[:, :N] = f
        flat = pad.reshape(19, nb, 8192).transpose(1, 0, 2).reshape(-1)
    with np.errstate(over="ignore"):
        bits = flat.view(np.uint64)
        idx = np.arange(flat.size, dtype=np.uint64) * np.uint64(0x9e3779b97f4a7c15)
        return int(_mix64(bits ^ idx).sum(dtype=np.uint64))


def _oracle_request(r):
    """The request run by the oracle: final float64 state(s) and probe series."""
    import oracle_lib as ol
    from make_golden import CS, DX, RHO0
    cx, cy = serve.pulse_centre(r)
    probe = [serve.probe_point(r)]
    two = serve.two_level(r)
    p = ol.make_params(r["nx"], r["ny"], r["nz"], op=r["op"], order=3, periodic=3 if two else 0,
                       mrt_mode=1, tau=serve.TAU, sigma=0.98, cs=CS, spacing=DX, rho0=RHO0)
    if two:
        x0, x1 = serve.fine_region(r)
        o = ol.OracleRefined(p, x0, x1)
        o.set_filter("NOFI")
        if r.get("scheme") == "direct":
            o.set_coupling("DC1", "Mf", "Mc", "hMf")
        o.levels[0].set_state(ol.pulse_ic(p, r["amp"], cx, cy, r["halfwidth"]))
        o.levels[1].set_state(ol.pulse_ic(o.levels[1].p, r["amp"], 2 * (cx - x0), 2 * cy, 2 * r["halfwidth"]))
        series = o.step(r["steps"], probe)[:, 0]
        return [o.levels[0].state(), o.levels[1].state()], series
    o = ol.OracleCase(p)
    o.set_state(ol.pulse_ic(p, r["amp"], cx, cy, r["halfwidth"]))
    series = o.step(r["steps"], probe)[:, 0]
    return [o.state()], series


def test_served_results_match_the_oracle(gpu):
    """Every case of a served mix (one- and two-level, float64 and
    float32-shifted) against the oracle's own run of the same request: float64
    cases give the oracle's probe series to the bit and its final state to the
    bit (the served checksum equals the checksum of the oracle's state, summed
    over both levels of a two-level case) and its mass to 1e-13; float32-shifted
    cases follow the oracle's probe series within 2e-4 of the pulse amplitude."""
    reqs = _small_mix(12, seed=5)
    assert any(serve.two_level(r) for r in reqs) and any(not serve.two_level(r) for r in reqs)
    assert {r["scheme"] for r in reqs if serve.two_level(r)} == {"overlap", "direct"}
    assert {r["precision"] for r in reqs} == {"F64", "F32-SHIFTED"}
    results, summary = serve.run_mix(reqs, hbm_budget_bytes=1 << 30, max_active=4, graph_chunk=8, chunks_ahead=2,
                                     timeout_s=120)
    assert summary["failed"] == 0
    for res, q in zip(results, reqs):
        states, series = _oracle_request(q)
        print("case", q["tag"], q["op"], q["precision"], "two-level" if serve.two_level(q) else "one-level",
              flush=True)
        if q["precision"] == "F64":
            assert np.array_equal(res["probe"], series), q
            csum = sum(layout_checksum64(s) for s in states) % (1 << 64)
            assert res["checksum"] == csum, q
            assert res["mass"] == pytest.approx(states[0].sum(), rel=1e-13, abs=0), q
        else:
            rho0 = 1.17621
            err = np.max(np.abs(res["probe"] - series)) / (q["amp"] * rho0)
            assert err <= 2e-4, (q, err)
