"""Packed stepping (acp_pack_*) on the GPU.

A pack advances several resident one-level cases with shared launches that
walk a device-resident member table.  Every comparison here is exact: a
member's populations (both arrays), probe series, step counter and mass equal
those of the same case stepped alone with acp_case_step, whatever else is in
the pack, when it joined, and whether its steps ran from captured graphs.
"""
import numpy as np
import pytest

import acpulse_vclbm_amd as acp
from acpulse_vclbm_amd import _capi, serve

pytestmark = pytest.mark.gpu

# members of mixed operators, sizes, LODI phase counts (nz = 3: four launches
# per step; 3D: two) and step counts; probes on some
MEMBERS = [
    dict(shape=(40, 36, 3), op="SRT", steps=37, probes=[(20, 18, 1), (5, 30, 1)]),
    dict(shape=(48, 30, 3), op="HRR", transverse=True, steps=50, probes=[(10, 12, 1)]),
    dict(shape=(20, 18, 9), op="MRT-RM", mrt_mode="RELAXED", steps=23, probes=[]),
    dict(shape=(26, 22, 5), op="RR", transverse=True, steps=64, probes=[(13, 11, 2)]),
    dict(shape=(33, 17, 3), op="MRT-GS", steps=41, probes=[(1, 8, 1)]),
    dict(shape=(300, 7, 3), op="SRT", steps=19, probes=[]),
]


def _case(m, precision, order=3):
    nx, ny, nz = m["shape"]
    d = acp.make_desc(nx, ny, nz, op=m["op"], order=order, transverse=m.get("transverse", False),
                      precision=precision, mrt_mode=m.get("mrt_mode", "REFERENCE"), tau=0.53)
    c = acp.Case(d)
    c.init_pulse(1e-3, nx / 2 - 0.5, ny / 2 + 0.25, 3.0)
    if m["probes"]:
        c.set_probes(m["probes"], capacity=m["steps"] + 5)
    return c


def _alone(m, precision, order=3):
    with _case(m, precision, order) as c:
        c.step(m["steps"])
        c.sync()
        probes = c.read_probes() if m["probes"] else None
        return c.populations(), c.previous_populations(), probes, c.info().steps_done, c.mass()


def _check(c, m, ref):
    f, fprev, probes, steps, mass = ref
    assert np.array_equal(c.populations(), f), m
    assert np.array_equal(c.previous_populations(), fprev), m
    assert c.info().steps_done == steps == m["steps"]
    assert c.mass() == mass
    if m["probes"]:
        p = c.read_probes()
        assert p.shape == (m["steps"], len(m["probes"]))
        assert np.array_equal(p, probes), m


@pytest.mark.parametrize("precision,graphs", [("F64", 0), ("F64", 8), ("F32-SHIFTED", 0), ("F32-SHIFTED", 6),
                                               ("F16-SHIFTED", 4)])
def test_pack_members_equal_alone(gpu, precision, graphs):
    """Members join at different times (three at once, the rest after 10
    steps), take their own step counts (odd and even, none a multiple of the
    chunk) and leave when done; each ends bit-identical to its run alone."""
    refs = [_alone(m, precision) for m in MEMBERS]
    cases = [_case(m, precision) for m in MEMBERS]
    try:
        with acp.Pack(precision, 3) as pk:
            if graphs:
                pk.enable_graphs(graphs)
            for c, m in zip(cases[:3], MEMBERS[:3]):
                pk.add(c, m["steps"])
            pk.step(10)
            for c, m in zip(cases[3:], MEMBERS[3:]):
                pk.add(c, m["steps"])
            assert pk.members() == len(MEMBERS)
            with pytest.raises(_capi.AcpError) as e:
                cases[0].step(1)
            assert e.value.code == _capi.ACP_E_STATE
            done = set()
            plan = [7, 16, 3, 24, 8, 1, 40]
            for n in plan:
                pk.step(n)
                for i, (c, m) in enumerate(zip(cases, MEMBERS)):
                    if i not in done and pk.steps_left(c) == 0:
                        pk.remove(c)
                        done.add(i)
                        _check(c, m, refs[i])
            assert done == set(range(len(MEMBERS)))
            assert pk.members() == 0
    finally:
        for c in cases:
            c.close()


def test_pack_member_continues_alone(gpu):
    """A member removed before its steps are done keeps its state: stepping it
    alone for the rest equals stepping it alone throughout."""
    m = dict(MEMBERS[1])
    ref = _alone(m, "F64")
    with _case(m, "F64") as c, _case(MEMBERS[0], "F64") as other, acp.Pack("F64", 3) as pk:
        pk.add(c, m["steps"])
        pk.add(other, MEMBERS[0]["steps"])
        pk.step(21)
        pk.remove(c)
        c.step(m["steps"] - 21)
        _check(c, m, ref)


def test_pack_refuses_ineligible_cases(gpu):
    with acp.Pack("F64", 3) as pk:
        for spec in (dict(boundary="PERIODIC"), dict(cubic=True, op="RR"), dict(boundary="LODI-X")):
            d = acp.make_desc(24, 20, 3, tau=0.6, **spec)
            with acp.Case(d) as c:
                c.init_pulse(1e-3, 12.0, 10.0, 3.0)
                with pytest.raises(_capi.AcpError) as e:
                    pk.add(c, 10)
                assert e.value.code == _capi.ACP_E_UNSUPPORTED
        for spec in (dict(precision="F32-SHIFTED"), dict(order=2)):
            d = acp.make_desc(24, 20, 3, tau=0.6, **spec)
            with acp.Case(d) as c:
                c.init_pulse(1e-3, 12.0, 10.0, 3.0)
                with pytest.raises(_capi.AcpError) as e:
                    pk.add(c, 10)
                assert e.value.code == _capi.ACP_E_ARG
        d = acp.make_desc(24, 20, 3, tau=0.6)
        with acp.Case(d) as c:
            with pytest.raises(_capi.AcpError) as e:  # not initialised
                pk.add(c, 10)
            assert e.value.code == _capi.ACP_E_STATE
            c.init_pulse(1e-3, 12.0, 10.0, 3.0)
            c.set_probes([(3, 3, 1)], capacity=5)
            with pytest.raises(_capi.AcpError) as e:  # probe buffer too small
                pk.add(c, 10)
            assert e.value.code == _capi.ACP_E_STATE
            pk.add(c, 5)
            with pytest.raises(_capi.AcpError):
                pk.add(c, 5)
            pk.step(5)
            pk.remove(c)
            assert c.read_probes().shape == (5, 1)


def _small_mix(n, seed):
    reqs = serve.make_requests(n, seed=seed, nxy=(24, 72), steps=(20, 90))
    for i, r in enumerate(reqs):
        r["arrival"] = 0.002 * i
    return reqs


def test_packed_server_results_equal_alone(gpu):
    """The serving mix with packing on for every size (one-level cases of both
    tiers in packs, two-level cases on their own streams): each result equals
    the same request run alone, bit for bit."""
    reqs = _small_mix(16, seed=5)
    assert sum(not serve.two_level(r) for r in reqs) >= 6
    alone = [serve.run_alone(r) for r in reqs]
    results, summary = serve.run_mix(reqs, hbm_budget_bytes=1 << 30, max_active=8, graph_chunk=8, chunks_ahead=2,
                                     timeout_s=120, pack_max_nodes=1 << 30)
    assert summary["failed"] == 0, [(r["tag"], r["status"], r["error"]) for r in results if r["status"]]
    for r, a, q in zip(results, alone, reqs):
        assert r["steps"] == q["steps"]
        assert r["checksum"] == a["checksum"], q
        assert r["mass"] == a["mass"], q
        assert np.array_equal(r["probe"], a["probe"]), q
