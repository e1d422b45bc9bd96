"""Direct coupling (Coupling-Approach DC1 / DC2) on the GPU: both levels bit for
bit against the oracle's restatement in float64, the shifted tiers within a
stated tolerance of it, graph replays equal to plain launches."""
import numpy as np
import pytest

import acpulse_vclbm_amd as acp
import oracle_lib as ol
from make_golden import CS, DX, RHO0

pytestmark = pytest.mark.gpu

LINKSETS = [("Mf", "Mc", "hMf"), ("MfP", "Mc", "hMfP"), ("Mf", "McP", "hAll")]


def _pair(op, order, mrt, links, filt, periodic=True, x0=8, x1=16, nx=24, ny=10, nz=3, cx=6.0, prec="F64"):
    p = ol.make_params(nx, ny, nz, op=op, order=order, periodic=True if periodic else 3, tau=0.6, mrt_mode=mrt,
                       cs=CS, spacing=DX, rho0=RHO0, transverse=not periodic)
    o = ol.OracleRefined(p, x0, x1)
    o.set_coupling("DC1", *links)
    o.set_filter(filt)
    amp, cy, b = 1e-3, 5.0, 2.5
    o.levels[0].set_state(ol.pulse_ic(p, amp, cx, cy, b))
    o.levels[1].set_state(ol.pulse_ic(o.levels[1].p, amp, 2 * (cx - x0), 2 * cy, 2 * b))
    d = acp.make_desc(nx, ny, nz, op=op, order=order, tau=0.6, boundary="PERIODIC" if periodic else "LODI-X",
                      transverse=not periodic, mrt_mode="RELAXED" if mrt else "REFERENCE", precision=prec,
                      f16_scale=1000.0)
    r = acp.Refined(d, x0, x1)
    r.set_coupling("DC1", *links)
    r.set_filter(filt)
    r.init_pulse(amp, cx, cy, b)
    return o, r


@pytest.mark.parametrize("links", LINKSETS, ids=["-".join(s) for s in LINKSETS])
@pytest.mark.parametrize("op,order,mrt,filt", [("SRT", 3, 0, "NOFI"), ("RR", 3, 0, "LAG"), ("RR", 4, 0, "NOFI"),
                                               ("MRT-RM", 3, 1, "NOFI"), ("MRT-GS", 2, 1, "LAG"),
                                               ("HRR", 3, 0, "NOFI"), ("HRR", 4, 0, "LAG")])
def test_direct_coupling_matches_oracle_bitexact(gpu, op, order, mrt, filt, links):
    print("case", op, order, filt, links, flush=True)
    o, r = _pair(op, order, mrt, links, filt)
    with r:
        for lv in (0, 1):
            assert np.array_equal(r.level(lv).populations(), o.levels[lv].state())
        for n in (1, 4, 11):
            r.step(n)
            o.step(n)
            for lv in (0, 1):
                assert np.array_equal(r.level(lv).populations(), o.levels[lv].state()), (n, lv)


@pytest.mark.parametrize("x1", [17, 23], ids=["inside", "to-lodi-face"])
@pytest.mark.parametrize("op", ["RR", "HRR"])
def test_direct_coupling_lodi_x_matches_oracle_bitexact(gpu, op, x1):
    """LODI x faces on the coarse level; the fine region inside it (two
    coupled faces) or reaching its high-x face (one)."""
    print("case lodi-x", op, x1, flush=True)
    o, r = _pair(op, 3, 0, ("Mf", "Mc", "hMf"), "LAG", periodic=False, x1=x1, cx=12.0)
    with r:
        for n in (1, 4, 11):
            r.step(n)
            o.step(n)
            for lv in (0, 1):
                assert np.array_equal(r.level(lv).populations(), o.levels[lv].state()), (n, lv)


@pytest.mark.parametrize("op,prec,coupling,tol", [("RR", "F32-SHIFTED", "F64", 1e-4),
                                                  ("RR", "F32-SHIFTED", "F32-SHIFTED", 1e-3),
                                                  ("RR", "F16-SHIFTED", "F64", 2e-2),
                                                  ("HRR", "F32-SHIFTED", "F64", 1e-4)])
def test_direct_coupling_shifted_tiers_follow_float64(gpu, op, prec, coupling, tol):
    """Coarse density after 30 coarse steps within `tol` of the largest density
    deviation of the float64 oracle."""
    print("case", op, prec, coupling, flush=True)
    o, r = _pair(op, 3, 0, ("MfP", "Mc", "hMfP"), "NOFI", prec=prec)
    o.step(30)
    r64 = o.levels[0].state().sum(axis=0)
    with r:
        r.set_coupling_precision(coupling)
        r.step(30)
        rho, _ = r.level(0).macros()
    assert np.max(np.abs(rho - r64)) <= tol * np.max(np.abs(r64 - RHO0))


@pytest.mark.parametrize("op", ["RR", "HRR"])
def test_direct_coupling_graphs_equal_plain_launches(gpu, op):
    states = []
    for chunk in (0, 4):
        _, r = _pair(op, 3, 0, ("Mf", "McP", "hMf"), "LAG")
        with r:
            r.enable_graphs(chunk)
            r.step(12)
            states.append([r.level(lv).populations() for lv in (0, 1)])
    for lv in (0, 1):
        assert np.array_equal(states[0][lv], states[1][lv])


def test_direct_coupling_refusals(gpu):
    d = acp.make_desc(24, 10, 3, op="RR", tau=0.6, boundary="PERIODIC")
    with acp.Refined(d, 8, 16) as r:
        with pytest.raises(acp.AcpError):
            r.set_coupling("DC1", "MfP", "McP", "hMf")
        with pytest.raises(acp.AcpError):
            r.set_coupling("DC2", "fAll", "Mc", "hMf")
        with pytest.raises(acp.AcpError):
            r.set_coupling("STD", "Mf", "McP", "hAll")  # McP belongs to the direct coupling
        with pytest.raises(acp.AcpError):
            r.set_coupling("DC3")
        r.set_coupling("DC2", "Mf", "McP", "hMfP")
        r.set_coupling("STD", "Mf", "Mc", "hMf")
        r.set_coupling("STD")
    with acp.Refined(d, 8, 10) as r:
        with pytest.raises(acp.AcpError):
            r.set_coupling("STD", "Mf", "Mc", "hMf")  # one grid-coupling column inside each face


def test_hrr_switches_between_couplings_bitexact(gpu):
    """HRR moves its handed strain rates between the coarse grid-coupling
    column (standard) and the interface column (direct): a run that switches
    from one coupling to the other and back equals the oracle's doing the
    same."""
    o, r = _pair("HRR", 3, 0, ("Mf", "Mc", "hMf"), "NOFI")
    with r:
        for scheme in ("STD", "DC1", "STD", "DC2"):
            o.set_coupling(scheme, *(("fAll", "cAll", "hAll") if scheme == "STD" else ("Mf", "Mc", "hMf")))
            r.set_coupling(scheme, *(("fAll", "cAll", "hAll") if scheme == "STD" else ("Mf", "Mc", "hMf")))
            r.step(3)
            o.step(3)
            for lv in (0, 1):
                assert np.array_equal(r.level(lv).populations(), o.levels[lv].state()), (scheme, lv)


def test_refined_pulse_case_as_specified_matches_oracle(gpu):
    """The refined acoustic-pulse case as specified, at full size: coarse
    256 x 128 x 3 with LODI x faces and periodic y, z, fine level (ratio 2)
    over coarse x in [128, 255] ending on a LODI face, direct coupling, RR,
    F2C filter LAG, pulse 1e-3 of half-width 8 at (64, 64), probe at
    (96, 64, 1).  Its first 40 coarse steps: the probe series and both levels'
    states equal the oracle's bit for bit (the oracle restates the reference's
    arithmetic at a few MLUPS, so the 2000-step run is the GPU's alone)."""
    import os

    from acpulse_vclbm_amd.refinement import TAU
    steps = 40
    p = ol.make_params(256, 128, 3, op="RR", order=3, periodic=3, tau=TAU, mrt_mode=1, cs=CS, spacing=DX, rho0=RHO0,
                       nthreads=min(16, os.cpu_count() or 1))
    o = ol.OracleRefined(p, 128, 255)
    o.set_coupling("DC1", "Mf", "Mc", "hMf")
    o.set_filter("LAG")
    o.levels[0].set_state(ol.pulse_ic(p, 1e-3, 64.0, 64.0, 8.0))
    o.levels[1].set_state(ol.pulse_ic(o.levels[1].p, 1e-3, 2 * (64.0 - 128), 128.0, 16.0))
    print("oracle", steps, "steps", flush=True)
    po = o.step(steps, probes=[(96, 64, 1)])[:, 0]
    d = acp.make_desc(256, 128, 3, op="RR", order=3, tau=TAU, boundary="LODI-X", mrt_mode="RELAXED")
    with acp.Refined(d, 128, 255) as r:
        r.set_coupling("DC1", "Mf", "Mc", "hMf")
        r.set_filter("LAG")
        r.init_pulse(1e-3, 64.0, 64.0, 8.0)
        c = r.level(0)
        c.set_probes([(96, 64, 1)], capacity=steps)
        r.step(steps)
        pg = c.read_probes()[:, 0].copy()
        assert np.array_equal(pg, po)
        for lv in (0, 1):
            assert np.array_equal(r.level(lv).populations(), o.levels[lv].state()), lv
