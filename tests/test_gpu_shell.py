"""The reference's two-level geometry on the GPU (acp_refined_create_shell): a
fine shell around a coarse core, both levels bit for bit against the
oracle's restatement in float64."""
import numpy as np
import pytest

import acpulse_vclbm_amd as acp
import oracle_lib as ol
from make_golden import CS, DX, RHO0

pytestmark = pytest.mark.gpu

N, LAYERS = (21, 17, 13), (3, 6)


def _pair(op, order, mrt, cubic, links, filt, transverse=False):
    p = ol.make_params(*N, op=op, order=order, periodic=False, tau=0.6, mrt_mode=mrt, cubic=cubic, cs=CS, spacing=DX,
                       rho0=RHO0, transverse=transverse, sigma=0.7)
    o = ol.OracleRefined(p, 0, 0, shell=LAYERS)
    o.set_coupling("STD", *links)
    o.set_filter(filt)
    d = acp.make_desc(*N, op=op, order=order, tau=0.6, sigma=0.7, cubic=cubic, boundary="LODI", transverse=transverse,
                      mrt_mode="RELAXED" if mrt else "REFERENCE")
    r = acp.Refined(d, shell=LAYERS)
    r.set_coupling("STD", *links)
    r.set_filter(filt)
    return o, r, p


def _pulse(lp, amp, c, b, U=0.0):
    """A Gaussian density pulse (centre c, half-width b in coarse nodes) in a
    uniform flow U along x, at a level's nodes, as feq."""
    h = lp.spacing / DX
    x = np.tile(np.arange(lp.nx), lp.ny * lp.nz) * h
    y = np.tile(np.repeat(np.arange(lp.ny), lp.nx), lp.nz) * h
    z = np.repeat(np.arange(lp.nz), lp.nx * lp.ny) * h
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    rho = RHO0 * (1 + amp * np.exp(-np.log(2) * r2 / b ** 2))
    u = np.vstack([np.full(x.size, U), np.zeros(x.size), np.zeros(x.size)])
    return ol.equilibrium_field(lp, rho, u)


CASES = [("SRT", 3, 0, False, ("fAll", "cAll", "hAll"), "NOFI"),
         ("RR", 3, 0, False, ("Mf", "Mc", "hMf"), "LAG"),
         ("HRR", 3, 0, False, ("fAll", "cAll", "hAll"), "NOFI"),
         ("HRR", 4, 0, False, ("MfP", "Mc", "hMfP"), "TOU"),
         ("MRT-RM", 3, 1, False, ("MfP", "cAll", "hMf"), "NOFI"),
         ("MRT-GS", 2, 1, False, ("fAll", "Mc", "hAll"), "LAG"),
         ("RR", 4, 0, True, ("fAll", "cAll", "hAll"), "NOFI"),
         ("HRR", 3, 0, True, ("Mf", "Mc", "hMf"), "LAG")]


@pytest.mark.parametrize("op,order,mrt,cubic,links,filt", CASES,
                         ids=[f"{c[0]}-o{c[1]}-{'-'.join(c[4])}-{c[5]}" + ("-cubic" if c[3] else "") for c in CASES])
def test_shell_matches_oracle_bitexact(gpu, op, order, mrt, cubic, links, filt):
    """A pulse in a mean flow crossing the interface at its faces, edges and
    corners: both levels' populations (and psi with the cubic-Mach terms) bit
    for bit after 1, 4 and 11 coarse steps."""
    print("case shell", op, order, links, filt, cubic, flush=True)
    o, r, p = _pair(op, order, mrt, cubic, links, filt, transverse=op == "HRR")
    U = 0.0 if op.startswith("MRT") else 0.05 * CS
    f0 = [_pulse(lv.p, 1e-3, (9.0, 8.0, 6.0), 2.0, U) for lv in o.levels]
    with r:
        for lv in (0, 1):
            o.levels[lv].set_state(f0[lv])
            r.level(lv).set_populations(f0[lv])
        for n in (1, 4, 11):
            r.step(n)
            o.step(n)
            for lv in (0, 1):
                assert np.array_equal(r.level(lv).populations(), o.levels[lv].state()), (n, lv)
                if cubic:
                    assert np.array_equal(r.level(lv).psi(), o.levels[lv].psi()), (n, lv)


@pytest.mark.parametrize("filt", ["NOFI", "LAG", "TOU"])
def test_shell_restriction_alone_matches_oracle(gpu, filt):
    o, r, p = _pair("SRT", 3, 0, False, ("fAll", "Mc", "hAll"), filt)
    rng = np.random.default_rng(5)
    W = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])
    fs = [RHO0 * W[:, None] * (1.0 + 1e-3 * rng.standard_normal((19, lv.N))) for lv in o.levels]
    with r:
        for lv in (0, 1):
            o.levels[lv].set_state(fs[lv])
            r.level(lv).set_populations(fs[lv])
        o.restrict()
        r.restrict()
        assert np.array_equal(r.level(0).populations(), o.levels[0].state())


def test_shell_graphs_equal_plain_launches(gpu):
    states = []
    for chunk in (0, 4):
        _, r, p = _pair("HRR", 3, 0, False, ("fAll", "cAll", "hAll"), "LAG")
        with r:
            r.init_pulse(1e-3, 9.0, 8.0, 2.0)
            r.enable_graphs(chunk)
            r.step(12)
            states.append([r.level(lv).populations() for lv in (0, 1)])
    for lv in (0, 1):
        assert np.array_equal(states[0][lv], states[1][lv])


def test_shell_refusals(gpu):
    kw = dict(op="RR", tau=0.6, boundary="LODI")
    for desc, layers in ((acp.make_desc(*N, **dict(kw, boundary="PERIODIC")), LAYERS),
                         (acp.make_desc(*N, **dict(kw, precision="F32-SHIFTED")), LAYERS),
                         (acp.make_desc(*N, **kw), (1, 6)), (acp.make_desc(*N, **kw), (6, 6)),
                         (acp.make_desc(*N, **dict(kw, op="SRT", cubic=True)), LAYERS)):
        with pytest.raises(acp.AcpError):
            acp.Refined(desc, shell=layers)
    with acp.Refined(acp.make_desc(*N, **kw), shell=LAYERS) as r:
        with pytest.raises(acp.AcpError):
            r.set_coupling("DC1", "Mf", "Mc", "hMf")
        with pytest.raises(acp.AcpError):
            r.set_coupling("STD", "Mf", "McP", "hMf")


@pytest.mark.parametrize("cubic", [False, True])
def test_shell_vtk_is_the_host_writer_of_the_oracle_states(gpu, tmp_path, cubic):
    """acp_refined_write_vtk of a shell case after 4 coarse steps (Non-Reflecting
    output, the cubic-Mach terms in the density) is byte for byte the host
    writer's file of the oracle's states."""
    o, r, p = _pair("HRR" if cubic else "RR", 3, 0, cubic, ("fAll", "cAll", "hAll"), "NOFI")
    f0 = [_pulse(lv.p, 1e-3, (9.0, 8.0, 6.0), 2.0, 0.05 * CS) for lv in o.levels]
    with r:
        for lv in (0, 1):
            o.levels[lv].set_state(f0[lv])
            r.level(lv).set_populations(f0[lv])
        r.step(4)
        o.step(4)
        r.write_vtk(str(tmp_path / "gpu"))
    d = acp.make_desc(*N, op="HRR" if cubic else "RR", order=3, tau=0.6, sigma=0.7, cubic=cubic, boundary="LODI")
    assert d.nonreflecting
    D = acp._capi.dptr
    st = [(lv.state(), lv.prev(), lv.psi() if cubic else None) for lv in o.levels]
    acp._capi.check(acp.lib().acp_write_vtk_shell_host(d, LAYERS[0], LAYERS[1], D(st[0][0]), D(st[0][1]),
                                                        D(st[0][2]) if cubic else None, D(st[1][0]), D(st[1][1]),
                                                        D(st[1][2]) if cubic else None,
                                                        str(tmp_path / "host").encode()))
    a = open(tmp_path / "gpu.vtk", "rb").read()
    b = open(tmp_path / "host.vtk", "rb").read()
    assert len(a) > 100000 and a == b


@pytest.mark.parametrize("op,order,cubic,links", [("HRR", 3, False, ("Mf", "Mc", "hMf")),
                                                  ("RR", 4, True, ("fAll", "cAll", "hAll"))])
def test_shell_leaves_the_hole_unread(gpu, op, order, cubic, links):
    """NaN planted in the fine nodes strictly inside the hole box (which the
    fine level does not step): both levels equal the oracle's bit for bit,
    NaN where the oracle holds it, and no NaN outside the hole."""
    print("case hole", op, order, cubic, flush=True)
    o, r, p = _pair(op, order, 0, cubic, links, "LAG", transverse=op == "HRR")
    f0 = [_pulse(lv.p, 1e-3, (9.0, 8.0, 6.0), 2.0, 0.05 * CS) for lv in o.levels]
    lp = o.levels[1].p
    i = np.tile(np.arange(lp.nx), lp.ny * lp.nz)
    j = np.tile(np.repeat(np.arange(lp.ny), lp.nx), lp.nz)
    k = np.repeat(np.arange(lp.nz), lp.nx * lp.ny)
    lyz, lx = LAYERS
    b = (2 * lx, 2 * (N[0] - 1 - lyz), 2 * lyz, 2 * (N[1] - 1 - lyz), 2 * lyz, 2 * (N[2] - 1 - lyz))
    hole = (i > b[0]) & (i < b[1]) & (j > b[2]) & (j < b[3]) & (k > b[4]) & (k < b[5])
    f0[1][:, hole] = np.nan
    with r:
        for lv in (0, 1):
            o.levels[lv].set_state(f0[lv])
            r.level(lv).set_populations(f0[lv])
        for n in (1, 6):
            r.step(n)
            o.step(n)
            for lv in (0, 1):
                g = r.level(lv).populations()
                assert np.array_equal(g, o.levels[lv].state(), equal_nan=True), (n, lv)
                assert not np.isnan(g[:, ~hole] if lv else g).any(), (n, lv)
