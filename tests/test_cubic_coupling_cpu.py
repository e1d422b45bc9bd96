"""The cubic-Mach terms across a two-level interface, in the oracle (CPU).

The reference's partner functions (Node.cpp:1892-2490) evaluate the strain
rate and the cubic-Mach psi of an interface node from one stencil and hand
both to the grid-coupling node (Node.cpp:2069), whose own strain-rate function
returns at once (Node.cpp:1790-1793): it collides with the handed psi in its
moments, Pi1 and result.  The partner functions are pinned against the
reference build (test_partner_kat_cpu.py); the interpolation onto the fine face
lives in the absent GridCoupling.cpp and is restated like the strain rate's.
These tests pin the rest by what the scheme fixes:
  - the handed collision with the node's own stencil values is the node's own
    collision, bit for bit (the latter is pinned against the reference);
  - on a field whose cubic products are at most quadratic along each
    derivative, every handed psi is the exact one (a neighbour at the wrong
    distance, on the wrong side or from the wrong level fails it);
  - uniform flows stay uniform; a pulse crossing the interface reflects no
    more than without the cubic terms.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

L = ol.lib()
DP = ctypes.POINTER(ctypes.c_double)
L.acpo_collide_handed.argtypes = [ctypes.POINTER(ol.Consts), ctypes.POINTER(ol.Params), DP, DP, DP]
L.acpo_refined_transfer.argtypes = [ctypes.c_void_p, ctypes.c_int, DP, DP, DP]
W = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])


def D(a):
    return a.ctypes.data_as(DP)


@pytest.mark.parametrize("order", [2, 3, 4])
@pytest.mark.parametrize("op", ["RR", "HRR"])
def test_handed_collision_with_own_values_is_the_collision(op, order):
    """A node whose stored psi is what its own stencil gives, handed that psi and
    that strain rate, collides exactly as it would by itself; the handed psi
    and S both matter (zero psi, or another S for HRR, changes the result)."""
    p = ol.make_params(8, 8, 8, op=op, order=order, cubic=True, periodic=True, tau=0.56, sigma=0.6)
    k = ol.consts(p)
    rng = np.random.default_rng(order + 10 * (op == "HRR"))
    for trial in range(5):
        u0 = 0.05 * p.cs * rng.uniform(-1, 1, 3)
        f = ol.feq(p, order, p.rho0 * (1 + 0.01 * rng.uniform(-1, 1)), u0)
        f = f + p.rho0 * W * 2e-3 * rng.uniform(-1, 1, 19)
        nb = np.empty((6, 4))
        nb[:, 0] = p.rho0 * (1 + 0.01 * rng.uniform(-1, 1, 6))
        nb[:, 1:] = u0 + 0.02 * p.cs * rng.uniform(-1, 1, (6, 3))
        S, psi = np.empty(6), np.empty(19)
        L.acpo_strain_psi(ctypes.byref(k), D(nb), D(S), D(psi))
        assert np.max(np.abs(psi)) > 0
        own, ps = f.copy(), psi.copy()
        L.acpo_collide(ctypes.byref(k), ctypes.byref(p), D(own), D(ps), D(nb))
        assert np.array_equal(ps, psi)  # its stencil gives the same psi again
        handed = f.copy()
        L.acpo_collide_handed(ctypes.byref(k), ctypes.byref(p), D(handed), D(S), D(psi.copy()))
        assert np.array_equal(handed, own), trial
        zero = f.copy()
        L.acpo_collide_handed(ctypes.byref(k), ctypes.byref(p), D(zero), D(S), D(np.zeros(19)))
        assert np.max(np.abs(zero - own)) > 0
        if op == "HRR":
            other = f.copy()
            L.acpo_collide_handed(ctypes.byref(k), ctypes.byref(p), D(other), D(S * 1.5), D(psi.copy()))
            assert np.max(np.abs(other - own)) > 0


def _exact_psi(k, p, X, U, g):
    """psi of the field u = (U, g X, 0), uniform density, at coarse position X:
    every cubic product is at most quadratic in x and constant in y and z, so
    the central difference over +-1 coarse node gives the exact derivatives."""
    nb = np.empty((6, 4))
    for m, dx in enumerate((1, 0, -1, 0, 0, 0)):
        nb[m] = (p.rho0, U, g * (X + dx), 0.0)
    S, psi = np.empty(6), np.empty(19)
    L.acpo_strain_psi(ctypes.byref(k), D(nb), D(S), D(psi))
    return psi


@pytest.mark.parametrize("op", ["RR", "HRR"])
@pytest.mark.parametrize("scheme", ["STD", "DC1"])
@pytest.mark.parametrize("half", [0, 1])
def test_handed_psi_is_exact_on_a_quadratic_field(op, scheme, half):
    """u = (U, g X, 0) on both levels (X in coarse nodes from the centre), the
    coarse level equal at t and t + 1: every handed psi (the coarse partners',
    the level-1 direct-coupling stencil's, their spread over the fine faces and
    the fine partners') equals the exact psi at its node's position."""
    nx, ny, nz, x0, x1 = 24, 6, 3, 8, 16
    p = ol.make_params(nx, ny, nz, op=op, cubic=True, periodic=True, tau=0.6, nthreads=2)
    r = ol.OracleRefined(p, x0, x1)
    if scheme != "STD":
        r.set_coupling(scheme, "Mf", "Mc", "hMf")
    U, g = 0.05 * p.cs, 2e-3 * p.cs
    for lv, xs in ((0, np.arange(nx, dtype=float)), (1, x0 + 0.5 * np.arange(2 * (x1 - x0) + 1))):
        lp = r.levels[lv].p
        X = np.tile(xs, lp.ny * lp.nz) - 12.0
        u = np.vstack([np.full_like(X, U), g * X, np.zeros_like(X)])
        r.levels[lv].set_state(ol.equilibrium_field(lp, np.full(X.size, p.rho0), u))
    nv, cp, fp = 25, ny * nz, 4 * ny * nz
    col, gc, face = np.full(2 * nv * cp, np.nan), np.full(2 * nv * cp, np.nan), np.full(2 * nv * fp, np.nan)
    L.acpo_refined_transfer(r.h, half, D(col), D(gc), D(face))
    k = ol.consts(p)
    at = {"face": (x0, x1), "col": (x0, x1), "gc": (x0 + 1, x1 - 1) if scheme == "STD" else (x0, x1)}
    planes = {"face": face.reshape(2, nv, fp)}
    if half == 0:
        planes["col"] = col.reshape(2, nv, cp)
        if scheme == "STD":
            planes["gc"] = gc.reshape(2, nv, cp)
    elif scheme != "STD":
        planes["gc"] = gc.reshape(2, nv, cp)
    else:
        planes["col"] = col.reshape(2, nv, cp)
    checked = 0
    for name, P in planes.items():
        for s in range(2):
            want = _exact_psi(k, p, at[name][s] - 12.0, U, g)
            got = P[s, 6:]
            scale = np.max(np.abs(want))
            assert scale > 0
            for q in range(19):
                assert np.allclose(got[q], want[q], rtol=0, atol=1e-9 * scale), (name, s, q, got[q].min(), want[q])
            checked += 1
    assert checked >= 4


@pytest.mark.parametrize("op,scheme,filt", [("RR", "STD", "NOFI"), ("HRR", "STD", "LAG"), ("RR", "DC1", "LAG"),
                                            ("HRR", "DC1", "NOFI"), ("HRR", "STD", "TOU")])
def test_cubic_uniform_flow_stays_uniform(op, scheme, filt):
    """A uniform moving equilibrium is a fixed point on both levels with the
    cubic-Mach terms (psi of a uniform field is zero) and through the handed
    values."""
    p = ol.make_params(24, 10, 3, op=op, cubic=True, periodic=True, tau=0.6, nthreads=4)
    r = ol.OracleRefined(p, 8, 16)
    r.set_filter(filt)
    if scheme != "STD":
        r.set_coupling(scheme, "Mf", "Mc", "hMf")
    u = np.array([0.02 * p.cs, -0.01 * p.cs, 0.005 * p.cs])
    for lv in r.levels:
        lv.set_state(ol.equilibrium_field(lv.p, np.full(lv.N, p.rho0), np.repeat(u[:, None], lv.N, axis=1)))
    r.step(6)
    ref = r.levels[0].state()[:, 0]
    for lv in r.levels:
        assert np.max(np.abs(lv.state() - ref[:, None])) <= 1e-14 * np.max(np.abs(ref))
        assert np.max(np.abs(lv.psi())) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("op", ["RR", "HRR"])
def test_cubic_interface_reflection_is_small(op):
    """A pulse in a mean flow crossing the interface, seen by a coarse probe in
    front of it, against the same pulse on the uniform coarse grid (both with
    the cubic-Mach terms): identical until the echo can arrive, and the
    reflection no more than 10 % above that of the same run without them."""
    nx, ny, x0, x1, steps = 64, 32, 32, 64, 60
    res = {}
    for cubic in (False, True):
        p = ol.make_params(nx, ny, 3, op=op, cubic=cubic, periodic=True, tau=0.55, nthreads=8, sigma=0.5)
        r = ol.OracleRefined(p, x0, x1)
        amp, cx, cy, b, U = 1e-3, 16.0, 16.0, 4.0, 0.1 * p.cs
        states = []
        for lv, (lp, xs, sc) in enumerate(((p, np.arange(nx, dtype=float), 1.0),
                                            (r.levels[1].p, x0 + 0.5 * np.arange(2 * (x1 - x0) + 1), 2.0))):
            f = ol.pulse_ic(lp, amp, (cx - (x0 if lv else 0)) * sc, cy * sc, b * sc)
            rho, _ = ol.macros(lp, f)
            N = lp.nx * lp.ny * lp.nz
            u = np.vstack([np.full(N, U), np.zeros(N), np.zeros(N)])
            states.append(ol.equilibrium_field(lp, rho, u))
        r.levels[0].set_state(states[0])
        r.levels[1].set_state(states[1])
        un = ol.OracleCase(p)
        un.set_state(states[0])
        lc = 24 + nx * (16 + ny * 1)
        pr, pu = [], []
        for _ in range(steps):
            r.step(1)
            pr.append(r.levels[0].state()[:, lc].sum())
            pu.append(un.step(1, probes=[(24, 16, 1)])[0, 0])
        pr, pu = np.array(pr), np.array(pu)
        sig = np.max(np.abs(pu - pu[0]))
        assert sig > 0.05 * amp * p.rho0
        res[cubic] = (np.max(np.abs(pr[:20] - pu[:20])) / sig, np.max(np.abs(pr - pu)) / sig,
                      r.levels[1].psi())
    early, late, psi = res[True]
    assert early < 1e-3
    assert late < 0.05
    assert late <= 1.1 * res[False][1]
    assert np.max(np.abs(psi)) > 0  # the cubic terms are live on the fine level
