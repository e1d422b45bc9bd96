"""Direct coupling (Coupling-Approach DC1 / DC2): invariants of the oracle's
restatement (CPU).

The reference declares the coupling (GridCoupling.h:28-31, 45) and calls it
(Lattice.cpp:14018-14058 at t + 1/2, 14699-14739 at t + 1) but its source,
GridCoupling.cpp, is absent, so these tests pin the restatement by what the
scheme fixes:
  - a pair of coincident nodes that already hold one state (the same
    equilibrium, the fine non-equilibrium a times the coarse one) is a fixed
    point of the pair solve, for every link set, face and Hermite order;
  - the solved state satisfies its defining equation: the coarse node
    completed with the rescaled fine-side links has the solved density and
    momentum;
  - uniform equilibria stay uniform through whole refined steps;
  - a pulse crossing the interface: no signal before the echo can arrive, a
    small echo, the coarse level's mass kept;
  - the combinations the case file forbids are refused.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

L = ol.lib()
KP = ctypes.POINTER(ol.Consts)
DP = ctypes.POINTER(ctypes.c_double)
L.acpo_dc_matrix.argtypes = [KP, ctypes.c_double, ctypes.c_int, ctypes.c_int, DP]
L.acpo_dc_state.argtypes = [KP, ctypes.c_int, ctypes.c_double, DP, ctypes.c_int, ctypes.c_int, DP, DP, DP]

E = np.array([(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 0, 1), (1, 0, -1), (0, 1, 1),
              (0, 1, -1), (-1, 0, 1), (-1, 0, -1), (0, -1, 1), (0, -1, -1), (-1, 1, 0), (1, -1, 0), (1, 1, 0),
              (-1, -1, 0), (0, 0, 0)])


def fine_side(set_, n):
    """Links the coarse node takes from the fine side on the face with
    outward fine-region normal -n: those arriving from inside the fine region
    (e_x = -n), with the common ones (e_x = 0) for McP."""
    return (E[:, 0] == -n) | ((set_ == 1) & (E[:, 0] == 0))


def pair_state(p, order, a, set_, n, cv, fv):
    k = ol.consts(p)
    m = np.empty(6)
    L.acpo_dc_matrix(ctypes.byref(k), a, set_, n, ol.dptr(m))
    e = np.empty(19)
    L.acpo_dc_state(ctypes.byref(k), order, a, ol.dptr(m), set_, n, ol.dptr(np.ascontiguousarray(cv)),
                    ol.dptr(np.ascontiguousarray(fv)), ol.dptr(e))
    return e


def moments(p, f):
    xi = np.array(ol.consts(p).xi)
    return np.array([f.sum(), (xi[:, 0] * f).sum(), (xi[:, 1] * f).sum(), (xi[:, 2] * f).sum()])


def zero_moment_noneq(p, rng, scale):
    """A non-equilibrium part with no mass and no momentum."""
    xi = np.array(ol.consts(p).xi)
    A = np.vstack([np.ones(19), xi.T])  # 4 x 19
    v = rng.standard_normal(19)
    v -= A.T @ np.linalg.solve(A @ A.T, A @ v)
    return scale * v


PARAMS = dict(tau=0.55, cs=1. / np.sqrt(3.), spacing=1.0, rho0=1.0)


@pytest.mark.parametrize("order", [2, 3, 4])
@pytest.mark.parametrize("set_", [0, 1])
@pytest.mark.parametrize("n", [1, -1])
def test_consistent_pair_is_a_fixed_point(order, set_, n):
    p = ol.make_params(8, 8, 3, order=order, **PARAMS)
    rng = np.random.default_rng(7 + order + 3 * set_ + n)
    for a in (0.5454545, 0.8, 0.95):
        for _ in range(8):
            rho = 1.0 + 0.01 * rng.standard_normal()
            u = 0.05 * p.cs * rng.standard_normal(3)
            eq = ol.feq(p, order, rho, u)
            neq = zero_moment_noneq(p, rng, 1e-4 * rho)
            e = pair_state(p, order, a, set_, n, eq + neq, eq + a * neq)
            assert np.max(np.abs(e - eq)) <= 4e-15 * rho, (a, rho, u)


@pytest.mark.parametrize("order", [2, 3, 4])
@pytest.mark.parametrize("set_", [0, 1])
@pytest.mark.parametrize("n", [1, -1])
def test_pair_state_solves_the_completion(order, set_, n):
    """Arbitrary (inconsistent) pairs: the coarse node completed with
    feq + (1/a)(fv - feq) on the fine-side links has the mass and momentum of
    the returned equilibrium (whose own moments are rho and rho u)."""
    p = ol.make_params(8, 8, 3, order=order, **PARAMS)
    rng = np.random.default_rng(100 + order)
    C = fine_side(set_, n)
    for a in (0.5454545, 0.9):
        for _ in range(8):
            base = ol.feq(p, order, 1.0, 0.03 * p.cs * rng.standard_normal(3))
            cv = base + 1e-3 * rng.standard_normal(19) * base
            fv = base + 1e-3 * rng.standard_normal(19) * base
            e = pair_state(p, order, a, set_, n, cv, fv)
            g = np.where(C, e + (1.0 / a) * (fv - e), cv)
            me, mg = moments(p, e), moments(p, g)
            assert np.max(np.abs(me - mg)) <= 2e-15 * max(1.0, np.max(np.abs(me))), (me, mg)
            # and the plain, unscaled mix is not a solution (the rescaling is live)
            g0 = np.where(C, fv, cv)
            assert np.max(np.abs(moments(p, g0) - me)) > 1e-9


def refined(op="SRT", nx=24, ny=10, nz=3, x0=8, x1=16, periodic=True):
    p = ol.make_params(nx, ny, nz, op=op, periodic=periodic, tau=0.6, mrt_mode=1, nthreads=4)
    return ol.OracleRefined(p, x0, x1), p


LINKSETS = [("Mf", "Mc", "hMf"), ("MfP", "Mc", "hMfP"), ("Mf", "McP", "hMf"), ("Mf", "Mc", "hAll")]


@pytest.mark.parametrize("links", LINKSETS, ids=["-".join(s) for s in LINKSETS])
@pytest.mark.parametrize("op,moving", [("SRT", True), ("RR", True), ("RR", False), ("MRT-RM", False),
                                       ("MRT-GS", False), ("HRR", True)])
def test_uniform_state_stays_uniform(op, moving, links):
    r, p = refined(op)
    r.set_coupling("DC1", *links)
    r.set_filter("LAG")
    u = np.array([0.02 * p.cs, -0.01 * p.cs, 0.0]) if moving else np.zeros(3)
    for lv in r.levels:
        N = lv.N
        lv.set_state(ol.equilibrium_field(lv.p, np.full(N, p.rho0), np.repeat(u[:, None], N, axis=1)))
    r.step(6)
    ref = r.levels[0].state()[:, 0]
    for lv in r.levels:
        assert np.max(np.abs(lv.state() - ref[:, None])) <= 1e-14 * np.max(np.abs(ref))


def test_forbidden_combinations_are_refused():
    r, _ = refined("RR")
    with pytest.raises(ValueError):
        r.set_coupling("DC1", "MfP", "McP", "hMf")  # EITHER MfP-Mc OR Mf-McP
    with pytest.raises(ValueError):
        r.set_coupling("DC1", "fAll", "Mc", "hMf")  # fAll with STD only
    with pytest.raises(ValueError):
        r.set_coupling("STD", "Mf", "McP", "hAll")  # McP with DC only
    with pytest.raises(ValueError):
        refined("RR", x1=10)[0].set_coupling("STD", "Mf", "Mc", "hAll")  # one grid-coupling column per face
    r.set_coupling("DC2", "Mf", "McP", "hMfP")
    r.set_coupling("STD", "MfP", "Mc", "hMf")


def _pulse(scheme, links=("Mf", "Mc", "hMf"), op="RR"):
    """Pulse at x = 16 in a periodic 64 x 32 x 3 duct refined over x in
    [32, 64]; probe at x = 24.  Returns the refined and the uniform coarse
    probe signals and the coarse level's total mass after each step."""
    nx, ny, x0, x1, steps = 64, 32, 32, 64, 60
    p = ol.make_params(nx, ny, 3, op=op, periodic=True, tau=0.55, nthreads=8, sigma=0.5)
    r = ol.OracleRefined(p, x0, x1)
    if scheme != "STD":
        r.set_coupling(scheme, *links)
    amp, cx, cy, b = 1e-3, 16.0, 16.0, 4.0
    r.levels[0].set_state(ol.pulse_ic(p, amp, cx, cy, b))
    r.levels[1].set_state(ol.pulse_ic(r.levels[1].p, amp, 2 * (cx - x0), 2 * cy, 2 * b))
    u = ol.OracleCase(p)
    u.set_state(ol.pulse_ic(p, amp, cx, cy, b))
    lc = 24 + nx * (16 + ny * 1)
    pr, pu, mass = [], [], []
    for _ in range(steps):
        r.step(1)
        st = r.levels[0].state()
        pr.append(st[:, lc].sum())
        mass.append(st.sum())
        pu.append(u.step(1, probes=[(24, 16, 1)])[0, 0])
    excess = ol.pulse_ic(p, amp, cx, cy, b).sum() - p.rho0 * p.nx * p.ny * p.nz
    return np.array(pr), np.array(pu), np.array(mass), excess, p


@pytest.mark.parametrize("op,links", [("RR", LINKSETS[0]), ("RR", LINKSETS[1]), ("RR", LINKSETS[2]),
                                      ("HRR", LINKSETS[0])], ids=["RR-Mf-Mc", "RR-MfP-Mc", "RR-Mf-McP", "HRR-Mf-Mc"])
def test_interface_reflection_is_small(op, links):
    """As for the standard coupling: identical physics at the probe until the
    interface's echo can arrive (the echo travels 24 nodes, about 42 steps),
    then an echo below half a per cent of the pulse (the standard coupling's
    is 0.27 %); the coarse level's mass (its nodes inside the fine region
    restricted from it) changes by less than 0.25 % of the pulse's excess
    mass.  HRR (sigma 0.5) hands its strain rates across the interface."""
    pr, pu, mass, excess, p = _pulse("DC1", links, op)
    sig = np.max(np.abs(pu - p.rho0))
    assert sig > 0.05 * 1e-3 * p.rho0
    assert np.max(np.abs(pr[:25] - pu[:25])) / sig < 1e-3
    assert np.max(np.abs(pr - pu)) / sig < 0.005
    m0 = ol.pulse_ic(p, 1e-3, 16.0, 16.0, 4.0).sum()
    assert np.max(np.abs(mass - m0)) < 0.0025 * excess


def test_direct_coupling_is_not_the_standard_one():
    dc, pu, _, _, p = _pulse("DC1")
    std = _pulse("STD")[0]
    sig = np.max(np.abs(pu - p.rho0))
    assert np.max(np.abs(dc - std)) / sig > 1e-4


L.acpo_refined_transfer.argtypes = [ctypes.c_void_p, ctypes.c_int, DP, DP, DP]


@pytest.mark.parametrize("scheme", ["STD", "DC1"])
@pytest.mark.parametrize("half", [0, 1])
def test_hrr_transfer_is_exact_on_a_linear_shear(scheme, half):
    """A velocity linear in x (u_y = g x, u_z = h x, uniform density) on both
    levels, the coarse level equal at t and t + 1: every handed strain rate
    (the coarse partners', the level-1 direct-coupling stencil's, their
    spread over the fine faces and the fine partners') is the exact
    S_xy = g / 2, S_xz = h / 2 (per metre), the others zero.  A neighbour
    taken one fine node away at the coarse distance, a wrong distance or a
    missing neighbour read from the wrong side fails it."""
    nx, ny, nz, x0, x1 = 24, 6, 3, 8, 16
    p = ol.make_params(nx, ny, nz, op="HRR", periodic=True, tau=0.6, nthreads=2)
    r = ol.OracleRefined(p, x0, x1)
    if scheme != "STD":
        r.set_coupling(scheme, "Mf", "Mc", "hMf")
    g, h = 2e-3 * p.cs, -1e-3 * p.cs  # velocity change per coarse node
    for lv, xs in ((0, np.arange(nx, dtype=float)), (1, x0 + 0.5 * np.arange(2 * (x1 - x0) + 1))):
        lp = r.levels[lv].p
        X = np.tile(xs, lp.ny * lp.nz) - 12.0
        u = np.vstack([np.zeros_like(X), g * X, h * X])
        f = ol.equilibrium_field(lp, np.full(X.size, p.rho0), u)
        r.levels[lv].set_state(f)  # both arrays: the coarse level equal at t and t + 1
    cp, fp = ny * nz, 4 * ny * nz
    col, gc, face = np.full(2 * 6 * cp, np.nan), np.full(2 * 6 * cp, np.nan), np.full(2 * 6 * fp, np.nan)
    ol.lib().acpo_refined_transfer(r.h, half, ol.dptr(col), ol.dptr(gc), ol.dptr(face))
    want = np.array([0., 0., 0., 0.5 * g / p.spacing, 0.5 * h / p.spacing, 0.])
    planes = [face.reshape(2, 6, fp)]
    if half == 0:
        planes.append(col.reshape(2, 6, cp))
        if scheme == "STD":
            planes.append(gc.reshape(2, 6, cp))
    elif scheme != "STD":
        planes.append(gc.reshape(2, 6, cp))
    else:
        planes.append(col.reshape(2, 6, cp))
    for P in planes:
        for v in range(6):
            assert np.allclose(P[:, v], want[v], rtol=1e-9, atol=1e-12 * abs(g / p.spacing)), (v, P[:, v].min(),
                                                                                             P[:, v].max())
