"""Two-level refinement: invariants of the oracle's restatement (CPU).

The reference's coupling source (GridCoupling.cpp) is absent, so the
refinement part of the oracle cannot be checked against the reference; these
tests pin it by what the scheme fixes: the non-equilibrium scale factor, exact
preservation of uniform states through every coupling operation, the open-x
push rule and agreement of a refined run with a uniform fine run.
"""
import numpy as np
import pytest

import oracle_lib as ol

W = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])


def refined(op="SRT", tau=0.6, nx=24, ny=10, nz=3, x0=8, x1=16):
    p = ol.make_params(nx, ny, nz, op=op, periodic=True, tau=tau, mrt_mode=1, nthreads=4)
    return ol.OracleRefined(p, x0, x1), p


@pytest.mark.parametrize("tau", [0.5003, 0.6, 1.0])
def test_scale_factor_is_tau_ratio(tau):
    r, _ = refined(tau=tau)
    tau_f = 2.0 * tau - 0.5  # same viscosity at half the spacing and time step
    assert r.scale() == pytest.approx(tau_f / (2.0 * tau), rel=1e-12)


@pytest.mark.parametrize("op,moving,filt", [("SRT", True, "NOFI"), ("RR", True, "NOFI"), ("SRT", False, "NOFI"),
                                            ("RR", False, "NOFI"), ("MRT-RM", False, "NOFI"),
                                            ("MRT-GS", False, "NOFI"), ("SRT", True, "LAG"), ("RR", False, "LAG"),
                                            ("HRR", True, "NOFI"), ("HRR", False, "LAG")])
def test_uniform_state_stays_uniform(op, moving, filt):
    """A uniform equilibrium is a fixed point of the step on each level and of
    the coupling (time interpolation of equal values, rescaling of a zero
    non-equilibrium part).  Moving equilibria only for the Hermite-equilibrium
    operators: the raw-moment and Gram-Schmidt MRT relax towards their own
    equilibria, of which a moving feq_3 is not one."""
    r, p = refined(op)
    r.set_filter(filt)
    u = np.array([0.02 * p.cs, -0.01 * p.cs, 0.0]) if moving else np.zeros(3)
    for lv in r.levels:
        N = lv.N
        rho = np.full(N, p.rho0)
        lv.set_state(ol.equilibrium_field(lv.p, rho, np.repeat(u[:, None], N, axis=1)))
    r.step(6)
    ref = r.levels[0].state()[:, 0]
    for lv in r.levels:
        f = lv.state()
        assert np.max(np.abs(f - ref[:, None])) <= 1e-14 * np.max(np.abs(ref))


def test_open_x_drops_exactly_the_links_leaving_in_x():
    p = ol.make_params(6, 5, 4, op="SRT", periodic=2, tau=0.6, nthreads=1)
    c = ol.OracleCase(p)
    rows = c.geometry()
    E = [(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 0, 1), (1, 0, -1), (0, 1, 1),
         (0, 1, -1), (-1, 0, 1), (-1, 0, -1), (0, -1, 1), (0, -1, -1), (-1, 1, 0), (1, -1, 0), (1, 1, 0), (-1, -1, 0)]
    for l in range(c.N):
        i, j, k = l % 6, (l // 6) % 5, l // 30
        assert rows[l, 26] == 0  # every node collides
        for q, (ex, ey, ez) in enumerate(E):
            if (i == 5 and ex > 0) or (i == 0 and ex < 0):
                assert rows[l, q] == -1
            else:
                assert rows[l, q] == (i + ex) + 6 * (((j + ey) % 5) + 5 * ((k + ez) % 4))


@pytest.mark.parametrize("op", ["SRT", "HRR"])
def test_interface_reflection_is_small(op):
    """A pulse crossing the coarse-to-fine interface, seen by a coarse-level probe
    in front of it, against the same pulse on the uniform coarse grid: identical
    physics until the interface's echo can arrive (pulse at x = 16, probe at
    x = 24, interface at x = 32: the echo travels 24 nodes, about 42 steps at
    c_s = 1/sqrt(3)); afterwards the difference is the reflection, a few per
    cent of the pulse for this ratio-2 interface.  HRR (sigma 0.5, the strain
    rate weighs half) couples its strain rates through the partner transfer."""
    nx, ny, x0, x1, steps = 64, 32, 32, 64, 60
    p = ol.make_params(nx, ny, 3, op=op, periodic=True, tau=0.55, nthreads=8, sigma=0.5)
    r = ol.OracleRefined(p, x0, x1)
    amp, cx, cy, b = 1e-3, 16.0, 16.0, 4.0
    r.levels[0].set_state(ol.pulse_ic(p, amp, cx, cy, b))
    fp = r.levels[1].p
    r.levels[1].set_state(ol.pulse_ic(fp, amp, 2 * (cx - x0), 2 * cy, 2 * b))
    u = ol.OracleCase(p)
    u.set_state(ol.pulse_ic(p, amp, cx, cy, b))
    lc = 24 + nx * (16 + ny * 1)
    pr, pu = [], []
    for _ in range(steps):
        r.step(1)
        pr.append(r.levels[0].state()[:, lc].sum())
        pu.append(u.step(1, probes=[(24, 16, 1)])[0, 0])
    pr, pu = np.array(pr), np.array(pu)
    sig = np.max(np.abs(pu - p.rho0))
    assert sig > 0.05 * amp * p.rho0  # the pulse reaches the probe
    early = np.max(np.abs(pr[:25] - pu[:25])) / sig
    late = np.max(np.abs(pr - pu)) / sig
    assert early < 1e-3
    assert late < 0.05


def _restricted(filt, fine_state, p, x0, x1):
    r = ol.OracleRefined(p, x0, x1)
    r.set_filter(filt)
    r.levels[0].set_state(np.zeros((19, p.nx * p.ny * p.nz)))
    r.levels[1].set_state(fine_state)
    r.restrict()
    return r.levels[0].state().reshape(19, p.nz, p.ny, p.nx), r.scale()


@pytest.mark.parametrize("axis", ["x", "y"])
def test_lag_is_the_19_point_mean_of_the_nonequilibrium(axis):
    """LAG hands the coarse node the arithmetic mean of the fine
    non-equilibria over the coincident node and its 18 D3Q19 neighbours.  For
    a non-equilibrium part h * (alpha + beta s^2) that varies quadratically
    along one axis s (fine node units; h carries no mass or momentum, so every
    node has the same density, velocity and equilibrium) the mean is
    h * (alpha + beta (s^2 + 10/19)): 10 of the 19 links have a non-zero
    component along s.  NOFI takes the coincident node's value, so
    LAG - NOFI = (1/a) beta (10/19) h exactly in real arithmetic (a 27-point
    mean would give 18/27)."""
    nx, ny, nz, x0, x1 = 20, 8, 3, 6, 14
    p = ol.make_params(nx, ny, nz, op="SRT", periodic=True, tau=0.6, nthreads=4)
    fp = ol.make_params(2 * (x1 - x0) + 1, 2 * ny, 2 * nz, op="SRT", periodic=2, tau=0.6, nthreads=4)
    E = np.array([(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 0, 1), (1, 0, -1),
                  (0, 1, 1), (0, 1, -1), (-1, 0, 1), (-1, 0, -1), (0, -1, 1), (0, -1, -1), (-1, 1, 0), (1, -1, 0),
                  (1, 1, 0), (-1, -1, 0), (0, 0, 0)])
    h = p.rho0 * W * (E[:, 0] ** 2 - E[:, 1] ** 2) * 1e-3
    fnx, fny, fnz = fp.nx, fp.ny, fp.nz
    k, j, i = np.meshgrid(np.arange(fnz), np.arange(fny), np.arange(fnx), indexing="ij")
    s = (i if axis == "x" else j).astype(float)
    alpha, beta = 0.3, 0.01
    feq0 = ol.equilibrium_field(fp, np.full(1, p.rho0), np.zeros((3, 1)))[:, 0]
    f = feq0[:, None] + h[:, None] * (alpha + beta * s.reshape(-1)[None, :] ** 2)
    lag, a = _restricted("LAG", f, p, x0, x1)
    nofi, _ = _restricted("NOFI", f, p, x0, x1)
    want = h / a * beta * 10.0 / 19.0
    # coarse nodes that were restricted; along y keep clear of the periodic wrap
    Js = range(1, ny - 1) if axis == "y" else range(ny)
    for I in range(x0 + 1, x1):
        for J in Js:
            for K in range(nz):
                d = lag[:, K, J, I] - nofi[:, K, J, I]
                assert np.max(np.abs(d - want)) <= 1e-9 * np.max(np.abs(want)), (I, J, K)
    # coarse nodes outside the fine region are left alone
    assert not np.any(lag[:, :, :, : x0 + 1]) and not np.any(lag[:, :, :, x1:])


E19 = np.array([(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 0, 1), (1, 0, -1),
                (0, 1, 1), (0, 1, -1), (-1, 0, 1), (-1, 0, -1), (0, -1, 1), (0, -1, -1), (-1, 1, 0), (1, -1, 0),
                (1, 1, 0), (-1, -1, 0), (0, 0, 0)])


def _tou_case():
    nx, ny, nz, x0, x1 = 20, 8, 3, 6, 14
    p = ol.make_params(nx, ny, nz, op="SRT", periodic=True, tau=0.6, nthreads=4)
    fp = ol.make_params(2 * (x1 - x0) + 1, 2 * ny, 2 * nz, op="SRT", periodic=2, tau=0.6, nthreads=4)
    return p, fp, x0, x1


@pytest.mark.parametrize("axis,shear,factor", [("x", (0, 1), 0.5), ("z", (0, 1), 0.0), ("y", (1, 2), 0.5),
                                               ("x", (1, 2), 0.0)])
def test_tou_filters_each_population_along_its_own_direction(axis, shear, factor):
    """TOU: a non-equilibrium part h * beta * (s - S)^2 quadratic along one axis
    s, h a pure shear (h_q = w_q xi_a xi_b, no mass, no momentum, so every
    node has the same equilibrium).  At the coarse node (s = S) the fine value
    is zero; the three-point mean along e_q gives beta h_q e_qs^2 / 2, which is
    h_q beta / 2 on the shear's links when they all move along s (xy along x,
    yz along y) and zero when none does (xy along z, yz along x), and carries
    no mass or momentum.  So TOU - NOFI = factor * beta * h / a exactly in
    real arithmetic; LAG (isotropic) would give 10/19 in both."""
    p, fp, x0, x1 = _tou_case()
    a_, b_ = shear
    h = p.rho0 * W * (E19[:, a_] * E19[:, b_]) * 1e-3
    fnx, fny, fnz = fp.nx, fp.ny, fp.nz
    k, j, i = np.meshgrid(np.arange(fnz), np.arange(fny), np.arange(fnx), indexing="ij")
    s = {"x": i, "y": j, "z": k}[axis].astype(float)
    feq0 = ol.equilibrium_field(fp, np.full(1, p.rho0), np.zeros((3, 1)))[:, 0]
    beta = 0.01
    I, J, K = x0 + 2, 2, 1  # the curvature is centred on this coarse node's fine partner
    S = {"x": 2 * (I - x0), "y": 2 * J, "z": 2 * K}[axis]
    f = feq0[:, None] + h[:, None] * beta * (s.reshape(-1)[None, :] - S) ** 2
    tou, a = _restricted("TOU", f, p, x0, x1)
    nofi, _ = _restricted("NOFI", f, p, x0, x1)
    d = tou[:, K, J, I] - nofi[:, K, J, I]
    want = factor * beta * h / a
    assert np.max(np.abs(d - want)) <= 1e-9 * np.max(np.abs(h / a))


def test_tou_keeps_the_coincident_density_and_velocity():
    """Random fine states: every restricted coarse node has the density and
    velocity of its coincident fine node (the filtered non-equilibrium carries
    no mass or momentum), and TOU differs from NOFI and from LAG."""
    p, fp, x0, x1 = _tou_case()
    rng = np.random.default_rng(11)
    FN = fp.nx * fp.ny * fp.nz
    f = p.rho0 * W[:, None] * (1.0 + 1e-3 * rng.standard_normal((19, FN)))
    tou, _ = _restricted("TOU", f, p, x0, x1)
    nofi, _ = _restricted("NOFI", f, p, x0, x1)
    lag, _ = _restricted("LAG", f, p, x0, x1)
    fr = f.reshape(19, fp.nz, fp.ny, fp.nx)
    for I in range(x0 + 1, x1):
        for J in range(p.ny):
            for K in range(p.nz):
                mc = ol.macros(p, tou[:, K, J, I][:, None])
                mf = ol.macros(fp, fr[:, 2 * K, 2 * J, 2 * (I - x0)][:, None])
                assert np.allclose(np.array(mc[0]), np.array(mf[0]), rtol=1e-14, atol=0)
                for a_ in range(3):
                    assert abs(mc[1][a_][0] - mf[1][a_][0]) <= 1e-13 * p.cs
    inside = (slice(None), slice(None), slice(None), slice(x0 + 1, x1))
    assert not np.allclose(tou[inside], nofi[inside], rtol=1e-12, atol=0)
    assert not np.allclose(tou[inside], lag[inside], rtol=1e-12, atol=0)


@pytest.mark.parametrize("op,moving", [("SRT", True), ("RR", False), ("HRR", True)])
def test_tou_uniform_state_stays_uniform(op, moving):
    r, p = refined(op)
    r.set_filter("TOU")
    u = np.array([0.02 * p.cs, -0.01 * p.cs, 0.0]) if moving else np.zeros(3)
    for lv in r.levels:
        lv.set_state(ol.equilibrium_field(lv.p, np.full(lv.N, p.rho0), np.repeat(u[:, None], lv.N, axis=1)))
    r.step(6)
    ref = r.levels[0].state()[:, 0]
    for lv in r.levels:
        assert np.max(np.abs(lv.state() - ref[:, None])) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("filt", ["LAG", "TOU"])
def test_filtered_interface_reflection_is_small(filt):
    """The interface-reflection check of the NOFI coupling, with LAG and TOU."""
    nx, ny, x0, x1, steps = 64, 32, 32, 64, 60
    p = ol.make_params(nx, ny, 3, op="SRT", periodic=True, tau=0.55, nthreads=8)
    r = ol.OracleRefined(p, x0, x1)
    r.set_filter(filt)
    amp, cx, cy, b = 1e-3, 16.0, 16.0, 4.0
    r.levels[0].set_state(ol.pulse_ic(p, amp, cx, cy, b))
    r.levels[1].set_state(ol.pulse_ic(r.levels[1].p, amp, 2 * (cx - x0), 2 * cy, 2 * b))
    u = ol.OracleCase(p)
    u.set_state(ol.pulse_ic(p, amp, cx, cy, b))
    lc = 24 + nx * (16 + ny * 1)
    pr, pu = [], []
    for _ in range(steps):
        r.step(1)
        pr.append(r.levels[0].state()[:, lc].sum())
        pu.append(u.step(1, probes=[(24, 16, 1)])[0, 0])
    pr, pu = np.array(pr), np.array(pu)
    sig = np.max(np.abs(pu - p.rho0))
    assert np.max(np.abs(pr[:25] - pu[:25])) / sig < 1e-3
    assert np.max(np.abs(pr - pu)) / sig < 0.05


def test_open_x_hrr_strain_is_second_order_one_sided():
    """HRR on a grid with open x faces (the fine level of a refined case): the
    FD strain at a face node takes the missing neighbour's macros by quadratic
    extrapolation, which makes the central difference the second-order
    one-sided one, exact for a velocity linear in x.  So a grid whose face
    lies at x = 0 evolves like a wider grid in which the same nodes are
    interior: after one HRR step (sigma 0.5, the strain weighs half) the nodes
    that received pushes from the face node agree to rounding.  A zero-gradient
    ghost (u(-1) = u(0)) would halve the face strain and miss by orders of
    magnitude."""
    ny, nz, shift = 4, 3, 2
    kw = dict(op="HRR", order=3, periodic=2, tau=0.6, sigma=0.5, nthreads=1)
    pa = ol.make_params(6, ny, nz, **kw)
    pb = ol.make_params(6 + 2 * shift, ny, nz, **kw)
    G = 0.02 * pa.cs / 6.0  # u_x = G * x (x in nodes)

    def state(p, x0):
        N = p.nx * p.ny * p.nz
        i = np.arange(N) % p.nx
        u = np.zeros((3, N))
        u[0] = G * (i - x0)
        return ol.equilibrium_field(p, np.full(N, p.rho0), u)

    a, b = ol.OracleCase(pa), ol.OracleCase(pb)
    a.set_state(state(pa, 0))
    b.set_state(state(pb, shift))
    a.step(1)
    b.step(1)
    fa = a.state().reshape(19, nz, ny, pa.nx)
    fb = b.state().reshape(19, nz, ny, pb.nx)
    # nodes 1..4 of the narrow grid: every push they receive comes from nodes
    # whose state and strain match the wide grid's nodes 3..6
    da = fa[..., 1:5]
    db = fb[..., 1 + shift:5 + shift]
    assert np.max(np.abs(da - db)) <= 1e-12 * np.max(np.abs(db))


def test_refined_hrr_with_sigma_one_is_refined_rr():
    """HRR with sigma = 1 gives the strain rate no weight, so the refined HRR
    step, whose interface columns collide with the strain rates handed over by
    the other level, equals the refined RR step bit for bit: the handed-over
    path changes nothing but the strain rate."""
    for periodic, x0, x1 in ((1, 6, 14), (3, 6, 19)):
        out = []
        for op in ("RR", "HRR"):
            p = ol.make_params(20, 8, 3, op=op, order=3, periodic=periodic, tau=0.56, sigma=1.0, nthreads=4)
            r = ol.OracleRefined(p, x0, x1)
            r.levels[0].set_state(ol.pulse_ic(p, 1e-2, 8.0, 4.0, 2.5))
            fp = r.levels[1].p
            r.levels[1].set_state(ol.pulse_ic(fp, 1e-2, 2 * (8.0 - x0), 8.0, 5.0))
            r.step(7)
            out.append((r.levels[0].state(), r.levels[1].state()))
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])

