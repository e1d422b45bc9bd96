"""Pins of the oracle that need no reference output: closed forms of the
lattice, moment identities of the equilibria, conservation, special-case
reductions between operators, transform round trips, and symmetries of the
periodic mode (which the reference does not have)."""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

L = ol.lib()
D = ol.dptr
E = np.array([[1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 0, 1], [1, 0, -1],
              [0, 1, 1], [0, 1, -1], [-1, 0, 1], [-1, 0, -1], [0, -1, 1], [0, -1, -1], [-1, 1, 0], [1, -1, 0],
              [1, 1, 0], [-1, -1, 0], [0, 0, 0]])


def _p(**kw):
    return ol.make_params(9, 8, 7, **kw)


def test_lattice_closed_forms():
    k = ol.consts(_p())
    w = np.array(k.w)
    xi = np.array(k.xi)
    cs2 = k.cs2
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.abs((w[:, None] * xi).sum(0)).max() < 1e-12 * k.c
    assert np.allclose((w[:, None, None] * xi[:, :, None] * xi[:, None, :]).sum(0), cs2 * np.eye(3),
                       rtol=1e-14, atol=0)
    # the direction order and the opposite map (FluidNode.h:45-62)
    assert np.array_equal(np.sign(xi).astype(int), E)
    opp = [2, 3, 0, 1, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 15, 14, 17, 16, 18]
    assert np.array_equal(E[opp], -E)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_equilibrium_moments(order):
    """sum feq = rho, sum xi feq = rho u, sum (xi xi) feq = rho (u u + cs^2 I).

    Order 4 keeps the first two exactly but not the third: the reference forms
    its fourth-order polynomials as H_xxyy + (1/6) H_zz in SI units
    (Node.cpp:1040-1052), adding m^2/s^2 to m^4/s^4, so the orthogonalisation
    against second order is lost and O(u^4) terms leak into the stress
    (about 1e-6 relative at u = 0.05 c_s)."""
    p = _p(order=order)
    k = ol.consts(p)
    xi = np.array(k.xi)
    rng = np.random.default_rng(order)
    for _ in range(20):
        rho = p.rho0 * (1 + 0.01 * rng.uniform(-1, 1))
        u = 0.05 * p.cs * rng.uniform(-1, 1, 3)
        f = ol.feq(p, order, rho, u)
        assert abs(f.sum() - rho) < 1e-14 * rho
        assert np.allclose(xi.T @ f, rho * u, rtol=1e-12, atol=1e-13 * rho * p.cs)
        P = (xi[:, :, None] * xi[:, None, :] * f[:, None, None]).sum(0)
        tol = 1e-12 if order < 4 else 1e-5
        assert np.allclose(P, rho * (np.outer(u, u) + k.cs2 * np.eye(3)), rtol=tol, atol=tol * rho * k.cs2)


def _collide(p, f, nb=None):
    k = ol.consts(p)
    g = np.ascontiguousarray(f.copy())
    ps = np.zeros(19)
    nbp = None if nb is None else D(np.ascontiguousarray(nb))
    L.acpo_collide(ctypes.byref(k), ctypes.byref(p), D(g), D(ps), nbp)
    return g


@pytest.mark.parametrize("op,mrt", [("SRT", 0), ("RR", 0), ("HRR", 0), ("MRT-RM", 1), ("MRT-GS", 1)])
def test_collisions_conserve_and_fix_rest_state(op, mrt):
    p = _p(op=op, order=3, mrt_mode=mrt, tau=0.57)
    rest_nb = np.tile([p.rho0, 0.0, 0.0, 0.0], (6, 1))
    feq0 = ol.feq(p, 3, p.rho0, np.zeros(3))
    if op != "MRT-RM":  # the raw-moment equilibrium is a different polynomial
        assert np.allclose(_collide(p, feq0, rest_nb), feq0, rtol=1e-15, atol=0)
    k = ol.consts(p)
    xi = np.array(k.xi)
    rng = np.random.default_rng(7)
    w = np.array(k.w)
    for _ in range(10):
        f = ol.feq(p, 3, p.rho0 * 1.01, 0.03 * p.cs * rng.uniform(-1, 1, 3)) + p.rho0 * w * 0.002 * rng.uniform(-1, 1, 19)
        nb = np.column_stack([p.rho0 * (1 + 0.01 * rng.uniform(-1, 1, 6)), 0.02 * p.cs * rng.uniform(-1, 1, (6, 3))])
        g = _collide(p, f, nb)
        assert abs(g.sum() - f.sum()) < 1e-14 * f.sum()
        assert np.allclose(xi.T @ g, xi.T @ f, rtol=0, atol=1e-13 * f.sum() * k.c)


def test_hrr_with_sigma_one_is_rr():
    """a = sigma Pi1 + (1 - sigma)(...) reduces to Pi1 for sigma = 1: HRR == RR bit for bit."""
    rng = np.random.default_rng(3)
    for order in (2, 3, 4):
        pr = _p(op="RR", order=order, sigma=1.0, tau=0.6)
        ph = _p(op="HRR", order=order, sigma=1.0, tau=0.6)
        k = ol.consts(pr)
        w = np.array(k.w)
        f = ol.feq(pr, 3, pr.rho0, 0.03 * pr.cs * rng.uniform(-1, 1, 3)) + pr.rho0 * w * 0.002 * rng.uniform(-1, 1, 19)
        nb = np.column_stack([pr.rho0 * np.ones(6), 0.02 * pr.cs * rng.uniform(-1, 1, (6, 3))])
        assert np.array_equal(_collide(pr, f), _collide(ph, f, nb))


def test_strain_of_linear_field_is_exact():
    """Central differences are exact for u = G x: S = sym(G)."""
    p = _p()
    k = ol.consts(p)
    G = np.array([[0.3, -0.2, 0.1], [0.05, -0.4, 0.25], [0.15, 0.2, 0.1]]) * 10.0
    h = p.spacing
    nb = np.zeros((6, 4))
    for m in range(6):
        x = E[m] * h
        nb[m] = [p.rho0, *(G @ x)]
    S = np.empty(6)
    psi = np.empty(19)
    L.acpo_strain_psi(ctypes.byref(k), D(nb), D(S), D(psi))
    sym = 0.5 * (G + G.T)
    ref = [sym[0, 0], sym[1, 1], sym[2, 2], sym[0, 1], sym[0, 2], sym[1, 2]]
    assert np.allclose(S, ref, rtol=1e-12, atol=1e-12)


def test_rr_keeps_second_order_nonequilibrium():
    """If f - feq is a pure second-order Hermite term, RR (order 2) returns
    feq + (1 - omega)(f - feq)."""
    p = _p(op="RR", order=2, tau=0.7)
    k = ol.consts(p)
    xi = np.array(k.xi)
    w = np.array(k.w)
    A = np.array([[1.0, 0.3, -0.2], [0.3, -0.5, 0.1], [-0.2, 0.1, -0.5]]) * 1e-3 * k.cs2
    H = xi[:, :, None] * xi[:, None, :] - k.cs2 * np.eye(3)
    fneq = w * np.einsum("qab,ab->q", H, A) / (2 * k.cs2 ** 2)
    u = np.array([0.01, -0.02, 0.005]) * p.cs
    feq = ol.feq(p, 2, p.rho0, u)
    g = _collide(p, feq + fneq)
    assert np.allclose(g, feq + (1 - k.omega) * fneq, rtol=1e-10, atol=1e-12 * p.rho0)


def test_mrt_round_trips_and_equilibrium_limit():
    p = _p(op="MRT-RM", mrt_mode=1, tau=1.0)  # omega = 1: every moment relaxes to equilibrium
    k = ol.consts(p)
    kb = ctypes.byref(k)
    rng = np.random.default_rng(5)
    w = np.array(k.w)
    for _ in range(10):
        f = np.ascontiguousarray(ol.feq(p, 3, p.rho0, 0.03 * p.cs * rng.uniform(-1, 1, 3)) +
                                 p.rho0 * w * 0.002 * rng.uniform(-1, 1, 19))
        m = np.empty(19)
        back = np.empty(19)
        L.acpo_rm_moments(kb, D(f), D(m))
        L.acpo_rm_inverse(kb, D(m), D(back))
        assert np.allclose(back, f, rtol=1e-13, atol=0)
        L.acpo_gs_moments(kb, D(f), D(m))
        L.acpo_gs_inverse(kb, D(m), D(back))
        assert np.allclose(back, f, rtol=1e-13, atol=0)
        m = np.empty(19)
        L.acpo_rm_moments(kb, D(f), D(m))
        u = m[1:4] / m[0]
        eq = np.empty(19)
        L.acpo_rm_feq4(kb, m[0], D(np.ascontiguousarray(u)), D(eq))
        g = _collide(p, f)
        assert np.allclose(g, eq, rtol=1e-13, atol=0)


def _random_state(p, seed):
    rng = np.random.default_rng(seed)
    N = p.nx * p.ny * p.nz
    rho = p.rho0 * (1 + 0.01 * rng.uniform(-1, 1, N))
    u = 0.03 * p.cs * rng.uniform(-1, 1, (3, N))
    return ol.equilibrium_field(p, rho, u)


def test_periodic_streaming_is_a_permutation():
    """MRT as the reference runs it collides nothing: on a periodic grid n steps
    shift each population plane by n c_q exactly."""
    p = ol.make_params(10, 9, 8, op="MRT-RM", periodic=True)
    f0 = _random_state(p, 1)
    c = ol.OracleCase(p)
    c.set_state(f0)
    c.step(5)
    f = c.state().reshape(19, 8, 9, 10)
    g = f0.reshape(19, 8, 9, 10)
    for q in range(19):
        assert np.array_equal(f[q], np.roll(g[q], (5 * E[q][2], 5 * E[q][1], 5 * E[q][0]), axis=(0, 1, 2)))


@pytest.mark.parametrize("op", ["SRT", "HRR"])
def test_periodic_translation_and_mass(op):
    p = ol.make_params(12, 10, 8, op=op, periodic=True, tau=0.6)
    f0 = _random_state(p, 2).reshape(19, 8, 10, 12)
    a, b = ol.OracleCase(p), ol.OracleCase(p)
    a.set_state(f0)
    b.set_state(np.roll(f0, (1, -2, 3), axis=(1, 2, 3)))
    a.step(15)
    b.step(15)
    fa = a.state().reshape(19, 8, 10, 12)
    assert np.array_equal(np.roll(fa, (1, -2, 3), axis=(1, 2, 3)), b.state().reshape(19, 8, 10, 12))
    assert abs(fa.sum() - f0.sum()) < 1e-13 * f0.sum()


def test_lodi_and_periodic_agree_inside_the_light_cone():
    """Boundary influence travels at most one node per step: after n steps the
    LODI and periodic runs agree bit for bit on nodes more than n+2 from a face."""
    nx, ny, nz, n = 24, 22, 20, 4
    pl = ol.make_params(nx, ny, nz, op="SRT", tau=0.6)
    pp = ol.make_params(nx, ny, nz, op="SRT", tau=0.6, periodic=True)
    f0 = _random_state(pl, 4)
    a, b = ol.OracleCase(pl), ol.OracleCase(pp)
    a.set_state(f0)
    b.set_state(f0)
    a.step(n)
    b.step(n)
    m = n + 2
    fa = a.state().reshape(19, nz, ny, nx)[:, m:-m, m:-m, m:-m]
    fb = b.state().reshape(19, nz, ny, nx)[:, m:-m, m:-m, m:-m]
    assert np.array_equal(fa, fb)


def test_planted_mistake_is_caught():
    """A wrong LODI sign convention in a copy of the inputs (K1 of opposite sign)
    must change the boundary result: guards against a test that cannot fail."""
    pa = ol.make_params(9, 7, 3, op="SRT", k1=1.0)
    pb = ol.make_params(9, 7, 3, op="SRT", k1=-1.0)
    f0 = ol.pulse_ic(pa, 1e-2, 4.0, 3.0, 2.0)
    a, b = ol.OracleCase(pa), ol.OracleCase(pb)
    a.set_state(f0)
    b.set_state(f0)
    a.step(3)
    b.step(3)
    assert not np.array_equal(a.state(), b.state())
