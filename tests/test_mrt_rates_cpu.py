"""Closed-form pins of the relaxation-rate tables of the two MRT operators.

The relaxed MRT collision is affine in the moments at fixed density and
momentum: m* = m - S (m - m_eq(rho, j)).  Perturbing a state f along the
population direction v_r that changes moment r alone (v_r = M^-1 e_r, M the
moment basis) leaves rho, j and m_eq unchanged when r is not conserved, so the
post-collision state moves by exactly (1 - s_r) eps v_r: moment r changes by
(1 - s_r) eps and every other moment stays.

The bases are built here from their defining polynomials in the lattice
velocity (not from the oracle's coefficient tables):
  Gram-Schmidt (calcAllMoments_GSbasis, Node.cpp:2934-3004): rho, e, eps, jx,
  qx, jy, qy, jz, qz, 3pxx, 3pixx, pww, piww, pxy, pyz, pxz, mx, my, mz;
  raw moments (calcAllMoments_RMbasis, Node.cpp:2693-2762).
The rates are those the reference assigns:
  MRT-GS (Node.cpp:2803-2827): 0 for {0, 3, 5, 7}, omega for
  {1, 9, 11, 13, 14, 15}, 1 otherwise;
  MRT-RM (Node.cpp:2576-2584): omega for 5..9, 1 otherwise.
A planted mistake (omega on moment 10 instead of 9, or moment 15 left at 1),
made in a copy of the oracle's collision assembled from its own moment,
equilibrium and inverse functions, must fail the same check.
"""
import ctypes

import numpy as np
import pytest

import oracle_lib as ol

L = ol.lib()
D = ol.dptr
E = np.array([[1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 0, 1], [1, 0, -1],
              [0, 1, 1], [0, 1, -1], [-1, 0, 1], [-1, 0, -1], [0, -1, 1], [0, -1, -1], [-1, 1, 0], [1, -1, 0],
              [1, 1, 0], [-1, -1, 0], [0, 0, 0]], dtype=float)

GS_RATE_OMEGA = {1, 9, 11, 13, 14, 15}
GS_CONSERVED = {0, 3, 5, 7}
RM_RATE_OMEGA = {5, 6, 7, 8, 9}
RM_CONSERVED = {0, 1, 2, 3}


def gs_basis():
    """Rows of the Gram-Schmidt basis in lattice units, from the moments' polynomials."""
    cx, cy, cz = E.T
    c2 = cx * cx + cy * cy + cz * cz
    g = 2 * c2 - 3
    rows = [np.ones(19), c2 - 1, 3 * c2 * c2 - 6 * c2 + 1,
            cx, (3 * c2 - 5) * cx, cy, (3 * c2 - 5) * cy, cz, (3 * c2 - 5) * cz,
            3 * cx * cx - c2, g * (3 * cx * cx - c2), cy * cy - cz * cz, g * (cy * cy - cz * cz),
            cx * cy, cy * cz, cx * cz,
            (cy * cy - cz * cz) * cx, (cz * cz - cx * cx) * cy, (cx * cx - cy * cy) * cz]
    return np.array(rows)


def rm_basis():
    """Rows of the raw-moment basis in lattice units, in the reference's order
    (Node.cpp:2701-2760): M000, M100, M010, M001, the trace M200+M020+M002,
    M200-M002, M020-M002, M110, M101, M011, M201, M102, M021, M012, M210,
    M120, M202, M220, M022."""
    cx, cy, cz = E.T
    return np.array([np.ones(19), cx, cy, cz, cx * cx + cy * cy + cz * cz, cx * cx - cz * cz, cy * cy - cz * cz,
                     cx * cy, cx * cz, cy * cz,
                     cx * cx * cz, cx * cz * cz, cy * cy * cz, cy * cz * cz, cx * cx * cy, cx * cy * cy,
                     cx * cx * cz * cz, cx * cx * cy * cy, cy * cy * cz * cz])


def _params(op):
    return ol.make_params(9, 8, 7, op=op, order=3, mrt_mode=1, tau=0.5003)


def _collide_oracle(p, f):
    k = ol.consts(p)
    g = np.ascontiguousarray(f.copy())
    ps = np.zeros(19)
    L.acpo_collide(ctypes.byref(k), ctypes.byref(p), D(g), D(ps), None)
    return g


def _collide_gs_with_rates(p, f, rates):
    """The oracle's Gram-Schmidt collision assembled from its own moment,
    equilibrium-moment and inverse functions, with a given rate table."""
    k = ol.consts(p)
    kb = ctypes.byref(k)
    m = np.empty(19)
    meq = np.empty(19)
    df = np.empty(19)
    L.acpo_gs_moments(kb, D(np.ascontiguousarray(f)), D(m))
    L.acpo_gs_eqmoments(m[0], m[3], m[5], m[7], D(meq))
    dm = np.array([0.0 if rates[r] == 0.0 else rates[r] * (meq[r] - m[r]) for r in range(19)])
    L.acpo_gs_inverse(kb, D(dm), D(df))
    return f + df


def _base_state(p, seed):
    k = ol.consts(p)
    w = np.array(k.w)
    rng = np.random.default_rng(seed)
    u = 0.03 * p.cs * rng.uniform(-1, 1, 3)
    return ol.feq(p, 3, p.rho0 * 1.004, u) + p.rho0 * w * 0.002 * rng.uniform(-1, 1, 19)


def _rate_response(collide, p, M, r, seed=0):
    """Moment change per unit perturbation along v_r after one collision, in
    the lattice-unit basis M: (1 - s_r) e_r for a non-conserved moment r."""
    v = np.linalg.solve(M, np.eye(19)[r])
    f = _base_state(p, seed)
    eps = 1e-4 * p.rho0 / np.abs(v).max()
    d = collide(p, f + eps * v) - collide(p, f)
    return (M @ d) / eps


def _check_table(collide, p, M, omega_set, conserved):
    om = ol.consts(p).omega
    for r in range(19):
        if r in conserved:
            continue
        s = om if r in omega_set else 1.0
        resp = _rate_response(collide, p, M, r)
        want = (1.0 - s) * np.eye(19)[r]
        if not np.allclose(resp, want, rtol=0, atol=1e-6):
            return False, r, resp
    return True, None, None


def test_gs_basis_is_the_reference_basis():
    """The polynomial basis reproduces the oracle's moments up to the c^n
    factor of each row (so the directions v_r are the reference's)."""
    p = _params("MRT-GS")
    k = ol.consts(p)
    M = gs_basis()
    pw = [0, 2, 4, 1, 3, 1, 3, 1, 3, 2, 4, 2, 4, 2, 2, 2, 3, 3, 3]
    f = _base_state(p, 3)
    m = np.empty(19)
    L.acpo_gs_moments(ctypes.byref(k), D(np.ascontiguousarray(f)), D(m))
    assert np.allclose(m, (M @ f) * k.c ** np.array(pw, dtype=float), rtol=1e-12, atol=1e-12 * p.rho0)


def test_gs_relaxation_rates_match_node_cpp():
    p = _params("MRT-GS")
    ok, r, resp = _check_table(_collide_oracle, p, gs_basis(), GS_RATE_OMEGA, GS_CONSERVED)
    assert ok, f"moment {r}: response {resp}"


def test_gs_conserved_moments_are_not_relaxed():
    p = _params("MRT-GS")
    M = gs_basis()
    for seed in range(5):
        f = _base_state(p, 10 + seed)
        g = _collide_oracle(p, f)
        for r in sorted(GS_CONSERVED):
            assert abs(M[r] @ (g - f)) < 1e-13 * p.rho0


def test_rm_relaxation_rates_match_node_cpp():
    p = _params("MRT-RM")
    ok, r, resp = _check_table(_collide_oracle, p, rm_basis(), RM_RATE_OMEGA, RM_CONSERVED)
    assert ok, f"moment {r}: response {resp}"


def test_assembled_gs_collision_is_the_oracle():
    """The assembled copy with the reference's table equals the oracle's
    collision bit for bit, so the planted mistakes below differ from the oracle
    in their rate table only."""
    p = _params("MRT-GS")
    om = ol.consts(p).omega
    rates = [0.0 if r in GS_CONSERVED else (om if r in GS_RATE_OMEGA else 1.0) for r in range(19)]
    for seed in range(4):
        f = _base_state(p, 20 + seed)
        assert np.array_equal(_collide_gs_with_rates(p, f, rates), _collide_oracle(p, f))


@pytest.mark.parametrize("mistake", ["omega_on_10_not_9", "moment_15_at_one", "omega_on_2"])
def test_planted_rate_mistake_is_caught(mistake):
    p = _params("MRT-GS")
    om = ol.consts(p).omega
    rates = [0.0 if r in GS_CONSERVED else (om if r in GS_RATE_OMEGA else 1.0) for r in range(19)]
    if mistake == "omega_on_10_not_9":
        rates[9], rates[10] = 1.0, om
    elif mistake == "moment_15_at_one":
        rates[15] = 1.0
    else:
        rates[2] = om

    def mutant(pp, f):
        return _collide_gs_with_rates(pp, f, rates)

    ok, _, _ = _check_table(mutant, p, gs_basis(), GS_RATE_OMEGA, GS_CONSERVED)
    assert not ok
