"""Scaling-Style of the standard coupling in the oracle (CPU).

SRTSTYLE rescales the non-equilibrium part of a population set handed between
the levels by one factor, a = tau_f / (2 tau_c) (1/a fine to coarse).
MRTSTYLE rescales it moment by moment in the Gram-Schmidt basis of the MRT-GS
operator, each moment by its own relaxation-time ratio: a for the moments the
operator relaxes with omega (1, 9, 11, 13, 14, 15), 1/2 for those it relaxes
at rate 1 on both levels.  The reference's arithmetic for either is absent
(GridCoupling.cpp), so these tests pin the restatement by what the definition
fixes, with the Gram-Schmidt basis built here from its polynomials
(independent of the oracle's tables):
  - each moment's factor, measured by scaling a pure moment direction (a
    rate table with two moments swapped, or a factor on a conserved moment,
    fails);
  - the result equals M^-1 diag(factors) M n for a generic non-equilibrium
    part n (with the oracle's own moment functions it would be circular);
  - coarse to fine and back is the identity;
  - SRTSTYLE is the single factor, bit for bit;
  - a uniform state stays uniform through a refined run with MRTSTYLE.
"""
import numpy as np
import pytest

import oracle_lib as ol
from test_mrt_rates_cpu import GS_CONSERVED, GS_RATE_OMEGA, gs_basis


def _refined(tau=0.56, op="SRT"):
    p = ol.make_params(24, 10, 3, op=op, periodic=True, tau=tau, mrt_mode=1, nthreads=4)
    return ol.OracleRefined(p, 8, 16), p


def _factors(a):
    c2f = np.array([0.0 if r in GS_CONSERVED else (a if r in GS_RATE_OMEGA else 0.5) for r in range(19)])
    f2c = np.array([0.0 if r in GS_CONSERVED else (1.0 / a if r in GS_RATE_OMEGA else 2.0) for r in range(19)])
    return c2f, f2c


def _neq(p, seed):
    """A non-equilibrium part: random populations minus their own equilibrium."""
    k = ol.consts(p)
    rng = np.random.default_rng(seed)
    u = 0.03 * p.cs * rng.uniform(-1, 1, 3)
    f = ol.feq(p, 3, p.rho0 * 1.002, u) + p.rho0 * np.array(k.w) * 0.002 * rng.uniform(-1, 1, 19)
    rho = f.sum()
    uu = (np.array(k.xi).T @ f) / rho
    return f - ol.feq(p, 3, rho, uu)


@pytest.mark.parametrize("f2c", [False, True])
def test_mrtstyle_factor_of_each_moment(f2c):
    r, p = _refined()
    a = r.scale()
    assert abs(a - 0.5) > 0.03  # the two factor groups differ
    r.set_scaling("MRTSTYLE")
    M = gs_basis()
    want = _factors(a)[1 if f2c else 0]
    for m in range(19):
        if m in GS_CONSERVED:
            continue
        v = np.linalg.solve(M, np.eye(19)[m]) * 1e-3 * p.rho0
        got = M @ r.scale_neq(v, f2c)
        ref = want[m] * (M @ v)
        assert np.allclose(got, ref, rtol=0, atol=1e-12 * np.abs(M @ v).max()), (m, got[m] / (M @ v)[m])


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_mrtstyle_is_the_moment_space_scaling(seed):
    r, p = _refined(tau=0.56)
    r.set_scaling("MRTSTYLE")
    a = r.scale()
    M = gs_basis()
    n = _neq(p, seed)
    for f2c, fac in zip((False, True), _factors(a)):
        want = np.linalg.solve(M, fac * (M @ n))
        got = r.scale_neq(n, f2c)
        assert np.allclose(got, want, rtol=0, atol=1e-13 * np.abs(n).max() / min(a, 0.5))
        # it differs from the single SRT factor: the rate-1 moments are scaled by 1/2, not a
        assert np.max(np.abs(got - (a if not f2c else 1 / a) * n)) > 1e-3 * np.abs(n).max()


def test_mrtstyle_round_trip_and_srtstyle_bits():
    r, p = _refined()
    a = r.scale()
    n = _neq(p, 4)
    r.set_scaling("MRTSTYLE")
    back = r.scale_neq(r.scale_neq(n, False), True)
    # to the rounding of the populations n came from (its conserved moments,
    # a few 1e-17 of rho0, are dropped on the way)
    assert np.allclose(back, n, rtol=0, atol=1e-15 * p.rho0)
    r.set_scaling("SRTSTYLE")
    assert np.array_equal(r.scale_neq(n, False), a * n)
    assert np.array_equal(r.scale_neq(n, True), (1.0 / a) * n)


@pytest.mark.parametrize("op,filt", [("SRT", "NOFI"), ("RR", "LAG"), ("MRT-GS", "TOU")])
def test_mrtstyle_uniform_state_stays_uniform(op, filt):
    r, p = _refined(op=op)
    r.set_scaling("MRTSTYLE")
    r.set_filter(filt)
    u = np.array([0.02 * p.cs, -0.01 * p.cs, 0.0]) if op != "MRT-GS" else np.zeros(3)
    for lv in r.levels:
        N = lv.N
        lv.set_state(ol.equilibrium_field(lv.p, np.full(N, p.rho0), np.repeat(u[:, None], N, axis=1)))
    r.step(6)
    ref = r.levels[0].state()[:, 0]
    for lv in r.levels:
        assert np.max(np.abs(lv.state() - ref[:, None])) <= 1e-14 * np.max(np.abs(ref))
