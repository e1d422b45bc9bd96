"""The matrix-core form of the relaxed MRT-RM collision (ACP_BULK_MFMA).

f* = f + C (feq - f) with the case's collision matrix C = M^-1 S M on
v_mfma_f64_16x16x4f64.  In exact arithmetic this is the raw-moment collision
of the default form (collide_mrt_rm64, the oracle's MRT-RM with the
relaxation applied); the matrix product rounds in its own order, so the
comparison is to a tolerance: 1e-13 of the largest population after 6 steps
(a step's rounding is a few 1e-16 relative), mass conserved to 1e-13.  A
dropped or misplaced term of C (a wrong rate, a transposed operand, a
mis-laid MFMA lane) moves populations by the non-equilibrium size, about
1e-3 of f here, ten orders above the bound.
"""
import numpy as np
import pytest

import acpulse_vclbm_amd as acp
import oracle_lib as ol
from acpulse_vclbm_amd import _capi

pytestmark = pytest.mark.gpu

W = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])


def _random_f(p, seed):
    rng = np.random.default_rng(seed)
    N = p.nx * p.ny * p.nz
    rho = p.rho0 * (1.0 + 0.01 * rng.uniform(-1, 1, N))
    u = 0.05 * p.cs * rng.uniform(-1, 1, (3, N))
    f = ol.equilibrium_field(p, rho, u)
    return np.ascontiguousarray(f + p.rho0 * W[:, None] * 0.003 * rng.uniform(-1, 1, (19, N)))


@pytest.mark.parametrize("n,tau", [((256, 12, 10), 0.56), ((512, 6, 4), 0.5003)])
def test_mfma_form_follows_the_oracle(gpu, n, tau):
    nx, ny, nz = n
    p = ol.make_params(nx, ny, nz, op="MRT-RM", order=3, periodic=1, mrt_mode=1, tau=tau)
    f0 = _random_f(p, 81)
    o = ol.OracleCase(p)
    o.set_state(f0)
    o.step(6)
    ref = o.state()
    d = acp.make_desc(nx, ny, nz, op="MRT-RM", boundary="PERIODIC", mrt_mode="RELAXED", tau=tau)
    with acp.Case(d) as c, acp.Case(d) as a:
        c.set_bulk_variant("mfma")
        for x in (c, a):
            x.set_populations(f0)
            x.step(6)
        f, fa = c.populations(), a.populations()
        m0, m = a.mass(), c.mass()
    assert np.array_equal(fa, ref)  # the default form is the oracle's, bit for bit
    err = np.max(np.abs(f - ref)) / np.max(np.abs(ref))
    assert err < 1e-13, err
    neq = np.max(np.abs(ref - f0)) / np.max(np.abs(ref))
    assert neq > 1e-4  # the steps moved the state far beyond the tolerance
    assert abs(m - m0) <= 1e-13 * m0


def test_mfma_form_refuses_other_cases(gpu):
    for spec in (dict(boundary="LODI"), dict(nx=200), dict(op="SRT"), dict(precision="F32-SHIFTED"),
                 dict(mrt_mode="REFERENCE")):
        kw = dict(nx=256, ny=4, nz=4, op="MRT-RM", boundary="PERIODIC", mrt_mode="RELAXED")
        kw.update(spec)
        nx, ny, nz = kw.pop("nx"), kw.pop("ny"), kw.pop("nz")
        with acp.Case(acp.make_desc(nx, ny, nz, tau=0.6, **kw)) as c:
            assert _capi.lib().acp_case_set_bulk_variant(c.h, 4) == _capi.ACP_E_UNSUPPORTED, spec
