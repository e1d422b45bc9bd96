"""acp_case_get_macros_noalloc against acp_case_get_macros, through the C ABI.

The variant stages the fields in the library's fixed device buffer instead of
a buffer allocated per call; its results must be the allocating entry point's
to the bit, on two calls in a row (the second on a smaller case, so that the
buffer holds the first call's larger field when the second runs), in the
float64 and float32-shifted tiers.  A case beyond the buffer is refused.
"""
import numpy as np
import pytest

import acpulse_vclbm_amd as acp
from acpulse_vclbm_amd import _capi

pytestmark = pytest.mark.gpu


def _both(c):
    N = c.nx * c.ny * c.nz
    L = _capi.lib()
    rho_a, u_a = np.empty(N), np.empty(3 * N)
    rho_b, u_b = np.full(N, np.nan), np.full(3 * N, np.nan)
    _capi.check(L.acp_case_get_macros(c.h, _capi.dptr(rho_a), _capi.dptr(u_a)))
    _capi.check(L.acp_case_get_macros_noalloc(c.h, _capi.dptr(rho_b), _capi.dptr(u_b)))
    return (rho_a, u_a), (rho_b, u_b)


def test_noalloc_macros_equal_the_allocating_entry_point(gpu):
    big = acp.make_desc(96, 80, 5, op="HRR", transverse=True, tau=0.53)
    small = acp.make_desc(40, 30, 3, op="SRT", tau=0.6, precision="F32-SHIFTED")
    print("case big float64 HRR 96x80x5, then small float32-shifted SRT 40x30x3", flush=True)
    with acp.Case(big) as a, acp.Case(small) as b:
        a.init_pulse(1e-3, 40.0, 33.0, 6.0)
        b.init_pulse(2e-3, 15.0, 12.0, 4.0)
        a.step(7)
        b.step(5)
        for c in (a, b, a):
            (ra, ua), (rb, ub) = _both(c)
            assert np.array_equal(ra, rb)
            assert np.array_equal(ua, ub)
            assert np.isfinite(rb).all() and (rb > 0).all()


def test_noalloc_macros_refuse_a_case_beyond_the_buffer(gpu):
    d = acp.make_desc(2048, 1024, 3, op="SRT", tau=0.6, precision="F16-SHIFTED")
    with acp.Case(d) as c:
        N = 2048 * 1024 * 3
        rho, u = np.empty(N), np.empty(3 * N)
        rc = _capi.lib().acp_case_get_macros_noalloc(c.h, _capi.dptr(rho), _capi.dptr(u))
        assert rc == _capi.ACP_E_ARG
        assert b"scratch buffer" in _capi.lib().acp_last_error()
