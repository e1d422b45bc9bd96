"""Serving loop, host side: request mix, footprint accounting (no GPU needed)."""
import ctypes

import numpy as np
import pytest

import acpulse_vclbm_amd as acp
from acpulse_vclbm_amd import _capi, serve


def test_request_mix_is_reproducible_and_in_range():
    a = serve.make_requests(256, seed=1, rate=2.0)
    b = serve.make_requests(256, seed=1, rate=2.0)
    assert a == b
    assert serve.make_requests(256, seed=2, rate=2.0) != a
    nx = np.array([r["nx"] for r in a])
    st = np.array([r["steps"] for r in a])
    amp = np.array([r["amp"] for r in a])
    hw = np.array([r["halfwidth"] for r in a])
    arr = np.array([r["arrival"] for r in a])
    assert nx.min() >= 128 and nx.max() <= 1024
    assert st.min() >= 1000 and st.max() <= 8000
    assert amp.min() >= 1e-4 and amp.max() <= 1e-2
    assert hw.min() >= 4 and hw.max() <= 20
    assert np.all(np.diff(arr) >= 0)
    # Poisson arrivals at rate 2/s: mean gap 0.5 s (256 samples: within 20 %)
    assert 0.4 < np.diff(arr).mean() < 0.6
    assert {r["op"] for r in a} == set(serve.OPS)
    assert {r["precision"] for r in a} == {"F64", "F32-SHIFTED"}
    f = np.mean([r["refined"] for r in a])
    assert 0.35 < f < 0.65


def test_rate_only_scales_arrivals():
    a = serve.make_requests(50, seed=1, rate=1.0)
    b = serve.make_requests(50, seed=1, rate=4.0)
    for x, y in zip(a, b):
        assert x["nx"] == y["nx"] and x["op"] == y["op"] and x["steps"] == y["steps"]
        assert y["arrival"] == pytest.approx(x["arrival"] / 4.0)


def _fp(d, nprobe=0, cap=0):
    return acp.footprint(d, nprobe, cap)


@pytest.mark.parametrize("n", [(3, 3, 3), (7, 5, 3), (40, 36, 3), (17, 9, 12)])
def test_footprint_lodi_term_is_the_schedule(n):
    nx, ny, nz = n
    d_lodi = acp.make_desc(nx, ny, nz, op="SRT", tau=0.6)
    d_per = acp.make_desc(nx, ny, nz, op="SRT", tau=0.6, boundary="PERIODIC")
    count = len(acp.lodi_schedule(nx, ny, nz)[0])
    assert count == nx * ny * nz - (nx - 2) * (ny - 2) * (nz - 2)
    diff = _fp(d_lodi) - _fp(d_per)
    assert diff == (count * 64 + 255) // 256 * 256


def _elems(N, precision):
    """Elements of one population array: blocks of 8192 nodes (19 planes each),
    the last one padded, for float64 and float32; whole planes for float16 and
    for grids of at most one block."""
    B = 8192
    if precision == "F16-SHIFTED" or N <= B:
        return 19 * N
    return -(-N // B) * 19 * B


@pytest.mark.parametrize("n", [(64, 48, 3), (16, 16, 16), (40, 40, 9)])
def test_footprint_tiers_and_probes(n):
    f64 = _fp(acp.make_desc(*n, op="SRT", tau=0.6, boundary="PERIODIC"))
    f32 = _fp(acp.make_desc(*n, op="SRT", tau=0.6, boundary="PERIODIC", precision="F32-SHIFTED"))
    f16 = _fp(acp.make_desc(*n, op="SRT", tau=0.6, boundary="PERIODIC", precision="F16-SHIFTED"))
    N = n[0] * n[1] * n[2]
    r = lambda b: (b + 255) // 256 * 256
    # float64: constants + 2 population arrays of float64; float32-shifted: both constant
    # blocks (one 256-byte unit each) + 2 arrays of float32; the staging and reduction
    # buffers are common to every tier
    assert f64 - f32 == 2 * r(_elems(N, "F64") * 8) - 2 * r(_elems(N, "F32-SHIFTED") * 4) - 256
    assert f32 - f16 == 2 * r(_elems(N, "F32-SHIFTED") * 4) - 2 * r(_elems(N, "F16-SHIFTED") * 2)
    base = acp.make_desc(*n, op="HRR", tau=0.6)
    assert _fp(base, 1, 100) - _fp(base) == 256 + r(800) + 256
    assert _fp(base, 3, 1000) - _fp(base) == r(24) + r(24000) + 256
    cub = acp.make_desc(*n, op="HRR", tau=0.6, cubic=True)
    assert _fp(cub) - _fp(base) == r(19 * N * 8)


def test_footprint_rejects_what_create_rejects():
    d = acp.make_desc(16, 16, 3, op="SRT", tau=0.6)
    d.max_level = 1
    with pytest.raises(_capi.AcpError) as e:
        acp.footprint(d)
    assert e.value.code == _capi.ACP_E_UNSUPPORTED
    d = acp.make_desc(16, 16, 3, op="SRT", tau=0.6, precision="F32-SHIFTED", cubic=True)
    with pytest.raises(_capi.AcpError):
        acp.footprint(d)


def test_refined_footprint_is_both_levels():
    """A refined case reserves its fine level (half spacing, 2(x1-x0)+1 x 2ny x
    2nz, open x faces) plus its coarse level with the probes."""
    d = acp.make_desc(64, 40, 3, op="RR", tau=0.56, boundary="PERIODIC")
    fd = acp.make_desc(65, 80, 6, op="RR", nu=d.nu, spacing=d.spacing / 2, boundary="OPEN-X")
    assert acp.refined_footprint(d, 32, 64, 1, 500) == _fp(fd) + _fp(d, 1, 500)
    for bad in ((0, 32), (10, 11), (1, 64), (40, 65)):
        with pytest.raises(_capi.AcpError):
            acp.refined_footprint(d, *bad)
    with pytest.raises(_capi.AcpError) as e:
        acp.refined_footprint(acp.make_desc(64, 40, 3, op="RR", tau=0.56), 32, 64)
    assert e.value.code == _capi.ACP_E_UNSUPPORTED
    # LODI x faces: a region reaching the high-x face gives a fine level whose
    # far face is a LODI face (one record per node of that face)
    dl = acp.make_desc(64, 40, 3, op="RR", tau=0.56, boundary="LODI-X")
    fl = acp.make_desc(63, 80, 6, op="RR", nu=d.nu, spacing=d.spacing / 2, boundary="OPEN-LODI-X")
    assert acp.refined_footprint(dl, 32, 63, 1, 500) == _fp(fl) + _fp(dl, 1, 500)


def test_refined_requests_in_the_mix():
    """Half of the requests are flagged for refinement; they become two-level
    cases on a coarse grid with LODI x faces, with the fine level over the
    upper half in x up to the high-x face and the pulse in the lower half."""
    reqs = serve.make_requests(256, seed=1)
    c = serve.mix_counts(reqs)
    assert c["refined_flag"] == c["two_level"]
    assert 0.35 < c["two_level"] / len(reqs) < 0.65
    for r in reqs:
        d = serve.request_desc(r)
        if serve.two_level(r):
            assert d.boundary == _capi.BOUNDARY["LODI-X"]
            x0, x1 = serve.fine_region(r)
            assert (x0, x1) == (r["nx"] // 2, r["nx"] - 1)
            assert serve.pulse_centre(r)[0] < x0 <= serve.probe_point(r)[0]
            assert serve.footprint(r) == acp.refined_footprint(d, x0, x1, 1, r["steps"])
        else:
            assert d.boundary == _capi.BOUNDARY["LODI"]
            assert serve.footprint(r) == _fp(d, 1, r["steps"])


def test_server_refuses_a_tag_in_flight():
    """The native server writes a finished case's probe series into the
    buffer registered under its tag: a second request with a tag still in
    flight is refused before anything reaches the server, and the first
    request's buffer stays registered."""
    srv = serve.Server.__new__(serve.Server)
    srv.h = None
    first = np.zeros(8)
    srv._probe_bufs = {5: first}
    req = serve.make_requests(1, seed=1)[0]
    req["tag"] = 5
    with pytest.raises(ValueError):
        srv.submit(req)
    assert srv._probe_bufs[5] is first
