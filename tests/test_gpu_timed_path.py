"""The kernel forms the large measured grids take, against the oracle.

The bulk kernels come in addressing and streaming forms chosen by array size:
32-bit offsets from a uniform base while every per-node offset stays below
4 GiB, 64-bit offsets beyond (the 512^3 float64 and float32 grids), and
non-temporal loads and stores once the two population arrays exceed the
256 MB last-level cache (every grid above about 530^2 x 3 nodes in float64).
Here every form runs on small grids, forced through acp_case_set_bulk_variant,
bit for bit against the oracle (float64) or against the default form (shifted
tiers, which the other suites hold to the float64 oracle); the 512^3 grid of
the throughput measurement runs as a tiling of a 128^3 period and must stay
the period's own run tiled, bit for bit; and the 1024 x 1024 x 3 serving grid
with LODI faces runs against the oracle.
"""
import numpy as np
import pytest

import acpulse_vclbm_amd as acp
import oracle_lib as ol

pytestmark = pytest.mark.gpu

W = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])


def _random_f(p, seed, amp_u=0.05):
    rng = np.random.default_rng(seed)
    N = p.nx * p.ny * p.nz
    rho = p.rho0 * (1.0 + 0.01 * rng.uniform(-1, 1, N))
    u = amp_u * p.cs * rng.uniform(-1, 1, (3, N))
    f = ol.equilibrium_field(p, rho, u)
    return np.ascontiguousarray(f + p.rho0 * W[:, None] * 0.003 * rng.uniform(-1, 1, (19, N)))


BND_MODE = {"LODI": 0, "PERIODIC": 1, "OPEN-X": 2, "LODI-X": 3, "OPEN-LODI-X": 4}

# (op, order, cubic, transverse, boundary, grid): 64^3 grids take the LDS row
# stores (rows of 64 nodes, at least 4096 rows), 41 x 29 x 17 the direct stores
VARIANT_CASES = [
    ("MRT-RM", 3, False, False, "PERIODIC", (64, 64, 64)),
    ("SRT", 3, False, False, "PERIODIC", (64, 64, 64)),
    ("MRT-GS", 3, False, False, "LODI", (64, 64, 64)),
    ("RR", 4, True, False, "LODI", (64, 64, 64)),
    ("HRR", 3, False, True, "LODI", (41, 29, 17)),
    ("HRR", 4, True, True, "LODI-X", (41, 29, 17)),
    ("RR", 3, False, False, "OPEN-LODI-X", (41, 29, 17)),
    ("SRT", 2, False, False, "LODI", (41, 29, 17)),
    # periodic planes of eight blocks: the last row's wrapped pushes reach
    # nearly two planes away, which the block-relative form's window must hold
    ("MRT-RM", 3, False, False, "PERIODIC", (256, 256, 6)),
    ("HRR", 3, False, False, "PERIODIC", (256, 256, 6)),
]


@pytest.mark.parametrize("variant", ["wide_nt", "narrow_nt", "narrow_plain", "relative_nt"])
@pytest.mark.parametrize("op,order,cubic,tr,bnd,n", VARIANT_CASES,
                         ids=[f"{a}-o{b}-{e}-{f[0]}" + ("-cubic" if c else "") for a, b, c, d, e, f in VARIANT_CASES])
def test_float64_bulk_forms_match_oracle_bitexact(gpu, variant, op, order, cubic, tr, bnd, n):
    nx, ny, nz = n
    relaxed = op.startswith("MRT")
    p = ol.make_params(nx, ny, nz, op=op, order=order, cubic=cubic, transverse=tr, periodic=BND_MODE[bnd],
                       mrt_mode=int(relaxed), tau=0.56)
    f0 = _random_f(p, 71)
    print("case", variant, op, order, bnd, n, flush=True)
    o = ol.OracleCase(p)
    o.set_state(f0)
    probes = [(1, 2, 3), (nx - 2, ny // 2, nz // 2)]
    po = o.step(4, probes)
    d = acp.make_desc(nx, ny, nz, op=op, order=order, cubic=cubic, transverse=tr, boundary=bnd,
                      mrt_mode="RELAXED" if relaxed else "REFERENCE", tau=0.56)
    with acp.Case(d) as c:
        c.set_bulk_variant(variant)
        c.set_populations(f0)
        c.set_probes(probes, 4)
        c.step(4)
        f = c.populations()
        assert np.array_equal(f, o.state()), f"max |df| = {np.abs(f - o.state()).max():.3e}"
        assert np.array_equal(c.read_probes(), po)


SHIFT_CASES = [
    ("MRT-RM", 3, "PERIODIC", (128, 24, 8)),   # two nodes per thread, several rows per workgroup
    ("HRR", 3, "PERIODIC", (128, 24, 8)),      # two nodes per thread, one row per workgroup
    ("SRT", 3, "LODI", (64, 64, 64)),          # LDS row stores with boundary rows
    ("RR", 4, "LODI-X", (41, 29, 17)),         # direct stores
    ("MRT-GS", 3, "PERIODIC", (64, 64, 64)),
]


@pytest.mark.parametrize("prec", ["F32-SHIFTED", "F16-SHIFTED"])
@pytest.mark.parametrize("op,order,bnd,n", SHIFT_CASES, ids=[f"{a}-{c}-{d[0]}" for a, b, c, d in SHIFT_CASES])
def test_shifted_bulk_forms_are_identical(gpu, prec, op, order, bnd, n):
    """Every addressing and streaming form gives the default form's bits, and
    the default form follows the float64 oracle (2e-4 / 2e-2 of the largest
    density deviation for float32 / float16)."""
    nx, ny, nz = n
    relaxed = op.startswith("MRT")
    p = ol.make_params(nx, ny, nz, op=op, order=order, periodic=BND_MODE[bnd], mrt_mode=int(relaxed), tau=0.56)
    f0 = _random_f(p, 72, amp_u=0.02)
    print("case", prec, op, bnd, n, flush=True)
    d = acp.make_desc(nx, ny, nz, op=op, order=order, boundary=bnd, precision=prec, f16_scale=1000.0,
                      mrt_mode="RELAXED" if relaxed else "REFERENCE", tau=0.56)
    out = {}
    for variant in ("auto", "wide_nt", "narrow_nt", "narrow_plain"):
        with acp.Case(d) as c:
            c.set_bulk_variant(variant)
            c.set_populations(f0)
            c.step(5)
            out[variant] = (c.checksum(), c.populations())
    for variant in ("wide_nt", "narrow_nt", "narrow_plain"):
        assert out[variant][0] == out["auto"][0], variant
        assert np.array_equal(out[variant][1], out["auto"][1]), variant
    o = ol.OracleCase(p)
    o.set_state(f0)
    o.step(5)
    r64 = o.state().sum(axis=0)
    rho = out["auto"][1].sum(axis=0)
    tol = 2e-4 if prec == "F32-SHIFTED" else 2e-2
    assert np.max(np.abs(rho - r64)) <= tol * np.max(np.abs(r64 - p.rho0))


def test_variant_argument_checks(gpu):
    from acpulse_vclbm_amd import _capi
    with acp.Case(acp.make_desc(10, 10, 3, tau=0.6)) as c:
        assert _capi.lib().acp_case_set_bulk_variant(c.h, 7) == _capi.ACP_E_ARG
        assert _capi.lib().acp_case_set_bulk_variant(c.h, -1) == _capi.ACP_E_ARG
        c.set_bulk_variant("wide_nt")
        c.set_bulk_variant("auto")


def _tile_compare(big, small, reps):
    """big [19][N_big] equals small [19][N_small] tiled reps times per axis (z, y, x)."""
    nz, ny, nx = small.shape[1:]
    for q in range(19):
        b = big[q].reshape(reps[0], nz, reps[1], ny, reps[2], nx)
        if not (b == small[q][None, :, None, :, None, :]).all():
            return q
    return None


@pytest.mark.parametrize("prec", ["F64", "F32-SHIFTED"])
def test_512_cube_is_its_128_period_tiled(gpu, prec):
    """The throughput grid itself: 512^3 periodic raw-moment MRT (relaxed), the
    64-bit-offset non-temporal kernel of the measurement.  Its initial state
    is a 128^3 random state tiled 4 x 4 x 4; after 3 steps it must be the
    128^3 run tiled, to the bit.  The 128^3 float64 run must equal the oracle
    to the bit, so the 512^3 float64 state equals the oracle's tiled state."""
    n, big = 128, 512
    p = ol.make_params(n, n, n, op="MRT-RM", order=3, periodic=True, mrt_mode=1, tau=0.503)
    f0 = _random_f(p, 5, amp_u=0.01)
    d_small = acp.make_desc(n, n, n, op="MRT-RM", boundary="PERIODIC", mrt_mode="RELAXED", tau=0.503,
                            precision=prec)
    print("case", prec, "128^3", flush=True)
    with acp.Case(d_small) as c:
        c.set_populations(f0)
        c.step(3)
        small = c.populations()
    if prec == "F64":
        o = ol.OracleCase(p)
        o.set_state(f0)
        o.step(3)
        assert np.array_equal(small, o.state())
        del o
    small = small.reshape(19, n, n, n)
    print("case", prec, "512^3 tiled", flush=True)
    tiled = np.tile(f0.reshape(19, n, n, n), (1, 4, 4, 4)).reshape(19, -1)
    del f0
    d_big = acp.make_desc(big, big, big, op="MRT-RM", boundary="PERIODIC", mrt_mode="RELAXED", tau=0.503,
                          precision=prec)
    with acp.Case(d_big) as c:
        c.set_populations(tiled)
        del tiled
        c.step(3)
        out = c.populations()
    q = _tile_compare(out, small, (4, 4, 4))
    assert q is None, f"population {q} differs from the tiled period"


Q2D = [("SRT", 3, False, "LODI"), ("HRR", 3, True, "LODI"), ("RR", 4, False, "LODI"), ("MRT-GS", 3, False, "LODI"),
       ("HRR", 3, True, "LODI-X")]


@pytest.mark.parametrize("op,order,tr,bnd", Q2D, ids=[f"{a}-o{b}-{d}" for a, b, c, d in Q2D])
def test_serving_grid_1024_matches_oracle(gpu, op, order, tr, bnd):
    """The largest quasi-2D serving grid, 1024 x 1024 x 3 (non-temporal bulk
    kernel, two workgroups per row, LODI phases of a million nodes): 8 steps of
    a pulse, float64 bit for bit against the oracle with the probe series, and
    the float32-shifted tier within 2e-4 of the largest density deviation."""
    nx, ny, nz = 1024, 1024, 3
    relaxed = op.startswith("MRT")
    p = ol.make_params(nx, ny, nz, op=op, order=order, transverse=tr, periodic=BND_MODE[bnd], mrt_mode=int(relaxed),
                       tau=0.5003, nthreads=16)
    f0 = ol.pulse_ic(p, 1e-3, 700.0, 400.0, 12.0)
    f0 += p.rho0 * W[:, None] * 1e-4 * np.random.default_rng(3).uniform(-1, 1, f0.shape)
    probes = [(0, 400, 1), (700, 400, 1), (1023, 1023, 2), (1, 0, 0)]
    print("case", op, order, bnd, flush=True)
    o = ol.OracleCase(p)
    o.set_state(f0)
    po = o.step(8, probes)
    fo = o.state()
    del o
    kw = dict(op=op, order=order, transverse=tr, boundary=bnd, mrt_mode="RELAXED" if relaxed else "REFERENCE",
              tau=0.5003)
    with acp.Case(acp.make_desc(nx, ny, nz, **kw)) as c:
        c.set_populations(f0)
        c.set_probes(probes, 8)
        c.step(8)
        f = c.populations()
        assert np.array_equal(f, fo), f"max |df| = {np.abs(f - fo).max():.3e}"
        assert np.array_equal(c.read_probes(), po)
    with acp.Case(acp.make_desc(nx, ny, nz, precision="F32-SHIFTED", **kw)) as c:
        c.set_populations(f0)
        c.step(8)
        rho, _ = c.macros()
    r64 = fo.sum(axis=0)
    assert np.max(np.abs(rho - r64)) <= 2e-4 * np.max(np.abs(r64 - p.rho0))
