"""HIP path (libacpulse.so via the C ABI) against the reference fixtures and the oracle.

Float64 tier: bit-identical results are required (np.array_equal), the
reference and the oracle agree bit for bit and the kernels keep the reference's
floating-point association with FMA contraction disabled.
Shifted tiers: tolerances stated per test, against the float64 oracle.
"""
import glob
import hashlib
import json
import math
import os

import numpy as np
import pytest

import acpulse_vclbm_amd as acp
import oracle_lib as ol
from make_golden import CS, DX, RHO0, params_of

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu


def desc_of(spec, **over):
    kw = dict(op=spec["op"], order=spec["order"], cubic=spec.get("cubic", False),
              transverse=spec.get("transverse", False), rho0=RHO0, cs=CS, spacing=spec.get("spacing", DX),
              mach=spec.get("mach", 0.0), k1=spec.get("k1", 1.0), sigma=spec.get("sigma", 0.98))
    if spec.get("nu") is not None:
        kw["nu"] = spec["nu"]
    else:
        kw["tau"] = spec.get("tau", 0.6)
    kw.update(over)
    return acp.make_desc(spec["nx"], spec["ny"], spec["nz"], **kw)


STEP_CASES = sorted(glob.glob(os.path.join(GOLD, "step_*.npz")))


@pytest.mark.parametrize("path", STEP_CASES, ids=[os.path.basename(p)[5:-4] for p in STEP_CASES])
def test_reference_step_fixtures_bitexact(gpu, path):
    z = np.load(path)
    spec = json.loads(str(z["spec"]))
    print("case", os.path.basename(path), flush=True)
    with acp.Case(desc_of(spec)) as c:
        c.set_populations(z["f0"])
        c.set_probes([tuple(p) for p in z["probes"]], int(z["nsteps"]))
        c.step(int(z["nsteps"]))
        f = c.populations()
        pr = c.read_probes()
    assert np.array_equal(f, z["f"]), f"max |df| = {np.abs(f - z['f']).max():.3e}"
    assert np.array_equal(pr, z["probe"])


def _big(name):
    return np.load(os.path.join(GOLD, f"big_{name}.npz"))


def test_config1r_bgk_500_steps(gpu):
    """200x200x3 BGK, tau = 0.5003, pulse 1e-3 at (100,100), LODI faces: probe series
    and the full final state equal the reference's (SHA-256 of the population array)."""
    z = _big("cfg1r_srt")
    spec = json.loads(str(z["spec"]))
    p = params_of(spec)
    with acp.Case(desc_of(spec)) as c:
        c.init_pulse(1e-3, 100.0, 100.0, 10.0)
        c.set_probes([tuple(x) for x in z["probes"]], 500)
        c.step(500)
        f = c.populations()
        pr = c.read_probes()
    assert np.array_equal(pr, z["probe"])
    assert hashlib.sha256(f.tobytes()).hexdigest() == str(z["sha256"])
    rho = f.sum(axis=0).reshape(p.nz, p.ny, p.nx)
    assert np.array_equal(rho[1], z["rho_plane"])


def test_config1r_hrr_transverse_500_steps(gpu):
    z = _big("cfg1r_hrr_tr")
    spec = json.loads(str(z["spec"]))
    with acp.Case(desc_of(spec)) as c:
        c.init_pulse(1e-3, 100.0, 100.0, 10.0)
        c.set_probes([tuple(x) for x in z["probes"]], 500)
        c.step(500)
        f = c.populations()
        pr = c.read_probes()
    assert np.array_equal(pr, z["probe"])
    assert hashlib.sha256(f.tobytes()).hexdigest() == str(z["sha256"])


def test_reference_initialization_pulse(gpu):
    """Lattice::initialization's hard-coded pulse computed by the library, 60 BGK steps."""
    z = _big("refpulse_srt")
    spec = json.loads(str(z["spec"]))
    with acp.Case(desc_of(spec)) as c:
        c.init_reference()
        c.set_probes([tuple(x) for x in z["probes"]], 60)
        c.step(60)
        f = c.populations()
        pr = c.read_probes()
    assert np.array_equal(pr, z["probe"])
    assert hashlib.sha256(f.tobytes()).hexdigest() == str(z["sha256"])


def test_config3r_mrtrm_48cubed(gpu):
    z = _big("cfg3r_mrtrm")
    spec = json.loads(str(z["spec"]))
    p = params_of(spec)
    N = 48 ** 3
    rng = np.random.default_rng(0)
    rho = RHO0 * (1.0 + 1e-3 * rng.standard_normal(N))
    u = 1e-3 * CS * math.sqrt(3.0) * rng.standard_normal((3, N))
    ic = ol.equilibrium_field(p, rho, u)
    with acp.Case(desc_of(spec)) as c:
        c.set_populations(ic)
        c.set_probes([tuple(x) for x in z["probes"]], 50)
        c.step(50)
        f = c.populations()
        pr = c.read_probes()
    assert np.array_equal(pr, z["probe"])
    assert hashlib.sha256(f.tobytes()).hexdigest() == str(z["sha256"])


def _random_f(p, seed, amp_u=0.05):
    rng = np.random.default_rng(seed)
    N = p.nx * p.ny * p.nz
    rho = p.rho0 * (1.0 + 0.01 * rng.uniform(-1, 1, N))
    u = amp_u * p.cs * rng.uniform(-1, 1, (3, N))
    f = ol.equilibrium_field(p, rho, u)
    w = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])
    return np.ascontiguousarray(f + p.rho0 * w[:, None] * 0.003 * rng.uniform(-1, 1, (19, N)))


NEW_MODES = [
    ("SRT", 3, "PERIODIC", "REFERENCE"), ("RR", 4, "PERIODIC", "REFERENCE"), ("HRR", 3, "PERIODIC", "REFERENCE"),
    ("MRT-RM", 3, "PERIODIC", "RELAXED"), ("MRT-GS", 3, "PERIODIC", "RELAXED"),
    ("MRT-RM", 3, "LODI", "RELAXED"), ("MRT-GS", 2, "LODI", "RELAXED"), ("HRR", 2, "PERIODIC", "REFERENCE"),
]


@pytest.mark.parametrize("op,order,bnd,mrt", NEW_MODES, ids=[f"{a}-o{b}-{c}-{d}" for a, b, c, d in NEW_MODES])
def test_new_modes_match_oracle_bitexact(gpu, op, order, bnd, mrt):
    """Periodic domains and relaxed MRT have no reference output; the oracle
    restates them with the same arithmetic, so the GPU must match it exactly."""
    nx, ny, nz = 16, 12, 10
    periodic = bnd == "PERIODIC"
    relaxed = mrt == "RELAXED"
    p = ol.make_params(nx, ny, nz, op=op, order=order, periodic=periodic, mrt_mode=int(relaxed), tau=0.56,
                       transverse=True)
    f0 = _random_f(p, 11)
    with acp.Case(acp.make_desc(nx, ny, nz, op=op, order=order, boundary=bnd, mrt_mode=mrt, tau=0.56,
                                transverse=True)) as c:
        c.set_populations(f0)
        c.step(20)
        f = c.populations()
    o = ol.OracleCase(p)
    o.set_state(f0)
    o.step(20)
    assert np.array_equal(f, o.state()), f"max |df| = {np.abs(f - o.state()).max():.3e}"


ROW_LDS = [("SRT", 3, 64, 6, 5), ("RR", 4, 64, 6, 5), ("HRR", 3, 128, 4, 4), ("MRT-RM", 3, 64, 7, 3),
           ("MRT-GS", 3, 128, 3, 4), ("HRR-cubic", 3, 64, 5, 4)]


@pytest.mark.parametrize("op,order,nx,ny,nz", ROW_LDS, ids=[f"{a}-nx{c}" for a, b, c, d, e in ROW_LDS])
def test_periodic_full_row_workgroups_match_oracle_bitexact(gpu, op, order, nx, ny, nz):
    """Periodic grids whose rows fit one workgroup (nx a multiple of 64 up to
    512) take the path that realigns x-shifted links through LDS before the
    store; the result must be the same bits as the oracle's push."""
    cubic = op.endswith("-cubic")
    op = op.replace("-cubic", "")
    p = ol.make_params(nx, ny, nz, op=op, order=order, periodic=True, mrt_mode=1, tau=0.56, cubic=cubic)
    f0 = _random_f(p, 5)
    with acp.Case(acp.make_desc(nx, ny, nz, op=op, order=order, boundary="PERIODIC", mrt_mode="RELAXED", tau=0.56,
                                cubic=cubic)) as c:
        c.set_populations(f0)
        c.step(9)
        f = c.populations()
    o = ol.OracleCase(p)
    o.set_state(f0)
    o.step(9)
    assert np.array_equal(f, o.state()), f"max |df| = {np.abs(f - o.state()).max():.3e}"


@pytest.mark.parametrize("prec,tol", [("F32-SHIFTED", 2e-4), ("F16-SHIFTED", 2e-2)])
def test_periodic_full_row_workgroups_shifted_tiers(gpu, prec, tol):
    """The same LDS path in the shifted tiers, against the float64 oracle
    (pulse of amplitude 1e-3, 40 steps; tolerance relative to the pulse)."""
    nx, ny, nz = 64, 48, 3
    p = ol.make_params(nx, ny, nz, op="SRT", periodic=True, tau=0.56)
    f0 = ol.pulse_ic(p, 1e-3, 20.0, 24.0, 5.0)
    o = ol.OracleCase(p)
    o.set_state(f0)
    o.step(40)
    rho64 = o.state().sum(axis=0)
    with acp.Case(acp.make_desc(nx, ny, nz, op="SRT", boundary="PERIODIC", tau=0.56, precision=prec)) as c:
        c.init_pulse(1e-3, 20.0, 24.0, 5.0)
        c.step(40)
        rho, _ = c.macros()
    dev = rho64 - p.rho0
    assert np.max(np.abs(rho - rho64)) <= tol * np.max(np.abs(dev))


BLOCK_GRIDS = [("HRR", 3, "LODI", (41, 29, 17), True), ("RR", 4, "PERIODIC", (64, 40, 9), False),
               ("SRT", 2, "LODI", (23, 31, 13), False), ("MRT-RM", 3, "PERIODIC", (128, 20, 7), False)]


@pytest.mark.parametrize("op,order,bnd,n,tr", BLOCK_GRIDS, ids=[f"{a}-{c}-{d[0]}x{d[1]}x{d[2]}" for a, b, c, d, e
                                                                in BLOCK_GRIDS])
def test_block_layout_grids_match_oracle_bitexact(gpu, op, order, bnd, n, tr):
    """Grids of two to three layout blocks of 8192 nodes with a padded last
    block (float64 and float32 store populations in blocks): the state after
    upload, after 7 steps and the probe series equal the oracle's bit for bit,
    and a float32-shifted run follows it (2e-4 of the largest density
    deviation)."""
    nx, ny, nz = n
    periodic = bnd == "PERIODIC"
    assert 8192 < nx * ny * nz < 3 * 8192 and (nx * ny * nz) % 8192
    mrt = "RELAXED" if op.startswith("MRT") else "REFERENCE"
    p = ol.make_params(nx, ny, nz, op=op, order=order, periodic=periodic, mrt_mode=int(op.startswith("MRT")),
                       tau=0.56, transverse=tr)
    f0 = _random_f(p, 21)
    probes = [(1, 2, 3), (nx - 2, ny // 2, nz // 2), (nx // 2, ny - 1, nz - 1)]
    o = ol.OracleCase(p)
    o.set_state(f0)
    po = o.step(7, probes)
    d = acp.make_desc(nx, ny, nz, op=op, order=order, boundary=bnd, mrt_mode=mrt, tau=0.56, transverse=tr)
    with acp.Case(d) as c:
        c.set_populations(f0)
        assert np.array_equal(c.populations(), f0)
        c.set_probes(probes, 7)
        c.step(7)
        assert np.array_equal(c.populations(), o.state())
        assert np.array_equal(c.read_probes(), po)
        m64 = c.mass()
    assert m64 == pytest.approx(o.state().sum(), rel=1e-13)
    with acp.Case(acp.make_desc(nx, ny, nz, op=op, order=order, boundary=bnd, mrt_mode=mrt, tau=0.56,
                                transverse=tr, precision="F32-SHIFTED")) as c:
        c.set_populations(f0)
        c.step(7)
        rho, _ = c.macros()
    r64 = o.state().sum(axis=0)
    assert np.max(np.abs(rho - r64)) <= 2e-4 * np.max(np.abs(r64 - p.rho0))


LODI_X = [("SRT", 2, "LODI-X", False), ("HRR", 3, "LODI-X", False), ("RR", 4, "LODI-X", True),
          ("MRT-RM", 3, "LODI-X", False), ("MRT-GS", 3, "LODI-X", False), ("HRR", 4, "LODI-X", True),
          ("SRT", 3, "OPEN-LODI-X", False), ("HRR", 3, "OPEN-LODI-X", False)]


@pytest.mark.parametrize("n", [(16, 12, 10), (40, 24, 12)], ids=["one-block", "two-blocks"])
@pytest.mark.parametrize("op,order,bnd,cubic", LODI_X,
                         ids=[f"{a}-o{b}-{c}" + ("-cubic" if d else "") for a, b, c, d in LODI_X])
def test_lodi_x_faces_match_oracle_bitexact(gpu, op, order, bnd, cubic, n):
    """LODI x faces on grids wrapped in y and z (the oracle's modes 3 and 4,
    which the CPU suite pins to the reference's all-face LODI by causality):
    9 steps with transverse terms, the state, probes on both x faces and the
    mass bit for bit / to 1e-13 against the oracle."""
    nx, ny, nz = n
    mode = {"LODI-X": 3, "OPEN-LODI-X": 4}[bnd]
    relaxed = op.startswith("MRT")
    p = ol.make_params(nx, ny, nz, op=op, order=order, cubic=cubic, periodic=mode, mrt_mode=int(relaxed), tau=0.56,
                       transverse=True)
    f0 = _random_f(p, 31)
    probes = [(0, 1, 2), (nx - 1, ny - 1, nz - 1), (nx // 2, 0, 0)]
    print("case", op, order, bnd, cubic, n, flush=True)
    o = ol.OracleCase(p)
    o.set_state(f0)
    po = o.step(9, probes)
    d = acp.make_desc(nx, ny, nz, op=op, order=order, cubic=cubic, boundary=bnd,
                      mrt_mode="RELAXED" if relaxed else "REFERENCE", tau=0.56, transverse=True)
    with acp.Case(d) as c:
        c.set_populations(f0)
        c.set_probes(probes, 9)
        c.step(9)
        f = c.populations()
        assert np.array_equal(f, o.state()), f"max |df| = {np.abs(f - o.state()).max():.3e}"
        assert np.array_equal(c.read_probes(), po)
        m64 = c.mass()
        info = c.info()
    assert m64 == pytest.approx(o.state().sum(), rel=1e-13)
    assert info.lodi_nodes == (2 if mode == 3 else 1) * ny * nz and info.lodi_launches == 1


@pytest.mark.parametrize("prec,tol", [("F32-SHIFTED", 2e-4), ("F16-SHIFTED", 2e-2)])
@pytest.mark.parametrize("bnd,nx", [("LODI-X", 128), ("LODI-X", 50), ("OPEN-LODI-X", 50)])
@pytest.mark.parametrize("op", ["SRT", "HRR"])
def test_lodi_x_faces_shifted_tiers(gpu, prec, tol, bnd, nx, op):
    """The shifted tiers with LODI x faces against the float64 oracle: a pulse
    (amplitude 1e-3) next to the high-x face, 40 steps, density within the
    tolerances of the other shifted-tier tests (of the largest density
    deviation).  nx = 128 is a row length the two-nodes-per-thread kernel
    takes on fully periodic grids; with LODI faces the one-node kernel runs."""
    ny, nz = 20, 3
    mode = {"LODI-X": 3, "OPEN-LODI-X": 4}[bnd]
    p = ol.make_params(nx, ny, nz, op=op, order=3, periodic=mode, tau=0.56, sigma=0.98, transverse=True)
    f0 = ol.pulse_ic(p, 1e-3, nx - 9.0, 9.0, 5.0)
    o = ol.OracleCase(p)
    o.set_state(f0)
    o.step(40)
    rho64, _ = ol.macros(p, o.state())
    d = acp.make_desc(nx, ny, nz, op=op, order=3, boundary=bnd, tau=0.56, sigma=0.98, transverse=True,
                      precision=prec, f16_scale=1000.0)
    with acp.Case(d) as c:
        c.set_populations(f0)
        c.step(40)
        rho, _ = c.macros()
    dev = rho64 - p.rho0
    assert np.max(np.abs(rho - rho64)) <= tol * np.max(np.abs(dev))


ALL_ORDERS = [("SRT", 2), ("SRT", 3), ("SRT", 4), ("RR", 2), ("RR", 3), ("RR", 4), ("HRR", 3), ("HRR", 4),
              ("MRT-RM", 3), ("MRT-GS", 3)]


@pytest.mark.parametrize("n", [(16, 12, 10), (40, 24, 12)], ids=["one-block", "two-blocks"])
@pytest.mark.parametrize("op,order", ALL_ORDERS, ids=[f"{a}-o{b}" for a, b in ALL_ORDERS])
def test_float32_tier_follows_float64_on_moving_fields(gpu, op, order, n):
    """Every operator and equilibrium order in the float32-shifted tier
    against the float64 tier on a random field with velocities of 0.05 c_s and
    random non-equilibrium parts (periodic, 7 steps): density within 2e-6 of
    the largest density deviation (measured 2e-7 to 5e-7).  At N_eq = 4 this
    pins the reference's mixing of fourth- and second-order Hermite terms in
    SI units, which a lattice-unit restatement must carry as 1 / (6 c^2)."""
    nx, ny, nz = n
    mrt = "RELAXED" if op.startswith("MRT") else "REFERENCE"
    p = ol.make_params(nx, ny, nz, op=op, order=order, periodic=True, mrt_mode=int(op.startswith("MRT")), tau=0.56)
    f0 = _random_f(p, 21)
    out = {}
    for prec in ("F64", "F32-SHIFTED"):
        with acp.Case(acp.make_desc(nx, ny, nz, op=op, order=order, boundary="PERIODIC", mrt_mode=mrt, tau=0.56,
                                    precision=prec)) as c:
            c.set_populations(f0)
            c.step(7)
            out[prec] = c.macros()[0]
    r64 = out["F64"]
    err = np.max(np.abs(out["F32-SHIFTED"] - r64)) / np.max(np.abs(r64 - p.rho0))
    assert err <= 2e-6, err


PAIR_ROWS = [("SRT", 3), ("RR", 4), ("HRR", 3), ("MRT-RM", 3), ("MRT-GS", 3)]


@pytest.mark.parametrize("prec,tol", [("F32-SHIFTED", 2e-4), ("F16-SHIFTED", 2e-2)])
@pytest.mark.parametrize("op,order", PAIR_ROWS, ids=[a for a, _ in PAIR_ROWS])
def test_periodic_pair_rows_shifted_tiers(gpu, prec, tol, op, order):
    """Periodic rows of 128 to 1024 nodes (multiples of 128) take the shifted
    kernel with two nodes per thread.  Against the float64 oracle on a pulse
    off the grid's centre (amplitude 1e-3, 40 steps), with the same tolerances
    as the one-node-per-thread kernels, for every collision operator."""
    nx, ny, nz = 128, 20, 3
    mrt = "RELAXED" if op.startswith("MRT") else "REFERENCE"
    p = ol.make_params(nx, ny, nz, op=op, order=order, periodic=True, mrt_mode=int(op.startswith("MRT")), tau=0.56,
                       sigma=0.98)
    f0 = ol.pulse_ic(p, 1e-3, 41.0, 9.0, 5.0)
    o = ol.OracleCase(p)
    o.set_state(f0)
    o.step(40)
    rho64, _ = ol.macros(p, o.state())
    d = acp.make_desc(nx, ny, nz, op=op, order=order, boundary="PERIODIC", mrt_mode=mrt, tau=0.56, sigma=0.98,
                      precision=prec, f16_scale=1000.0)
    with acp.Case(d) as c:
        c.set_populations(f0)
        c.step(40)
        rho, _ = c.macros()
    dev = rho64 - p.rho0
    err = np.max(np.abs(rho - rho64)) / np.max(np.abs(dev))
    print(f"{prec} {op}: relative density error {err:.3e}")
    assert err <= tol


@pytest.mark.parametrize("prec", ["F32-SHIFTED", "F16-SHIFTED"])
def test_periodic_pair_rows_wrap_exactly(gpu, prec):
    """A state of period 128 in x on a 256-wide periodic grid stays the
    128-wide grid's state tiled twice, to the bit: the realignment of x-links
    through LDS wraps at the row's end exactly like the neighbour offsets."""
    ny, nz = 10, 4
    p = ol.make_params(128, ny, nz, op="RR", order=3, periodic=True, tau=0.56)
    f0 = _random_f(p, 17).reshape(19, nz, ny, 128)
    out = {}
    for nx in (128, 256):
        f = np.concatenate([f0] * (nx // 128), axis=3).reshape(19, -1)
        with acp.Case(acp.make_desc(nx, ny, nz, op="RR", order=3, boundary="PERIODIC", tau=0.56,
                                    precision=prec, f16_scale=1000.0)) as c:
            c.set_populations(f)
            c.step(12)
            out[nx] = c.populations().reshape(19, nz, ny, nx)
    assert np.array_equal(out[256][..., :128], out[128])
    assert np.array_equal(out[256][..., 128:], out[128])


def test_periodic_mass_and_translation(gpu):
    """Periodic BGK: exact shift equivariance, mass conserved to round-off."""
    nx, ny, nz = 16, 12, 10
    p = ol.make_params(nx, ny, nz, op="SRT", periodic=True, tau=0.7)
    f0 = _random_f(p, 3).reshape(19, nz, ny, nx)
    shifted = np.roll(f0, (2, -3, 5), axis=(1, 2, 3))
    d = acp.make_desc(nx, ny, nz, op="SRT", boundary="PERIODIC", tau=0.7)
    with acp.Case(d) as a, acp.Case(d) as b:
        a.set_populations(f0.reshape(19, -1))
        b.set_populations(shifted.reshape(19, -1))
        m0 = a.mass()
        a.step(30)
        b.step(30)
        fa = a.populations().reshape(19, nz, ny, nx)
        fb = b.populations().reshape(19, nz, ny, nx)
        m1 = a.mass()
    assert np.array_equal(np.roll(fa, (2, -3, 5), axis=(1, 2, 3)), fb)
    assert abs(m1 - m0) / m0 < 1e-13


def test_init_random_statistics(gpu):
    """Philox/Box-Muller field: density and velocity fluctuations have the requested
    standard deviation and the populations are the equilibrium of (rho, u)."""
    d = acp.make_desc(64, 64, 64, op="MRT-RM", boundary="PERIODIC", mrt_mode="RELAXED", tau=0.503)
    with acp.Case(d) as c:
        c.init_random(0, 1e-3, 1e-3)
        rho, u = c.macros()
    cs_lu = CS * math.sqrt(3.0)
    drho = rho / RHO0 - 1.0
    assert abs(drho.std() - 1e-3) < 2e-5 and abs(drho.mean()) < 2e-5
    for a in range(3):
        ua = u[a] / cs_lu
        assert abs(ua.std() - 1e-3) < 2e-5 and abs(ua.mean()) < 2e-5


@pytest.mark.parametrize("prec,tol", [("F32-SHIFTED", 2e-4), ("F16-SHIFTED", 2e-2)])
@pytest.mark.parametrize("op", ["SRT", "HRR", "MRT-RM"])
def test_shifted_tiers_against_f64(gpu, prec, tol, op):
    """Quasi-2D pulse (amplitude 1e-3) with LODI faces, 300 steps: the density
    perturbation of the shifted tier against the float64 oracle, relative to its
    maximum: max|drho_tier - drho_64| / max|drho_64| <= tol (2e-4 for float32,
    2e-2 for float16 at scale 1e3)."""
    nx, ny, nz, steps = 96, 80, 3, 300
    mrt = "RELAXED" if op == "MRT-RM" else "REFERENCE"
    p = ol.make_params(nx, ny, nz, op=op, order=3, mrt_mode=int(op == "MRT-RM"), tau=0.52, transverse=True,
                       sigma=0.98)
    f0 = ol.pulse_ic(p, 1e-3, 48.0, 40.0, 6.0)
    o = ol.OracleCase(p)
    o.set_state(f0)
    o.step(steps)
    r64, _ = ol.macros(p, o.state())
    d = acp.make_desc(nx, ny, nz, op=op, order=3, precision=prec, mrt_mode=mrt, tau=0.52, transverse=True,
                      f16_scale=1000.0)
    with acp.Case(d) as c:
        c.set_populations(f0)
        c.step(steps)
        rho, _ = c.macros()
    d64 = r64 / RHO0 - 1.0
    dt = rho / RHO0 - 1.0
    err = np.abs(dt - d64).max() / np.abs(d64).max()
    print(f"{prec} {op}: relative density error {err:.3e}")
    assert err <= tol


def test_probe_and_mass_f64(gpu):
    nx, ny, nz = 20, 18, 6
    p = ol.make_params(nx, ny, nz, op="RR", order=3, tau=0.6)
    f0 = _random_f(p, 5)
    with acp.Case(acp.make_desc(nx, ny, nz, op="RR", tau=0.6)) as c:
        c.set_populations(f0)
        c.set_probes([(3, 4, 2), (10, 9, 3)], 10)
        c.step(10)
        pr = c.read_probes()
        f = c.populations()
        m = c.mass()
    o = ol.OracleCase(p)
    o.set_state(f0)
    po = o.step(10, [(3, 4, 2), (10, 9, 3)])
    assert np.array_equal(pr, po)
    assert abs(m - f.sum()) <= 1e-12 * abs(m)


VTK_CASES = sorted(glob.glob(os.path.join(GOLD, "vtk_*.npz")))


@pytest.mark.parametrize("path", VTK_CASES, ids=[os.path.basename(p)[4:-4] for p in VTK_CASES])
def test_case_vtk_output_matches_reference_bytes(gpu, tmp_path, path):
    """acp_case_write_vtk after GPU steps equals the reference's writeResultsVTK file."""
    z = np.load(path)
    spec = json.loads(str(z["spec"]))
    d = desc_of(spec)
    d.nonreflecting = int(z["nonreflecting"])
    stem = str(tmp_path / "gpu")
    with acp.Case(d) as c:
        c.set_populations(z["f0"])
        c.step(int(z["nsteps"]))
        acp._capi.check(acp.lib().acp_case_write_vtk(c.h, stem.encode()))
    assert open(stem + ".vtk", "rb").read() == z["vtk"].tobytes()


def test_acp_run_driver_end_to_end(gpu, tmp_path):
    """The command-line driver: the reference's main() loop on a case file, writing
    <Save-as>_<t>.vtk at t = 0 and every WriteInterval; the last file equals the
    oracle state written by the same writer."""
    import subprocess
    from make_golden import bc_text
    nx, ny, nz, steps = 40, 30, 3, 6
    txt = bc_text(nx, ny, nz, op="SRT", order=3, tau=0.55, spacing=0.02, steps=steps)
    txt = txt.replace("WriteInterval 1000000", "WriteInterval 3").replace("Save-as /tmp/acp_golden",
                                                                            f"Save-as {tmp_path}/run")
    (tmp_path / "bc.txt").write_text(txt)
    exe = os.path.join(os.path.dirname(acp._capi.LIB_PATH), "acp_run")
    r = subprocess.run([exe, str(tmp_path / "bc.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "LB acoustic pulse solver" in r.stdout and "given maximum number of time steps reached" in r.stdout
    assert "solved 100.00% of the case" in r.stdout
    for t in (0, 3, 6):
        assert os.path.exists(tmp_path / f"run_{t}.vtk")
    p = ol.make_params(nx, ny, nz, op="SRT", order=3, tau=0.55, spacing=0.02)
    o = ol.OracleCase(p)
    o.set_state(ol.reference_ic(p, (nx - 0.5) * 0.02, (ny - 0.5) * 0.02))
    o.step(steps)
    cur, prev = o.state(), o.prev()
    d = acp.read_case_file(str(tmp_path / "bc.txt"))
    acp._capi.check(acp.lib().acp_write_vtk_host(d, acp._capi.dptr(cur), acp._capi.dptr(prev), None,
                                                 str(tmp_path / "oracle").encode()))
    assert open(tmp_path / "run_6.vtk", "rb").read() == open(tmp_path / "oracle.vtk", "rb").read()


def test_lattice_interface_both_arrays(gpu, tmp_path):
    """The Control / Lattice interface (the reference's main() loop, Lattice.h:102-120):
    initialization, solve with the caller's ping-pong flag, and getDistribution from
    either array, against the oracle's current and previous arrays bit for bit; a
    flag naming the previous array is refused by solve."""
    from make_golden import bc_text
    nx, ny, nz, steps = 12, 10, 5, 5
    (tmp_path / "bc.txt").write_text(bc_text(nx, ny, nz, op="RR", order=3, tau=0.56, spacing=0.02, steps=steps,
                                             transverse=True))
    bc = acp.Control(str(tmp_path / "bc.txt"))
    lat = acp.Lattice(bc)
    lat.initialization(bc)
    alt = True
    for _ in range(steps):
        lat.solve(alt)
        alt = not alt
    with pytest.raises(acp.AcpError):
        lat.solve(not alt)
    p = ol.make_params(nx, ny, nz, op="RR", order=3, tau=0.56, spacing=0.02, transverse=True)
    o = ol.OracleCase(p)
    o.set_state(ol.reference_ic(p, (nx - 0.5) * 0.02, (ny - 0.5) * 0.02))
    o.step(steps)
    cur, prev = o.state(), o.prev()
    for (q, i, j, k) in [(0, 0, 0, 0), (5, 3, 4, 2), (18, 11, 9, 4), (13, 6, 0, 3)]:
        l = i + nx * (j + ny * k)
        assert lat.getDistribution(alt, q, 0, i, j, k) == cur[q, l]
        assert lat.getDistribution(not alt, q, 0, i, j, k) == prev[q, l]
    assert lat.getDistribution(alt, 19, 0, 0, 0, 0) == 0.0
    assert np.array_equal(lat.case.previous_populations(), prev)


LDS_ROWS = [("SRT", 3, "LODI", False, False), ("HRR", 3, "LODI", True, True), ("RR", 4, "LODI-X", False, True),
            ("SRT", 2, "OPEN-X", False, False), ("MRT-GS", 3, "LODI", False, False)]


@pytest.mark.parametrize("op,order,bnd,cubic,tr", LDS_ROWS, ids=[f"{a}-{c}" for a, b, c, d, e in LDS_ROWS])
def test_boundary_rows_through_lds_match_oracle_bitexact(gpu, op, order, bnd, cubic, tr):
    """Grids of at least 4096 rows of 64 nodes store through the LDS row
    exchange also when their rows hold boundary nodes (LODI faces, edges and
    corners, LODI or open x faces): each realigned position is written only if
    its source node pushes the link.  64^3 nodes, 3 steps, bit for bit against
    the oracle, and the float32- and float16-shifted tiers through the same path
    within 2e-4 and 2e-2 of the largest density deviation."""
    n = 64
    mode = {"LODI": 0, "LODI-X": 3, "OPEN-X": 2}[bnd]
    relaxed = op.startswith("MRT")
    p = ol.make_params(n, n, n, op=op, order=order, cubic=cubic, transverse=tr, periodic=mode,
                       mrt_mode=int(relaxed), tau=0.56)
    f0 = _random_f(p, 41)
    print("case", op, order, bnd, flush=True)
    o = ol.OracleCase(p)
    o.set_state(f0)
    o.step(3)
    kw = dict(op=op, order=order, cubic=cubic, transverse=tr, boundary=bnd,
              mrt_mode="RELAXED" if relaxed else "REFERENCE", tau=0.56)
    with acp.Case(acp.make_desc(n, n, n, **kw)) as c:
        c.set_populations(f0)
        c.step(3)
        f = c.populations()
    assert np.array_equal(f, o.state()), f"max |df| = {np.abs(f - o.state()).max():.3e}"
    if not cubic:
        r64 = o.state().sum(axis=0)
        for prec, tol in (("F32-SHIFTED", 2e-4), ("F16-SHIFTED", 2e-2)):
            with acp.Case(acp.make_desc(n, n, n, precision=prec, f16_scale=1000.0, **kw)) as c:
                c.set_populations(f0)
                c.step(3)
                rho, _ = c.macros()
            assert np.max(np.abs(rho - r64)) <= tol * np.max(np.abs(r64 - p.rho0)), prec


@pytest.mark.parametrize("coupling", ["STD", "DC1"])
def test_acp_run_driver_two_level(gpu, tmp_path, coupling):
    """The driver on a MaxLevel 1 case file (this engine's refined geometry:
    Boundary LODI-X, Refined-Region 12 23, the fine level reaching the high-x
    face): the reference's two-level summary, <Save-as>_<t>.vtk at t = 0 and
    every WriteInterval, and the last file equal to the two-level writer's
    output of the oracle's refined run of the same case (the reference's
    initial pulse on both levels)."""
    import subprocess
    from make_golden import bc_text
    nx, ny, nz, steps, x0, x1 = 24, 10, 3, 4, 12, 23
    txt = bc_text(nx, ny, nz, op="RR", order=3, tau=0.56, spacing=0.02, steps=steps).replace("MaxLevel 0", "MaxLevel 1")
    txt = txt.replace("WriteInterval 1000000", "WriteInterval 2").replace("Save-as /tmp/acp_golden",
                                                                            f"Save-as {tmp_path}/run")
    links = ("Mf", "Mc", "hMf") if coupling != "STD" else ("fAll", "cAll", "hAll")
    txt = (txt.replace("Coupling-Approach STD", f"Coupling-Approach {coupling}")
           .replace("Coarse-to-Fine fAll", f"Coarse-to-Fine {links[0]}")
           .replace("Fine-to-Coarse cAll", f"Fine-to-Coarse {links[1]}")
           .replace("Hanging-Node-Reconstruction hAll", f"Hanging-Node-Reconstruction {links[2]}")
           .replace("F2C-Filter NOFI", "F2C-Filter LAG"))
    txt += f"Boundary LODI-X\nRefined-Region {x0} {x1}\n"
    (tmp_path / "bc.txt").write_text(txt)
    exe = os.path.join(os.path.dirname(acp._capi.LIB_PATH), "acp_run")
    r = subprocess.run([exe, str(tmp_path / "bc.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "dimensionless collision frequency on fine grid" in r.stdout
    assert f"coupling approach = {coupling}" in r.stdout and "f2c filtering = LAG" in r.stdout
    assert "solved 100.00% of the case" in r.stdout
    for t in (0, 2, 4):
        assert os.path.exists(tmp_path / f"run_{t}.vtk")
    d = acp.read_case_file(str(tmp_path / "bc.txt"))
    d.max_level = 0
    p0 = ol.make_params(nx, ny, nz, op="RR", order=3, tau=0.56, spacing=0.02, periodic=3)
    o = ol.OracleRefined(p0, x0, x1)
    o.set_filter("LAG")
    o.set_coupling(coupling, *links)
    p1 = o.levels[1].p
    o.levels[0].set_state(ol.reference_ic_level(p0, 0, 0, 0.02, d.domain[0], d.domain[1]))
    o.levels[1].set_state(ol.reference_ic_level(p1, 1, 2 * x0, 0.02, d.domain[0], d.domain[1]))
    o.step(steps)
    D = acp._capi.dptr
    st = [(o.levels[lv].state(), o.levels[lv].prev()) for lv in (0, 1)]
    acp._capi.check(acp.lib().acp_write_vtk2_host(d, D(st[0][0]), D(st[0][1]), x0, x1, D(st[1][0]), D(st[1][1]),
                                                   str(tmp_path / "oracle").encode()))
    assert open(tmp_path / "run_4.vtk", "rb").read() == open(tmp_path / "oracle.vtk", "rb").read()
    # and the library's own two-level writer on the GPU state gives the same bytes
    with acp.Refined(d, x0, x1) as rc:
        rc.set_filter("LAG")
        rc.set_coupling(coupling, *links)
        rc.init_reference()
        rc.step(steps)
        rc.write_vtk(str(tmp_path / "gpu"))
    assert open(tmp_path / "gpu.vtk", "rb").read() == open(tmp_path / "oracle.vtk", "rb").read()


def test_acp_run_driver_reference_two_level_geometry(gpu, tmp_path):
    """The driver on a MaxLevel 1 case file with the reference's keys alone
    (its own geometry, a fine shell around a coarse core: RefinementLayerYZ 3,
    RefinementLayerX 6, Overlap 1, HRR with the standard coupling): the
    two-level summary, <Save-as>_<t>.vtk at t = 0 and every WriteInterval, and
    the last file equal to the shell writer's output of the oracle's run of
    the same case (the reference's initial pulse on both levels)."""
    import subprocess
    from make_golden import bc_text
    nx, ny, nz, steps, ly, lx = 21, 17, 13, 4, 3, 6
    txt = (bc_text(nx, ny, nz, op="HRR", order=3, tau=0.56, spacing=0.02, steps=steps)
           .replace("MaxLevel 0", "MaxLevel 1").replace("RefinementLayerYZ 4", f"RefinementLayerYZ {ly}")
           .replace("RefinementLayerX 75", f"RefinementLayerX {lx}")
           .replace("WriteInterval 1000000", "WriteInterval 2").replace("Save-as /tmp/acp_golden",
                                                                        f"Save-as {tmp_path}/run"))
    (tmp_path / "bc.txt").write_text(txt)
    exe = os.path.join(os.path.dirname(acp._capi.LIB_PATH), "acp_run")
    r = subprocess.run([exe, str(tmp_path / "bc.txt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "dimensionless collision frequency on fine grid" in r.stdout
    assert "coupling approach = STD" in r.stdout and "solved 100.00% of the case" in r.stdout
    assert "error" not in r.stderr.lower(), r.stderr
    for t in (0, 2, 4):
        assert os.path.exists(tmp_path / f"run_{t}.vtk")
    d = acp.read_case_file(str(tmp_path / "bc.txt"))
    d.max_level = 0
    p0 = ol.make_params(nx, ny, nz, op="HRR", order=3, tau=0.56, spacing=0.02, periodic=0)
    o = ol.OracleRefined(p0, 0, 0, shell=(ly, lx))
    for lv in (0, 1):
        o.levels[lv].set_state(ol.reference_ic_level(o.levels[lv].p, lv, 0, 0.02, d.domain[0], d.domain[1]))
    o.step(steps)
    D = acp._capi.dptr
    st = [(o.levels[lv].state(), o.levels[lv].prev()) for lv in (0, 1)]
    acp._capi.check(acp.lib().acp_write_vtk_shell_host(d, ly, lx, D(st[0][0]), D(st[0][1]), None, D(st[1][0]),
                                                        D(st[1][1]), None, str(tmp_path / "oracle").encode()))
    assert open(tmp_path / "run_4.vtk", "rb").read() == open(tmp_path / "oracle.vtk", "rb").read()
