"""The CPU oracle against data produced by the reference solver itself
(tests/golden/, written by tests/golden/make_golden.py from the reference's own
uniform-grid sources).  Everything is compared bit for bit."""
import glob
import hashlib
import json
import math
import os

import numpy as np
import pytest

import oracle_lib as ol
from make_golden import CS, DX, RHO0, params_of

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "geom_*.npz"))),
                         ids=lambda p: os.path.basename(p)[5:-4])
def test_connectivity_matches_setConnectivity(path):
    """Push targets, LODI normal/second-normal/tangent neighbours, normalDir and
    isCorner of every node (Lattice.cpp:237-1394)."""
    z = np.load(path)
    nx, ny, nz = (int(v) for v in z["dims"])
    c = ol.OracleCase(ol.make_params(nx, ny, nz))
    assert np.array_equal(c.geometry(), z["rows"])


def _kat_outputs(order):
    z = np.load(os.path.join(GOLD, f"kat_o{order}.npz"))
    spec = json.loads(str(z["spec"]))
    return z, params_of(spec)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_per_function_vectors(order):
    """feq_{2,3,4}, Pi1, RR off-equilibria, collideSRT/RR/HRR, the HRR strain and
    psi, and the raw-moment / Gram-Schmidt MRT building blocks."""
    import ctypes
    z, p = _kat_outputs(order)
    k = ol.consts(p)
    kb = ctypes.byref(k)
    L = ol.lib()
    D = ol.dptr
    ns = len(z["in_rho"])
    out = np.empty(19)
    got = {n: [] for n in ["feq2", "feq3", "feq4", "pi1", "rroffeq2", "rroffeq3", "rroffeq4", "collide_srt",
                           "collide_rr", "collide_hrr", "hrr_strain", "hrr_psi", "rm_moments", "rm_feq4",
                           "rm_inverse", "gs_moments", "gs_eqmoments", "gs_inverse"]}
    for s in range(ns):
        rho = float(z["in_rho"][s])
        u = np.ascontiguousarray(z["in_u"][3 * s:3 * s + 3])
        f = np.ascontiguousarray(z["in_f"][19 * s:19 * s + 19])
        nb = np.ascontiguousarray(z["in_nb"][114 * s:114 * s + 114].reshape(6, 19))
        for o in (2, 3, 4):
            L.acpo_feq(kb, o, rho, D(u), D(out))
            got[f"feq{o}"].append(out.copy())
        r = ctypes.c_double()
        v = np.empty(3)
        L.acpo_moments(kb, D(f), None, ctypes.byref(r), D(v))
        eq = np.empty(19)
        L.acpo_feq(kb, order, r.value, D(v), D(eq))
        pi = np.empty(6)
        L.acpo_pi1(kb, D(f), D(eq), None, D(pi))
        got["pi1"].append(np.array([pi[0], pi[3], pi[4], pi[1], pi[5], pi[2]]))
        for o in (2, 3, 4):
            L.acpo_offeq(kb, o, D(pi), D(v), D(out))
            got[f"rroffeq{o}"].append(out.copy())
        for op, name in (("SRT", "collide_srt"), ("RR", "collide_rr")):
            pp = ol.make_params(p.nx, p.ny, p.nz, op=op, order=order, nu=p.nu, cs=p.cs, spacing=p.spacing,
                                rho0=p.rho0, sigma=p.sigma)
            g = f.copy()
            ps = np.zeros(19)
            L.acpo_collide(kb, ctypes.byref(pp), D(g), D(ps), None)
            got[name].append(g)
        mac = np.empty((6, 4))
        for m in range(6):
            rr = ctypes.c_double()
            vv = np.empty(3)
            L.acpo_moments(kb, D(np.ascontiguousarray(nb[m])), None, ctypes.byref(rr), D(vv))
            mac[m] = [rr.value, *vv]
        pp = ol.make_params(p.nx, p.ny, p.nz, op="HRR", order=order, nu=p.nu, cs=p.cs, spacing=p.spacing,
                            rho0=p.rho0, sigma=p.sigma)
        g = f.copy()
        ps = np.zeros(19)
        L.acpo_collide(kb, ctypes.byref(pp), D(g), D(ps), D(mac))
        got["collide_hrr"].append(g)
        got["hrr_psi"].append(ps.copy())
        S = np.empty(6)
        L.acpo_strain_psi(kb, D(mac), D(S), D(ps))
        got["hrr_strain"].append(np.array([S[0], S[3], S[4], S[1], S[5], S[2]]))
        m = np.empty(19)
        L.acpo_rm_moments(kb, D(f), D(m))
        got["rm_moments"].append(m.copy())
        L.acpo_rm_feq4(kb, rho, D(u), D(out))
        got["rm_feq4"].append(out.copy())
        L.acpo_rm_inverse(kb, D(m), D(out))
        got["rm_inverse"].append(out.copy())
        L.acpo_gs_moments(kb, D(f), D(m))
        got["gs_moments"].append(m.copy())
        meq = np.empty(19)
        L.acpo_gs_eqmoments(m[0], m[3], m[5], m[7], D(meq))
        got["gs_eqmoments"].append(meq)
        L.acpo_gs_inverse(kb, D(m), D(out))
        got["gs_inverse"].append(out.copy())
    upper = np.array([1, 1, 1, 0, 1, 1, 0, 0, 1], bool)  # xx xy xz . yy yz . . zz
    for name, vals in got.items():
        a = np.concatenate(vals)
        b = z[name]
        if name in ("pi1", "hrr_strain"):
            b = b.reshape(-1, 9)[:, upper].reshape(-1)
        assert np.array_equal(a, b), f"{name}: max |diff| {np.abs(a - b).max():.3e}"


STEP_CASES = sorted(glob.glob(os.path.join(GOLD, "step_*.npz")))


@pytest.mark.parametrize("path", STEP_CASES, ids=lambda p: os.path.basename(p)[5:-4])
def test_whole_steps(path):
    """12 uniform steps (Lattice::timeStepping*, Lattice.cpp:1496-2155) from a random
    non-equilibrium state: every population of the current array and the probe."""
    z = np.load(path)
    spec = json.loads(str(z["spec"]))
    c = ol.OracleCase(params_of(spec))
    c.set_state(z["f0"])
    pr = c.step(int(z["nsteps"]), [tuple(x) for x in z["probes"]])
    assert np.array_equal(c.state(), z["f"])
    assert np.array_equal(pr, z["probe"])


def _run_big(name, ic_fn, threads=8):
    z = np.load(os.path.join(GOLD, f"big_{name}.npz"))
    spec = json.loads(str(z["spec"]))
    p = params_of(spec)
    p.nthreads = threads
    c = ol.OracleCase(p)
    c.set_state(ic_fn(p, spec))
    pr = c.step(int(z["nsteps"]), [tuple(x) for x in z["probes"]])
    f = c.state()
    assert np.array_equal(pr, z["probe"])
    assert hashlib.sha256(f.tobytes()).hexdigest() == str(z["sha256"])


def test_reference_pulse_initialization():
    """Lattice::initialization's hard-coded pulse (Lattice.cpp:1415-1433), 60 steps."""
    sp = 0.02
    _run_big("refpulse_srt", lambda p, s: ol.reference_ic(p, (s["nx"] - 0.5) * sp, (s["ny"] - 0.5) * sp))


def test_config3r_mrtrm():
    def ic(p, s):
        N = 48 ** 3
        rng = np.random.default_rng(0)
        rho = RHO0 * (1.0 + 1e-3 * rng.standard_normal(N))
        u = 1e-3 * CS * math.sqrt(3.0) * rng.standard_normal((3, N))
        return ol.equilibrium_field(p, rho, u)
    _run_big("cfg3r_mrtrm", ic)


@pytest.mark.slow
@pytest.mark.parametrize("name", ["cfg1r_srt", "cfg1r_hrr_tr"])
def test_config1r_500_steps(name):
    """200x200x3, tau 0.5003, pulse 1e-3, 500 steps, LODI on all faces."""
    _run_big(name, lambda p, s: ol.pulse_ic(p, 1e-3, 100.0, 100.0, 10.0))
