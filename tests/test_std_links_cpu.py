"""Link sets of the standard coupling in the oracle (CPU).

BoundaryConditions.txt:73-86 lets the standard coupling (STD) replace only
some links: Coarse-to-Fine Mf (the fine face node's missing links, those
arriving from outside the fine region), MfP (also the common ones, e_x = 0)
or fAll; Fine-to-Coarse Mc (the coarse grid-coupling node's missing links,
those arriving from inside the fine region) or cAll; Hanging-Node-
Reconstruction hMf, hMfP or hAll.  The arithmetic is in the absent
GridCoupling.cpp; these tests pin the restatement by what the keys fix:
  - coarse to fine and fine to coarse alone, on a level of zeros, write exactly
    the links of their sets at exactly the coupled nodes (a link set taken for
    the other face, or a hanging node written outside its set, fails);
  - fAll / cAll / hAll given explicitly are the default, bit for bit;
  - uniform states stay uniform; a pulse crossing the interface reflects
    about as much as with every link replaced.
"""
import numpy as np
import pytest

import oracle_lib as ol

E_X = np.array([1, 0, -1, 0, 0, 0, 1, 1, 0, 0, -1, -1, 0, 0, -1, 1, 1, -1, 0])
SETS = {"Mf": 0, "MfP": 1, "fAll": 2, "hMf": 0, "hMfP": 1, "hAll": 2, "Mc": 0, "cAll": 2}


def _links(name, n):
    """Directions of a set on a face with normal n (+1 at x0, -1 at x1)."""
    s = SETS[name]
    return (E_X == n) | ((s >= 1) & (E_X == 0)) | (s == 2)


def _refined(op="RR", x1=16, periodic=True, nz=4):
    p = ol.make_params(24, 6, nz, op=op, periodic=1 if periodic else 3, tau=0.6, nthreads=4)
    return ol.OracleRefined(p, 8, x1), p


@pytest.mark.parametrize("c2f,hang", [("Mf", "hMf"), ("MfP", "hMf"), ("Mf", "hAll"), ("fAll", "hMfP")])
@pytest.mark.parametrize("x1", [16, 23], ids=["two-faces", "to-lodi-face"])
def test_coarse_to_fine_writes_exactly_its_links(c2f, hang, x1):
    r, p = _refined(x1=x1, periodic=x1 == 16)
    r.set_coupling("STD", c2f, "cAll", hang)
    rng = np.random.default_rng(3)
    c, f = r.levels
    c.set_state(p.rho0 * (1 + 0.1 * rng.uniform(size=(19, c.N))) / 19)
    f.set_state(np.zeros((19, f.N)))
    r.prolong(0)
    F = f.state().reshape(19, f.p.nz, f.p.ny, f.p.nx)
    faces = [(0, 1)] + ([(f.p.nx - 1, -1)] if x1 == 16 else [])
    touched = np.zeros(f.p.nx, bool)
    for fi, n in faces:
        touched[fi] = True
        col = F[:, :, :, fi]
        coinc = np.zeros(col.shape[1:], bool)
        coinc[::2, ::2] = True
        for q in range(19):
            want_c = _links(c2f, n)[q]
            want_h = _links(hang, n)[q]
            assert np.all((col[q][coinc] != 0) == want_c), (fi, q)
            # a hanging node's link is written when its set holds it and its
            # coincident neighbours hold a value there
            assert np.all((col[q][~coinc] != 0) == (want_c and want_h)), (fi, q)
    assert not np.any(F[:, :, :, ~touched])


@pytest.mark.parametrize("f2c", ["Mc", "cAll"])
@pytest.mark.parametrize("x1", [16, 23], ids=["two-faces", "to-lodi-face"])
@pytest.mark.parametrize("filt", ["NOFI", "LAG", "TOU"])
def test_fine_to_coarse_writes_exactly_its_links(f2c, x1, filt):
    r, p = _refined(x1=x1, periodic=x1 == 16)
    r.set_coupling("STD", "fAll", f2c, "hAll")
    r.set_filter(filt)
    rng = np.random.default_rng(4)
    c, f = r.levels
    c.set_state(np.zeros((19, c.N)))
    f.set_state(p.rho0 * (1 + 0.1 * rng.uniform(size=(19, f.N))) / 19)
    r.restrict()
    C = c.state().reshape(19, p.nz, p.ny, p.nx)
    gc = {9: -1} if x1 == 23 else {9: -1, 15: 1}
    for I in range(p.nx):
        inside = 8 < I < x1
        for q in range(19):
            written = np.all(C[q, :, :, I] != 0)
            if not inside:
                assert not np.any(C[q, :, :, I]), (I, q)
            elif f2c == "Mc" and I in gc:
                assert written == bool(E_X[q] == gc[I]), (I, q)
                if not written:
                    assert not np.any(C[q, :, :, I]), (I, q)
            else:
                assert written, (I, q)


def test_explicit_all_sets_are_the_default():
    runs = []
    for explicit in (False, True):
        r, p = _refined(op="HRR")
        if explicit:
            r.set_coupling("STD", "fAll", "cAll", "hAll")
        for lv in r.levels:
            lv.set_state(ol.pulse_ic(lv.p, 1e-3, 6.0 * (2 if lv is r.levels[1] else 1) - (16 if lv is r.levels[1] else 0),
                                     3.0 * (2 if lv is r.levels[1] else 1), 2.0))
        r.step(5)
        runs.append([lv.state() for lv in r.levels])
    for a, b in zip(*runs):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("op,links", [("RR", ("Mf", "Mc", "hMf")), ("SRT", ("MfP", "Mc", "hMfP")),
                                      ("HRR", ("Mf", "cAll", "hAll")), ("MRT-RM", ("MfP", "Mc", "hMf"))])
def test_uniform_state_stays_uniform(op, links):
    p = ol.make_params(24, 10, 3, op=op, periodic=True, tau=0.6, mrt_mode=1, nthreads=4)
    r = ol.OracleRefined(p, 8, 16)
    r.set_coupling("STD", *links)
    u = np.zeros(3) if op == "MRT-RM" else np.array([0.02 * p.cs, -0.01 * p.cs, 0.0])
    for lv in r.levels:
        lv.set_state(ol.equilibrium_field(lv.p, np.full(lv.N, p.rho0), np.repeat(u[:, None], lv.N, axis=1)))
    r.step(6)
    ref = r.levels[0].state()[:, 0]
    for lv in r.levels:
        assert np.max(np.abs(lv.state() - ref[:, None])) <= 1e-14 * np.max(np.abs(ref))


@pytest.mark.parametrize("links", [("Mf", "Mc", "hMf"), ("MfP", "cAll", "hMfP"), ("fAll", "Mc", "hAll")])
def test_interface_reflection_is_small(links):
    """The pulse of the standard coupling's reflection check (pulse at x = 16,
    probe at x = 24, interface at x = 32): identical to the uniform coarse run
    until the echo can arrive, a reflection under 5 % of the pulse after, and
    within a factor 2 of the echo with every link replaced."""
    nx, ny, x0, x1, steps = 64, 32, 32, 64, 60
    res = {}
    for key in ("all", "partial"):
        p = ol.make_params(nx, ny, 3, op="RR", periodic=True, tau=0.55, nthreads=8)
        r = ol.OracleRefined(p, x0, x1)
        if key == "partial":
            r.set_coupling("STD", *links)
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
        res[key] = (np.max(np.abs(pr[:25] - pu[:25])) / sig, np.max(np.abs(pr - pu)) / sig)
    early, late = res["partial"]
    assert early < 1e-3
    assert late < 0.05
    assert late < 2 * res["all"][1] + 1e-4
