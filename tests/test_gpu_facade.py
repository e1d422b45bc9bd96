"""The reference's Lattice / Node interface on the GPU (include/acp_lattice.hpp).

tests/_bin/lattice_check is a program written against the reference's public
interface (Lattice.h:102-120, Node.h:38-99): it calls createLattice,
setConnectivity, initialization, the timeStepping member of the case file's
operator, solve(flag), the nestedTimeStepping member and solve(flagL0,
flagL1) of a two-level lattice, and the NodeView getters.  What it prints is
compared with the engine's own entry points (bit for bit) and the node
relations with the oracle's connectivity, which the CPU suite pins to the
reference's setConnectivity."""
import os
import subprocess

import numpy as np
import pytest

import acpulse_vclbm_amd as acp
import oracle_lib as ol
from make_golden import CS, DX, RHO0, bc_text

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_bin", "lattice_check")


def _run(path, mode):
    assert os.path.exists(EXE), "tests/_bin/lattice_check is not built (__graft_entry__.build())"
    return subprocess.run([EXE, str(path), mode], capture_output=True, text=True, timeout=120)


def _parse(out):
    rec = {}
    for line in out.splitlines():
        t = line.split()
        if t[0] in ("pop", "popprev", "psi", "macro", "push", "lodi", "node"):
            rec.setdefault(t[0], {})[tuple(int(x) for x in t[1:5])] = t[5:]
        elif t[0] == "get":
            rec.setdefault("get", {})[tuple(int(x) for x in t[1:6])] = [float(x) for x in t[6:]]
        elif t[0] == "partner":
            rec.setdefault("partner", {})[tuple(int(x) for x in t[1:5])] = tuple(int(x) for x in t[6:10])
    return rec


UNIFORM = [("SRT", 3, False, False), ("MRT-GS", 3, False, False), ("MRT-RM", 2, False, False),
           ("RR", 4, True, False), ("HRR", 3, False, True)]


@pytest.mark.parametrize("op,order,cubic,tr", UNIFORM, ids=[u[0] for u in UNIFORM])
def test_uniform_lattice_program_matches_the_engine(gpu, tmp_path, op, order, cubic, tr):
    nx, ny, nz, steps = 12, 10, 7, 4
    f = tmp_path / "bc.txt"
    f.write_text(bc_text(nx, ny, nz, op=op, order=order, cubic=cubic, transverse=tr, tau=0.56, steps=steps))
    print("case", op, order, flush=True)
    r = _run(f, "uniform")
    assert r.returncode == 0, r.stderr
    rec = _parse(r.stdout)
    with acp.Case(acp.read_case_file(str(f))) as c:
        c.init_reference()
        c.step(steps)
        cur, prev = c.populations(), c.previous_populations()
        rho, u = c.macros()
        psi = np.zeros((19, nx * ny * nz))
        acp._capi.check(acp.lib().acp_case_get_psi(c.h, acp._capi.dptr(psi)))
    p = ol.make_params(nx, ny, nz, op=op, order=order, cubic=cubic, transverse=tr, tau=0.56, cs=CS, spacing=DX,
                       rho0=RHO0)
    rows = ol.OracleCase(p).geometry()
    for key, vals in rec["pop"].items():
        lv, i, j, k = key
        l = i + nx * (j + ny * k)
        assert [float(v) for v in vals] == list(cur[:, l]), key
        assert [float(v) for v in rec["popprev"][key]] == list(prev[:, l]), key
        assert [float(v) for v in rec["psi"][key]] == list(psi[:, l]), key
        mv = rec["macro"][key]  # rho <v> u <ux> <uy> <uz>
        m = [float(mv[1]), float(mv[3]), float(mv[4]), float(mv[5])]
        assert m == [float(rho[l]), float(u[0, l]), float(u[1, l]), float(u[2, l])], key
        # Node tags, neighbours and LODI relations against the oracle's connectivity
        node = rec["node"][key]
        boundary = rows[l][26] == 1
        assert node[1] == ("boundary" if boundary else "fluid") and int(node[3]) == int(boundary), key
        if boundary:
            assert (int(node[5]), int(node[7])) == (rows[l][24], rows[l][25]), key
        assert float(node[9]) == i * DX and float(node[11]) == j * DX and float(node[13]) == k * DX
        push = rec["push"][key]
        for q in range(18):
            d, tgt = push[q].split(":")
            want = rows[l][q]
            if tgt == "-":
                assert want == -1, (key, q)
            else:
                a, b, cc = (int(x) for x in tgt.split(","))
                assert a + nx * (b + ny * cc) == want, (key, q)
        lodi = rec["lodi"][key]
        for t in range(6):
            want = rows[l][18 + t] if boundary else -1
            got = -1 if lodi[t] == "-" else (lambda a, b, cc: a + nx * (b + ny * cc))(*(int(x) for x in lodi[t].split(",")))
            assert got == want, (key, t)
    centre = (nx // 2) + nx * ((ny // 2) + ny * (nz // 2))
    assert rec["get"][(1, 0, nx // 2, ny // 2, nz // 2)] == list(cur[:, centre])
    assert rec["get"][(0, 0, nx // 2, ny // 2, nz // 2)] == list(prev[:, centre])


def test_stepping_member_of_another_operator_exits(gpu, tmp_path):
    f = tmp_path / "bc.txt"
    f.write_text(bc_text(10, 9, 5, op="SRT", tau=0.6, steps=2))
    r = _run(f, "mismatch")
    assert r.returncode == 1
    assert "timeSteppingRR" in r.stderr and "one operator per case" in r.stderr


def _refined_text(nx, ny, nz, op, region, steps, coupling=None, cubic=False):
    txt = bc_text(nx, ny, nz, op=op, tau=0.56, steps=steps, cubic=cubic).replace("MaxLevel 0", "MaxLevel 1")
    if coupling:
        scheme, c2f, f2c, hang, filt = coupling[:5]
        scaling = coupling[5] if len(coupling) > 5 else "SRTSTYLE"
        txt = (txt.replace("Coupling-Approach STD", f"Coupling-Approach {scheme}")
               .replace("Coarse-to-Fine fAll", f"Coarse-to-Fine {c2f}").replace("Fine-to-Coarse cAll", f"Fine-to-Coarse {f2c}")
               .replace("Hanging-Node-Reconstruction hAll", f"Hanging-Node-Reconstruction {hang}")
               .replace("F2C-Filter NOFI", f"F2C-Filter {filt}")
               .replace("Scaling-Style SRTSTYLE", f"Scaling-Style {scaling}"))
    return txt + f"Boundary LODI-X\nRefined-Region {region[0]} {region[1]}\n"


COUPLINGS = [None, ("DC1", "Mf", "Mc", "hMf", "LAG"), ("DC2", "MfP", "Mc", "hMfP", "NOFI"),
             ("STD", "fAll", "cAll", "hAll", "TOU"), ("STD", "fAll", "cAll", "hAll", "LAG", "MRTSTYLE"),
             ("STD", "Mf", "Mc", "hMf", "NOFI")]


@pytest.mark.parametrize("coupling", COUPLINGS, ids=["STD-NOFI", "DC1-LAG", "DC2-MfP", "STD-TOU", "STD-LAG-MRTSTYLE",
                                                     "STD-Mf-Mc"])
@pytest.mark.parametrize("op", ["RR", "HRR"])
def test_two_level_lattice_program_matches_the_engine(gpu, tmp_path, op, coupling):
    """MaxLevel 1 with this engine's refined duct geometry: the reference's
    initial pulse on both levels (against the oracle's, bit for bit), one
    nestedTimeStepping step and then solve(flagL0, flagL1) steps, equal to
    acp_refined_step's, with getDistribution on both levels and the
    coarse/fine partners."""
    _two_level_program(tmp_path, op, coupling, False)


@pytest.mark.parametrize("op,coupling", [("RR", COUPLINGS[0]), ("HRR", COUPLINGS[1])], ids=["RR-STD", "HRR-DC1"])
def test_two_level_lattice_program_with_cubic_mach_terms(gpu, tmp_path, op, coupling):
    """The same with Cubic-Mach-Correction yes: the case file's two-level run
    equals acp_refined_step's with the cubic-Mach terms handed across the
    interface."""
    _two_level_program(tmp_path, op, coupling, True)


def _two_level_program(tmp_path, op, coupling, cubic):
    nx, ny, nz, steps, x0, x1 = 24, 8, 3, 4, 12, 23
    f = tmp_path / "bc.txt"
    f.write_text(_refined_text(nx, ny, nz, op, (x0, x1), steps, coupling, cubic))
    d = acp.read_case_file(str(f))
    assert d.cubic == int(cubic)
    d.max_level = 0
    print("case refined", op, coupling, "cubic" if cubic else "", flush=True)
    with acp.Refined(d, x0, x1) as rc:
        if coupling:
            rc.set_coupling(*coupling[:4])
            rc.set_filter(coupling[4])
            if len(coupling) > 5:
                rc.set_scaling(coupling[5])
        rc.init_reference()
        p0 = ol.make_params(nx, ny, nz, op=op, tau=0.56, cs=CS, spacing=DX, rho0=RHO0, periodic=3)
        p1 = ol.make_params(2 * (x1 - x0) + 1, 2 * ny, 2 * nz, op=op, tau=0.56, cs=CS, spacing=DX / 2, rho0=RHO0,
                            periodic=4)
        assert np.array_equal(rc.level(0).populations(), ol.reference_ic_level(p0, 0, 0, DX, d.domain[0], d.domain[1]))
        assert np.array_equal(rc.level(1).populations(),
                              ol.reference_ic_level(p1, 1, 2 * x0, DX, d.domain[0], d.domain[1]))
        rc.step(steps)
        lv = [rc.level(0).populations(), rc.level(1).populations()]
    r = _run(f, "nested")
    assert r.returncode == 0, r.stderr
    rec = _parse(r.stdout)
    fnx, fny = 2 * (x1 - x0) + 1, 2 * ny
    for key, vals in rec["pop"].items():
        level, i, j, k = key
        X, Y = (nx, ny) if level == 0 else (fnx, fny)
        assert [float(v) for v in vals] == list(lv[level][:, i + X * (j + Y * k)]), key
    mid = (x0 + x1) // 2
    assert rec["get"][(1, 0, mid, ny // 2, nz // 2)] == list(lv[0][:, mid + nx * (ny // 2 + ny * (nz // 2))])
    assert rec["get"][(1, 1, 2 * x0 + 3, ny, nz)] == list(lv[1][:, 3 + fnx * (ny + fny * nz)])
    assert rec["partner"][(0, mid, ny // 2, nz // 2)] == (1, 2 * (mid - x0), 2 * (ny // 2), 2 * (nz // 2))
    assert rec["node"][(1, 0, ny, nz)][1] == "gridcoupling"
    assert rec["node"][(1, fnx - 1, 3, 1)][1] == "boundary"


# even step counts: the program prints the array the flag true names, the current one after an even count
@pytest.mark.parametrize("n,layers,steps,ref_steps", [((41, 41, 21), (4, 10), 4, 3), ((151, 151, 21), (4, 75), 2, 1)],
                         ids=["41x41x21", "as-shipped-151x151x21"])
def test_reference_two_level_geometry_through_the_facade(gpu, tmp_path, n, layers, steps, ref_steps):
    """The reference's own MaxLevel 1 geometry from its keys alone (a fine
    shell around a coarse core, RefinementLayerYZ / RefinementLayerX, Overlap
    1, HRR with the standard coupling, as the shipped BoundaryConditions.txt
    asks; the second case at its sizes): the compiled program's
    nestedTimeStepping / solve run equals acp_refined_create_shell's at the
    core box's surface, the hole's faces, edges and corners and inside both
    levels, and the engine's run equals the oracle's bit for bit."""
    nx, ny, nz = n
    txt = (bc_text(nx, ny, nz, op="HRR", tau=0.56, steps=steps).replace("MaxLevel 0", "MaxLevel 1")
           .replace("RefinementLayerYZ 4", f"RefinementLayerYZ {layers[0]}")
           .replace("RefinementLayerX 75", f"RefinementLayerX {layers[1]}"))
    f = tmp_path / "bc.txt"
    f.write_text(txt)
    d = acp.read_case_file(str(f))
    assert d.boundary == acp.BOUNDARY["LODI"]
    d.max_level = 0
    print("case shell", n, layers, flush=True)
    with acp.Refined(d, shell=layers) as rc:
        rc.init_reference()
        p0 = ol.make_params(nx, ny, nz, op="HRR", tau=0.56, cs=CS, spacing=DX, rho0=RHO0, periodic=0,
                            nthreads=min(16, os.cpu_count() or 1))
        o = ol.OracleRefined(p0, 0, 0, shell=layers)
        for lv in (0, 1):
            init = ol.reference_ic_level(o.levels[lv].p, lv, 0, DX, d.domain[0], d.domain[1])
            assert np.array_equal(rc.level(lv).populations(), init), lv
            o.levels[lv].set_state(init)
        rc.step(ref_steps)
        o.step(ref_steps)
        for lv in (0, 1):
            assert np.array_equal(rc.level(lv).populations(), o.levels[lv].state()), lv
        rc.step(steps - ref_steps)
        lv = [rc.level(0).populations(), rc.level(1).populations()]
    r = _run(f, "nested")
    assert r.returncode == 0, r.stderr
    rec = _parse(r.stdout)
    dims = [(nx, ny), (2 * nx - 1, 2 * ny - 1)]
    assert len(rec["pop"]) >= 14
    for key, vals in rec["pop"].items():
        level, i, j, k = key
        X, Y = dims[level]
        assert [float(v) for v in vals] == list(lv[level][:, i + X * (j + Y * k)]), key
    lx, ly = layers[1], layers[0]
    assert rec["node"][(0, lx - 1, ny // 2, nz // 2)][1] == "gridcoupling"
    assert rec["node"][(0, nx // 2, ny // 2, nz // 2)][1] == "fluid"
    assert rec["node"][(1, 2 * lx, 2 * ly, 2 * ly)][1] == "gridcoupling"
    assert rec["node"][(1, 2 * lx - 2, ny - 1, nz - 1)][1] == "fluid"
    assert rec["partner"][(0, lx - 1, ny // 2, nz // 2)] == (1, 2 * (lx - 1), 2 * (ny // 2), 2 * (nz // 2))


@pytest.mark.parametrize("key,value", [("F2C-Filter NOFI", "F2C-Filter XYZ"), ("Scaling-Style SRTSTYLE", "Scaling-Style XYZ"),
                                       ("Coupling-Approach STD", "Coupling-Approach XYZ"),
                                       ("Fine-to-Coarse cAll", "Fine-to-Coarse McP")])
def test_two_level_coupling_keys_the_engine_lacks_end_the_program(gpu, tmp_path, key, value):
    """A two-level case file asking for what the engine does not provide (an
    unknown filter, an unknown scaling style, an unknown scheme, or the
    standard coupling with the direct coupling's McP) ends the program naming
    the key."""
    f = tmp_path / "bc.txt"
    f.write_text(_refined_text(24, 8, 3, "RR", (12, 23), 2).replace(key, value))
    r = _run(f, "nested")
    assert r.returncode == 1
    assert value.split()[0] in r.stderr, r.stderr
