"""The VTK writer of libacpulse.so against files written by the reference's own
Lattice::writeResultsVTK (Lattice.cpp:4094-4349), byte for byte, on states
computed by the oracle."""
import glob
import json
import os

import numpy as np
import pytest

import acpulse_vclbm_amd as acp
import oracle_lib as ol
from make_golden import params_of

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(glob.glob(os.path.join(GOLD, "vtk_*.npz")))


def _desc(spec, nonreflecting):
    d = acp.make_desc(spec["nx"], spec["ny"], spec["nz"], op=spec["op"], order=spec["order"],
                      cubic=spec.get("cubic", False), transverse=spec.get("transverse", False),
                      tau=spec.get("tau", 0.6))
    d.nonreflecting = int(nonreflecting)
    return d


@pytest.mark.parametrize("path", CASES, ids=lambda p: os.path.basename(p)[4:-4])
def test_vtk_bytes_match_reference(tmp_path, path):
    z = np.load(path)
    spec = json.loads(str(z["spec"]))
    c = ol.OracleCase(params_of(spec))
    c.set_state(z["f0"])
    c.step(int(z["nsteps"]))
    cur, prev, psi = c.state(), c.prev(), c.psi()
    d = _desc(spec, int(z["nonreflecting"]))
    stem = str(tmp_path / "out")
    acp._capi.check(acp.lib().acp_write_vtk_host(d, acp._capi.dptr(cur), acp._capi.dptr(prev),
                                                 acp._capi.dptr(psi) if spec.get("cubic") else None,
                                                 stem.encode()))
    got = open(stem + ".vtk", "rb").read()
    ref = z["vtk"].tobytes()
    assert got == ref, _first_diff(got, ref)


def _first_diff(a, b):
    for n, (x, y) in enumerate(zip(a.splitlines(), b.splitlines())):
        if x != y:
            return f"line {n}: {x!r} != {y!r}"
    return f"lengths {len(a)} != {len(b)}"


def _parse_vtk(path):
    """Sections of a legacy ASCII VTK file written by the engine."""
    lines = open(path).read().split("\n")
    out, i = {}, 0
    while i < len(lines):
        t = lines[i].split()
        if t and t[0] == "POINTS":
            n = int(t[1])
            out["points"] = np.array([[float(v) for v in lines[i + 1 + m].split()] for m in range(n)])
            i += n
        elif t and t[0] == "CELLS":
            n = int(t[1])
            assert int(t[2]) == 9 * n
            out["cells"] = np.array([[int(v) for v in lines[i + 1 + m].split()] for m in range(n)])
            i += n
        elif t and t[0] == "CELL_TYPES":
            out["cell_types"] = [lines[i + 1 + m] for m in range(int(t[1]))]
        elif t and t[0] == "POINT_DATA":
            out["n"] = int(t[1])
        elif t and t[0] == "SCALARS":
            n = out["n"]
            out[t[1]] = np.array([float(v) for v in lines[i + 2:i + 2 + n]])
            i += n + 1
        elif t and t[0] == "VECTORS":
            n = out["n"]
            out[t[1]] = np.array([[float(v) for v in lines[i + 1 + m].split()] for m in range(n)])
            i += n
        i += 1
    return out


@pytest.mark.parametrize("boundary,x1", [("LODI-X", 11), ("LODI-X", 15), ("PERIODIC", 16)],
                         ids=["lodi-x-inside", "lodi-x-to-face", "periodic-wrap"])
def test_two_level_vtk_layout(tmp_path, boundary, x1):
    """Two-level output from oracle states (acp_write_vtk2_host): the fine
    nodes then the coarse columns 0..x0 and x1..nx-1 at their coordinates,
    VOXEL cells that join neighbouring printed nodes of one level, and per
    point the level's Omega, NodeID, pressure, density and velocity of that
    level's populations (LODI nodes with i > 0 from the other array under
    Non-Reflecting)."""
    nx, ny, nz, x0 = 16, 6, 3, 5
    d = acp.make_desc(nx, ny, nz, op="RR", tau=0.6, boundary=boundary)
    fd = acp.CaseDesc()
    acp._capi.check(acp.lib().acp_refined_fine_desc(d, x0, x1, fd))
    periodic = True if boundary == "PERIODIC" else 3
    p = ol.make_params(nx, ny, nz, op="RR", periodic=periodic, tau=0.6, cs=d.cs, spacing=d.spacing, rho0=d.rho0)
    o = ol.OracleRefined(p, x0, x1)
    o.levels[0].set_state(ol.pulse_ic(p, 1e-3, 4.0, 3.0, 2.0))
    o.levels[1].set_state(ol.pulse_ic(o.levels[1].p, 1e-3, 2 * (4.0 - x0), 6.0, 4.0))
    o.step(3)
    lv = [(o.levels[0].state(), o.levels[0].prev()), (o.levels[1].state(), o.levels[1].prev())]
    stem = str(tmp_path / "two")
    D = acp._capi.dptr
    acp._capi.check(acp.lib().acp_write_vtk2_host(d, D(lv[0][0]), D(lv[0][1]), x0, x1, D(lv[1][0]), D(lv[1][1]),
                                                   stem.encode()))
    v = _parse_vtk(stem + ".vtk")
    fnx, fny, fnz = fd.nx, fd.ny, fd.nz
    cols = list(range(0, x0 + 1)) + list(range(x1, nx))
    FN, CP = fnx * fny * fnz, len(cols) * ny * nz
    assert v["n"] == FN + CP == len(v["points"])
    sp = d.spacing
    k, j, i = np.meshgrid(np.arange(fnz), np.arange(fny), np.arange(fnx), indexing="ij")
    fine_xyz = np.stack([(2 * x0 + i.ravel()) * sp / 2, j.ravel() * sp / 2, k.ravel() * sp / 2], 1)
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.array(cols), indexing="ij")
    coarse_xyz = np.stack([I.ravel() * sp, J.ravel() * sp, K.ravel() * sp], 1)
    assert np.allclose(v["points"], np.vstack([fine_xyz, coarse_xyz]), rtol=1e-7, atol=1e-12)
    # cells: unit voxels of one level (side sp/2 among fine points, sp among coarse)
    assert set(v["cell_types"]) == {"11"} and np.all(v["cells"][:, 0] == 8)
    nfc = (fnx - 1) * (fny - 1) * (fnz - 1)
    runs = sum(1 for a in range(len(cols) - 1) if cols[a + 1] == cols[a] + 1)
    assert len(v["cells"]) == nfc + runs * (ny - 1) * (nz - 1)
    for c, cell in enumerate(v["cells"][:, 1:]):
        P = v["points"][cell]
        side = sp / 2 if c < nfc else sp
        assert np.all(cell < FN) if c < nfc else np.all(cell >= FN)
        assert np.allclose(P.max(0) - P.min(0), side, rtol=1e-6), c
    # fields against the oracle's moments of each level
    for level, rows in ((1, range(FN)), (0, range(FN, FN + CP))):
        st, pv = lv[level]
        pp = o.levels[level].p
        dd = fd if level else d
        rho, u = ol.macros(pp, st)
        rho_o, u_o = ol.macros(pp, pv)
        idx = np.arange(FN) if level else (I.ravel() + nx * (J.ravel() + ny * K.ravel()))
        ii = idx % dd.nx
        lodi = (ii == dd.nx - 1) & (dd.boundary in (acp.BOUNDARY["LODI-X"], acp.BOUNDARY["OPEN-LODI-X"]))
        lodi |= (ii == 0) & (dd.boundary == acp.BOUNDARY["LODI-X"])
        other = lodi & (ii > 0) & bool(d.nonreflecting)
        want_rho = np.where(other, rho_o[idx], rho[idx])
        rows = np.array(list(rows))
        assert np.allclose(v["Density"][rows], want_rho, rtol=5e-8, atol=0), level
        assert np.allclose(v["Pressure"][rows], want_rho * d.cs ** 2, rtol=5e-8, atol=0), level
        want_u = np.where(other[None, :], u_o[:, idx], u[:, idx]).T
        assert np.allclose(v["Velocity"][rows], want_u, rtol=5e-8, atol=1e-12 * d.cs), level
        gc = (ii == 0) | ((ii == dd.nx - 1) & (dd.boundary == acp.BOUNDARY["OPEN-X"])) if level else np.zeros_like(lodi)
        want_id = np.where(lodi, 1, np.where(gc, 2, 0))
        assert np.array_equal(v["NodeID"][rows], want_id), level
    omega = v["Omega"]
    assert len(set(omega[:FN])) == 1 and len(set(omega[FN:])) == 1 and omega[0] != omega[-1]


@pytest.mark.parametrize("nonreflecting", [False, True])
def test_shell_two_level_vtk_layout(tmp_path, nonreflecting):
    """Two-level output of the reference's own geometry from oracle states
    (acp_write_vtk_shell_host): the points of the level-1 bottom, top, left
    and right slabs and inlet and outlet blocks, then the level-0 core, in
    that order and at their coordinates (planes where two parts meet printed
    by both); VOXEL cells that join neighbouring points of each fine part,
    and the core's cells with the level-1 plane size as z stride as the
    reference's writer computes it; per point the level's Omega, NodeID (1 on
    the domain faces of level 1, 2 on the hole's surface), pressure, density
    and velocity (LODI nodes with i > 0 from the other array under
    Non-Reflecting)."""
    nx, ny, nz, ly, lx = 21, 17, 13, 3, 6
    d = acp.make_desc(nx, ny, nz, op="RR", tau=0.6, boundary="LODI")
    d.nonreflecting = int(nonreflecting)
    p = ol.make_params(nx, ny, nz, op="RR", periodic=False, tau=0.6, cs=d.cs, spacing=d.spacing, rho0=d.rho0)
    o = ol.OracleRefined(p, 0, 0, shell=(ly, lx))
    for lv in o.levels:
        lv.set_state(ol.pulse_ic(lv.p, 1e-3, 9.0 * lv.p.nx / nx, 8.0 * lv.p.ny / ny, 2.0 * lv.p.nx / nx))
    o.step(3)
    lv = [(o.levels[0].state(), o.levels[0].prev()), (o.levels[1].state(), o.levels[1].prev())]
    stem = str(tmp_path / "shell")
    D = acp._capi.dptr
    acp._capi.check(acp.lib().acp_write_vtk_shell_host(d, ly, lx, D(lv[0][0]), D(lv[0][1]), None, D(lv[1][0]),
                                                        D(lv[1][1]), None, stem.encode()))
    v = _parse_vtk(stem + ".vtk")
    fnx, fny, fnz = 2 * nx - 1, 2 * ny - 1, 2 * nz - 1
    imax, jmax, kmax, R, RX = fnx - 1, fny - 1, fnz - 1, 2 * ly, 2 * lx
    parts = [(1, 0, imax, 0, jmax, 0, R), (1, 0, imax, 0, jmax, kmax - R, kmax), (1, 0, imax, 0, R, R, kmax - R),
             (1, 0, imax, jmax - R, jmax, R, kmax - R), (1, 0, RX, R, jmax - R, R, kmax - R),
             (1, imax - R, imax, R, jmax - R, R, kmax - R), (0, lx, nx - 1 - ly, ly, ny - 1 - ly, ly, nz - 1 - ly)]
    order = []  # (level, i, j, k) of every printed point
    for lev, i0, i1, j0, j1, k0, k1 in parts:
        K, J, I = np.meshgrid(np.arange(k0, k1 + 1), np.arange(j0, j1 + 1), np.arange(i0, i1 + 1), indexing="ij")
        order.append(np.stack([np.full(I.size, lev), I.ravel(), J.ravel(), K.ravel()], 1))
    order = np.vstack(order)
    P = len(order)
    assert v["n"] == P == len(v["points"])
    h = np.where(order[:, 0] == 1, d.spacing / 2, d.spacing)
    assert np.allclose(v["points"], order[:, 1:] * h[:, None], rtol=1e-7, atol=1e-12)
    # cells: unit voxels within each fine part; the core part's stride quirk
    assert set(v["cell_types"]) == {"11"} and np.all(v["cells"][:, 0] == 8)
    sizes = [(i1 - i0) * (j1 - j0) * (k1 - k0) for _, i0, i1, j0, j1, k0, k1 in parts]
    assert len(v["cells"]) == sum(sizes)
    nfine = sum(sizes[:6])
    start = np.cumsum([0] + [(i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1) for _, i0, i1, j0, j1, k0, k1 in parts])
    part_of = np.searchsorted(start, np.arange(P), side="right") - 1
    for c, cell in enumerate(v["cells"][:nfine, 1:]):
        pts = v["points"][cell]
        assert np.all(part_of[cell] == part_of[cell[0]]), c
        assert np.allclose(pts.max(0) - pts.min(0), d.spacing / 2, rtol=1e-6), c
    core = v["cells"][nfine:, 1:]
    W = (nx - 1 - ly) - lx + 1
    assert np.array_equal(core[:, 2] - core[:, 0], np.full(len(core), W))
    assert np.array_equal(core[:, 4] - core[:, 0], np.full(len(core), fnx * fny))  # the level-1 plane size
    assert core[0, 0] == start[6]
    # fields against the oracle's moments of each level
    for level in (0, 1):
        rows = np.where(order[:, 0] == level)[0]
        I, J, K = order[rows, 1], order[rows, 2], order[rows, 3]
        pp = o.levels[level].p
        nxl, nyl, nzl = pp.nx, pp.ny, pp.nz
        idx = I + nxl * (J + nyl * K)
        st, pv = lv[level]
        rho, u = ol.macros(pp, st)
        rho_o, u_o = ol.macros(pp, pv)
        face = (I == 0) | (J == 0) | (K == 0) | (I == nxl - 1) | (J == nyl - 1) | (K == nzl - 1)
        other = face & (I > 0) & nonreflecting
        want_rho = np.where(other, rho_o[idx], rho[idx])
        assert np.allclose(v["Density"][rows], want_rho, rtol=5e-8, atol=0), level
        assert np.allclose(v["Pressure"][rows], want_rho * d.cs ** 2, rtol=5e-8, atol=0), level
        want_u = np.where(other[None, :], u_o[:, idx], u[:, idx]).T
        assert np.allclose(v["Velocity"][rows], want_u, rtol=5e-8, atol=1e-12 * d.cs), level
        if level == 1:
            hole = ((I >= RX) & (I <= imax - R) & (J >= R) & (J <= jmax - R) & (K >= R) & (K <= kmax - R))
            inner = ((I > RX) & (I < imax - R) & (J > R) & (J < jmax - R) & (K > R) & (K < kmax - R))
            want_id = np.where(face, 1, np.where(hole & ~inner, 2, 0))
        else:
            assert not face.any()
            want_id = np.zeros(len(rows), int)
        assert np.array_equal(v["NodeID"][rows], want_id), level
    assert len(set(v["Omega"][order[:, 0] == 1])) == 1 and len(set(v["Omega"][order[:, 0] == 0])) == 1
