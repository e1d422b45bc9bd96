"""The C++ interface of the reference's Control and Lattice (include/acp_lattice.hpp), CPU.

A small program is compiled against the header and libacpulse.so and reads a
case file through acp::Control: every getter must return what the reference's
Control derives from the same file (Control.cpp:8-131; the values below are
restated from the file and the reference's formulas), the keyword readers must
keep first-match semantics, and a missing keyword or MaxLevel > 0 must exit
with status 1 as the reference does.  Nothing here touches a GPU.
"""
import math
import os
import shutil
import subprocess

import pytest

import acpulse_vclbm_amd as acp
from make_golden import CS, RHO0, bc_text

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.dirname(acp._capi.LIB_PATH)

PROBE = r'''
#include <cstdio>
#include "acp_lattice.hpp"
using namespace acp;
int main(int argc, char** argv) {
    Control bc(argv[1]);
    if (argc > 2) { Lattice lat(bc); return 0; }
    std::printf("nodes %d %d %d\n", bc.getNumberOfNodesX(), bc.getNumberOfNodesY(), bc.getNumberOfNodesZ());
    std::printf("ints %d %d %d %d %d %d %d\n", bc.getLevelMax(), bc.getTimeStepMax(), bc.getWriteInterval(),
                bc.getNumberOfThreads(), bc.getRefinementLayer(), bc.getRefinementLayerX(), bc.getOrderOfEquilibrium());
    std::printf("dbl %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", bc.getSpacing(),
                bc.getHybridParameter(), bc.getDensity(), bc.getMachNumber(), bc.getMolecularVelocity(),
                bc.getSpeedOfSound(), bc.getKinematicViscosity(), bc.getTimeStep(), bc.getCollisionFrequency());
    std::printf("dom %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", bc.getChannelSizeX_(), bc.getChannelSizeY_(),
                bc.getChannelSizeZ_(), bc.getLODIK1relaxation(), bc.getLODIK2relaxation(), bc.getOverlap(),
                bc.getReynoldsNumber());
    std::printf("str %s|%s|%s|%s|%s|%s|%s|%s|%s|%s|%s\n", bc.getCaseName().c_str(), bc.getCollisionModel().c_str(),
                bc.getCouplingReplacementC2F().c_str(), bc.getCouplingReplacementF2C().c_str(),
                bc.getHangingNodeReconstruction().c_str(), bc.getScalingStyle().c_str(),
                bc.getReflectionHandling().c_str(), bc.getTransverseInclusion().c_str(),
                bc.getCubicMachCorrection().c_str(), bc.getF2Cfilter().c_str(), bc.getCouplingProcedure().c_str());
    for (int q = 0; q < 19; ++q) {
        int x = 0, y = 1, z = 2;
        std::printf("xsi %d %.17g %.17g %.17g %.17g\n", q, bc.getXsi(q, x), bc.getXsi(q, y), bc.getXsi(q, z),
                    bc.getWeightFactors()[q]);
    }
    double v[3];
    bc.readVector(argv[1], "Vec", v);
    std::printf("vec %.17g %.17g %.17g\n", v[0], v[1], v[2]);
    std::printf("first %.17g\n", bc.readValue(argv[1], "Dup"));
    bc.setTime_(7.5);
    std::printf("time %.17g\n", bc.getTime_());
    return 0;
}
'''


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    d = tmp_path_factory.mktemp("facade")
    src = d / "probe.cpp"
    src.write_text(PROBE)
    exe = d / "probe"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT}/include", str(src), f"-L{LIBDIR}", "-lacpulse",
                    f"-Wl,-rpath,{LIBDIR}", "-Wl,--allow-shlib-undefined", "-o", str(exe)], check=True,
                   capture_output=True, text=True)
    return exe


def _run(exe, path, *extra):
    return subprocess.run([str(exe), str(path), *extra], capture_output=True, text=True, timeout=60)


def test_control_getters_match_the_reference_derivations(probe, tmp_path):
    nx, ny, nz, dx, tau = 40, 30, 3, 0.02, 0.55
    txt = bc_text(nx, ny, nz, op="HRR", order=4, tau=tau, spacing=dx, steps=77, transverse=True) + \
        "Vec 1.5 -2 3e-3\nDup 1.25\nDup 9\n"
    f = tmp_path / "bc.txt"
    f.write_text(txt)
    r = _run(probe, f)
    assert r.returncode == 0, r.stderr
    out = {}
    xsi = []
    for line in r.stdout.splitlines():
        k, rest = line.split(" ", 1)
        if k == "xsi":
            xsi.append([float(t) for t in rest.split()[1:]])
        else:
            out[k] = rest
    assert out["nodes"] == f"{nx} {ny} {nz}"
    assert out["ints"] == "0 77 1000000 8 4 75 4"
    sp, sig, rho, ma, c, cs, nu, dt, om = (float(t) for t in out["dbl"].split())
    c_ref = CS * math.sqrt(3.0)
    dt_ref = dx / c_ref
    assert (sp, sig, rho, ma, cs) == (dx, 0.98, RHO0, 0.0, CS)
    assert c == c_ref and dt == dt_ref
    assert om == (CS ** 2 * dt_ref) / (nu + 0.5 * CS ** 2 * dt_ref)  # Control.cpp:54
    assert om == pytest.approx(1.0 / tau, rel=1e-12)
    lx, ly, lz, k1, k2, ov, re = (float(t) for t in out["dom"].split())
    assert (lx, ly, lz) == ((nx - 0.5) * dx, (ny - 0.5) * dx, (nz - 0.5) * dx)
    assert (k1, k2, ov, re) == (1.0, 1.0, 1.0, 0.0)
    assert out["str"] == "/tmp/acp_golden|HRR|fAll|cAll|hAll|SRTSTYLE|yes|yes|no|NOFI|STD"
    # D3Q19 in the reference's order, scaled by c; weights 1/18, 1/36, 1/3
    e = [(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 0, 1), (1, 0, -1), (0, 1, 1),
         (0, 1, -1), (-1, 0, 1), (-1, 0, -1), (0, -1, 1), (0, -1, -1), (-1, 1, 0), (1, -1, 0), (1, 1, 0),
         (-1, -1, 0), (0, 0, 0)]
    assert len(xsi) == 19
    for q, (row, ev) in enumerate(zip(xsi, e)):
        assert row[:3] == [a * c_ref for a in ev]
        assert row[3] == (1 / 18 if q < 6 else (1 / 36 if q < 18 else 1 / 3))
    assert out["vec"] == "1.5 -2 0.0030000000000000001"
    assert out["first"] == "1.25"  # first match wins
    assert out["time"] == "7.5"


def test_missing_keyword_and_refinement_exit_like_the_reference(probe, tmp_path):
    txt = bc_text(20, 20, 3)
    f = tmp_path / "bc.txt"
    f.write_text(txt.replace("Coupling-Approach STD\n", ""))
    r = _run(probe, f)
    assert r.returncode == 1 and 'No Value for keyword "Coupling-Approach"' in r.stderr
    r = _run(probe, tmp_path / "absent.txt")
    assert r.returncode == 1 and "Error reading boundary conditions file" in r.stderr
    # MaxLevel 1 on a 20 x 20 x 3 grid with the reference's layers (4, 75):
    # the fine shell leaves no coarse core
    f.write_text(txt.replace("MaxLevel 0", "MaxLevel 1"))
    r = _run(probe, f, "lattice")
    assert r.returncode == 1 and "shell geometry" in r.stderr and "core" in r.stderr


def test_keyword_readers_through_the_c_abi(tmp_path):
    f = tmp_path / "k.txt"
    f.write_text("Alpha 2.5 //!note\nName first second\nPath  /a b/c d \nAlpha 7\n")
    L = acp.lib()
    import ctypes
    v = ctypes.c_double()
    assert L.acp_control_read_value(str(f).encode(), b"Alpha", ctypes.byref(v)) == 0 and v.value == 2.5
    buf = ctypes.create_string_buffer(64)
    assert L.acp_control_read_string(str(f).encode(), b"Name", buf, 64) == 0 and buf.value == b"first"
    assert L.acp_control_read_filename(str(f).encode(), b"Path", buf, 64) == 0 and buf.value == b"/a b/c d "
    assert L.acp_control_read_string(str(f).encode(), b"Name", buf, 3) == acp._capi.ACP_E_ARG
    assert L.acp_control_read_value(str(f).encode(), b"Beta", ctypes.byref(v)) == acp._capi.ACP_E_CONFIG


def test_reference_two_level_geometry_refusals_name_their_reason(tmp_path):
    """The reference's own MaxLevel 1 geometry (a fine shell around a coarse
    core, Lattice.cpp:6-213, selected by its keys alone): layers that leave no
    core, layers under 2 and Overlap 0 (the direct coupling, not available on
    the shell) end the Lattice constructor with status 1 and the reason,
    before it touches a device."""
    exe = os.path.join(ROOT, "tests", "_bin", "lattice_check")
    assert os.path.exists(exe), "tests/_bin/lattice_check is not built (__graft_entry__.build())"
    f = tmp_path / "bc.txt"
    base = bc_text(20, 20, 7, op="HRR").replace("MaxLevel 0", "MaxLevel 1")
    for text, why in ((base, "no coarse core"),
                      (base.replace("RefinementLayerX 75", "RefinementLayerX 6").replace("RefinementLayerYZ 4",
                                                                                          "RefinementLayerYZ 1"),
                       "at least 2"),
                      (base.replace("RefinementLayerX 75", "RefinementLayerX 6").replace("RefinementLayerYZ 4",
                                                                                          "RefinementLayerYZ 2")
                       .replace("Overlap 1", "Overlap 0"), "Overlap 0")):
        f.write_text(text)
        r = subprocess.run([exe, str(f), "create"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 1, text
        assert why in r.stderr, r.stderr
