"""Host-side checks of libacpulse.so that need no GPU: it loads, exports every
function include/acp.h declares, reads case files with the reference's
semantics (Control.cpp:8-273), and builds the LODI schedule the reference's
pass order implies."""
import ctypes
import os

import numpy as np
import pytest

import acpulse_vclbm_amd as acp
from acpulse_vclbm_amd import _capi
import oracle_lib as ol
from make_golden import bc_text


def test_library_exports_every_header_symbol():
    L = acp.lib()
    names = _capi.header_symbols()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def _write(tmp_path, text, name="bc.txt"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def test_case_file_semantics(tmp_path):
    txt = bc_text(20, 12, 3, op="HRR", order=4, transverse=True, cubic=True, tau=0.55, mach=0.1, k1=0.5,
                  sigma=0.9)
    # commented-out alternatives before the real key, trailing comments and ';'
    txt = "//Collision-Model SRT\n#Order-Of-Equilibrium 2\n" + txt
    txt = txt.replace("Density 1.17621", "Density 1.17621;")
    txt = txt.replace("SpeedOfSound 347.3", "SpeedOfSound 347.3; //!air")
    txt = txt.replace("Save-as /tmp/acp_golden", "Save-as   /tmp/my results/case one")
    d = acp.read_case_file(_write(tmp_path, txt))
    assert (d.nx, d.ny, d.nz) == (20, 12, 3)
    assert d.op == _capi.OP["HRR"] and d.order == 4 and d.cubic == 1 and d.transverse == 1
    assert d.rho0 == 1.17621 and d.cs == 347.3 and abs(d.mach - 0.1) < 1e-15 and d.k1 == 0.5 and d.sigma == 0.9
    assert d.save_as.decode() == "/tmp/my results/case one"
    assert d.precision == _capi.PREC["F64"] and d.boundary == _capi.BOUNDARY["LODI"]
    c = acp.Control(_write(tmp_path, txt, "b2.txt"))
    assert c.getCollisionModel() == "HRR" and c.getNumberOfNodesX() == 20 and c.getTransverseInclusion() == "yes"


def test_first_match_wins(tmp_path):
    txt = bc_text(10, 10, 3) + "Collision-Model RR\nMax-Time-Step 7\n"
    d = acp.read_case_file(_write(tmp_path, txt))
    assert d.op == _capi.OP["SRT"] and d.timesteps == 10


def test_node_count_rounding(tmp_path):
    """numberOfNodes = ceil(L / dx + 1e-8) (Control.cpp:43-45)."""
    txt = bc_text(10, 10, 3).replace("Domain-Size-X 0.095", "Domain-Size-X 0.1")
    d = acp.read_case_file(_write(tmp_path, txt))
    assert d.nx == 11


def test_engine_keys(tmp_path):
    txt = bc_text(10, 10, 3) + "Precision F32-SHIFTED\nBoundary PERIODIC\nMRT-Mode RELAXED\nFP16-Scale 500\n"
    d = acp.read_case_file(_write(tmp_path, txt))
    assert d.precision == 1 and d.boundary == 1 and d.mrt_mode == 1 and d.f16_scale == 500.0


@pytest.mark.parametrize("bad,code", [
    ("Collision-Model SRT", _capi.ACP_E_CONFIG),          # key removed -> missing
    ("Order-Of-Equilibrium 3", _capi.ACP_E_CONFIG),
])
def test_missing_keys_are_errors(tmp_path, bad, code):
    txt = bc_text(10, 10, 3).replace(bad + "\n", "")
    with pytest.raises(acp.AcpError) as e:
        acp.read_case_file(_write(tmp_path, txt))
    assert e.value.code == code


def test_bad_values_are_errors(tmp_path):
    with pytest.raises(acp.AcpError):
        acp.read_case_file(_write(tmp_path, bc_text(10, 10, 3).replace("Order-Of-Equilibrium 3", "Order-Of-Equilibrium 5")))
    with pytest.raises(acp.AcpError):
        acp.read_case_file(_write(tmp_path, bc_text(10, 10, 3).replace("Collision-Model SRT", "Collision-Model XYZ")))
    with pytest.raises(acp.AcpError):
        acp.read_case_file(str(tmp_path / "does-not-exist.txt"))


@pytest.mark.parametrize("dims", [(9, 7, 3), (8, 3, 6), (7, 6, 5), (4, 5, 4), (3, 3, 3), (33, 17, 3)])
def test_lodi_schedule_matches_oracle(dims):
    """Same boundary nodes in the same phases as the oracle's schedule (which
    reproduces the reference bit for bit); phases in order, and within a phase
    the engine visits nodes in ascending index order (nodes of one phase are
    independent, so their order does not change the result)."""
    nodes, ph = acp.lodi_schedule(*dims)
    c = ol.OracleCase(ol.make_params(*dims))
    onode, oph = c.lodi_schedule()
    assert np.all(np.diff(ph) >= 0)
    assert dict(zip(nodes.tolist(), ph.tolist())) == dict(zip(onode.tolist(), oph.tolist()))
    for p in np.unique(ph):
        sel = nodes[ph == p]
        assert np.all(np.diff(sel) > 0)
        assert np.array_equal(sel, np.sort(onode[oph == p]))
    nx, ny, nz = dims
    assert len(nodes) == nx * ny * nz - (nx - 2) * (ny - 2) * (nz - 2)


def test_create_rejects_bad_descriptions():
    for kw, code in [(dict(nx=2), _capi.ACP_E_ARG), (dict(order=5), _capi.ACP_E_CONFIG)]:
        d = acp.make_desc(kw.get("nx", 8), 8, 8, tau=0.6)
        if "order" in kw:
            d.order = kw["order"]
        h = ctypes.c_void_p()
        rc = acp.lib().acp_case_create(ctypes.byref(d), None, ctypes.byref(h))
        assert rc == code and not h.value
    d = acp.make_desc(8, 8, 8, tau=0.6, precision="F32-SHIFTED", cubic=True)
    h = ctypes.c_void_p()
    assert acp.lib().acp_case_create(ctypes.byref(d), None, ctypes.byref(h)) == _capi.ACP_E_UNSUPPORTED
    d = acp.make_desc(8, 8, 8, tau=0.6)
    d.max_level = 1
    assert acp.lib().acp_case_create(ctypes.byref(d), None, ctypes.byref(h)) == _capi.ACP_E_UNSUPPORTED


def test_null_arguments_do_not_crash():
    L = acp.lib()
    assert L.acp_control_read(None, None) == _capi.ACP_E_ARG
    assert L.acp_case_step(None, 1) == _capi.ACP_E_ARG
    assert L.acp_case_info_get(None, None) == _capi.ACP_E_ARG
    L.acp_case_destroy(None)


@pytest.mark.parametrize("planes,elems,block", [(2, 100, 0), (19, 0, 0), (19, 100, 3), (19, 100, -8), (1, 100, 64)])
def test_copy_probe_rejects_bad_arguments(planes, elems, block):
    """Argument checks run before any device call: planes 1 or 19, positive
    sizes, blocks a power of two and only with 19 planes."""
    ms = ctypes.c_double()
    rc = acp.lib().acp_copy_probe(planes, elems, block, 1, None, ctypes.byref(ms))
    assert rc == _capi.ACP_E_ARG


def _pass_position(nx, ny, nz):
    """Position of every boundary node in the reference's sequential LODI pass
    (Lattice.cpp:1582-1624: x faces by (j, k), then z faces by (i, j), then y
    faces by (i, k), each face pair low then high)."""
    pos, t = {}, 0
    lin = lambda i, j, k: i + nx * (j + ny * k)
    for j in range(ny):
        for k in range(nz):
            for i in (0, nx - 1):
                pos[lin(i, j, k)] = t
                t += 1
    for i in range(1, nx - 1):
        for j in range(ny):
            for k in (0, nz - 1):
                pos[lin(i, j, k)] = t
                t += 1
    for i in range(1, nx - 1):
        for k in range(1, nz - 1):
            for j in (0, ny - 1):
                pos[lin(i, j, k)] = t
                t += 1
    return pos


@pytest.mark.parametrize("dims", [(3, 3, 3), (5, 4, 3), (7, 5, 3), (9, 7, 3), (4, 5, 4), (8, 3, 6), (7, 6, 5),
                                  (33, 17, 3), (12, 9, 4), (40, 30, 3)])
def test_lodi_launch_order_keeps_the_reference_reads(dims):
    """Every LODI update reads the B entries of its normal neighbours.  In the
    reference's sequential pass a read sees a boundary node's new entries when
    that node comes earlier in the pass, its old ones otherwise.  The engine
    runs launches one after another and the records of a launch in parallel.
    Each read must see the same version as in the reference and never race
    with a write in the same launch; each boundary node is updated exactly
    once."""
    nx, ny, nz = dims
    rec, starts = acp.lodi_launches(*dims)
    n = len(rec)
    assert n == nx * ny * nz - (nx - 2) * (ny - 2) * (nz - 2)
    pos = _pass_position(nx, ny, nz)
    assert sorted(pos) == sorted(rec["node"].tolist())
    assert starts[0] == 0 and starts[-1] == n and np.all(np.diff(starts) > 0)
    launch = {}  # node -> launch
    for p in range(len(starts) - 1):
        for r in range(int(starts[p]), int(starts[p + 1])):
            assert rec["node"][r] not in launch
            launch[int(rec["node"][r])] = p
    assert len(launch) == n
    for r in range(n):
        a = int(rec["node"][r])
        for x in (int(rec["n1"][r]), int(rec["n2"][r])):
            if x not in pos:
                continue  # an interior node: its B entries are final before the LODI pass
            assert launch[x] != launch[a], f"node {a} reads node {x} written in the same launch"
            assert (launch[x] < launch[a]) == (pos[x] < pos[a]), (a, x)


STRUCTS = [("acp_case_desc", _capi.CaseDesc), ("acp_case_info", _capi.CaseInfo), ("acp_server_desc", _capi.ServerDesc),
           ("acp_request", _capi.Request), ("acp_result", _capi.Result), ("acp_server_stats", _capi.ServerStats)]


def test_ctypes_structures_match_the_header(tmp_path):
    """Size and every field offset of the binding's structures against the C
    compiler's layout of include/acp.h (a mismatch would hand the library
    misplaced fields)."""
    import subprocess
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "acp.h"', "int main(void) {"]
    for cname, py in STRUCTS:
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'    printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{inc}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l}
    for cname, py in STRUCTS:
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
