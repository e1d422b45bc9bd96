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
