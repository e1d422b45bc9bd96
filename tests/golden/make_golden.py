#!/usr/bin/env python3
"""Writes the golden fixtures under tests/golden/ from the reference solver.

Runs oracle/_ref/acp_ref_driver (built by `make -C oracle ref` from the
reference's own sources) on small cases and stores its outputs:

  geom_<nx>x<ny>x<nz>.npz   connectivity of Lattice::setConnectivity
  kat_o<N>.npz              per-function known-answer vectors
  kat_partner.npz           known answers of the two-level strain-rate partner
                            functions (Node.cpp:1892-2490)
  step_<name>.npz           initial populations, final populations and probe
                            series after whole uniform steps
  big_<name>.npz            larger runs: probe series, final density plane and
                            the SHA-256 of the final population array

Only this container has the reference; the fixtures are what travels.
"""
import argparse
import hashlib
import json
import math
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle_lib as ol  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "acp_ref_driver")

CS = 347.3
DX = 0.01
RHO0 = 1.17621


def nu_for_tau(tau, cs=CS, dx=DX):
    dt = dx / (cs * math.sqrt(3.0))
    return (tau - 0.5) * cs * cs * dt


def bc_text(nx, ny, nz, op="SRT", order=3, cubic=False, transverse=False, tau=0.6, mach=0.0, k1=1.0,
            sigma=0.98, threads=8, spacing=DX, nu=None, steps=10, nonreflecting=True):
    if nu is None:
        nu = nu_for_tau(tau, dx=spacing)
    yn = lambda b: "yes" if b else "no"  # noqa: E731
    return "\n".join([
        f"Max-Time-Step {steps}", "WriteInterval 1000000", "Save-as /tmp/acp_golden", f"Threads {threads}",
        f"Density {RHO0!r}", f"KinematicViscosity {nu!r}", f"SpeedOfSound {CS!r}", f"MACH-Number {mach!r}",
        f"Order-Of-Equilibrium {order}", f"Collision-Model {op}", f"Hybrid-Parameter {sigma!r}",
        f"Cubic-Mach-Correction {yn(cubic)}", "Coupling-Approach STD", f"Non-Reflecting {yn(nonreflecting)}",
        f"Transverse {yn(transverse)}", f"K1-Relaxation {k1!r}", "K2-Relaxation 1.0", "F2C-Filter NOFI",
        "Coarse-to-Fine fAll", "Fine-to-Coarse cAll", "Hanging-Node-Reconstruction hAll", "Scaling-Style SRTSTYLE",
        f"Domain-Size-X {(nx - 0.5) * spacing!r}", f"Domain-Size-Y {(ny - 0.5) * spacing!r}",
        f"Domain-Size-Z {(nz - 0.5) * spacing!r}", f"Spacing {spacing!r}", "MaxLevel 0", "RefinementLayerYZ 4",
        "RefinementLayerX 75", "Overlap 1", ""])


def run_driver(args, cwd):
    r = subprocess.run([DRIVER] + args, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver failed: {args}\n{r.stderr}")
    return r


def params_of(spec):
    return ol.make_params(spec["nx"], spec["ny"], spec["nz"], op=spec["op"], order=spec["order"],
                          cubic=spec.get("cubic", False), transverse=spec.get("transverse", False),
                          tau=spec.get("tau", 0.6), mach=spec.get("mach", 0.0), k1=spec.get("k1", 1.0),
                          sigma=spec.get("sigma", 0.98), cs=CS, spacing=spec.get("spacing", DX),
                          rho0=RHO0, nu=spec.get("nu"))


def random_ic(spec, seed):
    """rho = rho0 (1 + a U), u = b cs U, f = feq_3 + rho0 w n U with U ~ uniform(-1, 1)."""
    p = params_of(spec)
    N = p.nx * p.ny * p.nz
    rng = np.random.default_rng(seed)
    rho = RHO0 * (1.0 + 0.01 * rng.uniform(-1, 1, N))
    u = 0.05 * CS * rng.uniform(-1, 1, (3, N))
    p3 = params_of(dict(spec, order=3))
    f = ol.equilibrium_field(p3, rho, u)
    w = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])
    f += RHO0 * w[:, None] * 0.003 * rng.uniform(-1, 1, (19, N))
    return np.ascontiguousarray(f)


def spec_json(spec):
    return json.dumps(spec, sort_keys=True)


def make_step_case(name, spec, nsteps, seed, tmp, outdir, probes=((2, 2, 1),)):
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    nx, ny, nz = spec["nx"], spec["ny"], spec["nz"]
    with open(os.path.join(d, "bc.txt"), "w") as fh:
        fh.write(bc_text(nx, ny, nz, op=spec["op"], order=spec["order"], cubic=spec.get("cubic", False),
                         transverse=spec.get("transverse", False), tau=spec.get("tau", 0.6),
                         mach=spec.get("mach", 0.0), k1=spec.get("k1", 1.0), sigma=spec.get("sigma", 0.98)))
    f0 = random_ic(spec, seed)
    f0.tofile(os.path.join(d, "ic.bin"))
    args = ["run", "bc.txt", str(nsteps), ".", "--ic", "ic.bin"]
    for pr in probes:
        args += ["--probe", ",".join(map(str, pr))]
    run_driver(args, d)
    state = np.fromfile(os.path.join(d, "state.bin")).reshape(19, nx * ny * nz)
    probe = np.fromfile(os.path.join(d, "probe.bin")).reshape(nsteps, len(probes))
    np.savez_compressed(os.path.join(outdir, f"step_{name}.npz"), spec=spec_json(spec), nsteps=nsteps,
                        probes=np.array(probes, dtype=np.int64), f0=f0, f=state, probe=probe)


def make_big_case(name, spec, nsteps, ic, tmp, outdir, probes, plane_k=1, spacing=DX, dom=None):
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    nx, ny, nz = spec["nx"], spec["ny"], spec["nz"]
    with open(os.path.join(d, "bc.txt"), "w") as fh:
        fh.write(bc_text(nx, ny, nz, op=spec["op"], order=spec["order"], cubic=spec.get("cubic", False),
                         transverse=spec.get("transverse", False), tau=spec.get("tau", 0.6),
                         mach=spec.get("mach", 0.0), k1=spec.get("k1", 1.0), sigma=spec.get("sigma", 0.98),
                         spacing=spacing, nu=spec.get("nu")))
    args = ["run", "bc.txt", str(nsteps), "."]
    if ic is not None:
        ic.tofile(os.path.join(d, "ic.bin"))
        args += ["--ic", "ic.bin"]
    for pr in probes:
        args += ["--probe", ",".join(map(str, pr))]
    run_driver(args, d)
    N = nx * ny * nz
    state = np.fromfile(os.path.join(d, "state.bin")).reshape(19, N)
    probe = np.fromfile(os.path.join(d, "probe.bin")).reshape(nsteps, len(probes))
    secs = open(os.path.join(d, "time.txt")).read().split()
    rho = state.sum(axis=0).reshape(nz, ny, nx)
    np.savez_compressed(os.path.join(outdir, f"big_{name}.npz"), spec=spec_json(spec), nsteps=nsteps,
                        probes=np.array(probes, dtype=np.int64), probe=probe, rho_plane=rho[plane_k],
                        plane_k=plane_k, sha256=hashlib.sha256(state.tobytes()).hexdigest(),
                        ref_seconds=float(secs[0]), ref_threads=int(secs[5]),
                        ic_kind="reference" if ic is None else "injected")


def make_vtk_case(name, spec, nsteps, seed, tmp, outdir, nonreflecting):
    """The reference's own writeResultsVTK output after nsteps (Lattice.cpp:4094-4349)."""
    d = os.path.join(tmp, name)
    os.makedirs(d, exist_ok=True)
    nx, ny, nz = spec["nx"], spec["ny"], spec["nz"]
    with open(os.path.join(d, "bc.txt"), "w") as fh:
        fh.write(bc_text(nx, ny, nz, op=spec["op"], order=spec["order"], cubic=spec.get("cubic", False),
                         transverse=spec.get("transverse", False), tau=spec.get("tau", 0.6),
                         nonreflecting=nonreflecting))
    f0 = random_ic(spec, seed)
    f0.tofile(os.path.join(d, "ic.bin"))
    run_driver(["vtk", "bc.txt", str(nsteps), "out", "--ic", "ic.bin"], d)
    txt = open(os.path.join(d, "out.vtk"), "rb").read()
    np.savez_compressed(os.path.join(outdir, f"vtk_{name}.npz"), spec=spec_json(spec), nsteps=nsteps, f0=f0,
                        nonreflecting=int(nonreflecting), vtk=np.frombuffer(txt, dtype=np.uint8))


def make_partner_kat(tmp, outdir, nsamples=48, seed=77):
    """Known answers of the strain-rate partner functions of the two-level
    step (Node.cpp:1892-2490) from the reference driver's `partner` mode."""
    d = os.path.join(tmp, "partner")
    os.makedirs(d)
    spec = dict(nx=8, ny=8, nz=8, op="HRR", order=3, tau=0.57)
    with open(os.path.join(d, "bc.txt"), "w") as fh:
        fh.write(bc_text(8, 8, 8, op="HRR", order=3, tau=0.57))
    run_driver(["partner", "bc.txt", ".", str(nsamples), str(seed)], d)
    arrs = {n[:-4]: np.fromfile(os.path.join(d, n)) for n in os.listdir(d) if n.endswith(".bin")}
    np.savez_compressed(os.path.join(outdir, "kat_partner.npz"), spec=spec_json(spec), nsamples=nsamples, **arrs)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=HERE)
    ap.add_argument("--skip-big", action="store_true")
    ap.add_argument("--only-vtk", action="store_true")
    ap.add_argument("--only-partner", action="store_true")
    a = ap.parse_args()
    if not os.path.exists(DRIVER):
        sys.exit("build the reference driver first: make -C oracle ref")
    out = a.out
    with tempfile.TemporaryDirectory() as tmp:
        make_partner_kat(tmp, out)
        print("partner kat", flush=True)
        if a.only_partner:
            return
        make_vtk_case("srt_q2d_nonrefl", dict(nx=9, ny=7, nz=3, op="SRT", order=3), 7, 101, tmp, out, True)
        make_vtk_case("hrr_3d_refl", dict(nx=7, ny=6, nz=5, op="HRR", order=3, transverse=True), 5, 102, tmp, out,
                      False)
        make_vtk_case("rr_cubic_3d_nonrefl", dict(nx=7, ny=6, nz=5, op="RR", order=3, cubic=True), 6, 103, tmp, out,
                      True)
        print("vtk cases", flush=True)
        if a.only_vtk:
            return
        # connectivity
        for (nx, ny, nz) in [(7, 6, 5), (9, 7, 3), (8, 3, 6), (4, 5, 4), (5, 4, 3), (6, 6, 6)]:
            d = os.path.join(tmp, f"g{nx}{ny}{nz}")
            os.makedirs(d)
            with open(os.path.join(d, "bc.txt"), "w") as fh:
                fh.write(bc_text(nx, ny, nz))
            run_driver(["geom", "bc.txt", "."], d)
            rows = np.fromfile(os.path.join(d, "geom.bin"), dtype=np.int64).reshape(-1, 27)
            np.savez_compressed(os.path.join(out, f"geom_{nx}x{ny}x{nz}.npz"), rows=rows, dims=np.array([nx, ny, nz]))
        # per-function vectors
        for order in (2, 3, 4):
            d = os.path.join(tmp, f"kat{order}")
            os.makedirs(d)
            with open(os.path.join(d, "bc.txt"), "w") as fh:
                fh.write(bc_text(7, 6, 5, op="HRR", order=order, tau=0.57, sigma=0.9))
            run_driver(["kat", "bc.txt", ".", "64", str(1000 + order)], d)
            arrs = {n[:-4]: np.fromfile(os.path.join(d, n)) for n in os.listdir(d) if n.endswith(".bin")}
            np.savez_compressed(os.path.join(out, f"kat_o{order}.npz"), order=order,
                                spec=spec_json(dict(nx=7, ny=6, nz=5, op="HRR", order=order, tau=0.57, sigma=0.9)),
                                **arrs)
        # whole-step cases
        cases = []
        g3 = dict(nx=7, ny=6, nz=5)
        for op in ("SRT", "RR", "HRR"):
            for order in (2, 3, 4):
                cases.append((f"{op.lower()}_o{order}_3d", dict(g3, op=op, order=order, transverse=(order == 3))))
        for op in ("MRT-GS", "MRT-RM"):
            cases.append((f"{op.lower().replace('-', '')}_o3_3d", dict(g3, op=op, order=3, transverse=True)))
        cases.append(("hrr_sigma05_o3_3d", dict(g3, op="HRR", order=3, sigma=0.5, tau=0.55)))
        cases.append(("rr_cubic_o3_3d", dict(g3, op="RR", order=3, cubic=True)))
        cases.append(("hrr_cubic_o3_3d", dict(g3, op="HRR", order=3, cubic=True, transverse=True)))
        cases.append(("hrr_cubic_o4_3d", dict(g3, op="HRR", order=4, cubic=True)))
        cases.append(("srt_mach_k1_3d", dict(g3, op="SRT", order=3, mach=0.1, k1=0.5, transverse=True)))
        q2 = dict(nx=9, ny=7, nz=3)
        for op in ("SRT", "RR", "HRR", "MRT-GS", "MRT-RM"):
            cases.append((f"{op.lower().replace('-', '')}_o3_q2d", dict(q2, op=op, order=3)))
        cases.append(("srt_o3_q2d_tr", dict(q2, op="SRT", order=3, transverse=True)))
        cases.append(("hrr_o4_q2d_tr", dict(q2, op="HRR", order=4, transverse=True)))
        cases.append(("srt_o3_thiny", dict(nx=8, ny=3, nz=6, op="SRT", order=3, transverse=True)))
        cases.append(("hrr_o3_thiny", dict(nx=8, ny=3, nz=6, op="HRR", order=3, transverse=True)))
        cases.append(("srt_o3_thinx", dict(nx=4, ny=6, nz=5, op="SRT", order=3, transverse=True)))
        for i, (name, spec) in enumerate(cases):
            make_step_case(name, spec, 12, 7 + i, tmp, out)
            print("step", name, flush=True)
        if not a.skip_big:
            # reference's own hard-coded pulse (Lattice.cpp:1415-1433) on a quasi-2D grid
            spec = dict(nx=64, ny=40, nz=3, op="SRT", order=3, tau=0.52, spacing=0.02)
            make_big_case("refpulse_srt", spec, 60, None, tmp, out, probes=((40, 20, 1), (10, 20, 1)), spacing=0.02)
            print("big refpulse", flush=True)
            # north-star config 1 restated with LODI faces (1R) and its HRR/transverse variant
            nu1 = 3e-4 * CS * DX / math.sqrt(3.0)
            for name, op, tr in (("cfg1r_srt", "SRT", False), ("cfg1r_hrr_tr", "HRR", True)):
                spec = dict(nx=200, ny=200, nz=3, op=op, order=3, transverse=tr, nu=nu1, sigma=0.98)
                p = params_of(spec)
                ic = ol.pulse_ic(p, 1e-3, 100.0, 100.0, 10.0)
                make_big_case(name, spec, 500, ic, tmp, out, probes=((150, 100, 1),))
                print("big", name, flush=True)
            # 3R: 48^3 MRT-RM as the reference runs it, random fluctuation IC
            spec = dict(nx=48, ny=48, nz=48, op="MRT-RM", order=3, tau=0.5 + 3e-4 * 10)
            p = params_of(spec)
            N = 48 ** 3
            rng = np.random.default_rng(0)
            rho = RHO0 * (1.0 + 1e-3 * rng.standard_normal(N))
            u = 1e-3 * CS * math.sqrt(3.0) * rng.standard_normal((3, N))
            ic = ol.equilibrium_field(p, rho, u)
            make_big_case("cfg3r_mrtrm", spec, 50, ic, tmp, out, probes=((24, 24, 24), (1, 1, 1)), plane_k=24)
            print("big cfg3r", flush=True)


if __name__ == "__main__":
    main()
