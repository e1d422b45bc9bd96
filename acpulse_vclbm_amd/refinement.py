"""Interface reflection of the two-level refinement (acp_refined_*).

Case (the refined acoustic-pulse configuration): coarse 256 x 128 x 3 grid
with LODI non-reflecting x faces (classical, K1 = 1) and periodic y and z, RR
collision, N_eq = 3, tau = 0.5 + 3e-4, SI parameters of the reference's case
file, the direct coupling (DC) or the standard one (STD), F2C filter LAG or
NOFI; fine level (ratio 2) over the right half, coarse x in [128, 255], which
reaches the high-x face, so the fine level ends on a LODI face there and is
coupled only at the interface x = 128; Gaussian pulse
rho0 (1 + A exp(-ln2 r^2 / 8^2)) at (64, 64), A = 1e-3; density probe at
(96, 64, 1) on the coarse level; 2000 coarse steps.

Arrivals at the probe (c_s = 1/sqrt(3) nodes per step, half-width 8 nodes):
the direct wave after ~55 steps (32 nodes), the interface's echo after ~166
(96 nodes), the low-x face's echo after ~277 (160 nodes; the same face in the
refined and the uniform run), the high-x face's echo not before ~580 (350
nodes, less the pulse's width; the refined run meets a fine-level LODI face
there, the uniform run a coarse one).  R is taken over the first 560 steps,
where the refined and the uniform coarse run differ only by what the
interface does; `R_all_steps` over the whole run adds the difference of the
two high-x faces.  Reference signals: the same pulse on the uniform coarse
grid (same resolution as the probe's level: the difference is the
interface's effect) and on a uniform grid at the fine resolution
(511 x 256 x 6 with LODI x faces, probe at (192, 128, 2)).
R = max_t |p - p_ref| / max_t |p_ref - rho0|.

`geometry="periodic"` selects the earlier fully periodic form of the case:
fine level over [128, 256] on a periodic coarse grid, two interfaces, 260
steps (the window before the second interface's echo).
"""
import time

import numpy as np

TAU = 0.5 + 3e-4

# coarse boundary, fine region end, steps, R window, uniform fine grid (nx, boundary)
GEOMETRY = {
    "duct": dict(boundary="LODI-X", x1=255, steps=2000, window=560, fine_nx=511),
    "periodic": dict(boundary="PERIODIC", x1=256, steps=260, window=260, fine_nx=512),
}


def _refined_run(acp, d, x1, f2c_filter, precision, coupling, amp, steps, graph_chunk=0, scheme="STD"):
    """One refined pulse run: (probe series at (96, 64, 1), seconds of the
    stepping, fine-level node count).  With `graph_chunk` > 0 the coarse steps
    are replayed from captured graphs of that many steps."""
    with acp.Refined(d, 128, x1) as r:
        r.set_coupling(scheme)
        r.set_filter(f2c_filter)
        if precision != "F64":
            r.set_coupling_precision(coupling)
        if graph_chunk:
            r.enable_graphs(graph_chunk)
        r.init_pulse(amp, 64.0, 64.0, 8.0)
        c = r.level(0)
        c.set_probes([(96, 64, 1)], capacity=steps)
        c.sync()
        t0 = time.perf_counter()
        r.step(steps)
        c.sync()
        dt = time.perf_counter() - t0
        return c.read_probes()[:, 0].copy(), dt, r.level(1).N


def interface_reflection(acp, op="RR", steps=None, amp=1e-3, precision="F64", coupling="F64", with_fine=True,
                         f2c_filter="NOFI", geometry="duct", graph_chunk=0, scheme="STD"):
    """Reflection of the refined pulse case for one tier; `coupling` selects the
    coupling arithmetic of the shifted tiers ("F64" or "F32-SHIFTED"),
    `f2c_filter` the fine-to-coarse filter ("NOFI" or "LAG"), `scheme` the
    Coupling-Approach ("STD", or "DC1" for the direct coupling with the
    missing links exchanged, Mf / Mc / hMf) and `geometry` the coarse level's
    x faces ("duct": LODI, "periodic").  With `graph_chunk`
    > 0 the run is repeated with the coarse steps replayed from captured graphs
    of that many steps; its timing is reported beside the plain launches' and
    its probe series must equal theirs bit for bit."""
    g = GEOMETRY[geometry]
    steps = g["steps"] if steps is None else steps
    win = min(g["window"], steps)
    kw = dict(op=op, order=3, boundary=g["boundary"], mrt_mode="RELAXED", precision=precision)
    d = acp.make_desc(256, 128, 3, tau=TAU, **kw)
    p_ref, dt, fine_nodes = _refined_run(acp, d, g["x1"], f2c_filter, precision, coupling, amp, steps, scheme=scheme)
    rho0 = d.rho0
    with acp.Case(d) as u:
        u.init_pulse(amp, 64.0, 64.0, 8.0)
        u.set_probes([(96, 64, 1)], capacity=steps)
        u.step(steps)
        p_uc = u.read_probes()[:, 0].copy()
    sig_c = np.max(np.abs(p_uc[:win] - rho0))
    node_updates = 256 * 128 * 3 + 2 * fine_nodes
    out = {
        "geometry": geometry, "coarse_boundary": g["boundary"], "fine_region": [128, g["x1"]],
        "operator": op, "precision": precision, "coupling": coupling if precision != "F64" else "F64",
        "f2c_filter": f2c_filter, "coupling_approach": scheme,
        "steps": steps, "window": win, "amplitude": amp,
        "R_vs_uniform_coarse": float(np.max(np.abs(p_ref[:win] - p_uc[:win])) / sig_c),
        "R_all_steps": float(np.max(np.abs(p_ref - p_uc)) / sig_c),
        "early_max_diff": float(np.max(np.abs(p_ref[:100] - p_uc[:100])) / sig_c),
        "ms_per_coarse_step": dt / steps * 1e3,
        "node_updates_per_coarse_step": node_updates,
        "mlups": node_updates * steps / dt / 1e6,
        "probe": p_ref,
    }
    if graph_chunk:
        p_g, dt_g, _ = _refined_run(acp, d, g["x1"], f2c_filter, precision, coupling, amp, steps, graph_chunk, scheme)
        if not np.array_equal(p_g, p_ref):
            raise RuntimeError("refined graph replays differ from plain launches")
        out["graph_chunk"] = graph_chunk
        out["ms_per_coarse_step_graph"] = dt_g / steps * 1e3
        out["mlups_graph"] = node_updates * steps / dt_g / 1e6
    if with_fine:
        # uniform grid at the fine resolution: same viscosity at half the spacing, two steps per coarse step
        df = acp.make_desc(g["fine_nx"], 256, 6, nu=d.nu, spacing=d.spacing / 2.0, **kw)
        with acp.Case(df) as u:
            u.init_pulse(amp, 128.0, 128.0, 16.0)
            u.set_probes([(192, 128, 2)], capacity=2 * win)
            u.step(2 * win)
            p_uf = u.read_probes()[1::2, 0].copy()
        sig_f = np.max(np.abs(p_uf - rho0))
        out["R_vs_uniform_fine"] = float(np.max(np.abs(p_ref[:win] - p_uf)) / sig_f)
        out["uniform_coarse_vs_fine"] = float(np.max(np.abs(p_uc[:win] - p_uf)) / sig_f)
    return out


def restriction_ablation(acp, op="RR", f2c_filter="LAG", scheme="DC1"):
    """Refined pulse case (by default as specified: direct coupling, RR, LAG)
    in the shifted tiers with the coupling arithmetic in float64 and in
    float32, against the float64 refined run: probe-series L2 error and the
    relative error of the interface reflection coefficient."""
    ref = interface_reflection(acp, op, with_fine=False, f2c_filter=f2c_filter, scheme=scheme)
    p64, r64 = ref["probe"], ref["R_vs_uniform_coarse"]
    rho0 = acp.make_desc(8, 8, 8, tau=0.6).rho0
    sig = np.linalg.norm(p64 - rho0)
    out = {"operator": op, "f2c_filter": f2c_filter, "coupling_approach": scheme, "R_f64": r64}
    for prec in ("F32-SHIFTED", "F16-SHIFTED"):
        for coupling in ("F64", "F32-SHIFTED"):
            x = interface_reflection(acp, op, precision=prec, coupling=coupling, with_fine=False,
                                     f2c_filter=f2c_filter, scheme=scheme)
            out[f"{prec} coupling {coupling}"] = {
                "probe_l2": float(np.linalg.norm(x["probe"] - p64) / sig),
                "R": x["R_vs_uniform_coarse"],
                "R_rel": abs(x["R_vs_uniform_coarse"] - r64) / r64,
            }
    return out


def main():
    import argparse
    import json

    import acpulse_vclbm_amd as acp

    ap = argparse.ArgumentParser(description="interface reflection of the refined acoustic-pulse case")
    ap.add_argument("--op", default="RR")
    ap.add_argument("--precision", default="F64")
    ap.add_argument("--coupling", default="F64")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--f2c-filter", default="NOFI", choices=["NOFI", "LAG", "TOU"])
    ap.add_argument("--scheme", default="STD", choices=["STD", "DC1", "DC2"])
    ap.add_argument("--geometry", default="duct", choices=sorted(GEOMETRY))
    ap.add_argument("--graph-chunk", type=int, default=0, help="also time graph replays of this many coarse steps (even; 0: no)")
    a = ap.parse_args()
    r = interface_reflection(acp, op=a.op, steps=a.steps, precision=a.precision, coupling=a.coupling,
                             f2c_filter=a.f2c_filter, geometry=a.geometry, graph_chunk=a.graph_chunk,
                             scheme=a.scheme)
    r.pop("probe", None)
    print(json.dumps(r))


if __name__ == "__main__":
    main()
