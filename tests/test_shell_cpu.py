"""The reference's two-level geometry in the oracle (CPU): a fine shell around
a coarse core (MaxLevel 1, createLattice, Lattice.cpp:6-213).

The coupling arithmetic is in the absent GridCoupling.cpp, so the oracle's
restatement is pinned by what the geometry and the scheme fix:
  - the grid-coupling nodes are the surfaces createLattice names (the hole box
    [2 RefinementLayerX, 2 (imax - RefinementLayerYZ)] x ... on level 1, the
    core box one coarse cell outside it on level 0), counted against the
    closed form of a box surface, and the missing links are those
    setConnectivity leaves without a neighbour (the links arriving from the
    hole, Lattice.cpp:237-1394);
  - coarse to fine and fine to coarse alone write exactly their link sets at
    exactly the grid-coupling and restricted nodes;
  - a uniform state stays uniform; the handed strain rates are exact on a
    linear velocity field (edges and corners of the interface included);
  - a pulse in the core reaches a probe exactly as on the uniform coarse grid
    until the interface's echo can arrive, and the echo is small.
"""
import numpy as np
import pytest

import oracle_lib as ol

E = np.array([(1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 0, 1), (1, 0, -1), (0, 1, 1),
              (0, 1, -1), (-1, 0, 1), (-1, 0, -1), (0, -1, 1), (0, -1, -1), (-1, 1, 0), (1, -1, 0), (1, 1, 0),
              (-1, -1, 0), (0, 0, 0)])


def _shell(op="SRT", n=(21, 17, 13), layers=(3, 6), cubic=False, tau=0.6, **kw):
    p = ol.make_params(*n, op=op, periodic=False, tau=tau, cubic=cubic, mrt_mode=1, nthreads=4, **kw)
    return ol.OracleRefined(p, 0, 0, shell=layers), p


def _surface(a, b, c):
    return a * b * c - max(a - 2, 0) * max(b - 2, 0) * max(c - 2, 0)


def test_grid_coupling_nodes_are_the_createlattice_surfaces():
    r, p = _shell()
    nx, ny, nz = p.nx, p.ny, p.nz
    lyz, lx = 3, 6
    L = r.shell_lists()
    nfs, nfc, ncr, ncs = L["counts"]
    imax, jmax, kmax = nx - 1, ny - 1, nz - 1
    # level 1 (Lattice.cpp:125-213): the surface of [2 lx, 2 (imax - lyz)] x [2 lyz, 2 (jmax - lyz)] x ...
    hole = (2 * (imax - lyz) - 2 * lx + 1, 2 * (jmax - lyz) - 2 * lyz + 1, 2 * (kmax - lyz) - 2 * lyz + 1)
    assert nfs == _surface(*hole)
    assert nfc == _surface(*[(h + 1) // 2 for h in hole])
    # level 0 (Lattice.cpp:38-108, Overlap 1): the surface of [lx - 1, imax - lyz + 1] x ...
    core = (imax - lyz + 1 - (lx - 1) + 1, jmax - lyz + 1 - (lyz - 1) + 1, kmax - lyz + 1 - (lyz - 1) + 1)
    assert ncs == _surface(*core)
    assert ncr == ncs + ((nx - 2) * (ny - 2) * (nz - 2) - np.prod(core))
    fnx, fny = 2 * nx - 1, 2 * ny - 1
    fs = L["fs"]
    i, j, k = fs % fnx, (fs // fnx) % fny, fs // (fnx * fny)
    # setConnectivity leaves a level-1 grid-coupling node without the
    # neighbours strictly inside the hole (the "front" node its +x one; a
    # "front_bottom" edge node keeps it, a bottom face node): a front face
    # node two or more nodes from the face's edges misses exactly the links
    # arriving with e_x = -1
    r.set_coupling("STD", "Mf", "cAll", "hMf")
    L = r.shell_lists()
    front = np.where((i == 2 * lx) & (j > 2 * lyz + 1) & (j < 2 * (jmax - lyz) - 1) & (k > 2 * lyz + 1)
                     & (k < 2 * (kmax - lyz) - 1))[0]
    assert front.size > 0
    assert np.all(L["fs_c2f"][front] == sum(1 << q for q in range(19) if E[q][0] == -1))
    # next to the bottom edge the diagonal link from the bottom face (a
    # stepped grid-coupling node) is not missing
    near = np.where((i == 2 * lx) & (j == 2 * (jmax - lyz) // 2 * 1) & (k == 2 * lyz + 1))[0]
    assert near.size == 1
    assert L["fs_c2f"][near[0]] == sum(1 << q for q in range(19) if E[q][0] == -1 and E[q][2] != 1)
    # a corner has no neighbour strictly inside the hole (D3Q19 has no
    # three-axis diagonal): nothing is missing; MfP hands it the links
    # arriving along the surface
    corner = np.where((i == 2 * lx) & (j == 2 * lyz) & (k == 2 * lyz))[0]
    assert corner.size == 1
    assert L["fs_c2f"][corner[0]] == 0
    r.set_coupling("STD", "MfP", "cAll", "hMf")
    L = r.shell_lists()
    assert L["fs_c2f"][corner[0]] == sum(1 << q for q in range(19) if (E[q] <= 0).all())


@pytest.mark.parametrize("c2f,hang", [("Mf", "hMf"), ("MfP", "hMfP"), ("fAll", "hAll"), ("Mf", "hAll")])
def test_coarse_to_fine_writes_exactly_its_links(c2f, hang):
    r, p = _shell()
    r.set_coupling("STD", c2f, "cAll", hang)
    c, f = r.levels
    rng = np.random.default_rng(1)
    c.set_state(p.rho0 * (1 + 0.1 * rng.uniform(size=(19, c.N))) / 19)
    f.set_state(np.zeros((19, f.N)))
    r.prolong(0)
    F = f.state()
    L = r.shell_lists()
    fs, nfc = L["fs"], L["counts"][1]
    others = np.ones(f.N, bool)
    others[fs] = False
    assert not np.any(F[:, others])
    bits = (L["fs_c2f"][:, None] >> np.arange(19)) & 1
    hbits = (L["fs_hang"][:, None] >> np.arange(19)) & 1
    got = (F[:, fs] != 0).T
    assert np.array_equal(got[:nfc], bits[:nfc].astype(bool))
    # hanging nodes: their set, where one of their coincident neighbours along
    # the surface (the odd axes +-1) holds a value
    fnx, fny = f.p.nx, f.p.ny
    pos = {int(l): t for t, l in enumerate(fs)}
    stride = (1, fnx, fnx * fny)
    for t in range(nfc, len(fs)):
        l = int(fs[t])
        x = (l % fnx, (l // fnx) % fny, l // (fnx * fny))
        odd = [d for d in range(3) if x[d] & 1]
        nbs = [l + sa * stride[odd[0]] + (sb * stride[odd[1]] if len(odd) > 1 else 0)
               for sa in (-1, 1) for sb in ((-1, 1) if len(odd) > 1 else (0,))]
        held = np.zeros(19, bool)
        for nb in nbs:
            assert pos[nb] < nfc
            held |= bits[pos[nb]].astype(bool)
        assert np.array_equal(got[t], hbits[t].astype(bool) & held), t
    assert bits.sum() > 0


@pytest.mark.parametrize("f2c", ["Mc", "cAll"])
@pytest.mark.parametrize("filt", ["NOFI", "LAG", "TOU"])
def test_fine_to_coarse_writes_exactly_its_links(f2c, filt):
    r, p = _shell()
    r.set_coupling("STD", "fAll", f2c, "hAll")
    r.set_filter(filt)
    c, f = r.levels
    rng = np.random.default_rng(2)
    c.set_state(np.zeros((19, c.N)))
    f.set_state(p.rho0 * (1 + 0.1 * rng.uniform(size=(19, f.N))) / 19)
    r.restrict()
    C = c.state()
    L = r.shell_lists()
    cr, ncs = L["cr"], L["counts"][3]
    others = np.ones(c.N, bool)
    others[cr] = False
    assert not np.any(C[:, others])
    bits = ((L["cr_mask"][:, None] >> np.arange(19)) & 1).astype(bool)
    assert np.array_equal((C[:, cr] != 0).T, bits)
    if f2c == "Mc":
        # the core box's surface takes the links arriving from outside it
        nx, ny = p.nx, p.ny
        I, J, K = cr[:ncs] % nx, (cr[:ncs] // nx) % ny, cr[:ncs] // (nx * ny)
        lo, hi = np.array([5, 2, 2]), np.array([nx - 1 - 3 + 1, ny - 1 - 3 + 1, p.nz - 1 - 3 + 1])
        for t in range(0, ncs, 7):
            X = np.array([I[t], J[t], K[t]])
            up = X[None, :] - E
            outside = ~np.all((up >= lo) & (up <= hi), axis=1)
            assert np.array_equal(bits[t], outside), t
        assert np.all(bits[ncs:])
    else:
        assert np.all(bits)


@pytest.mark.parametrize("op,moving,links", [("SRT", True, ("fAll", "cAll", "hAll")), ("RR", True, ("Mf", "Mc", "hMf")),
                                             ("HRR", True, ("fAll", "cAll", "hAll")), ("MRT-RM", False, ("MfP", "Mc", "hMfP")),
                                             ("MRT-GS", False, ("fAll", "cAll", "hAll"))])
def test_uniform_state_stays_uniform(op, moving, links):
    r, p = _shell(op=op)
    r.set_coupling("STD", *links)
    u = np.array([0.02 * p.cs, -0.01 * p.cs, 0.005 * p.cs]) if moving else np.zeros(3)
    if moving:
        # LODI faces drive a moving state towards rest; the interface is then
        # checked on a state at rest that the faces keep
        u = np.zeros(3)
    for lv in r.levels:
        lv.set_state(ol.equilibrium_field(lv.p, np.full(lv.N, p.rho0), np.repeat(u[:, None], lv.N, axis=1)))
    r.step(5)
    ref = r.levels[0].state()[:, 0]
    for lv in r.levels:
        assert np.max(np.abs(lv.state() - ref[:, None])) <= 1e-14 * np.max(np.abs(ref))


def _linear_field(lp, spacing0, U):
    """u = (a y, b z, c x) with positions in level-0 nodes (a divergence-free
    linear field), uniform density, as feq."""
    nx, ny, nz = lp.nx, lp.ny, lp.nz
    h = lp.spacing / spacing0
    x = np.tile(np.arange(nx), ny * nz) * h
    y = np.tile(np.repeat(np.arange(ny), nx), nz) * h
    z = np.repeat(np.arange(nz), nx * ny) * h
    a, b, c = U
    u = np.vstack([a * (y - 8), b * (z - 6), c * (x - 10)])
    return ol.equilibrium_field(lp, np.full(x.size, lp.rho0), u)


@pytest.mark.parametrize("half", [0, 1])
@pytest.mark.parametrize("cubic", [False, True])
def test_handed_strain_is_exact_on_a_linear_field(half, cubic):
    """Every handed strain rate (coarse partners of the fine grid-coupling
    nodes, their mean along the surface, fine partners of the core box's
    surface) is the exact sym(grad u) of a linear field on both levels, at
    faces, edges and corners of the interface."""
    r, p = _shell(op="HRR", cubic=cubic)
    g = 1e-3 * p.cs
    U = (g, -0.5 * g, 0.7 * g)
    for lv in r.levels:
        lv.set_state(_linear_field(lv.p, p.spacing, U))
    L = r.shell_lists()
    nfs, ncs = L["counts"][0], L["counts"][3]
    nv = 25 if cubic else 6
    face, gc = np.full(nv * nfs, np.nan), np.full(nv * ncs, np.nan)
    import ctypes
    D = ctypes.POINTER(ctypes.c_double)
    ol.lib().acpo_refined_transfer.argtypes = [ctypes.c_void_p, ctypes.c_int, D, D, D]
    ol.lib().acpo_refined_transfer(r.h, half, None, gc.ctypes.data_as(D), face.ctypes.data_as(D))
    a, b, c = U
    d = p.spacing
    want = np.array([0., 0., 0., 0.5 * a / d, 0.5 * c / d, 0.5 * b / d])  # xx yy zz xy xz yz
    planes = [face.reshape(nv, nfs)] + ([gc.reshape(nv, ncs)] if half == 0 else [])
    for P in planes:
        for v in range(6):
            assert np.allclose(P[v], want[v], rtol=1e-9, atol=1e-12 * g / d), (v, P[v].min(), P[v].max())


def _pulse3d(lp, spacing0, amp, c, b):
    """A Gaussian density pulse at rest, centre c and half-width b in level-0
    nodes, as feq."""
    h = lp.spacing / spacing0
    x = np.tile(np.arange(lp.nx), lp.ny * lp.nz) * h
    y = np.tile(np.repeat(np.arange(lp.ny), lp.nx), lp.nz) * h
    z = np.repeat(np.arange(lp.nz), lp.nx * lp.ny) * h
    r2 = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2
    rho = lp.rho0 * (1 + amp * np.exp(-np.log(2) * r2 / b ** 2))
    return ol.equilibrium_field(lp, rho, np.zeros((3, x.size)))


@pytest.mark.parametrize("op,links", [("SRT", ("fAll", "cAll", "hAll")), ("HRR", ("fAll", "cAll", "hAll")),
                                      ("RR", ("Mf", "Mc", "hMf"))])
def test_pulse_in_the_core_reflects_little(op, links):
    """A 3D pulse at the centre of the core (41^3 coarse nodes, the fine shell
    10 nodes deep at the low-x end and 4 elsewhere), probed 6 nodes towards
    the low-x interface, against the same pulse on the uniform coarse grid
    with LODI faces: identical until the interface's echo can arrive (the
    pulse's leading edge needs about 9 steps to the interface at x = 10 and 7
    more back to the probe), then a reflection of a few per cent at most."""
    n, layers = (41, 41, 41), (4, 10)
    r, p = _shell(op=op, n=n, layers=layers, tau=0.55, sigma=0.5)
    r.set_coupling("STD", *links)
    amp, c, b = 1e-3, (20.0, 20.0, 20.0), 2.5  # its tail at the interface: 1.5e-5 of amp
    for lv in r.levels:
        lv.set_state(_pulse3d(lv.p, p.spacing, amp, c, b))
    u = ol.OracleCase(p)
    u.set_state(_pulse3d(p, p.spacing, amp, c, b))
    probe = (14, 20, 20)
    lc = probe[0] + n[0] * (probe[1] + n[1] * probe[2])
    pr, pu = [], []
    for _ in range(36):
        r.step(1)
        pr.append(r.levels[0].state()[:, lc].sum())
        pu.append(u.step(1, probes=[probe])[0, 0])
    pr, pu = np.array(pr), np.array(pu)
    sig = np.max(np.abs(pu - p.rho0))
    assert sig > 0.02 * amp * p.rho0
    early = np.max(np.abs(pr[:12] - pu[:12])) / sig  # the leading edge (2 b ahead) echoes after about 13 steps
    late = np.max(np.abs(pr - pu)) / sig
    assert early < 1e-3, early
    assert late < 0.08, late


def _hole_interior(lp, layers, n, shrink=0):
    """Mask of the level-1 nodes strictly inside the hole box (2 RefinementLayerX,
    2 (imax - RefinementLayerYZ)) x ..., which the fine level leaves out (the
    box shrunk by `shrink` nodes on every side)."""
    lyz, lx = layers
    b = (2 * lx, 2 * (n[0] - 1 - lyz), 2 * lyz, 2 * (n[1] - 1 - lyz), 2 * lyz, 2 * (n[2] - 1 - lyz))
    b = [v + (shrink if a % 2 == 0 else -shrink) for a, v in enumerate(b)]
    i = np.tile(np.arange(lp.nx), lp.ny * lp.nz)
    j = np.tile(np.repeat(np.arange(lp.ny), lp.nx), lp.nz)
    k = np.repeat(np.arange(lp.nz), lp.nx * lp.ny)
    return (i > b[0]) & (i < b[1]) & (j > b[2]) & (j < b[3]) & (k > b[4]) & (k < b[5])


@pytest.mark.parametrize("op,order,cubic,links,filt", [("SRT", 3, False, ("fAll", "cAll", "hAll"), "NOFI"),
                                                       ("HRR", 3, False, ("Mf", "Mc", "hMf"), "LAG"),
                                                       ("RR", 4, True, ("MfP", "cAll", "hMfP"), "TOU"),
                                                       ("MRT-RM", 3, False, ("Mf", "Mc", "hAll"), "NOFI")])
def test_nothing_reads_the_hole(op, order, cubic, links, filt):
    """The fine nodes strictly inside the hole box are not stepped, and no
    coupling, restriction or stencil reads them: NaN planted there (in both
    population arrays) stays there (the box surface pushes into the first
    layer), and every other node of both levels is, bit for bit, what it is
    with ordinary values in the hole."""
    n, layers = (21, 17, 13), (3, 6)
    runs = []
    for plant in (0.0, np.nan):
        r, p = _shell(op=op, n=n, layers=layers, cubic=cubic, order=order, transverse=op == "HRR")
        r.set_coupling("STD", *links)
        r.set_filter(filt)
        for lv in r.levels:
            f = _pulse3d(lv.p, p.spacing, 1e-3, (9.0, 8.0, 6.0), 2.0)
            if lv is r.levels[1]:
                hole = _hole_interior(lv.p, layers, n)
                f[:, hole] = plant if np.isnan(plant) else f[:, hole] * 1.01
            lv.set_state(f)
        r.step(7)
        runs.append([lv.state() for lv in r.levels])
    (c0, f0), (c1, f1) = runs
    assert np.array_equal(c0, c1) and not np.isnan(c1).any()
    assert np.all(np.isnan(f1[:, _hole_interior(r.levels[1].p, layers, n, shrink=1)]))
    assert np.array_equal(f0[:, ~hole], f1[:, ~hole]) and not np.isnan(f1[:, ~hole]).any()
