"""LODI x faces on grids wrapped in y and z (oracle modes 3 and 4), CPU.

The reference has no periodic axes, so these modes have no reference output.
They are pinned to modes the reference fixtures do pin, through causality: a
node's state after K steps depends only on nodes within a few lattice spacings
per step (two for the push plus HRR's strain stencil, three for a LODI node's
second normal neighbour), so two grids that agree on a wide enough region
agree there bit for bit for K steps.

- Mode 3 (LODI on both x faces) against mode 0 (LODI on all six faces, the
  reference's geometry): the mode-0 grid holds several periods of a state
  periodic in y and z, and its central period must evolve exactly like the
  wrapped grid, x-face LODI nodes (with their transverse terms) included.
- Mode 4 (open face x = 0, LODI face x = nx - 1) against mode 2 (both faces
  open) near x = 0 and against mode 3 near x = nx - 1.
- A refined case whose fine level ends on a LODI face: that face relaxes over
  the level-0 domain length, with the fine spacing and time step, so it
  evolves exactly like the high-x face of a uniform fine grid whose own length
  equals that domain length.
"""
import numpy as np
import pytest

import oracle_lib as ol

W = np.array([1 / 18] * 6 + [1 / 36] * 12 + [1 / 3])


def _field(p, rng):
    """Random near-equilibrium state on the grid of p, as [19][nz][ny][nx]."""
    N = p.nx * p.ny * p.nz
    rho = p.rho0 * (1.0 + 0.01 * rng.uniform(-1, 1, N))
    u = 0.05 * p.cs * rng.uniform(-1, 1, (3, N))
    f = ol.equilibrium_field(p, rho, u)
    f = f + p.rho0 * W[:, None] * 0.003 * rng.uniform(-1, 1, (19, N))
    return f.reshape(19, p.nz, p.ny, p.nx)


CASES = [("SRT", 2, False, False), ("HRR", 3, True, False), ("RR", 4, True, True), ("MRT-RM", 3, True, False)]


@pytest.mark.parametrize("op,order,transverse,cubic", CASES, ids=[c[0] + f"-o{c[1]}" for c in CASES])
def test_lodi_x_matches_all_face_lodi_in_the_causal_cone(op, order, transverse, cubic):
    nx, ty, tz, steps = 10, 6, 3, 5
    my, mz = 11, 17  # periods in y and z of the all-face grid: margins of 30 and 24 nodes
    kw = dict(op=op, order=order, transverse=transverse, cubic=cubic, tau=0.56, mrt_mode=1, nthreads=4)
    pw = ol.make_params(nx, ty, tz, periodic=3, **kw)
    pa = ol.make_params(nx, ty * my, tz * mz, periodic=0, **kw)
    f = _field(pw, np.random.default_rng(7))
    wrapped = ol.OracleCase(pw)
    wrapped.set_state(f.reshape(19, -1))
    whole = ol.OracleCase(pa)
    whole.set_state(np.ascontiguousarray(np.tile(f, (1, mz, my, 1))).reshape(19, -1))
    wrapped.step(steps)
    whole.step(steps)
    a = wrapped.state().reshape(19, tz, ty, nx)
    b = whole.state().reshape(19, tz * mz, ty * my, nx)
    jc, kc = ty * (my // 2), tz * (mz // 2)
    centre = b[:, kc:kc + tz, jc:jc + ty, :]
    assert np.array_equal(a, centre), f"max |df| = {np.abs(a - centre).max():.3e}"
    # the LODI pass did act: the x faces differ from a plain periodic run
    per = ol.OracleCase(ol.make_params(nx, ty, tz, periodic=1, **kw))
    per.set_state(f.reshape(19, -1))
    per.step(steps)
    c = per.state().reshape(19, tz, ty, nx)
    assert not np.array_equal(a[..., 0], c[..., 0])


def test_lodi_x_schedule_is_the_x_faces_in_one_phase():
    p = ol.make_params(7, 5, 4, periodic=3, nthreads=1)
    node, ph = ol.OracleCase(p).lodi_schedule()
    want = sorted(i + 7 * (j + 5 * k) for k in range(4) for j in range(5) for i in (0, 6))
    assert sorted(node.tolist()) == want
    assert set(ph.tolist()) == {1}
    rows = ol.OracleCase(p).geometry()
    for l in want:
        i, j, k = l % 7, (l // 7) % 5, l // 35
        s = 1 if i == 0 else -1
        assert rows[l, 26] == 1  # boundary node: no collision
        assert rows[l, 18] == (i + s) + 7 * (j + 5 * k) and rows[l, 19] == (i + 2 * s) + 7 * (j + 5 * k)
        assert list(rows[l, 20:24]) == [i + 7 * ((j + 1) % 5 + 5 * k), i + 7 * ((j - 1) % 5 + 5 * k),
                                        i + 7 * (j + 5 * ((k + 1) % 4)), i + 7 * (j + 5 * ((k - 1) % 4))]
        assert rows[l, 24] == (0 if i == 0 else 2) and rows[l, 25] == 0
    # interior nodes next to the faces collide and push all 18 links, wrapped in y and z
    l = 1 + 7 * (0 + 5 * 0)
    assert rows[l, 26] == 0 and rows[l, 3] == 1 + 7 * (4 + 5 * 0) and rows[l, 2] == 0


def test_lodi_x_rejects_short_grids():
    with pytest.raises(ValueError):
        ol.OracleCase(ol.make_params(3, 5, 4, periodic=3, nthreads=1))


@pytest.mark.parametrize("op,order", [("SRT", 3), ("HRR", 3)])
def test_open_lodi_x_matches_open_and_lodi_faces_near_each(op, order):
    nx, ny, nz, steps = 44, 6, 3, 5
    kw = dict(op=op, order=order, transverse=True, tau=0.56, nthreads=4)
    f = _field(ol.make_params(nx, ny, nz, periodic=4, **kw), np.random.default_rng(3)).reshape(19, -1)
    out = {}
    for mode in (2, 3, 4):
        c = ol.OracleCase(ol.make_params(nx, ny, nz, periodic=mode, **kw))
        c.set_state(f)
        c.step(steps)
        out[mode] = c.state().reshape(19, nz, ny, nx)
    half = nx // 2  # 22 nodes from either face: beyond 3 nodes per step for 5 steps
    assert np.array_equal(out[4][..., :half], out[2][..., :half])
    assert np.array_equal(out[4][..., half:], out[3][..., half:])
    assert not np.array_equal(out[2][..., -1], out[3][..., -1])
    assert not np.array_equal(out[2][..., 0], out[3][..., 0])


@pytest.mark.parametrize("op", ["SRT", "HRR"])
def test_fine_lodi_face_uses_the_level0_domain_length(op):
    """Coarse 24 x 4 x 3 with LODI x faces, fine region [6, 23] reaching the
    high-x face.  The fine level (35 x 8 x 6) ends on a LODI face; against a
    uniform fine grid of 48 x 8 x 6 nodes (length 48 h/2, the coarse 24 h),
    aligned on that face and started from the same state, the nodes within 16
    of the face agree bit for bit after 2 coarse (4 fine) steps.  A fine level
    that took its own length (35 h/2) would differ at the face."""
    cnx, cny, cnz, x0 = 24, 4, 3, 6
    kw = dict(op=op, order=3, transverse=True, tau=0.56, nthreads=4)
    pc = ol.make_params(cnx, cny, cnz, periodic=3, **kw)
    r = ol.OracleRefined(pc, x0, cnx - 1)
    fp = r.levels[1].p
    assert fp.periodic == 4 and fp.nx == 35
    rng = np.random.default_rng(5)
    r.levels[0].set_state(_field(pc, rng).reshape(19, -1))
    ff = _field(fp, rng)
    r.levels[1].set_state(ff.reshape(19, -1))
    pu = ol.make_params(2 * cnx, fp.ny, fp.nz, periodic=3, spacing=fp.spacing, **kw)
    pu.nu = fp.nu
    u = ol.OracleCase(pu)
    fu = _field(pu, rng)
    off = pu.nx - fp.nx
    fu[..., off:] = ff
    u.set_state(np.ascontiguousarray(fu).reshape(19, -1))
    r.step(2)
    u.step(4)
    a = r.levels[1].state().reshape(19, fp.nz, fp.ny, fp.nx)[..., -16:]
    b = u.state().reshape(19, fp.nz, fp.ny, pu.nx)[..., -16:]
    assert np.array_equal(a, b), f"max |df| = {np.abs(a - b).max():.3e}"
    # the relaxation term matters at the face: its own length would change it
    pv = ol.make_params(fp.nx, fp.ny, fp.nz, periodic=4, spacing=fp.spacing, **kw)
    pv.nu = fp.nu
    v = ol.OracleCase(pv)
    v.set_state(ff.reshape(19, -1))
    v.step(4)
    c = v.state().reshape(19, fp.nz, fp.ny, fp.nx)[..., -1]
    assert not np.array_equal(c, a[..., -1])


@pytest.mark.parametrize("x1", [17, 23])
@pytest.mark.parametrize("op,filt", [("SRT", "NOFI"), ("RR", "LAG"), ("HRR", "NOFI")])
def test_refined_lodi_x_keeps_rest_state(op, filt, x1):
    """Fluid at rest at rho0 is a fixed point of every part of a refined step
    with LODI x faces: the LODI characteristic update of a resting node at rho0
    is that node's own equilibrium."""
    pc = ol.make_params(24, 4, 3, op=op, periodic=3, tau=0.6, mrt_mode=1, nthreads=4, transverse=True)
    r = ol.OracleRefined(pc, 6, x1)
    r.set_filter(filt)
    for lv in r.levels:
        lv.set_state(ol.equilibrium_field(lv.p, np.full(lv.N, pc.rho0), np.zeros((3, lv.N))))
    ref = r.levels[0].state()[:, 0]
    r.step(5)
    for lv in r.levels:
        f = lv.state()
        assert np.max(np.abs(f - ref[:, None])) <= 1e-14 * np.max(np.abs(ref))


def test_refined_lodi_x_rejects_regions_on_the_faces():
    pc = ol.make_params(24, 4, 3, periodic=3, nthreads=1)
    for x0, x1 in [(0, 10), (5, 24), (5, 6)]:
        with pytest.raises(ValueError):
            ol.OracleRefined(pc, x0, x1)


def test_capi_lodi_x_footprint_and_checks():
    """The C ABI (no GPU): a LODI_X case accounts one 64-byte schedule record
    per x-face node (both faces; the high face only for OPEN_LODI_X), rejects
    grids shorter than 4 nodes in x and the cubic-Mach terms at an open face,
    and a refined case with a LODI_X coarse level takes regions from x0 >= 1
    to x1 <= nx - 1."""
    import acpulse_vclbm_amd as acp

    def fp(**kw):
        return acp.footprint(acp.make_desc(40, 24, 6, op="SRT", tau=0.56, **kw))

    rec = 64
    assert fp(boundary="LODI-X") - fp(boundary="PERIODIC") == -(-2 * 24 * 6 * rec // 256) * 256
    assert fp(boundary="OPEN-LODI-X") - fp(boundary="OPEN-X") == -(-24 * 6 * rec // 256) * 256
    with pytest.raises(acp.AcpError):
        acp.footprint(acp.make_desc(3, 24, 6, op="SRT", tau=0.56, boundary="LODI-X"))
    with pytest.raises(acp.AcpError):
        acp.footprint(acp.make_desc(40, 24, 6, op="RR", cubic=True, tau=0.56, boundary="OPEN-LODI-X",
                                    precision="F32-SHIFTED"))
    acp.footprint(acp.make_desc(40, 24, 6, op="RR", cubic=True, tau=0.56, boundary="OPEN-LODI-X"))
    acp.footprint(acp.make_desc(40, 24, 6, op="RR", cubic=True, tau=0.56, boundary="LODI-X"))
    d = acp.make_desc(40, 24, 6, op="SRT", tau=0.56, boundary="LODI-X")
    for good in [(1, 3), (8, 39), (20, 30)]:
        assert acp.refined_footprint(d, *good) > acp.footprint(d)
    for bad in [(0, 10), (8, 40), (8, 9)]:
        with pytest.raises(acp.AcpError):
            acp.refined_footprint(d, *bad)
