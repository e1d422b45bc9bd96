"""ctypes binding of the CPU oracle (oracle/_build/libacp_oracle.so).

Test infrastructure: the oracle is the checker for the HIP path and is only
loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "_build", "libacp_oracle.so")

OPS = {"SRT": 0, "MRT-GS": 1, "MRT-RM": 2, "RR": 3, "HRR": 4}


class Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("op", ctypes.c_int32), ("order", ctypes.c_int32), ("cubic", ctypes.c_int32),
        ("transverse", ctypes.c_int32), ("periodic", ctypes.c_int32), ("mrt_mode", ctypes.c_int32),
        ("nthreads", ctypes.c_int32),
        ("rho0", ctypes.c_double), ("nu", ctypes.c_double), ("cs", ctypes.c_double),
        ("spacing", ctypes.c_double), ("mach", ctypes.c_double), ("k1", ctypes.c_double),
        ("sigma", ctypes.c_double),
    ]


class Consts(ctypes.Structure):
    _fields_ = [
        ("c", ctypes.c_double), ("dt", ctypes.c_double), ("omega", ctypes.c_double),
        ("w", ctypes.c_double * 19), ("xi", (ctypes.c_double * 3) * 19),
        ("cs2", ctypes.c_double), ("rcs2", ctypes.c_double), ("rcs4", ctypes.c_double),
        ("rcs6", ctypes.c_double), ("rcs8", ctypes.c_double),
        ("hx", ctypes.c_double * 19), ("hy", ctypes.c_double * 19), ("hz", ctypes.c_double * 19),
        ("hxx", ctypes.c_double * 19), ("hyy", ctypes.c_double * 19), ("hzz", ctypes.c_double * 19),
        ("hxy", ctypes.c_double * 19), ("hxz", ctypes.c_double * 19), ("hyz", ctypes.c_double * 19),
        ("pxx", ctypes.c_double * 19), ("pyy", ctypes.c_double * 19), ("pzz", ctypes.c_double * 19),
        ("h3", (ctypes.c_double * 19) * 6), ("h4", (ctypes.c_double * 19) * 3),
        ("rec_cs4_strain", ctypes.c_double), ("inv_pre", ctypes.c_double),
        ("nu", ctypes.c_double), ("sigma", ctypes.c_double), ("rho0", ctypes.c_double),
        ("spacing", ctypes.c_double), ("cs", ctypes.c_double),
        ("lodi_cs2", ctypes.c_double), ("dom_len", ctypes.c_double * 3), ("ma", ctypes.c_double),
        ("ma2", ctypes.c_double), ("k1", ctypes.c_double),
        ("c2", ctypes.c_double), ("c3", ctypes.c_double), ("c4", ctypes.c_double),
        ("ix", ctypes.c_double), ("ix2", ctypes.c_double), ("ix3", ctypes.c_double), ("ix4", ctypes.c_double),
        ("gs_pre", (ctypes.c_double * 19) * 19),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library missing: run __graft_entry__.build() (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER(Params)
        K = ctypes.POINTER(Consts)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int64)
        L.acpo_consts_init.argtypes = [K, P]
        L.acpo_feq.argtypes = [K, ctypes.c_int, ctypes.c_double, dp, dp]
        L.acpo_moments.argtypes = [K, dp, dp, dp, dp]
        L.acpo_pi1.argtypes = [K, dp, dp, dp, dp]
        L.acpo_offeq.argtypes = [K, ctypes.c_int, dp, dp, dp]
        L.acpo_strain_psi.argtypes = [K, dp, dp, dp]
        for n in ("acpo_rm_moments", "acpo_gs_moments", "acpo_rm_inverse", "acpo_gs_inverse"):
            getattr(L, n).argtypes = [K, dp, dp]
        L.acpo_rm_feq4.argtypes = [K, ctypes.c_double, dp, dp]
        L.acpo_gs_eqmoments.argtypes = [ctypes.c_double] * 4 + [dp]
        L.acpo_collide.argtypes = [K, P, dp, dp, dp]
        L.acpo_create.argtypes = [P]
        L.acpo_create.restype = ctypes.c_void_p
        L.acpo_destroy.argtypes = [ctypes.c_void_p]
        L.acpo_set_state.argtypes = [ctypes.c_void_p, dp]
        L.acpo_get_state.argtypes = [ctypes.c_void_p, dp]
        L.acpo_get_psi.argtypes = [ctypes.c_void_p, dp]
        L.acpo_get_prev.argtypes = [ctypes.c_void_p, dp]
        L.acpo_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, ctypes.c_int, dp]
        L.acpo_geometry.argtypes = [ctypes.c_void_p, ip]
        L.acpo_lodi_count.argtypes = [ctypes.c_void_p]
        L.acpo_lodi_count.restype = ctypes.c_int64
        L.acpo_lodi_schedule.argtypes = [ctypes.c_void_p, ip, ctypes.POINTER(ctypes.c_int32)]
        L.acpo_reference_ic.argtypes = [P, ctypes.c_double, ctypes.c_double, dp]
        L.acpo_reference_ic_level.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, dp]
        L.acpo_equilibrium_field.argtypes = [P, ctypes.c_int64, dp, dp, dp]
        L.acpo_pulse_ic.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, dp]
        L.acpo_refined_create.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.acpo_refined_create.restype = ctypes.c_void_p
        L.acpo_refined_create_shell.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.acpo_refined_create_shell.restype = ctypes.c_void_p
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.acpo_refined_shell_lists.argtypes = [ctypes.c_void_p, ip, ip, u32p, u32p, u32p, ip]
        L.acpo_refined_shell_lists.restype = ctypes.c_int64
        L.acpo_refined_destroy.argtypes = [ctypes.c_void_p]
        L.acpo_refined_level.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.acpo_refined_level.restype = ctypes.c_void_p
        L.acpo_refined_scale.argtypes = [ctypes.c_void_p]
        L.acpo_refined_scale.restype = ctypes.c_double
        L.acpo_refined_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.acpo_refined_step_probes.argtypes = [ctypes.c_void_p, ctypes.c_int, ip, ctypes.c_int, dp]
        L.acpo_refined_set_filter.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.acpo_refined_restrict.argtypes = [ctypes.c_void_p]
        L.acpo_refined_prolong.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.acpo_refined_set_coupling.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4
        L.acpo_refined_set_scaling.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.acpo_refined_scale_neq.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        _lib = L
    return _lib


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(nx, ny, nz, op="SRT", order=3, cubic=False, transverse=False, periodic=False,
                mrt_mode=0, nthreads=8, rho0=1.17621, nu=None, tau=0.6, cs=347.3, spacing=0.01,
                mach=0.0, k1=1.0, sigma=0.98):
    if nu is None:
        dt = spacing / (cs * 3 ** 0.5)
        nu = (tau - 0.5) * cs * cs * dt
    return Params(nx, ny, nz, OPS[op] if isinstance(op, str) else op, order, int(cubic), int(transverse),
                  int(periodic), int(mrt_mode), nthreads, rho0, nu, cs, spacing, mach, k1, sigma)


def consts(p):
    k = Consts()
    rc = lib().acpo_consts_init(ctypes.byref(k), ctypes.byref(p))
    assert rc == 0
    return k


class OracleCase:
    """Whole-lattice oracle run: reference step semantics on a [19][N] state."""

    def __init__(self, p):
        self.p = p
        self.N = p.nx * p.ny * p.nz
        self.h = lib().acpo_create(ctypes.byref(p))
        if not self.h:
            raise ValueError("oracle rejected the case parameters")

    def __del__(self):
        if getattr(self, "h", None):
            lib().acpo_destroy(self.h)
            self.h = None

    def set_state(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(19, self.N)
        lib().acpo_set_state(self.h, dptr(f))

    def state(self):
        f = np.empty((19, self.N))
        lib().acpo_get_state(self.h, dptr(f))
        return f

    def prev(self):
        f = np.empty((19, self.N))
        lib().acpo_get_prev(self.h, dptr(f))
        return f

    def psi(self):
        f = np.empty((19, self.N))
        lib().acpo_get_psi(self.h, dptr(f))
        return f

    def step(self, n, probes=()):
        idx = np.array([i + self.p.nx * (j + self.p.ny * k) for (i, j, k) in probes] or [0], dtype=np.int64)
        out = np.zeros(max(1, n * len(probes)))
        lib().acpo_step(self.h, n, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(probes), dptr(out))
        return out[: n * len(probes)].reshape(n, len(probes)) if probes else None

    def geometry(self):
        rows = np.empty((self.N, 27), dtype=np.int64)
        lib().acpo_geometry(self.h, rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))
        return rows

    def lodi_schedule(self):
        n = lib().acpo_lodi_count(self.h)
        node = np.empty(n, dtype=np.int64)
        ph = np.empty(n, dtype=np.int32)
        lib().acpo_lodi_schedule(self.h, node.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                 ph.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
        return node, ph


class _Level(OracleCase):
    """A level of an OracleRefined (owned by it)."""

    def __init__(self, h, p):
        self.p = p
        self.N = p.nx * p.ny * p.nz
        self.h = h

    def __del__(self):
        pass


# link sets of the coupling keys: 0 missing, 1 missing and common, 2 all
LINKS = {"Mf": 0, "Mc": 0, "hMf": 0, "MfP": 1, "McP": 1, "hMfP": 1, "fAll": 2, "cAll": 2, "hAll": 2}


class OracleRefined:
    """Two-level oracle run (coarse p periodic or with LODI x faces, fine over
    coarse x in [x0, x1])."""

    def __init__(self, p, x0, x1, shell=None):
        """shell = (layer_yz, layer_x): the reference's geometry, a fine shell
        around a coarse core (p with LODI faces); x0, x1 are then unused."""
        if shell is not None:
            self.h = lib().acpo_refined_create_shell(ctypes.byref(p), shell[0], shell[1])
            if not self.h:
                raise ValueError("oracle rejected the shell case")
            fp = Params.from_buffer_copy(p)
            fp.nx, fp.ny, fp.nz = 2 * p.nx - 1, 2 * p.ny - 1, 2 * p.nz - 1
            fp.spacing = p.spacing / 2.0
            self.levels = [_Level(lib().acpo_refined_level(self.h, 0), p),
                           _Level(lib().acpo_refined_level(self.h, 1), fp)]
            self.x0 = self.x1 = None
            return
        self.h = lib().acpo_refined_create(ctypes.byref(p), x0, x1)
        if not self.h:
            raise ValueError("oracle rejected the refined case")
        fp = Params.from_buffer_copy(p)
        fp.nx, fp.ny, fp.nz = 2 * (x1 - x0) + 1, 2 * p.ny, 2 * p.nz
        fp.spacing = p.spacing / 2.0
        fp.periodic = 4 if (p.periodic == 3 and x1 == p.nx - 1) else 2
        self.levels = [_Level(lib().acpo_refined_level(self.h, 0), p), _Level(lib().acpo_refined_level(self.h, 1), fp)]
        self.x0, self.x1 = x0, x1

    def __del__(self):
        if getattr(self, "h", None):
            lib().acpo_refined_destroy(self.h)
            self.h = None

    def scale(self):
        return lib().acpo_refined_scale(self.h)

    def step(self, n, probes=()):
        """n coarse steps; probes (coarse nodes) record the coarse density after
        the coarse level's own step, before the coupling writes into it."""
        if not probes:
            lib().acpo_refined_step(self.h, n)
            return None
        p = self.levels[0].p
        idx = np.array([i + p.nx * (j + p.ny * k) for (i, j, k) in probes], dtype=np.int64)
        out = np.zeros(n * len(probes))
        lib().acpo_refined_step_probes(self.h, n, idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), len(probes),
                                       dptr(out))
        return out.reshape(n, len(probes))

    def set_filter(self, name):
        """Fine-to-coarse filter: "NOFI", "LAG" or "TOU"."""
        if lib().acpo_refined_set_filter(self.h, {"NOFI": 0, "LAG": 1, "TOU": 2}[name]) != 0:
            raise ValueError(name)

    def set_coupling(self, scheme="STD", c2f="fAll", f2c="cAll", hanging="hAll"):
        """Coupling-Approach (STD, or DC1 / DC2 for the direct coupling) with the
        case file's link sets (Coarse-to-Fine, Fine-to-Coarse,
        Hanging-Node-Reconstruction)."""
        sc = {"STD": 0, "DC": 1, "DC1": 1, "DC2": 1}[scheme]
        rc = lib().acpo_refined_set_coupling(self.h, sc, LINKS[c2f], LINKS[f2c], LINKS[hanging])
        if rc != 0:
            raise ValueError((scheme, c2f, f2c, hanging))

    def restrict(self):
        """The fine-to-coarse restriction alone, on the levels' current states."""
        lib().acpo_refined_restrict(self.h)

    def shell_lists(self):
        """The shell geometry's node lists: fine grid-coupling nodes (coincident
        first) with their Coarse-to-Fine and hanging masks, the restricted
        coarse nodes (core-box surface first) with their masks, and the counts
        (fine, coincident, coarse, core-box surface)."""
        counts = np.zeros(4, dtype=np.int64)
        ip = ctypes.POINTER(ctypes.c_int64)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        lib().acpo_refined_shell_lists(self.h, None, None, None, None, None, counts.ctypes.data_as(ip))
        fs, cr = np.zeros(counts[0], np.int64), np.zeros(counts[2], np.int64)
        m1, m2, m3 = np.zeros(counts[0], np.uint32), np.zeros(counts[0], np.uint32), np.zeros(counts[2], np.uint32)
        lib().acpo_refined_shell_lists(self.h, fs.ctypes.data_as(ip), cr.ctypes.data_as(ip), m1.ctypes.data_as(u32p),
                                       m2.ctypes.data_as(u32p), m3.ctypes.data_as(u32p), counts.ctypes.data_as(ip))
        return {"fs": fs, "fs_c2f": m1, "fs_hang": m2, "cr": cr, "cr_mask": m3, "counts": counts}

    def prolong(self, half=0):
        """The standard coupling's coarse-to-fine alone (half: the coarse state
        at t + 1/2 from both of its arrays)."""
        lib().acpo_refined_prolong(self.h, int(half))

    def set_scaling(self, style):
        """Scaling-Style of the standard coupling: "SRTSTYLE" or "MRTSTYLE"."""
        if lib().acpo_refined_set_scaling(self.h, {"SRTSTYLE": 0, "MRTSTYLE": 1}[style]) != 0:
            raise ValueError(style)

    def scale_neq(self, n, f2c=False):
        """A non-equilibrium part rescaled coarse to fine (or fine to coarse)."""
        out = np.ascontiguousarray(n, dtype=np.float64).copy()
        lib().acpo_refined_scale_neq(self.h, int(bool(f2c)), dptr(out))
        return out


def feq(p, order, rho, u):
    k = consts(p)
    out = np.empty(19)
    uu = np.ascontiguousarray(u, dtype=np.float64)
    lib().acpo_feq(ctypes.byref(k), order, rho, dptr(uu), dptr(out))
    return out


def equilibrium_field(p, rho, u):
    """feq_N for every node; rho [N], u [3][N] -> [19][N]."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    N = rho.size
    f = np.empty((19, N))
    lib().acpo_equilibrium_field(ctypes.byref(p), N, dptr(rho), dptr(u), dptr(f))
    return f


def pulse_ic(p, amp, x0, y0, b):
    f = np.empty((19, p.nx * p.ny * p.nz))
    lib().acpo_pulse_ic(ctypes.byref(p), amp, x0, y0, b, dptr(f))
    return f


def reference_ic(p, domx, domy):
    f = np.empty((19, p.nx * p.ny * p.nz))
    lib().acpo_reference_ic(ctypes.byref(p), domx, domy, dptr(f))
    return f


def reference_ic_level(p, level, i0, spacing0, domx, domy):
    f = np.empty((19, p.nx * p.ny * p.nz))
    lib().acpo_reference_ic_level(ctypes.byref(p), level, i0, spacing0, domx, domy, dptr(f))
    return f


def macros(p, f):
    """(rho, ux, uy, uz) per node in the reference's units (no psi terms)."""
    k = consts(p)
    xi = np.array(k.xi)
    rho = np.zeros(f.shape[1])
    for q in range(19):
        rho += f[q]
    u = np.zeros((3, f.shape[1]))
    for a in range(3):
        for q in range(19):
            if xi[q][a] != 0.0:
                u[a] += xi[q][a] * f[q]
        u[a] /= rho
    return rho, u
