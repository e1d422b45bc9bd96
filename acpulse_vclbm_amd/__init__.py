"""MI355X D3Q19 lattice-Boltzmann engine: Python side of the C ABI (include/acp.h).

Mirrors the reference solver's driver interface (aleksandrShukman/acPulse_vcLBM,
main.cpp:46-157): ``Control(path)`` reads a BoundaryConditions.txt,
``Lattice(control)`` allocates the case, ``initialization(control)``,
``solve(alternating)`` and ``getDistribution(...)`` behave as there.  All work
runs in libacpulse.so; this module only marshals arguments.
"""
import ctypes
import math

import numpy as np

from . import _capi
from ._capi import (ACP_OK, AcpError, BOUNDARY, F2C_FILTER, CaseDesc, CaseInfo, MRT_MODE, OP, PREC, check, dptr, lib)

__all__ = ["Case", "Control", "Lattice", "Pack", "Refined", "CaseDesc", "AcpError", "make_desc", "lib", "lodi_schedule"]


def make_desc(nx, ny, nz, op="SRT", order=3, cubic=False, transverse=False, boundary="LODI",
              mrt_mode="REFERENCE", precision="F64", rho0=1.17621, cs=347.3, spacing=0.01, nu=None, tau=None,
              mach=0.0, k1=1.0, sigma=0.98, f16_scale=1000.0, timesteps=0, write_interval=1000000,
              save_as="acp_case"):
    """Case description with the reference's SI parameters.  Either ``nu`` (m^2/s)
    or ``tau`` (relaxation time, tau = 1/2 + nu/(cs^2 dt)) gives the viscosity."""
    d = CaseDesc()
    lib().acp_case_desc_default(ctypes.byref(d))
    d.nx, d.ny, d.nz = int(nx), int(ny), int(nz)
    d.op = OP[op] if isinstance(op, str) else int(op)
    d.order = int(order)
    d.cubic, d.transverse = int(bool(cubic)), int(bool(transverse))
    d.boundary = BOUNDARY[boundary] if isinstance(boundary, str) else int(boundary)
    d.mrt_mode = MRT_MODE[mrt_mode] if isinstance(mrt_mode, str) else int(mrt_mode)
    d.precision = PREC[precision] if isinstance(precision, str) else int(precision)
    d.rho0, d.cs, d.spacing = float(rho0), float(cs), float(spacing)
    if nu is None:
        if tau is None:
            raise ValueError("give nu or tau")
        dt = spacing / (cs * math.sqrt(3.0))
        nu = (tau - 0.5) * cs * cs * dt
    d.nu = float(nu)
    d.mach, d.k1, d.sigma, d.f16_scale = float(mach), float(k1), float(sigma), float(f16_scale)
    d.timesteps, d.write_interval = int(timesteps), int(write_interval)
    d.domain[0], d.domain[1], d.domain[2] = (nx - 0.5) * spacing, (ny - 0.5) * spacing, (nz - 0.5) * spacing
    d.save_as = save_as.encode()
    d.collision_model = (op if isinstance(op, str) else "SRT").encode()
    return d


def read_case_file(path):
    """acp_control_read: Control::Control semantics (Control.cpp:8-132)."""
    d = CaseDesc()
    check(lib().acp_control_read(str(path).encode(), ctypes.byref(d)))
    return d


class Case:
    """One device-resident case (acp_case_* of include/acp.h)."""

    def __init__(self, desc, stream=None):
        self.desc = desc
        self.h = ctypes.c_void_p()
        st = ctypes.c_void_p(int(stream)) if stream else None
        check(lib().acp_case_create(ctypes.byref(desc), st, ctypes.byref(self.h)))
        self.nx, self.ny, self.nz = desc.nx, desc.ny, desc.nz
        self.N = self.nx * self.ny * self.nz

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().acp_case_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def node(self, i, j, k):
        return i + self.nx * (j + self.ny * k)

    def info(self):
        inf = CaseInfo()
        check(lib().acp_case_info_get(self.h, ctypes.byref(inf)))
        return inf

    def init_reference(self):
        check(lib().acp_case_init_reference(self.h))

    def init_pulse(self, amp, x0, y0, b):
        check(lib().acp_case_init_pulse(self.h, amp, x0, y0, b))

    def init_density(self, rho):
        """Equilibrium state feq_N(rho, 0) from a density field of N values (SI)."""
        rho = np.ascontiguousarray(rho, dtype=np.float64).reshape(-1)
        if rho.size != self.info().nodes:
            raise ValueError("density field needs one value per node")
        check(lib().acp_case_init_density(self.h, dptr(rho)))

    def init_random(self, seed, rho_amp, u_amp):
        check(lib().acp_case_init_random(self.h, seed, rho_amp, u_amp))

    def set_populations(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(19, self.N)
        check(lib().acp_case_set_populations(self.h, dptr(f)))

    def step(self, n=1):
        check(lib().acp_case_step(self.h, int(n)))

    def sync(self):
        check(lib().acp_case_sync(self.h))

    def populations(self):
        f = np.empty((19, self.N))
        check(lib().acp_case_get_populations(self.h, dptr(f)))
        return f

    def psi(self):
        """The cubic-Mach state psi [19][N] (zero without the correction)."""
        f = np.empty((19, self.N))
        check(lib().acp_case_get_psi(self.h, dptr(f)))
        return f

    def previous_populations(self):
        """The other array of the ping-pong pair (the one the last step read)."""
        f = np.empty((19, self.N))
        check(lib().acp_case_get_previous_populations(self.h, dptr(f)))
        return f

    def macros(self):
        rho = np.empty(self.N)
        u = np.empty((3, self.N))
        check(lib().acp_case_get_macros(self.h, dptr(rho), dptr(u)))
        return rho, u

    def mass(self):
        m = ctypes.c_double()
        check(lib().acp_case_mass(self.h, ctypes.byref(m)))
        return m.value

    def set_probes(self, points, capacity):
        nodes = np.array([self.node(*p) for p in points], dtype=np.int64)
        check(lib().acp_case_set_probes(self.h, nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                        len(nodes), int(capacity)))
        self._nprobe = len(nodes)
        self._probe_cap = int(capacity)

    def read_probes(self):
        out = np.empty((max(self._probe_cap, 1), max(self._nprobe, 1)))
        n = ctypes.c_int32()
        check(lib().acp_case_read_probes(self.h, dptr(out), ctypes.byref(n)))
        return out[: n.value, : self._nprobe].copy()

    def enable_graphs(self, chunk):
        """Run every full chunk of `chunk` (even) steps as one captured-graph replay; 0 switches off."""
        check(lib().acp_case_enable_graphs(self.h, int(chunk)))

    BULK_VARIANTS = {"auto": 0, "wide_nt": 1, "narrow_nt": 2, "narrow_plain": 3, "mfma": 4, "relative_nt": 5}

    def set_bulk_variant(self, name):
        """Force the bulk kernels' addressing and streaming form ("auto",
        "wide_nt", "narrow_nt", "narrow_plain"); results are identical in every form."""
        check(lib().acp_case_set_bulk_variant(self.h, self.BULK_VARIANTS[name]))

    def checksum(self):
        """Order-independent 64-bit fingerprint of the current population array."""
        v = ctypes.c_uint64()
        check(lib().acp_case_checksum(self.h, ctypes.byref(v)))
        return v.value

    def write_vtk(self, stem):
        """Lattice::writeResultsVTK (Lattice.cpp:2212-2243, 4094-4349): <stem>.vtk."""
        check(lib().acp_case_write_vtk(self.h, str(stem).encode()))


class _LevelCase(Case):
    """A level of a Refined case: the acp_case API on a handle the refined case owns."""

    def __init__(self, h, desc):
        self.desc = desc
        self.h = h
        self.nx, self.ny, self.nz = desc.nx, desc.ny, desc.nz
        self.N = self.nx * self.ny * self.nz

    def close(self):
        self.h = ctypes.c_void_p()


# Coupling-Approach and the link sets of Coarse-to-Fine, Fine-to-Coarse and
# Hanging-Node-Reconstruction (BoundaryConditions.txt:52-86)
COUPLING = {"STD": 0, "DC1": 1, "DC2": 1}
LINKS = {"Mf": 0, "Mc": 0, "hMf": 0, "MfP": 1, "McP": 1, "hMfP": 1, "fAll": 2, "cAll": 2, "hAll": 2}


class Refined:
    """Two-level case (acp_refined_*): coarse level `desc` (PERIODIC or LODI-X
    boundary), fine level over coarse x in [x0, x1] at half the spacing; with a
    LODI-X coarse level and x1 = nx - 1 the fine level's high-x face is a LODI
    face and only the face at x0 is coupled.  level(0) / level(1) give the
    levels' cases for probes, read-back and output; step with Refined.step."""

    def __init__(self, desc, x0=0, x1=0, stream=None, shell=None):
        """shell = (layer_yz, layer_x): the reference's geometry, a fine shell
        around a coarse core (acp_refined_create_shell; `desc` with LODI faces,
        x0 and x1 unused)."""
        self.h = ctypes.c_void_p()
        st = ctypes.c_void_p(int(stream)) if stream else None
        if shell is not None:
            check(lib().acp_refined_create_shell(ctypes.byref(desc), int(shell[0]), int(shell[1]), st,
                                                 ctypes.byref(self.h)))
            fd = CaseDesc.from_buffer_copy(desc)
            fd.nx, fd.ny, fd.nz = 2 * desc.nx - 1, 2 * desc.ny - 1, 2 * desc.nz - 1
            fd.spacing = desc.spacing / 2.0
            self.x0 = self.x1 = 0
            self._levels = []
            for lv, d in ((0, desc), (1, fd)):
                hc = ctypes.c_void_p()
                check(lib().acp_refined_level(self.h, lv, ctypes.byref(hc)))
                self._levels.append(_LevelCase(hc, d))
            return
        check(lib().acp_refined_create(ctypes.byref(desc), int(x0), int(x1), st, ctypes.byref(self.h)))
        fd = CaseDesc.from_buffer_copy(desc)
        fd.nx, fd.ny, fd.nz = 2 * (x1 - x0) + 1, 2 * desc.ny, 2 * desc.nz
        fd.spacing = desc.spacing / 2.0
        lodi_end = desc.boundary == BOUNDARY["LODI-X"] and x1 == desc.nx - 1
        fd.boundary = BOUNDARY["OPEN-LODI-X" if lodi_end else "OPEN-X"]
        self.x0, self.x1 = int(x0), int(x1)
        self._levels = []
        for lv, d in ((0, desc), (1, fd)):
            hc = ctypes.c_void_p()
            check(lib().acp_refined_level(self.h, lv, ctypes.byref(hc)))
            self._levels.append(_LevelCase(hc, d))

    def level(self, i):
        return self._levels[i]

    def init_pulse(self, amp, x0, y0, b):
        check(lib().acp_refined_init_pulse(self.h, amp, x0, y0, b))

    def init_reference(self):
        """Lattice::initialization (Lattice.cpp:1397-1493) on both levels."""
        check(lib().acp_refined_init_reference(self.h))

    def set_coupling_precision(self, precision):
        """Shifted tiers: coupling arithmetic in "F64" (default) or "F32-SHIFTED"."""
        check(lib().acp_refined_set_coupling_precision(self.h, PREC[precision]))

    def set_filter(self, name):
        """Fine-to-coarse filter (F2C-Filter): "NOFI" (default), "LAG" or "TOU"."""
        if name not in F2C_FILTER:
            raise AcpError(_capi.ACP_E_UNSUPPORTED, f"fine-to-coarse filter {name!r}: NOFI, LAG or TOU")
        check(lib().acp_refined_set_filter(self.h, F2C_FILTER[name]))

    def set_scaling(self, style):
        """Scaling-Style of the standard coupling: "SRTSTYLE" or "MRTSTYLE"."""
        styles = {"SRTSTYLE": 0, "MRTSTYLE": 1}
        if style not in styles:
            raise AcpError(_capi.ACP_E_ARG, f"Scaling-Style {style!r}: SRTSTYLE or MRTSTYLE")
        check(lib().acp_refined_set_scaling(self.h, styles[style]))

    def set_coupling(self, scheme="STD", c2f=None, f2c=None, hanging=None):
        """Coupling-Approach "STD" (default) or "DC1" / "DC2" (direct coupling),
        with the case file's link sets: Coarse-to-Fine "Mf" / "MfP" / "fAll",
        Fine-to-Coarse "Mc" / "McP" / "cAll", Hanging-Node-Reconstruction
        "hMf" / "hMfP" / "hAll".  Unset link sets take "fAll" / "cAll" /
        "hAll" for STD and "Mf" / "Mc" / "hMf" for DC."""
        if scheme not in COUPLING:
            raise AcpError(_capi.ACP_E_UNSUPPORTED, f"Coupling-Approach {scheme!r}: STD, DC1 or DC2")
        dc = COUPLING[scheme] == 1
        names = (c2f or ("Mf" if dc else "fAll"), f2c or ("Mc" if dc else "cAll"), hanging or ("hMf" if dc else "hAll"))
        for n, allowed in zip(names, (("Mf", "MfP", "fAll"), ("Mc", "McP", "cAll"), ("hMf", "hMfP", "hAll"))):
            if n not in allowed:
                raise AcpError(_capi.ACP_E_ARG, f"link set {n!r}: one of {allowed}")
        check(lib().acp_refined_set_coupling(self.h, COUPLING[scheme], *(LINKS[n] for n in names)))

    def restrict(self):
        """The fine-to-coarse restriction alone, on the levels' current states."""
        check(lib().acp_refined_restrict(self.h))

    def enable_graphs(self, chunk):
        """Replay `chunk` coarse steps (even) from captured HIP graphs; 0 turns it off."""
        check(lib().acp_refined_enable_graphs(self.h, int(chunk)))

    def write_vtk(self, stem):
        """<stem>.vtk of both levels (Lattice::writeResultsVTK of a two-level lattice)."""
        check(lib().acp_refined_write_vtk(self.h, str(stem).encode()))

    def checksum(self):
        """Order-independent checksum of both levels' states (coarse + fine, mod 2^64)."""
        return (self._levels[0].checksum() + self._levels[1].checksum()) % (1 << 64)

    def step(self, n=1):
        check(lib().acp_refined_step(self.h, int(n)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            for lv in self._levels:
                lv.close()
            lib().acp_refined_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Pack:
    """Resident one-level cases of one tier and equilibrium order advanced by
    shared kernel launches (acp_pack_* of include/acp.h).  A member keeps its
    own results: they equal stepping it alone, bit for bit."""

    def __init__(self, precision="F64", order=3, stream=None):
        prec = PREC[precision] if isinstance(precision, str) else int(precision)
        self.h = ctypes.c_void_p()
        st = ctypes.c_void_p(int(stream)) if stream else None
        check(lib().acp_pack_create(prec, int(order), st, ctypes.byref(self.h)))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().acp_pack_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add(self, case, steps):
        check(lib().acp_pack_add(self.h, case.h, int(steps)))

    def step(self, n=1):
        check(lib().acp_pack_step(self.h, int(n)))

    def enable_graphs(self, chunk):
        check(lib().acp_pack_enable_graphs(self.h, int(chunk)))

    def steps_left(self, case):
        n = ctypes.c_int32()
        check(lib().acp_pack_steps_left(self.h, case.h, ctypes.byref(n)))
        return n.value

    def remove(self, case):
        check(lib().acp_pack_remove(self.h, case.h))

    def members(self):
        n = ctypes.c_int32()
        check(lib().acp_pack_members(self.h, ctypes.byref(n)))
        return n.value


def lodi_schedule(nx, ny, nz):
    """(nodes, phases) of the LODI pass in processing order."""
    L = lib()
    n = L.acp_lodi_schedule(nx, ny, nz, None, None, 0)
    nodes = np.empty(n, dtype=np.int64)
    ph = np.empty(n, dtype=np.int32)
    L.acp_lodi_schedule(nx, ny, nz, nodes.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                        ph.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), n)
    return nodes, ph


LODI_RECORD = np.dtype([("node", np.int64), ("n1", np.int64), ("n2", np.int64), ("t", np.int64, 4),
                        ("dir", np.int32), ("corner", np.int32)])


def lodi_launches(nx, ny, nz):
    """(records, launch_start) of the LODI pass as the kernels receive them
    (acp_lodi_launches); records is a structured array of LODI_RECORD."""
    L = lib()
    nl = ctypes.c_int32()
    n = L.acp_lodi_launches(nx, ny, nz, None, 0, None, 0, ctypes.byref(nl))
    rec = np.zeros(n, dtype=LODI_RECORD)
    starts = np.empty(nl.value + 1, dtype=np.int64)
    L.acp_lodi_launches(nx, ny, nz, rec.ctypes.data_as(ctypes.c_void_p), n,
                        starts.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), nl.value + 1, ctypes.byref(nl))
    return rec, starts


def footprint(desc, nprobe=0, probe_capacity=0):
    """Device bytes a case of `desc` holds with `nprobe` probes of `probe_capacity` steps (acp_case_footprint)."""
    b = ctypes.c_int64()
    check(lib().acp_case_footprint(ctypes.byref(desc), int(nprobe), int(probe_capacity), ctypes.byref(b)))
    return b.value


def refined_footprint(desc, x0, x1, nprobe=0, probe_capacity=0):
    """Device bytes of a refined case (both levels; probes on the coarse level),
    acp_refined_footprint."""
    b = ctypes.c_int64()
    check(lib().acp_refined_footprint(ctypes.byref(desc), int(x0), int(x1), int(nprobe), int(probe_capacity),
                                      ctypes.byref(b)))
    return b.value


def copy_probe(planes, plane_elems, reps=5, stream=None, block_nodes=0):
    """Mean milliseconds of one device copy of `planes` (1 or 19) float64
    planes of `plane_elems` elements, whole planes or (19 planes) in blocks of
    `block_nodes` nodes (acp_copy_probe)."""
    L = _capi.lib()
    ms = ctypes.c_double()
    _capi.check(L.acp_copy_probe(int(planes), int(plane_elems), int(block_nodes), int(reps), stream,
                                 ctypes.byref(ms)))
    return ms.value


class Control:
    """Control(string) of the reference (Control.h:17-75): case file reader and getters."""

    def __init__(self, filename):
        self.desc = read_case_file(filename)
        self.filename = filename
        self._time = 0.0

    def getNumberOfNodesX(self):
        return self.desc.nx

    def getNumberOfNodesY(self):
        return self.desc.ny

    def getNumberOfNodesZ(self):
        return self.desc.nz

    def getLevelMax(self):
        return self.desc.max_level

    def getTimeStepMax(self):
        return self.desc.timesteps

    def getWriteInterval(self):
        return self.desc.write_interval

    def getNumberOfThreads(self):
        return self.desc.threads

    def getOrderOfEquilibrium(self):
        return self.desc.order

    def getSpacing(self):
        return self.desc.spacing

    def getHybridParameter(self):
        return self.desc.sigma

    def getDensity(self):
        return self.desc.rho0

    def getMachNumber(self):
        return self.desc.mach

    def getSpeedOfSound(self):
        return self.desc.cs

    def getKinematicViscosity(self):
        return self.desc.nu

    def getMolecularVelocity(self):
        return self.desc.cs * math.sqrt(3.0)

    def getTimeStep(self):
        return self.desc.spacing / self.getMolecularVelocity()

    def getCollisionFrequency(self):
        cs2, dt = self.desc.cs ** 2, self.getTimeStep()
        return (cs2 * dt) / (self.desc.nu + 0.5 * cs2 * dt)

    def getReynoldsNumber(self):
        return 0.0

    def getChannelSizeX_(self):
        return self.desc.domain[0]

    def getChannelSizeY_(self):
        return self.desc.domain[1]

    def getChannelSizeZ_(self):
        return self.desc.domain[2]

    def getLODIK1relaxation(self):
        return self.desc.k1

    def getLODIK2relaxation(self):
        return self.desc.k2

    def getCaseName(self):
        return self.desc.save_as.decode()

    def getCollisionModel(self):
        return self.desc.collision_model.decode()

    def getReflectionHandling(self):
        return "yes" if self.desc.nonreflecting else "no"

    def getTransverseInclusion(self):
        return "yes" if self.desc.transverse else "no"

    def getCubicMachCorrection(self):
        return "yes" if self.desc.cubic else "no"

    def setTime_(self, t):
        self._time = float(t)

    def getTime_(self):
        return self._time


class Lattice:
    """Lattice(Control&) of the reference (Lattice.h:23-120) on the device.

    The reference keeps two population arrays and a ping-pong flag owned by the
    caller (main.cpp:55-56, 155-156); here the case tracks the current array and
    ``solve(alternating)`` checks that the caller's flag agrees with it."""

    def __init__(self, control, stream=None):
        self.bc = control
        self.case = Case(control.desc, stream=stream)
        self._alt = True

    def initialization(self, control=None):
        self.case.init_reference()
        self._alt = True

    def solve(self, alternating):
        if bool(alternating) != self._alt:
            raise AcpError(_capi.ACP_E_STATE, "ping-pong flag does not match the current array")
        self.case.step(1)
        self._alt = not self._alt

    def getDistribution(self, alternating, dir, level, i, j, k):
        if dir >= 19:
            return 0.0
        if level != 0:
            raise AcpError(_capi.ACP_E_UNSUPPORTED, "single level only")
        f = self.case.populations() if bool(alternating) == self._alt else self.case.previous_populations()
        return float(f[dir, self.case.node(i, j, k)])
