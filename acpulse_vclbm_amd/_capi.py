"""ctypes binding of libacpulse.so (include/acp.h).

The library is built in-tree by ``__graft_entry__.build()`` into
``acpulse_vclbm_amd/lib/libacpulse.so``.  Importing this module does not load
it; the first call of :func:`lib` does, and raises if it is missing, so a GPU
run never silently falls back to anything else.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libacpulse.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "acp.h")

ACP_OK, ACP_E_ARG, ACP_E_CONFIG, ACP_E_OOM, ACP_E_HIP, ACP_E_STATE, ACP_E_UNSUPPORTED, ACP_E_IO = range(8)
OP = {"SRT": 0, "MRT-GS": 1, "MRT-RM": 2, "RR": 3, "HRR": 4}
PREC = {"F64": 0, "F32-SHIFTED": 1, "F16-SHIFTED": 2}
BOUNDARY = {"LODI": 0, "PERIODIC": 1, "OPEN-X": 2, "LODI-X": 3, "OPEN-LODI-X": 4}
F2C_FILTER = {"NOFI": 0, "LAG": 1, "TOU": 2}
MRT_MODE = {"REFERENCE": 0, "RELAXED": 1}


class CaseDesc(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("op", ctypes.c_int32), ("order", ctypes.c_int32), ("cubic", ctypes.c_int32),
        ("transverse", ctypes.c_int32), ("nonreflecting", ctypes.c_int32), ("boundary", ctypes.c_int32),
        ("mrt_mode", ctypes.c_int32), ("precision", ctypes.c_int32), ("max_level", ctypes.c_int32),
        ("timesteps", ctypes.c_int32), ("write_interval", ctypes.c_int32), ("threads", ctypes.c_int32),
        ("rho0", ctypes.c_double), ("nu", ctypes.c_double), ("cs", ctypes.c_double),
        ("spacing", ctypes.c_double), ("mach", ctypes.c_double), ("k1", ctypes.c_double),
        ("k2", ctypes.c_double), ("sigma", ctypes.c_double), ("domain", ctypes.c_double * 3),
        ("f16_scale", ctypes.c_double), ("save_as", ctypes.c_char * 1024),
        ("collision_model", ctypes.c_char * 32),
    ]


class CaseInfo(ctypes.Structure):
    _fields_ = [
        ("nodes", ctypes.c_int64), ("device_bytes", ctypes.c_int64), ("bytes_per_node_step", ctypes.c_int64),
        ("lodi_nodes", ctypes.c_int64), ("lodi_launches", ctypes.c_int32), ("steps_done", ctypes.c_int32),
        ("omega", ctypes.c_double), ("dt", ctypes.c_double),
    ]


class ServerDesc(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("max_active", ctypes.c_int32),
                ("graph_chunk", ctypes.c_int32), ("launch_steps", ctypes.c_int32), ("chunks_ahead", ctypes.c_int32),
                ("admit_threads", ctypes.c_int32), ("pack_max_nodes", ctypes.c_int32)]


class Request(ctypes.Structure):
    _fields_ = [
        ("desc", CaseDesc), ("steps", ctypes.c_int32), ("init_kind", ctypes.c_int32),
        ("amp", ctypes.c_double), ("x0", ctypes.c_double), ("y0", ctypes.c_double), ("halfwidth", ctypes.c_double),
        ("seed", ctypes.c_uint64), ("rho_amp", ctypes.c_double), ("u_amp", ctypes.c_double),
        ("probe", ctypes.c_int32 * 3), ("probe_out", ctypes.POINTER(ctypes.c_double)),
        ("arrival", ctypes.c_double), ("tag", ctypes.c_uint64),
        ("refine", ctypes.c_int32 * 2), ("f2c_filter", ctypes.c_int32), ("coupling", ctypes.c_int32),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("tag", ctypes.c_uint64), ("status", ctypes.c_int32), ("steps", ctypes.c_int32),
        ("device_bytes", ctypes.c_int64), ("t_arrival", ctypes.c_double), ("t_submit", ctypes.c_double),
        ("t_admit", ctypes.c_double), ("t_start", ctypes.c_double), ("t_done", ctypes.c_double),
        ("mass", ctypes.c_double), ("checksum", ctypes.c_uint64), ("error", ctypes.c_char * 128),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("queued", ctypes.c_int64),
        ("active", ctypes.c_int32), ("peak_active", ctypes.c_int32), ("bytes_in_use", ctypes.c_int64),
        ("peak_bytes", ctypes.c_int64), ("budget_bytes", ctypes.c_int64), ("clock", ctypes.c_double),
    ]


INIT = {"REFERENCE": 0, "PULSE": 1, "RANDOM": 2}


class AcpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"acp error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libacpulse.so (raises when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() (make -C acpulse_vclbm_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64, dbl = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    D = ctypes.POINTER(CaseDesc)
    sig = {
        "acp_last_error": (ctypes.c_char_p, []),
        "acp_control_read": (ctypes.c_int, [ctypes.c_char_p, D]),
        "acp_control_read_value": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, dp]),
        "acp_control_read_string": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, i64]),
        "acp_control_read_filename": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, i64]),
        "acp_control_read_vector": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, dp]),
        "acp_case_desc_default": (None, [D]),
        "acp_case_create": (ctypes.c_int, [D, vp, ctypes.POINTER(vp)]),
        "acp_case_destroy": (None, [vp]),
        "acp_case_info_get": (ctypes.c_int, [vp, ctypes.POINTER(CaseInfo)]),
        "acp_case_init_reference": (ctypes.c_int, [vp]),
        "acp_case_init_pulse": (ctypes.c_int, [vp, dbl, dbl, dbl, dbl]),
        "acp_case_init_random": (ctypes.c_int, [vp, ctypes.c_uint64, dbl, dbl]),
        "acp_case_set_populations": (ctypes.c_int, [vp, dp]),
        "acp_case_init_density": (ctypes.c_int, [vp, dp]),
        "acp_case_step": (ctypes.c_int, [vp, i32]),
        "acp_case_sync": (ctypes.c_int, [vp]),
        "acp_case_get_populations": (ctypes.c_int, [vp, dp]),
        "acp_case_get_previous_populations": (ctypes.c_int, [vp, dp]),
        "acp_case_get_macros": (ctypes.c_int, [vp, dp, dp]),
        "acp_case_get_macros_noalloc": (ctypes.c_int, [vp, dp, dp]),
        "acp_case_get_psi": (ctypes.c_int, [vp, dp]),
        "acp_case_mass": (ctypes.c_int, [vp, dp]),
        "acp_case_set_probes": (ctypes.c_int, [vp, ctypes.POINTER(i64), i32, i32]),
        "acp_case_read_probes": (ctypes.c_int, [vp, dp, ctypes.POINTER(i32)]),
        "acp_case_write_vtk": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "acp_write_vtk_host": (ctypes.c_int, [D, dp, dp, dp, ctypes.c_char_p]),
        "acp_lodi_schedule": (ctypes.c_int64, [i32, i32, i32, ctypes.POINTER(i64), ctypes.POINTER(i32), i64]),
        "acp_lodi_launches": (ctypes.c_int64, [i32, i32, i32, vp, i64, ctypes.POINTER(i64), i32, ctypes.POINTER(i32)]),
        "acp_copy_probe": (ctypes.c_int, [i32, i64, i64, i32, vp, dp]),
        "acp_case_enable_graphs": (ctypes.c_int, [vp, i32]),
        "acp_case_footprint": (ctypes.c_int, [D, i32, i32, ctypes.POINTER(i64)]),
        "acp_case_create_in": (ctypes.c_int, [D, vp, vp, i64, ctypes.POINTER(vp)]),
        "acp_case_checksum": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_uint64)]),
        "acp_case_set_bulk_variant": (ctypes.c_int, [vp, ctypes.c_int32]),
        "acp_refined_create": (ctypes.c_int, [D, i32, i32, vp, ctypes.POINTER(vp)]),
        "acp_refined_create_shell": (ctypes.c_int, [D, i32, i32, vp, ctypes.POINTER(vp)]),
        "acp_refined_destroy": (None, [vp]),
        "acp_refined_level": (ctypes.c_int, [vp, i32, ctypes.POINTER(vp)]),
        "acp_refined_init_pulse": (ctypes.c_int, [vp, dbl, dbl, dbl, dbl]),
        "acp_refined_init_reference": (ctypes.c_int, [vp]),
        "acp_refined_step": (ctypes.c_int, [vp, i32]),
        "acp_refined_set_coupling_precision": (ctypes.c_int, [vp, i32]),
        "acp_refined_set_filter": (ctypes.c_int, [vp, i32]),
        "acp_refined_set_scaling": (ctypes.c_int, [vp, i32]),
        "acp_refined_set_coupling": (ctypes.c_int, [vp, i32, i32, i32, i32]),
        "acp_refined_write_vtk": (ctypes.c_int, [vp, ctypes.c_char_p]),
        "acp_write_vtk2_host": (ctypes.c_int, [ctypes.POINTER(CaseDesc), dp, dp, i32, i32, dp, dp, ctypes.c_char_p]),
        "acp_write_vtk_shell_host": (ctypes.c_int, [ctypes.POINTER(CaseDesc), i32, i32, dp, dp, dp, dp, dp, dp,
                                                    ctypes.c_char_p]),
        "acp_refined_fine_desc": (ctypes.c_int, [ctypes.POINTER(CaseDesc), i32, i32, ctypes.POINTER(CaseDesc)]),
        "acp_refined_restrict": (ctypes.c_int, [vp]),
        "acp_refined_footprint": (ctypes.c_int, [D, i32, i32, i32, i32, ctypes.POINTER(ctypes.c_int64)]),
        "acp_refined_create_in": (ctypes.c_int, [D, i32, i32, vp, vp, ctypes.c_int64, ctypes.POINTER(vp)]),
        "acp_refined_enable_graphs": (ctypes.c_int, [vp, i32]),
        "acp_pack_create": (ctypes.c_int, [i32, i32, vp, ctypes.POINTER(vp)]),
        "acp_pack_destroy": (None, [vp]),
        "acp_pack_add": (ctypes.c_int, [vp, vp, i32]),
        "acp_pack_step": (ctypes.c_int, [vp, i32]),
        "acp_pack_enable_graphs": (ctypes.c_int, [vp, i32]),
        "acp_pack_steps_left": (ctypes.c_int, [vp, vp, ctypes.POINTER(i32)]),
        "acp_pack_remove": (ctypes.c_int, [vp, vp]),
        "acp_pack_members": (ctypes.c_int, [vp, ctypes.POINTER(i32)]),
        "acp_server_create": (ctypes.c_int, [ctypes.POINTER(ServerDesc), ctypes.POINTER(vp)]),
        "acp_server_submit": (ctypes.c_int, [vp, ctypes.POINTER(Request)]),
        "acp_server_poll": (ctypes.c_int, [vp, ctypes.POINTER(Result), i32, i32, ctypes.POINTER(i32)]),
        "acp_server_get_stats": (ctypes.c_int, [vp, ctypes.POINTER(ServerStats)]),
        "acp_server_clock": (ctypes.c_double, [vp]),
        "acp_server_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != ACP_OK:
        raise AcpError(rc, lib().acp_last_error().decode(errors="replace"))


def header_symbols():
    """Names of the functions include/acp.h declares."""
    import re
    txt = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(acp_\w+)\s*\(", txt, flags=re.M)))


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
