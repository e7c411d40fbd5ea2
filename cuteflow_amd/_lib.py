"""ctypes binding of libcuteflow_hip.so (the C ABI of include/cuteflow_hip.h).

The library is the only compute path: if it is missing, or no GPU is
visible when a handle is created, calls raise instead of falling back to
anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .build import CGNS_LIB_PATH, LIB_PATH

CF_MAX_IO = 20
CF_MAX_HYDRO = 64

CF_OPT_KEEP_ORDER = 1
CF_OPT_NO_TILE_SKIP = 2
CF_OPT_TILE_128 = 4
CF_OPT_TILE_64 = 8
CF_OPT_XCD_CONTIG = 16
CF_OPT_FIXED_TILE_ORDER = 32
CF_TUNE_NCAND = 12

CF_ADJ_FREEZE_LIMITERS = 1

CF_OK = 0
CF_ERR_ARG = -1
CF_ERR_HIP = -2
CF_ERR_DT = -3
CF_ERR_MESH = -4
CF_ERR_NOMEM = -5
CF_ERR_IO = -6
CF_ERR_SIZE = -7


class CfParams(C.Structure):
    """Mirror of struct cf_params."""

    _fields_ = [
        ("gp", C.c_double),
        ("tolisec", C.c_double),
        ("tol", C.c_double),
        ("cfl", C.c_double),
        ("ts", C.c_double),
        ("constant_dt", C.c_double),
        ("override_manning", C.c_double),
        ("tol_reg_perm", C.c_double),
        ("iflux", C.c_int),
        ("ilimiteur", C.c_int),
        ("friction", C.c_int),
        ("fricimplic", C.c_int),
        ("inlet", C.c_int),
        ("outlet", C.c_int),
        ("is_dt_constant", C.c_int),
        ("is_override_manning", C.c_int),
        ("nombre_entrees", C.c_int),
        ("options", C.c_int),
        ("debitglob", C.c_double * CF_MAX_IO),
        ("h_sortie", C.c_double * CF_MAX_IO),
        ("long_entree", C.c_double * CF_MAX_IO),
        ("hydro_n", C.c_int * CF_MAX_IO),
        ("hydro_t", C.c_double * (CF_MAX_IO * CF_MAX_HYDRO)),
        ("hydro_q", C.c_double * (CF_MAX_IO * CF_MAX_HYDRO)),
    ]


class CfMesh(C.Structure):
    """Mirror of struct cf_mesh."""

    _fields_ = [
        ("nelt", C.c_int),
        ("nnt", C.c_int),
        ("conn", C.POINTER(C.c_int)),
        ("coords", C.POINTER(C.c_double)),
        ("boundary", C.POINTER(C.c_int)),
        ("manning", C.POINTER(C.c_double)),
    ]


# exported symbol -> (restype, argtypes)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_H = C.c_void_p
SIGNATURES = {
    "cf_build_boundary": (C.c_int, [C.c_int, C.c_int, _IP, C.c_int, _IP, _IP, C.c_int, _IP, _IP,
                                    C.c_int, _IP, _IP]),
    "cf_build_boundary_device": (C.c_int, [C.c_int, C.c_int, _IP, C.c_int, _IP, _IP, C.c_int, _IP, _IP,
                                           C.c_int, _IP, _IP]),
    "cf_build_boundary_device_noalloc": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                                   C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                                   C.c_void_p]),
    "cf_inlet_lengths": (C.c_int, [C.c_int, C.c_int, _IP, _DP, _IP, C.c_int, _DP]),
    "cf_create": (C.c_int, [C.POINTER(CfMesh), C.POINTER(CfParams), C.POINTER(_H)]),
    "cf_destroy": (None, [_H]),
    "cf_error_string": (C.c_char_p, [C.c_int]),
    "cf_set_state": (C.c_int, [_H, _DP]),
    "cf_get_state": (C.c_int, [_H, _DP]),
    "cf_set_state_device": (C.c_int, [_H, C.c_void_p]),
    "cf_get_state_device": (C.c_int, [_H, C.c_void_p]),
    "cf_initial_dt": (C.c_int, [_H, _DP]),
    "cf_get_time": (C.c_int, [_H, _DP, _DP]),
    "cf_step": (C.c_int, [_H, C.c_int, _DP, _IP]),
    "cf_step_async": (C.c_int, [_H, C.c_int]),
    "cf_step_finish": (C.c_int, [_H, _DP, _IP]),
    "cf_step_prepare": (C.c_int, [_H, C.c_int]),
    "cf_set_inflow": (C.c_int, [_H, _DP, C.c_int]),
    "cf_set_hydrograph": (C.c_int, [_H, C.c_int, C.c_int, _DP, _DP]),
    "cf_set_manning": (C.c_int, [_H, C.c_void_p]),
    "cf_get_geometry": (C.c_int, [_H, _DP, _DP, _DP, _DP]),
    "cf_last_run_stats": (C.c_int, [_H, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _IP]),
    "cf_get_stream": (C.c_void_p, [_H]),
    "cf_set_stream": (C.c_int, [_H, C.c_void_p]),
    "cf_batch_create": (C.c_int, [C.c_int, C.POINTER(CfMesh), C.POINTER(CfParams), C.POINTER(_H)]),
    "cf_batch_size": (C.c_int, [_H]),
    "cf_batch_set_state": (C.c_int, [_H, C.c_int, _DP]),
    "cf_batch_get_state": (C.c_int, [_H, C.c_int, _DP]),
    "cf_batch_set_state_device": (C.c_int, [_H, C.c_int, C.c_void_p]),
    "cf_batch_get_state_device": (C.c_int, [_H, C.c_int, C.c_void_p]),
    "cf_batch_get_time": (C.c_int, [_H, C.c_int, _DP, _DP]),
    "cf_batch_step": (C.c_int, [_H, C.c_int, _DP, _IP, _IP]),
    "cf_batch_step_finish": (C.c_int, [_H, _DP, _IP, _IP]),
    "cf_batch_set_inflow": (C.c_int, [_H, C.c_int, _DP, C.c_int]),
    "cf_batch_set_hydrograph": (C.c_int, [_H, C.c_int, C.c_int, C.c_int, _DP, _DP]),
    "cf_batch_set_manning": (C.c_int, [_H, C.c_int, C.c_void_p]),
    "cf_batch_get_geometry": (C.c_int, [_H, C.c_int, _DP, _DP, _DP, _DP]),
    "cf_set_discharge": (C.c_int, [_H, C.c_int]),
    "cf_batch_get_discharge": (C.c_int, [_H, C.c_int, _DP, _DP, _IP]),
    "cf_get_discharge": (C.c_int, [_H, _DP, _DP, _IP]),
    "cf_batch_autotune": (C.c_int, [C.c_int, C.POINTER(CfMesh), C.POINTER(CfParams), C.POINTER(C.c_void_p),
                                    C.c_int, _IP, _DP]),
    "cf_adj_set_flags": (C.c_int, [_H, C.c_int]),
    "cf_adj_set_gauges": (C.c_int, [_H, C.c_int, C.c_int, _IP]),
    "cf_adj_forward": (C.c_int, [_H, C.c_int, C.c_int, C.c_int, C.c_int]),
    "cf_adj_num_samples": (C.c_int, [_H]),
    "cf_adj_get_samples": (C.c_int, [_H, C.c_int, _DP, C.c_void_p]),
    "cf_adj_set_seed": (C.c_int, [_H, C.c_int, _DP, C.c_void_p, C.c_void_p]),
    "cf_adj_backward": (C.c_int, [_H]),
    "cf_adj_get_gradient": (C.c_int, [_H, C.c_int, C.c_void_p, C.c_void_p]),
    "cf_tlm_set_direction": (C.c_int, [_H, C.c_int, C.c_void_p, C.c_void_p]),
    "cf_tlm_run": (C.c_int, [_H]),
    "cf_tlm_set_mask_mode": (C.c_int, [_H, C.c_int]),
    "cf_tlm_get_flips": (C.c_int, [_H, C.c_int, C.c_int, _IP, C.c_int]),
    "cf_tlm_get": (C.c_int, [_H, C.c_int, _DP, _DP, C.c_void_p, C.c_void_p, C.c_void_p]),
    "cf_debug_pow": (C.c_int, [C.c_int, _DP, C.c_double, _DP]),
    "cf_debug_prepare": (C.c_int, [C.POINTER(CfMesh), C.POINTER(CfParams), C.c_int, C.c_int, C.c_int, _IP, _IP, _IP,
                                   _IP, _IP, C.c_int, _IP, C.c_void_p, _IP, _IP]),
}

_lib = None


class CuteflowError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        msg = load().cf_error_string(status).decode() if _lib is not None else str(status)
        super().__init__(f"{what}: {msg} (status {status})")


def load(path: str | None = None):
    """Load libcuteflow_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    # One HIP runtime per process: torch loads its own libamdhip64.so.7; when
    # it is loaded first the library binds to it (same soname), otherwise
    # torch later finds a foreign runtime under its soname and sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class CfCgnsInfo(C.Structure):
    """Mirror of struct cf_cgns_info (include/cuteflow_cgns.h)."""

    _fields_ = [(n, C.c_int) for n in ("nnt", "nelt", "ndi", "ndo", "nzones", "zone", "has_manning",
                                        "partitioned")]


_CP = C.c_char_p
CGNS_SIGNATURES = {
    "cf_cgns_mesh_info": (C.c_int, [_CP, C.POINTER(CfCgnsInfo)]),
    "cf_cgns_read_mesh": (C.c_int, [_CP, C.c_int, _CP, C.c_int, _CP, C.c_int, C.c_int, C.c_int,
                                    _DP, _IP, _DP, _IP, _IP, _IP, _IP]),
    "cf_cgns_write_mesh": (C.c_int, [_CP, C.c_int, C.c_int, _DP, _IP, C.c_int, C.c_int, _IP, _IP,
                                     C.c_int, C.c_int, _IP, _IP, C.c_int, _IP, _DP]),
    "cf_cgns_export_init": (C.c_int, [_CP, C.c_int, _DP]),
    "cf_cgns_export": (C.c_int, [_CP, C.c_int, C.c_int, _DP, _DP, _DP, C.c_double, C.c_double]),
    "cf_cgns_error": (C.c_char_p, []),
}

_cgns = None


def load_cgns(path: str | None = None):
    """Load libcuteflow_cgns.so (CGNS/HDF5 files, host only); raises if it
    has not been built (it needs an HDF5 installation at build time)."""
    global _cgns
    if _cgns is not None:
        return _cgns
    path = path or CGNS_LIB_PATH
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (cuteflow_amd.build.build_cgns(), "
                           "which needs HDF5 headers and library)")
    lib = C.CDLL(path)
    for name, (res, args) in CGNS_SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _cgns = lib
    return lib


def check_cgns(status: int, what: str) -> None:
    if status != CF_OK:
        raise CuteflowError(status, f"{what}: {load_cgns().cf_cgns_error().decode()}")


def check(status: int, what: str) -> None:
    if status != CF_OK:
        raise CuteflowError(status, what)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_IP)
