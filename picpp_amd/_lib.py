"""ctypes binding of libpicpp_amd.so (the C ABI declared in include/picpp_amd.h).

The library is loaded from picpp_amd/lib/ (built in-tree by picpp_amd.build). There is
no fallback: if the library is missing or fails to load, loading raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpicpp_amd.so")

PIC_OK = 0
ERRORS = {-1: "argument", -2: "device", -3: "memory", -4: "particle", -5: "solver"}
ERRORS_BY_NAME = {v: k for k, v in ERRORS.items()}

PRECISION_FP64, PRECISION_FP32 = 0, 1
STOP_REFERENCE, STOP_TRUE_RESIDUAL = 0, 1
ORTHO_MGS, ORTHO_CGS2 = 0, 1
SOLVER_FP64, SOLVER_MIXED = 0, 1
BOUNDARY_PERIODIC, BOUNDARY_WALL = 0, 1
PRESSURE_REFERENCE, PRESSURE_CELL = 0, 1
SCHEME_IMPLICIT, SCHEME_EXPLICIT = 0, 1

ARRAYS = {"E": 0, "Etheta": 1, "Enew": 2, "B": 3, "Bnew": 4, "J": 5, "mu": 6, "divPi": 7, "rhohat": 8, "Pi": 9,
          "density": 10, "rhs": 11, "solution": 12,
          "Ex": 13, "Ey": 14, "Ez": 15, "Bx": 16, "By": 17, "Bz": 18, "Jx": 19, "Jy": 20, "Jz": 21}

# Every symbol the header declares, with (restype, argtypes).
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_V = ctypes.c_void_p
_F = ctypes.POINTER(ctypes.c_float)


class PicConfig(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("dt", ctypes.c_double), ("nspecies", ctypes.c_int),
                ("mass", ctypes.c_double * 8), ("charge", ctypes.c_double * 8), ("weight", ctypes.c_double * 8),
                ("capacity", ctypes.c_longlong), ("precision", ctypes.c_int), ("stop_rule", ctypes.c_int),
                ("ortho", ctypes.c_int), ("gmres_tolerance", ctypes.c_double), ("gmres_max_iterations", ctypes.c_int),
                ("solver_precision", ctypes.c_int), ("boundary_x", ctypes.c_int), ("e_right", ctypes.c_double * 3),
                ("b_right", ctypes.c_double * 3), ("inject", ctypes.c_int), ("flow", ctypes.c_double * 3),
                ("ppc", ctypes.c_int * 8), ("temperature", ctypes.c_double * 8), ("seed", ctypes.c_ulonglong),
                ("pressure", ctypes.c_int), ("smoothing_count", ctypes.c_int), ("smoothing_parameter", ctypes.c_double),
                ("scheme", ctypes.c_int), ("divergence_cleanup", ctypes.c_int), ("filter_cut_wave_number", ctypes.c_int), ("run_kT", ctypes.c_double)]


class PicSolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("restarts", ctypes.c_int), ("estimated_residual", ctypes.c_double),
                ("true_residual", ctypes.c_double), ("seconds", ctypes.c_double), ("preconditioned", ctypes.c_int),
                ("fallbacks", ctypes.c_int)]


GENERAL_VALUES = 29  # PIC_GENERAL_VALUES


class PicBackupInfo(ctypes.Structure):
    _fields_ = [("input_type", ctypes.c_int), ("particles_per_bin", ctypes.c_int * 6),
                ("concentration", ctypes.c_double * 6), ("temperature", ctypes.c_double),
                ("plasma_period", ctypes.c_double), ("scale_factor", ctypes.c_double), ("time", ctypes.c_double),
                ("max_time", ctypes.c_double), ("current_iteration", ctypes.c_int), ("max_iteration", ctypes.c_int),
                ("size", ctypes.c_double * 3), ("debug_mode", ctypes.c_int), ("maxwell_matrix_size", ctypes.c_int),
                ("ext_j", ctypes.c_double), ("V0", ctypes.c_double * 3), ("E0", ctypes.c_double * 3),
                ("B0", ctypes.c_double * 3), ("species_type", ctypes.c_int * 8), ("particle_dx", ctypes.c_double)]


SIGNATURES = {
    "pic_create": (ctypes.c_int, [ctypes.POINTER(PicConfig), _V, ctypes.POINTER(_V)]),
    "pic_destroy": (None, [_V]),
    "pic_last_error": (ctypes.c_char_p, [_V]),
    "pic_upload_fields": (ctypes.c_int, [_V, _D, _D]),
    "pic_upload_particles": (ctypes.c_int, [_V, ctypes.c_longlong, _I, _I, _D, _D]),
    "pic_num_particles": (ctypes.c_longlong, [_V]),
    "pic_download_particles": (ctypes.c_int, [_V, _I, _I, _D, _D]),
    "pic_download_array": (ctypes.c_longlong, [_V, ctypes.c_int, _D]),
    "pic_device_array": (ctypes.c_int, [_V, ctypes.c_int, ctypes.POINTER(_D), ctypes.POINTER(ctypes.c_longlong)]),
    "pic_initialize": (ctypes.c_int, [_V]),
    "pic_rotation_and_moments": (ctypes.c_int, [_V]),
    "pic_field_rhs": (ctypes.c_int, [_V]),
    "pic_maxwell_solve": (ctypes.c_int, [_V, ctypes.POINTER(PicSolveInfo)]),
    "pic_update_b": (ctypes.c_int, [_V]),
    "pic_push": (ctypes.c_int, [_V]),
    "pic_commit_fields": (ctypes.c_int, [_V]),
    "pic_energy": (ctypes.c_int, [_V, _D]),
    "pic_step": (ctypes.c_int, [_V, ctypes.POINTER(PicSolveInfo)]),
    "pic_gmres_solve": (ctypes.c_int, [_V, _D, _D, ctypes.POINTER(PicSolveInfo)]),
    "pic_apply_operator": (ctypes.c_int, [_V, _D, _D]),
    "pic_apply_operator_factored": (ctypes.c_int, [_V, ctypes.c_int, _V, _V]),
    "pic_synchronize": (ctypes.c_int, [_V]),
    "pic_cleanup_divergence": (ctypes.c_int, [_V, ctypes.POINTER(PicSolveInfo)]),
    "pic_cleanup_divergence_magnetic": (ctypes.c_int, [_V, ctypes.POINTER(PicSolveInfo)]),
    "pic_filter_fields": (ctypes.c_int, [_V, ctypes.c_int]),
    "pic_energy_noalloc": (ctypes.c_int, [ctypes.POINTER(PicConfig), _V, _V, _V, ctypes.POINTER(ctypes.c_longlong), _V, _V]),
    "pic_philox": (ctypes.c_int, [_V] + [ctypes.c_ulonglong] * 6 + [ctypes.c_longlong, _V]),
    "pic_upload_particles_device": (ctypes.c_int, [_V, ctypes.POINTER(ctypes.c_longlong), _D, _D, _I]),
    "pic_profile_enable": (ctypes.c_int, [_V, ctypes.c_int]),
    "pic_profile_read": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_char_p, _D, ctypes.POINTER(ctypes.c_longlong), _D]),
    "pic_precond_create": (ctypes.c_int, [_V]),
    "pic_precond_param_count": (ctypes.c_longlong, []),
    "pic_precond_set_params": (ctypes.c_int, [_V, _F, ctypes.c_double]),
    "pic_precond_get_params": (ctypes.c_int, [_V, _F, _D]),
    "pic_precond_apply": (ctypes.c_int, [_V, _V, _V, _V]),
    "pic_precond_gradient": (ctypes.c_int, [_V, _V, _V, _V, _D, _F]),
    "pic_precond_log": (ctypes.c_int, [_V, ctypes.c_int]),
    "pic_precond_train": (ctypes.c_int, [_V, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, _D, _D]),
    "pic_precond_use": (ctypes.c_int, [_V, ctypes.c_int]),
    "pic_upload_yee": (ctypes.c_int, [_V, ctypes.POINTER(_D)]),
    "pic_explicit_step": (ctypes.c_int, [_V]),
    "pic_checkpoint_save": (ctypes.c_int, [_V, ctypes.c_char_p]),
    "pic_checkpoint_load": (ctypes.c_int, [_V, ctypes.c_char_p]),
    "pic_write_backup": (ctypes.c_int, [_V, ctypes.POINTER(PicBackupInfo)] + [ctypes.c_char_p] * 4),
    "pic_general": (ctypes.c_int, [_V, ctypes.POINTER(PicBackupInfo), ctypes.POINTER(ctypes.c_double)]),
    "pic_write_general": (ctypes.c_int, [_V, ctypes.POINTER(PicBackupInfo), ctypes.c_char_p]),
}

_lib = None


def load():
    """Load the HIP library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libpicpp_amd.so is not built (run python -m picpp_amd.build): %s" % LIB_PATH)
    _share_torch_hip_runtime()
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _share_torch_hip_runtime():
    """A process holds one HIP runtime (libamdhip64.so.7): the library's dependency binds to the
    copy already loaded. When torch is installed its bundled runtime is brought up first, so
    that torch and the library share it; the other order leaves torch without a usable device.
    Device buffers from torch are the library's inputs in tests and benchmarks."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


class PicError(RuntimeError):
    pass


def check(rc: int, ctx=None, what: str = ""):
    if rc < 0:
        msg = load().pic_last_error(ctx).decode() if ctx else ""
        raise PicError("%s failed (%s): %s" % (what, ERRORS.get(rc, rc), msg))
    return rc
