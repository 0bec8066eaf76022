"""ctypes binding of libwamgpu.so (include/wam_gpu.h).

The library is built in-tree by wam_amd/build.py.  Loading fails loudly if it
is missing: there is no CPU fallback for the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libwamgpu.so")


class WamConfig(ctypes.Structure):
    """struct wam_config of include/wam_gpu.h."""
    _fields_ = [
        ("kl", ctypes.c_int32), ("ml", ctypes.c_int32), ("fr1", ctypes.c_double), ("iref", ctypes.c_int32),
        ("iphys", ctypes.c_int32), ("shallow", ctypes.c_int32), ("spherical", ctypes.c_int32),
        ("refr_d", ctypes.c_int32), ("refr_c", ctypes.c_int32), ("iper", ctypes.c_int32),
        ("isnonlin", ctypes.c_int32), ("idelpro", ctypes.c_int32), ("idelt", ctypes.c_int32),
        ("ndepth", ctypes.c_int32), ("deptha", ctypes.c_double), ("depthd", ctypes.c_double),
        ("betamax", ctypes.c_double), ("zalp", ctypes.c_double), ("alpha", ctypes.c_double),
        ("cdfac", ctypes.c_double),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
        ("amosop", ctypes.c_int64), ("amonop", ctypes.c_int64), ("amowep", ctypes.c_int64),
        ("amoeap", ctypes.c_int64), ("xdella", ctypes.c_int64), ("xdello", ctypes.c_int64),
    ]


# Entry points declared in include/wam_gpu.h, with their ctypes signatures.
_P = ctypes.c_void_p
_I = ctypes.c_int32
SIGNATURES = {
    "wam_gpu_init": (_I, [ctypes.POINTER(_P), ctypes.POINTER(WamConfig), _P]),
    "wam_gpu_init_host": (_I, [ctypes.POINTER(_P), ctypes.POINTER(WamConfig), _P]),
    "wam_gpu_prepare_propagation": (_I, [_P, _P, _P, _P]),
    "wam_gpu_prepare_propagation_dev": (_I, [_P, _P, _P, _P]),
    "wam_gpu_propags": (_I, [_P, _I, _I, _I, _I, _P]),
    "wam_gpu_implsch": (_I, [_P, _I, _I, _I] + [_P] * 10 + [_I, _P, _P]),
    "wam_gpu_propags_dev": (_I, [_P, _P, _P]),
    "wam_gpu_implsch_dev": (_I, [_P] + [_P] * 8 + [_I, _P, _P, _P]),
    "wam_gpu_implsch_dev_noalloc": (_I, [_P] + [_P] * 8 + [_I, _P, _P, _P]),
    "wam_gpu_propags_ensemble_dev": (_I, [_P, _I, _P, _P]),
    "wam_gpu_implsch_ensemble_dev": (_I, [_P, _I] + [_P] * 8 + [_I, _P, _P, _P]),
    "wam_gpu_ensemble_run_dev": (_I, [_P, _I, _I] + [_P] * 9 + [_I]),
    "wam_gpu_autotune": (_I, [_P, _I] + [_P] * 8 + [_I]),
    "wam_gpu_get_tuning": (_I, [_P, _P]),
    "wam_gpu_set_tuning": (_I, [_P, _P]),
    "wam_gpu_get_tuning_log": (ctypes.c_int64, [_P, _P, ctypes.c_int64]),
    "wam_gpu_step_adjoint_dev": (_I, [_P, _I] + [_P] * 14),
    "wam_gpu_propags_adjoint_dev": (_I, [_P, _I, _P, _P]),
    "wam_gpu_wind_uv_dev": (_I, [_P] * 5 + [ctypes.c_int64]),
    "wam_gpu_wind_uv_adjoint_dev": (_I, [_P] * 7 + [ctypes.c_int64]),
    "wam_checkpoint_schedule": (ctypes.c_int64, [_I, _I, _P, ctypes.c_int64]),
    "wam_checkpoint_recomputed_steps": (ctypes.c_int64, [_I, _I]),
    "wam_gpu_multigrid_create": (_I, [ctypes.POINTER(_P), _I, _P]),
    "wam_gpu_multigrid_add_nest": (_I, [_P, _I, _I, _I, _P, _I, _P, _P, _P, _P]),
    "wam_gpu_multigrid_run": (_I, [_P, _I, _I]),
    "wam_gpu_multigrid_state": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "wam_gpu_multigrid_last_error": (ctypes.c_char_p, [_P]),
    "wam_gpu_multigrid_free": (None, [_P]),
    "wam_gpu_set_propagation_weights": (_I, [_P, _I]),
    "wam_gpu_set_source_options": (_I, [_P, _I, _I]),
    "wam_gpu_set_implsch_form": (_I, [_P, _I]),
    "wam_gpu_get_source_options": (_I, [_P, _P, _P]),
    "wam_gpu_set_stream": (_I, [_P, _P]),
    "wam_gpu_get_stream": (_I, [_P, _P]),
    "wam_gpu_output_parameters_dev": (_I, [_P] * 7 + [_I, _P]),
    "wam_gpu_output_parameters": (_I, [_P, _I, _I, _I] + [_P] * 6 + [_I, _P]),
    "wam_gpu_save_restart": (_I, [_P, ctypes.c_char_p, ctypes.c_char_p] + [_P] * 4),
    "wam_gpu_connect_restart": (_I, [_P, ctypes.c_char_p, _P] + [_P] * 7 + [_P]),
    "wam_restart_write": (_I, [ctypes.c_char_p, _I, _I, _I, ctypes.c_char_p] + [_P] * 7),
    "wam_restart_read": (_I, [ctypes.c_char_p, _I, _I, _I, _P] + [_P] * 7 + [_P]),
    "wam_gpu_boundary_output_dev": (_I, [_P, _I, _P, _P, _P, _P]),
    "wam_gpu_boundary_output": (_I, [_P, _I, _I, _I, _P, _I, _P, _P, _P]),
    "wam_gpu_boundary_input_dev": (_I, [_P, _I] + [_P] * 4 + [_I, ctypes.c_float, ctypes.c_float, _I] + [_P] * 7),
    "wam_gpu_boundary_input": (_I, [_P, _I, _I, _I, _P, _I] + [_P] * 4 + [_I, ctypes.c_float, ctypes.c_float, _I] +
                               [_P] * 6),
    "wam_boundary_write_header": (_I, [ctypes.c_char_p, _P]),
    "wam_boundary_append": (_I, [ctypes.c_char_p, _I, _I, _I, _P, _P, ctypes.c_char_p, _P, _P, _P]),
    "wam_boundary_read_header": (_I, [ctypes.c_char_p, _P, ctypes.POINTER(ctypes.c_int64)]),
    "wam_boundary_read": (_I, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), _I, _I, _I] + [_P] * 6),
    "wam_gpu_output_grids_dev": (_I, [_P, _P, _P, _P]),
    "wam_gpu_write_map": (_I, [_P, ctypes.c_char_p, _I, ctypes.c_char_p, ctypes.c_char_p, _P, _P]),
    "wam_map_write": (_I, [ctypes.c_char_p, _I, ctypes.c_char_p, ctypes.c_char_p, _I, _I] + [_P] * 5),
    "wam_gpu_set_timing": (_I, [_P, _I]),
    "wam_gpu_get_timing": (_I, [_P, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_int64)]),
    "wam_gpu_sync": (_I, [_P]),
    "wam_gpu_get_table": (ctypes.c_int64, [_P, ctypes.c_char_p, _P, ctypes.c_int64]),
    "wam_gpu_get_scalar": (ctypes.c_double, [_P, ctypes.c_char_p]),
    "wam_gpu_last_error": (ctypes.c_char_p, [_P]),
    "wam_gpu_finalize": (None, [_P]),
    "wam_gpu_eval_math": (_I, [_I, _P, _P, _P, ctypes.c_int64]),
}

_lib = None


def load():
    """Returns the loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch ships its own libamdhip64.so.7 / libhsa-runtime64.so; loading it
        # first makes this library bind to that same runtime (same SONAME), so
        # one process never holds two HIP runtimes competing for the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python wam_amd/build.py` (no CPU fallback exists)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
