"""ctypes binding of libpfhip.so (the HIP C ABI declared in include/periodicflow.h).

The library is loaded from this package directory.  There is no fallback:
if the shared object is missing or the HIP runtime cannot load it, every
entry point raises, so nothing silently runs elsewhere.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfhip.so")

PF_OK, PF_EINVAL, PF_ENOMEM, PF_EHIP, PF_EUNSUPPORTED = 0, -1, -2, -3, -4
PF_FP32, PF_FP64 = 0, 1
PF_DEALIAS_PAD32, PF_DEALIAS_TRUNC23 = 0, 1
INTEGRATOR_IDS = {"explicit_pc": 0, "implicit_pc": 1, "rk2": 2, "rk3": 3, "rk4": 4}
PF_NOISE_THERMAL, PF_NOISE_CORRELATED = 0, 1
PF_FORCING_NONE, PF_FORCING_CONSTANT_POWER = 0, 1
PASS_NAMES = ("inv_axis0_curl", "inv_axis1", "row_c2r_cross_r2c", "fwd_axis1", "fwd_axis0", "stage", "band",
              "fwd_axis0_stage")


class PfDesc(C.Structure):
    _fields_ = [
        ("dim", C.c_int32),
        ("n", C.c_int32 * 3),
        ("length", C.c_double * 3),
        ("precision", C.c_int32),
        ("dealias", C.c_int32),
        ("mask_nyquist", C.c_int32),
        ("integrator", C.c_int32),
        ("nonlinear", C.c_int32),
        ("noise_type", C.c_int32),
        ("forcing", C.c_int32),
        ("batch", C.c_int32),
        ("viscosity", C.c_double),
        ("dt", C.c_double),
        ("noise_mag", C.c_double),
        ("forcing_power", C.c_double),
        ("forcing_kmin", C.c_double),
        ("forcing_kmax", C.c_double),
    ]


class PfPlanInfo(C.Structure):
    _fields_ = [
        ("n_fourier", C.c_int64),
        ("n_physical", C.c_int64),
        ("padded", C.c_int32 * 3),
        ("noise_components", C.c_int32),
        ("workspace_bytes", C.c_int64),
        ("state_bytes", C.c_int64),
    ]


# name -> (restype, argtypes)
_P = C.c_void_p
_SIGS = {
    "pf_plan_create": (C.c_int, [C.POINTER(PfDesc), C.POINTER(C.c_void_p)]),
    "pf_plan_destroy": (C.c_int, [_P]),
    "pf_plan_get_info": (C.c_int, [_P, C.POINTER(PfPlanInfo)]),
    "pf_set_seeds": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "pf_set_step_counter": (C.c_int, [_P, C.c_int64]),
    "pf_get_step_counter": (C.c_int64, [_P]),
    "pf_add_step_counter": (C.c_int, [_P, C.c_int64]),
    "pf_set_member": (C.c_int, [_P, C.c_int32, C.c_uint64, C.c_int64]),
    "pf_set_active": (C.c_int, [_P, C.c_int32]),
    "pf_get_member_steps": (C.c_int, [_P, C.POINTER(C.c_int64)]),
    "pf_set_linear_operator": (C.c_int, [_P, _P, _P]),
    "pf_set_correlation": (C.c_int, [_P, _P, _P]),
    "pf_forward": (C.c_int, [_P, _P, _P, C.c_int32, _P]),
    "pf_backward": (C.c_int, [_P, _P, _P, C.c_int32, _P]),
    "pf_project": (C.c_int, [_P, _P, _P]),
    "pf_mask_zero_mode": (C.c_int, [_P, _P, _P]),
    "pf_rhs_nonlinear": (C.c_int, [_P, _P, _P, _P]),
    "pf_rhs_nonlinear_noalloc": (C.c_int, [C.POINTER(PfDesc), _P, _P, _P]),
    "pf_step": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "pf_step_accumulate_sk": (C.c_int, [_P, _P, C.c_int64, _P, _P]),
    "pf_step_multi": (C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_int32, C.c_int64,
                                C.POINTER(C.c_int32), _P]),
    "pf_compute_rhs": (C.c_int, [_P, _P, C.c_double, _P, _P, _P, _P, _P, _P]),
    "pf_stage_update": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P, _P]),
    "pf_draw_noise": (C.c_int, [_P, _P, _P]),
    "pf_diagnostics": (C.c_int, [_P, _P, _P, _P]),
    "pf_spectrum": (C.c_int, [_P, _P, _P, C.c_int32, _P]),
    "pf_accumulate_sk": (C.c_int, [_P, _P, _P, _P]),
    "pf_profile_enable": (C.c_int, [_P, C.c_int32]),
    "pf_profile_read": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32]),
    "pf_profile_select": (C.c_int, [_P, C.c_uint32]),
    "pf_set_fusion": (C.c_int, [_P, C.c_int32]),
    "pf_set_tiles": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "pf_get_tiles": (C.c_int, [_P, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "pf_autotune": (C.c_int, [_P, _P, C.c_int32, C.POINTER(C.c_double), _P]),
    "pf_last_error": (C.c_char_p, []),
    "pf_version": (C.c_char_p, []),
}

EXPORTED = tuple(_SIGS)

_lib = None


def load(path: str | None = None):
    """Load (once) and return the ctypes handle; raises if it cannot."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"periodicflow: HIP library not found at {p}; build it with "
            "`python periodic-flow_amd/build_lib.py`"
        )
    lib = C.CDLL(p, mode=C.RTLD_GLOBAL)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class PfError(RuntimeError):
    pass


def check(rc: int, what: str = ""):
    if rc == PF_OK:
        return
    msg = load().pf_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == PF_EINVAL:
        raise ValueError(text)
    if rc == PF_ENOMEM:
        raise MemoryError(text)
    if rc == PF_EUNSUPPORTED:
        raise NotImplementedError(text)
    raise PfError(text)


def call(name: str, *args):
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc
