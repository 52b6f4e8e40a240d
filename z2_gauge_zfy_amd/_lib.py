"""ctypes loader of the gfx950 engine libz2dqmc.so (C ABI in include/z2dqmc.h).

The library is the product path: when it is missing this module raises
instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libz2dqmc.so")

Z2_PREC_FP64 = 0
Z2_PREC_MIXED = 1
Z2_QR_BLOCKED = 0
Z2_QR_PIVOTED = 1
Z2_BETA_PER_CHAIN = -1
Z2_STAB_REFERENCE = 0
Z2_STAB_DSPLIT = 1
Z2_EXEC_EAGER = 0
Z2_EXEC_GRAPH = 1
Z2_FLIP_DELAYED = 0
Z2_FLIP_REGISTER = 1

# kernel classes for z2_profile_kernel
KERNEL_IDS = {"qrp": 1, "flip": 2, "kin_rows": 3, "kin_cols": 4, "gemm": 5, "orgqr": 6, "trsm": 7,
              "obs": 8, "obs_tau": 9, "qt_apply": 10}


class Z2Params(C.Structure):
    _fields_ = [("Nlx", C.c_int32), ("Nly", C.c_int32), ("Ltrot", C.c_int32), ("Nwrap", C.c_int32),
                ("Nsweep", C.c_int32), ("precision", C.c_int32)]


class Z2ChainDesc(C.Structure):
    _fields_ = [("RT", C.c_double), ("J", C.c_double), ("h", C.c_double), ("mu", C.c_double),
                ("Q", C.c_double), ("Beta", C.c_double), ("seed", C.c_int32),
                ("absolute", C.c_int32)]


class Z2BatchChain(C.Structure):
    _fields_ = [("Nlx", C.c_int32), ("Nly", C.c_int32), ("Ltrot", C.c_int32), ("Nwrap", C.c_int32),
                ("d", Z2ChainDesc)]


class Z2Error(RuntimeError):
    pass


_lib = None

# name -> (restype, argtypes)
_SIGS = {
    "z2_last_error": (C.c_char_p, []),
    "z2_ctx_create": (C.c_int, [C.POINTER(Z2Params), C.c_int32, C.POINTER(Z2ChainDesc),
                                C.POINTER(C.c_void_p)]),
    "z2_ctx_destroy": (C.c_int, [C.c_void_p]),
    "z2_ctx_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                              C.POINTER(C.c_int32)]),
    "z2_init_fields": (C.c_int, [C.c_void_p]),
    "z2_set_sigma": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_int32]),
    "z2_get_sigma": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "z2_get_seed": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "z2_pre": (C.c_int, [C.c_void_p]),
    "z2_sweep": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_enable_tau": (C.c_int, [C.c_void_p]),
    "z2_sweep_tau": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "z2_dyn_measure": (C.c_int, [C.c_void_p]),
    "z2_get_obs_tau": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "z2_get_dyn_drift": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                   C.POINTER(C.c_double)]),
    "z2_sweep_shadow": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "z2_sweep_shadow_tau": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                      C.POINTER(C.c_double)]),
    "z2_sweep_shadow_trace": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                        C.POINTER(C.c_int32), C.c_int32, C.POINTER(C.c_int32)]),
    "z2_obs_equal_ave": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_sweep_reset": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_prop_K_pre": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_prop_T_pre": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_prop_T_L": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_prop_T_R": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_propgr_K_R": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_propgr_T_R": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_localK_prop_L": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_localK_prop_R": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_dyn_reset": (C.c_int, [C.c_void_p]),
    "z2_obs_tau_reset": (C.c_int, [C.c_void_p]),
    "z2_obs_tau_calc": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_obs_tau_ave": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_wrap_tau": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_wrap_pre": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_wrap_L": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "z2_wrap_R": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "z2_slice_L": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_slice_R": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_obs_equal_calc": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_obs_reset": (C.c_int, [C.c_void_p]),
    "z2_sync": (C.c_int, [C.c_void_p]),
    "z2_get_green": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "z2_set_green": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "z2_get_drift": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                               C.POINTER(C.c_double)]),
    "z2_get_obs_equal": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "z2_get_counters": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]),
    "z2_get_lambda": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "z2_set_lambda": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]),
    "z2_set_sigma_therm": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32]),
    "z2_init_fields_therm": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "z2_set_is_beta": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "z2_set_stab_mode": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_get_is_beta": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_int32)]),
    "z2_global_sweep": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_get_global_counters": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64),
                                         C.POINTER(C.c_double)]),
    "z2_therm": (C.c_int, [C.c_void_p]),
    "z2_get_therm_counters": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_uint64)]),
    "z2_get_stream": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "z2_set_delay": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_set_flip_variant": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_set_qr_mode": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_profile_kernel": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_profile_read": (C.c_int, [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "z2_profile_work": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
    "z2_profile_spans": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64,
                                   C.POINTER(C.c_int64)]),
    "z2_set_kin_tiles": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "z2_set_u_precision": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_batch_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                  C.POINTER(C.c_void_p)]),
    "z2_batch_destroy": (C.c_int, [C.c_void_p]),
    "z2_batch_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "z2_batch_chain_ctx": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_int32)]),
    "z2_batch_group_ctx": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]),
    "z2_batch_init_fields": (C.c_int, [C.c_void_p]),
    "z2_batch_enable_tau": (C.c_int, [C.c_void_p]),
    "z2_batch_pre": (C.c_int, [C.c_void_p]),
    "z2_batch_sweep": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "z2_batch_sync": (C.c_int, [C.c_void_p]),
    "z2_batch_set_exec": (C.c_int, [C.c_void_p, C.c_int32]),
    "z2_batch_get_exec": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "z2_batch_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_char_p]),
    "z2_batch_load_tuning": (C.c_int, [C.c_void_p, C.c_char_p]),
    "z2_qdrp_batched": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_void_p]),
    "z2_qdrp_batched_noalloc": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p]),
    # kernel test hooks (include/z2dqmc_test.h)
    "z2_test_gemm": (C.c_int, [C.c_int32, C.POINTER(C.c_double), C.c_int32,
                               C.POINTER(C.c_double), C.c_int32, C.c_int32, C.c_int32,
                               C.POINTER(C.c_double)]),
    "z2_test_qrp": (C.c_int, [C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_int32), C.POINTER(C.c_double),
                              C.POINTER(C.c_double)]),
    "z2_test_qr_blocked": (C.c_int, [C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double)]),
    "z2_test_copy": (C.c_int, [C.c_int64, C.c_int32]),
    "z2_test_trsm": (C.c_int, [C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                               C.c_int32]),
}


def build() -> None:
    subprocess.run(["make", "-s", "-C", PKG], check=True)


def lib():
    """Load the engine; raises Z2Error when the HIP library is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise Z2Error(f"{LIB_PATH} is missing: build it with `make -C {PKG}` "
                          "(the engine has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise Z2Error(f"z2dqmc error {rc}: {lib().z2_last_error().decode()}")


def exported_symbols():
    return list(_SIGS)
