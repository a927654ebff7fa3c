"""ctypes binding of libdsht.so (declared in include/dsht.h).

torch is imported before the library is opened: torch ships its own HIP
runtime (libamdhip64.so, SONAME libamdhip64.so.7) and libdsht.so must bind to
that same runtime so torch's device pointers and streams are valid in it.
Loading fails loudly when the library has not been built -- there is no CPU
fallback.
"""
import ctypes as ct
import os
import threading

import torch  # noqa: F401  (must precede the CDLL: shared HIP runtime)

HERE = os.path.dirname(os.path.abspath(__file__))
# DSHT_LIBRARY selects an alternative build of the same C ABI (for A/B timing)
LIB_PATH = os.environ.get("DSHT_LIBRARY") or os.path.join(HERE, "libdsht.so")

DSHT_FP64, DSHT_FP32 = 0, 1
DSHT_MODE_CALL, DSHT_MODE_BASIC = 0, 1
DSHT_STATUS_EQU_RANGE, DSHT_STATUS_POLAR_RANGE = 1, 2
STAGES = ("points", "scatter", "fold", "contract", "rotate")
DSHT_OPT_SCATTER, DSHT_OPT_CHUNK, DSHT_OPT_CONTRACT = 1, 2, 3
DSHT_SCATTER_MATRIX, DSHT_SCATTER_VECTOR, DSHT_SCATTER_SPREAD, DSHT_SCATTER_AUTO = 0, 1, 2, 3
DSHT_CONTRACT_MATRIX, DSHT_CONTRACT_VECTOR = 0, 1
MAX_SETS = 64  # sets per batched call (dsht_transform_sets / dsht_transform_many)

# every entry point of include/dsht.h: name -> (restype, argtypes)
_dp = ct.POINTER(ct.c_double)
_vp = ct.c_void_p
SIGNATURES = {
    "make_table": (ct.c_int, [ct.c_int, ct.c_int, ct.c_double, _dp, _dp]),
    "do_transform": (ct.c_int, [ct.c_int, ct.c_int, ct.c_double, _dp, _dp, ct.c_int,
                                _dp, _dp, _dp, _dp, _dp]),
    "dsht_plan_create": (ct.c_int, [ct.c_int, ct.c_int, ct.c_double, ct.c_int, ct.POINTER(_vp)]),
    "dsht_plan_destroy": (ct.c_int, [_vp]),
    "dsht_plan_create_from_tables": (ct.c_int, [ct.c_int, ct.c_int, ct.c_double, ct.c_int, _vp,
                                                ct.POINTER(_vp)]),
    "dsht_plan_clone": (ct.c_int, [_vp, ct.POINTER(_vp)]),
    "dsht_plan_reserve": (ct.c_int, [_vp, ct.c_int64, ct.c_int]),
    "dsht_plan_reserve_for": (ct.c_int, [_vp, ct.c_int64, ct.c_int, ct.c_int]),
    "dsht_plan_footprint": (ct.c_int, [_vp, ct.POINTER(ct.c_int64)]),
    "dsht_plan_estimate": (ct.c_int, [ct.c_int, ct.c_int, ct.c_int, ct.c_int64,
                                      ct.POINTER(ct.c_int64), ct.POINTER(ct.c_int64)]),
    "dsht_plan_get_tables": (ct.c_int, [_vp, _vp, _vp]),
    "dsht_transform": (ct.c_int, [_vp, ct.c_int64, _vp, _vp, _vp, _vp, _vp, ct.c_int, _vp]),
    "dsht_transform_noalloc": (ct.c_int, [_vp, ct.c_int64, _vp, _vp, _vp, _vp, _vp, ct.c_int,
                                          _vp]),
    "dsht_transform_sets": (ct.c_int, [_vp, ct.c_int64, _vp, _vp, ct.c_int, _vp, ct.c_int64, _vp,
                                       _vp, ct.c_int, _vp]),
    "dsht_transform_many": (ct.c_int, [_vp, ct.c_int, ct.POINTER(ct.c_int64), ct.POINTER(_vp),
                                       ct.POINTER(_vp), ct.POINTER(_vp), _vp, _vp, ct.c_int, _vp]),
    "dsht_plan_configure": (ct.c_int, [_vp, ct.c_int, ct.c_int]),
    "dsht_plan_engine": (ct.c_int, [_vp, ct.c_int64, _vp]),
    "dsht_plan_set_timing": (ct.c_int, [_vp, ct.c_int]),
    "dsht_plan_get_timing": (ct.c_int, [_vp, ct.POINTER(ct.c_double), ct.POINTER(ct.c_int64)]),
    "dsht_plan_debug_rhs": (ct.c_int, [_vp, _vp, ct.c_int64]),
    "dsht_threej000_table": (ct.c_int, [ct.c_int, _vp, _vp]),
    "dsht_mode_coupling": (ct.c_int, [ct.c_int, _vp, ct.c_int64, _vp, _vp]),
    "dsht_synthesis_rings": (ct.c_int, [ct.c_int, _vp, ct.c_int, _vp, _vp, _vp]),
    "dsht_alm2cl": (ct.c_int, [ct.c_int, ct.c_int, _vp, ct.c_int64, _vp, ct.c_int, _vp]),
    "dsht_last_error": (ct.c_char_p, []),
}

_lock = threading.Lock()
_lib = None


def load():
    """Open libdsht.so once and declare its signatures."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is not built; run `python -m directsht_amd.build` "
                    "(the direct SHT has no CPU fallback)")
            lib = ct.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc, what):
    if rc != 0:
        msg = load().dsht_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (code {rc}): {msg}")
