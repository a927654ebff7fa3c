"""CPU checks of the C-ABI library: it loads, and it exports every function
include/dsht.h declares with the signatures the ctypes binding uses.  No
compute calls (there is no GPU here)."""
import ctypes as ct
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    src = open(os.path.join(ROOT, "include", "dsht.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s+(\w+)\s*\(", src, flags=re.M)
    return sorted(set(names))


def test_header_lists_expected_entry_points():
    names = _declared_functions()
    for n in ("make_table", "do_transform", "dsht_plan_create", "dsht_transform",
              "dsht_plan_destroy", "dsht_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from directsht_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libdsht.so not built (run python -m directsht_amd.build)")
    lib = ct.CDLL(_lib.LIB_PATH)
    for n in _declared_functions():
        assert hasattr(lib, n), n
    assert set(_declared_functions()) == set(_lib.SIGNATURES)
    _lib.load()  # declares argtypes for all of them


def test_last_error_callable_without_gpu():
    from directsht_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libdsht.so not built")
    assert isinstance(_lib.load().dsht_last_error(), bytes)


def test_estimate_without_gpu():
    """dsht_plan_estimate is pure host arithmetic (used for admission)."""
    from directsht_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libdsht.so not built")
    lib = _lib.load()
    tb, wb = ct.c_int64(), ct.c_int64()
    assert lib.dsht_plan_estimate(1000, 2048, _lib.DSHT_FP32, 10_000_000, ct.byref(tb), ct.byref(wb)) == 0
    # FP32 plans hold no Y table, only the recurrence's coefficients: 16 bytes
    # per (m, k < Nl - m + 32), plus the rotation constants (28 B per (l, m))
    # and the group metadata
    coef = 16 * (1000 * 1032 - 1000 * 999 // 2)
    assert coef + 500500 * 28 <= tb.value < coef + 500500 * 28 + 2_000_000
    assert wb.value > 10_000_000 * 52
    # FP64 plans: the Y-only table, Nlm * Nx * 8 bytes (+ small metadata)
    assert lib.dsht_plan_estimate(1000, 2048, _lib.DSHT_FP64, 10_000_000, ct.byref(tb), ct.byref(wb)) == 0
    assert abs(tb.value - 500500 * 2048 * 8) / (500500 * 2048 * 8) < 0.01
