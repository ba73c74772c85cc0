"""The HIP library on the CPU side: it loads, exports every symbol its header
declares, and rejects bad arguments without touching a GPU."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "spiral_pir.h")


def declared_functions() -> list[str]:
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sp_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("sp_ctx_create", "sp_ntt_fwd", "sp_ntt_inv", "sp_from_ntt", "sp_automorph", "sp_gadget_extprod",
                 "sp_db_mul", "sp_fold", "sp_expand", "sp_db_create", "sp_query_run", "sp_answer",
                 "sp_client_query", "sp_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from spiral_pir import _lib

    L = _lib.load()
    for name in declared_functions():
        assert hasattr(L, name), f"{name} declared in spiral_pir.h but not exported"
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True, check=True)
    exported = set(re.findall(r" T (sp_\w+)", out.stdout))
    assert set(declared_functions()) <= exported


def test_python_binding_covers_header():
    from spiral_pir import _lib

    bound = set(_lib.SIGNATURES) | set(_lib._RET)
    assert set(declared_functions()) <= bound


def test_version_string():
    from spiral_pir import version

    assert "gfx950" in version()


def test_expansion_depth_matches_reference_formula():
    from spiral_pir import _lib

    L = _lib.load()
    g, st = ctypes.c_int(), ctypes.c_int()
    for (nu1, nu2, eg, est) in [(6, 2, 7, 4), (9, 6, 10, 6), (10, 7, 11, 6), (8, 4, 9, 5)]:
        assert L.sp_expansion_depth(nu1, nu2, 8, ctypes.byref(g), ctypes.byref(st)) == 0
        assert (g.value, st.value) == (eg, est)
    assert L.sp_expansion_depth(11, 6, 8, ctypes.byref(g), ctypes.byref(st)) == _lib.SP_ERR_ARG


def test_null_arguments_are_errors_not_crashes():
    from spiral_pir import _lib

    L = _lib.load()
    assert L.sp_ctx_create(0, None) == _lib.SP_ERR_ARG
    assert b"NULL" in L.sp_last_error()
    assert L.sp_ntt_fwd(None, None, 1, 0, None) == _lib.SP_ERR_ARG
    assert L.sp_fold(None, None, 2, 8, None, None, None) == _lib.SP_ERR_ARG
    assert L.sp_db_mul(None, None, None, None, 6, 2, None) == _lib.SP_ERR_ARG
    assert L.sp_query_run(None, 7, None) == _lib.SP_ERR_ARG
    assert L.sp_ctx_destroy(None) == 0


def test_shape_helpers():
    from spiral_pir import Shape

    s = Shape(9, 6)
    assert (s.g, s.stop_round) == (10, 6)
    assert s.db_ntt_bytes == 2 ** 31
    assert s.db_device_bytes == 7 * 2 ** 28
    assert Shape(6, 2).db_device_bytes == Shape(6, 2).db_ntt_bytes
    assert s.records * s.record_bytes == 2 ** 28
