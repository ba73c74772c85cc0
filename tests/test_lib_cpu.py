"""The HIP library on the CPU side: it loads, exports every symbol its header
declares, and rejects bad arguments without touching a GPU."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np
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


class _FakeDb:
    """Stands in for a Database in argument checks that fail before any library call."""

    def __init__(self, shape):
        self.shape = shape
        self.handle = None
        self.calls = []

    def load_ntt(self, payload, check=True):
        self.calls.append(("ntt", payload.shape))

    def load_bytes(self, payload):
        self.calls.append(("bytes", payload.shape))


def _arrays(s):
    return (np.zeros(s.query_words, np.uint32), np.zeros(s.left_words, np.uint32),
            np.zeros(s.right_words, np.uint32), np.zeros(s.fold_words, np.uint32))


@pytest.mark.parametrize("which", [0, 1, 2, 3])
def test_host_inputs_of_the_wrong_size_are_refused(which):
    """Query.load / Batch.load / answer / Keys check every host array against the
    shape before the library copies the byte count the shape implies."""
    from spiral_pir import Shape
    from spiral_pir.server import Batch, Keys, Query, answer

    s = Shape(6, 2)
    db = _FakeDb(s)
    arrs = list(_arrays(s))
    arrs[which] = arrs[which][:-1]
    q = Query.__new__(Query)
    q.db, q._h = db, None
    b = Batch.__new__(Batch)
    b.db, b._h, b._hold = db, None, []
    names = ("query", "left", "right", "v_folding")
    with pytest.raises(ValueError, match=names[which]):
        q.load(*arrs)
    with pytest.raises(ValueError, match=names[which]):
        b.load(0, *arrs)
    with pytest.raises(ValueError, match=names[which]):
        answer(db, *arrs)
    if which in (1, 2):
        with pytest.raises(ValueError, match=names[which]):
            Keys(db, arrs[1], arrs[2])
    if which == 0:
        with pytest.raises(ValueError, match="query"):
            q.load_packed(arrs[0], None)
        with pytest.raises(ValueError, match="query"):
            b.load_packed(0, [arrs[0]], [None])
    if which in (0, 3):
        with pytest.raises(ValueError):
            b.load_many(0, [arrs[0]], [None], [arrs[3]])


def test_bulk_loads_need_open_keys():
    from spiral_pir import Shape
    from spiral_pir.server import Batch

    s = Shape(6, 2)
    b = Batch.__new__(Batch)
    b.db, b._h, b._hold = _FakeDb(s), None, []
    q, _, _, v = _arrays(s)
    with pytest.raises(ValueError, match="Keys"):
        b.load_many(0, [q], [object()], [v])
    with pytest.raises(ValueError, match="Keys"):
        b.load_packed(0, [q], [object()], asynchronous=True)
    assert b._hold == []  # nothing held for a load that never reached the library
