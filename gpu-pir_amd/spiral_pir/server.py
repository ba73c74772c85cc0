"""Server side of the answer path: context, resident database, per-query state.

Mirrors what the reference's test mains assemble by hand (gen_twiddles_gpu,
Database / FirstDim / FoldingQueryStorage structs, the db_mul and fold
launches; gpu-pir/db_mul_test.cu:131-212, fold_test.cu:30-94) as objects
backed by device memory owned by libspiral_hip.so.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .params import T_CONV, Shape

STAGE_EXPAND = 1
STAGE_DB_MUL = 2
STAGE_FOLD = 4
STAGES_ALL = 7

# indices for Context.set_variant
VARIANT_EXPAND = 0
VARIANT_FOLD = 1
VARIANT_DB_MUL = 2

READ_RESPONSE = 0
READ_EXPANDED = 1
READ_DB_MUL = 2
READ_RAW_ACCS = 3
READ_V_FOLDING = 4


def _ptr(a) -> ctypes.c_void_p:
    if a is None:
        return ctypes.c_void_p(0)
    assert a.flags["C_CONTIGUOUS"] and a.dtype in (np.uint32, np.uint8)
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """Device context: twiddle tables, constant polynomials, a HIP stream.

    Replaces gen_twiddles_gpu (gpu-pir/host.cu:126-138)."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _lib.call("sp_ctx_create", device, ctypes.byref(h))
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def synchronize(self) -> None:
        _lib.call("sp_ctx_synchronize", self._h)

    def set_variant(self, stage: int, variant: int) -> None:
        """Kernel choice per stage (STAGE_* index 0/1/2 = expand/fold/db_mul); 0 = automatic."""
        _lib.call("sp_ctx_set_variant", self._h, stage, variant)

    def close(self) -> None:
        if self._h:
            _lib.call("sp_ctx_destroy", self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Database:
    """A database resident in HBM in NTT form (n2 sub-databases)."""

    def __init__(self, ctx: Context, shape: Shape):
        self.ctx = ctx
        self.shape = shape
        h = ctypes.c_void_p()
        _lib.call("sp_db_create", ctx.handle, shape.nu1, shape.nu2, shape.n2, ctypes.byref(h))
        self._h = h

    @property
    def handle(self):
        return self._h

    def load_bytes(self, records: np.ndarray) -> None:
        """records: uint8 [2^(nu1+nu2)][n2][2048]; converted to NTT form on the GPU."""
        r = np.ascontiguousarray(records, dtype=np.uint8)
        assert r.size == self.shape.records * self.shape.record_bytes
        _lib.call("sp_db_load_bytes", self._h, _ptr(r))

    def load_ntt(self, db_ntt: np.ndarray) -> None:
        """Reference Database layout per sub-database: [n2][2^nu2][2^nu1][2][N] u32."""
        d = np.ascontiguousarray(db_ntt, dtype=np.uint32)
        assert d.size == self.shape.records * self.shape.n2 * 2 * 2048
        _lib.call("sp_db_load_ntt", self._h, _ptr(d))

    def load_file(self, path: str) -> None:
        """A database container (spiral_pir.dbfile): plaintext records or the NTT
        form, memory-mapped and streamed to the device.  Its shape must be this
        database's."""
        from . import dbfile

        shape, kind, payload = dbfile.read(path)
        if (shape.nu1, shape.nu2, shape.n2) != (self.shape.nu1, self.shape.nu2, self.shape.n2):
            raise ValueError(f"{path}: shape ({shape.nu1}, {shape.nu2}, n2 = {shape.n2}) is not this database's "
                             f"({self.shape.nu1}, {self.shape.nu2}, n2 = {self.shape.n2})")
        if kind == dbfile.KIND_RECORDS:
            self.load_bytes(payload)
        else:
            self.load_ntt(payload)

    @property
    def device_bytes(self) -> int:
        return int(_lib.load().sp_db_device_bytes(self._h))

    def close(self) -> None:
        if self._h:
            _lib.call("sp_db_destroy", self._h)
            self._h = None


class Query:
    """Device-resident state of one query: its ciphertext, keys and workspace."""

    def __init__(self, db: Database):
        self.db = db
        s = db.shape
        h = ctypes.c_void_p()
        _lib.call("sp_query_create", db.handle, s.t_gsw, s.t_exp, ctypes.byref(h))
        self._h = h

    def load(self, query: np.ndarray, left: np.ndarray, right: np.ndarray, v_folding: np.ndarray,
             v_folding_neg: np.ndarray | None = None) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (query, left, right, v_folding)]
        neg = None if v_folding_neg is None else np.ascontiguousarray(v_folding_neg, dtype=np.uint32)
        _lib.call("sp_query_load", self._h, *[_ptr(a) for a in arrs], _ptr(neg))

    def load_packed(self, query: np.ndarray, keys: "Keys") -> None:
        """A packed query (client.query_packed) with resident keys that carry a
        conversion key: the fold vectors are converted on the device."""
        q = np.ascontiguousarray(query, dtype=np.uint32)
        _lib.call("sp_query_load_packed", self._h, _ptr(q), keys.handle)

    def set_packing(self, q2_bits: int) -> None:
        """Response packing + modulus switching for later runs (0: off)."""
        _lib.call("sp_query_set_packing", self._h, q2_bits)
        self.q2_bits = q2_bits

    def read_packed(self) -> np.ndarray:
        """The compact response (needs set_packing and keys with packing keys)."""
        out = np.zeros(self.db.shape.packed_response_bytes(self.q2_bits), np.uint8)
        _lib.call("sp_query_read_packed", self._h, _ptr(out))
        return out

    def run(self, stages: int = STAGES_ALL, stream=None) -> None:
        _lib.call("sp_query_run", self._h, stages, ctypes.c_void_p(stream or 0))

    def read(self, what: int = READ_RESPONSE) -> np.ndarray:
        s = self.db.shape
        if what == READ_RESPONSE:
            out = np.zeros((s.n2, 2, 2, 2048), np.uint32)
        elif what == READ_EXPANDED:
            out = np.zeros((1 << s.g, 2, 2, 2048), np.uint32)
        elif what == READ_DB_MUL:
            out = np.zeros((s.n2, s.cols, 2, 2, 2048), np.uint32)
        elif what == READ_RAW_ACCS:
            out = np.zeros((s.n2, s.cols, 2, 2, 2048), np.uint32)
        elif what == READ_V_FOLDING:
            out = np.zeros((s.nu2, 2, 2 * s.t_gsw, 2, 2048), np.uint32)
        else:
            raise ValueError(what)
        _lib.call("sp_query_read", self._h, what, _ptr(out))
        return out

    def close(self) -> None:
        if self._h:
            _lib.call("sp_query_destroy", self._h)
            self._h = None

    @property
    def handle(self):
        return self._h


class Batch:
    """`capacity` query slots against one database, answered together: one launch per
    stage for the whole batch, optionally replayed from a captured hipGraph."""

    def __init__(self, db: Database, capacity: int):
        self.db = db
        self.capacity = capacity
        s = db.shape
        h = ctypes.c_void_p()
        _lib.call("sp_batch_create", db.handle, capacity, s.t_gsw, s.t_exp, ctypes.byref(h))
        self._h = h

    def load(self, slot: int, query, left, right, v_folding) -> None:
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (query, left, right, v_folding)]
        _lib.call("sp_batch_load", self._h, slot, *[_ptr(a) for a in arrs])

    def run(self, n: int, stages: int = STAGES_ALL, stream=None) -> None:
        _lib.call("sp_batch_run", self._h, n, stages, ctypes.c_void_p(stream or 0))

    def graph_launch(self, n: int, stream=None) -> None:
        _lib.call("sp_batch_graph_launch", self._h, n, ctypes.c_void_p(stream or 0))

    def synchronize(self) -> None:
        _lib.call("sp_batch_synchronize", self._h)

    def poll(self) -> bool:
        """True when the last run has finished (does not wait)."""
        d = ctypes.c_int()
        _lib.call("sp_batch_poll", self._h, ctypes.byref(d))
        return bool(d.value)

    def tune(self, n: int) -> dict:
        """Autotune the database multiply for n queries (after one expansion has run)."""
        a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        _lib.call("sp_batch_tune", self._h, n, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return {"valu_ms": a.value, "mfma_ms": b.value if b.value >= 0 else None,
                "chosen": {1: "valu", 2: "mfma"}[c.value]}

    def load_many(self, first: int, queries, keys: "list[Keys]", v_folding) -> None:
        """Slots [first, first + len(keys)): queries[i] ([2][2][2048] u32), resident keys[i],
        v_folding[i] (u32, fold_words).  Arrays in pinned memory make the transfers
        asynchronous DMA; they are not retained after the call."""
        n = len(keys)
        s = self.db.shape
        qs = [np.ascontiguousarray(q, dtype=np.uint32) for q in queries]
        vfs = [np.ascontiguousarray(v, dtype=np.uint32) for v in v_folding]
        if len(qs) != n or len(vfs) != n:
            raise ValueError("load_many: one query, key set and v_folding per slot")
        for q, v in zip(qs, vfs):
            if q.size != s.query_words or v.size != s.fold_words:
                raise ValueError("load_many: query or v_folding of the wrong size")
        P = ctypes.c_void_p
        qa = (P * n)(*[q.ctypes.data for q in qs])
        va = (P * n)(*[v.ctypes.data for v in vfs])
        ka = (P * n)(*[k.handle.value for k in keys])
        _lib.call("sp_batch_load_many", self._h, first, n, qa, ka, va)

    def load_packed(self, first: int, queries, keys: "list[Keys]") -> None:
        """Slots [first, first + len(keys)): packed queries (client.query_packed) with
        resident keys that carry a conversion key; each slot's fold vectors are
        converted from its expanded query on the device."""
        n = len(keys)
        s = self.db.shape
        qs = [np.ascontiguousarray(q, dtype=np.uint32) for q in queries]
        if len(qs) != n:
            raise ValueError("load_packed: one query and key set per slot")
        if any(q.size != s.query_words for q in qs):
            raise ValueError("load_packed: query of the wrong size")
        P = ctypes.c_void_p
        qa = (P * n)(*[q.ctypes.data for q in qs])
        ka = (P * n)(*[k.handle.value for k in keys])
        _lib.call("sp_batch_load_packed", self._h, first, n, qa, ka)

    def set_packing(self, q2_bits: int) -> None:
        """Response packing + modulus switching for later runs (0: off)."""
        _lib.call("sp_batch_set_packing", self._h, q2_bits)
        self.q2_bits = q2_bits

    def packed_responses(self, n: int) -> np.ndarray:
        """Compact responses of slots [0, n): uint8 [n][bytes]."""
        out = np.zeros((n, self.db.shape.packed_response_bytes(self.q2_bits)), np.uint8)
        _lib.call("sp_batch_read_packed", self._h, n, _ptr(out))
        return out

    def read(self, slot: int, what: int = READ_RESPONSE) -> np.ndarray:
        s = self.db.shape
        shapes = {READ_RESPONSE: (s.n2, 2, 2, 2048), READ_EXPANDED: (1 << s.g, 2, 2, 2048),
                  READ_DB_MUL: (s.n2, s.cols, 2, 2, 2048), READ_RAW_ACCS: (s.n2, s.cols, 2, 2, 2048),
                  READ_V_FOLDING: (s.nu2, 2, 2 * s.t_gsw, 2, 2048)}
        out = np.zeros(shapes[what], np.uint32)
        _lib.call("sp_batch_read", self._h, slot, what, _ptr(out))
        return out

    def responses(self, n: int) -> np.ndarray:
        s = self.db.shape
        out = np.zeros((n, s.n2, 2, 2, 2048), np.uint32)
        _lib.call("sp_batch_read_responses", self._h, n, _ptr(out))
        return out

    def close(self) -> None:
        if self._h:
            _lib.call("sp_batch_destroy", self._h)
            self._h = None

    @property
    def handle(self):
        return self._h


class Keys:
    """One client's expansion keys (and, for packed queries, its conversion key),
    resident on the device for one database shape."""

    def __init__(self, db: Database, left: np.ndarray, right: np.ndarray, conv: np.ndarray | None = None,
                 packing: np.ndarray | None = None):
        s = db.shape
        self.db = db
        h = ctypes.c_void_p()
        lft = np.ascontiguousarray(left, dtype=np.uint32)
        rgt = np.ascontiguousarray(right, dtype=np.uint32)
        _lib.call("sp_keys_create", db.handle, s.t_gsw, s.t_exp, _ptr(lft), _ptr(rgt), ctypes.byref(h))
        self._h = h
        self.packed = conv is not None
        if conv is not None:
            cv = np.ascontiguousarray(conv, dtype=np.uint32)
            if cv.size != s.conv_words:
                self.close()
                raise ValueError("Keys: conversion key of the wrong size")
            _lib.call("sp_keys_set_conversion", h, T_CONV, _ptr(cv))
        if packing is not None:
            pw = np.ascontiguousarray(packing, dtype=np.uint32)
            if pw.size != s.packing_words:
                self.close()
                raise ValueError("Keys: packing keys of the wrong size")
            _lib.call("sp_keys_set_packing", h, T_CONV, _ptr(pw))

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            _lib.call("sp_keys_destroy", self._h)
            self._h = None


class Group:
    """Several batches (each against its own database) answered by one graph launch
    whose branches run concurrently."""

    def __init__(self, batches: "list[Batch]"):
        self.batches = list(batches)
        h = ctypes.c_void_p()
        arr = (ctypes.c_void_p * len(self.batches))(*[b.handle.value for b in self.batches])
        _lib.call("sp_group_create", arr, len(self.batches), ctypes.byref(h))
        self._h = h

    def launch(self, ns: "list[int]", stream=None) -> None:
        arr = (ctypes.c_int * len(ns))(*ns)
        _lib.call("sp_group_launch", self._h, arr, ctypes.c_void_p(stream or 0))

    def synchronize(self) -> None:
        _lib.call("sp_group_synchronize", self._h)

    def close(self) -> None:
        if self._h:
            _lib.call("sp_group_destroy", self._h)
            self._h = None


def answer(db: Database, query, left, right, v_folding) -> np.ndarray:
    """One full answer (expansion, database multiply, fold) -> n2 raw ciphertexts."""
    s = db.shape
    out = np.zeros((s.n2, 2, 2, 2048), np.uint32)
    arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (query, left, right, v_folding)]
    _lib.call("sp_answer", db.handle, s.t_gsw, s.t_exp, *[_ptr(a) for a in arrs], _ptr(out))
    return out
