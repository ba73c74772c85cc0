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
from .params import MODULI, T_CONV, Shape

STAGE_EXPAND = 1
STAGE_DB_MUL = 2
STAGE_FOLD = 4
STAGES_ALL = 7

# indices for Context.set_variant
VARIANT_EXPAND = 0
VARIANT_FOLD = 1
VARIANT_DB_MUL = 2

# database-multiply variants (Context.set_variant(VARIANT_DB_MUL, v))
DBMUL_NAMES = {0: None, 1: "valu", 2: "mfma", 3: "valu_q2", 4: "valu_q4"}

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


def _sized(a, words: int, what: str) -> np.ndarray:
    """a as a contiguous u32 array of exactly `words` words (the byte count the
    library copies for it), else ValueError: a short array would make the
    library read past its end."""
    a = np.ascontiguousarray(a, dtype=np.uint32)
    if a.size != words:
        raise ValueError(f"{what}: {a.size} words, expected {words}")
    return a


def check_residues(ntt: np.ndarray, what: str, chunk: int = 1 << 24) -> None:
    """Raises ValueError unless every residue of an NTT-form array [..][2][2048]
    (residue m of q_m) is canonical; checked chunk by chunk (memory maps stay lazy)."""
    v = ntt.reshape(-1, 2, 2048)
    step = max(1, chunk // (2 * 2048))
    for k in range(0, v.shape[0], step):
        blk = v[k:k + step]
        if not ((blk[:, 0] < MODULI[0]).all() and (blk[:, 1] < MODULI[1]).all()):
            raise ValueError(f"{what}: residues out of range in polynomials {k}..{k + blk.shape[0] - 1}")


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

    def stream_read(self, buf_ptr: int, nbytes: int, sink_ptr: int, stream=None) -> None:
        """One launch of the HBM streaming-read ceiling kernel over nbytes of device
        memory at buf_ptr (sink: stream_sink_words() u32 words of device memory)."""
        _lib.call("sp_stream_read", self._h, ctypes.c_void_p(buf_ptr), nbytes, ctypes.c_void_p(sink_ptr),
                  ctypes.c_void_p(stream or 0))

    @staticmethod
    def stream_sink_words() -> int:
        return int(_lib.load().sp_stream_sink_words())

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

    def load_ntt(self, db_ntt: np.ndarray, check: bool = True) -> None:
        """Reference Database layout per sub-database: [n2][2^nu2][2^nu1][2][N] u32,
        canonical residues (checked unless check=False; the library reduces any
        word mod q_m in any case)."""
        d = _sized(db_ntt, self.shape.records * self.shape.n2 * 2 * 2048, "load_ntt")
        if check:
            check_residues(d, "load_ntt")
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
            check_residues(payload, path)
            self.load_ntt(payload, check=False)

    @property
    def device_bytes(self) -> int:
        return int(_lib.load().sp_db_device_bytes(self._h))

    @property
    def limb_bytes(self) -> int:
        """HBM the int8-limb copy for the MFMA multiply takes once a batch uses it
        (4 limb bytes per residue; only shapes with 2^nu1, 2^nu2 >= 32 have one)."""
        s = self.shape
        return s.db_ntt_bytes if (s.nu1 >= 5 and s.nu2 >= 5) else 0

    def build_limbs(self) -> int:
        """Build the int8-limb copy now (else on first MFMA use); returns its bytes."""
        n = ctypes.c_size_t()
        _lib.call("sp_db_build_limbs", self._h, ctypes.byref(n))
        return n.value

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
        arrs = _query_arrays(self.db.shape, "Query.load", query, left, right, v_folding)
        neg = None if v_folding_neg is None else _sized(v_folding_neg, self.db.shape.fold_words,
                                                       "Query.load: v_folding_neg")
        _lib.call("sp_query_load", self._h, *[_ptr(a) for a in arrs], _ptr(neg))

    def load_packed(self, query: np.ndarray, keys: "Keys") -> None:
        """A packed query (client.query_packed) with resident keys that carry a
        conversion key: the fold vectors are converted on the device."""
        q = _sized(query, self.db.shape.query_words, "Query.load_packed: query")
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
        self._hold: list = []  # host arrays an asynchronous load or readout still reads or writes

    def load(self, slot: int, query, left, right, v_folding) -> None:
        arrs = _query_arrays(self.db.shape, "Batch.load", query, left, right, v_folding)
        _lib.call("sp_batch_load", self._h, slot, *[_ptr(a) for a in arrs])

    def run(self, n: int, stages: int = STAGES_ALL, stream=None) -> None:
        _lib.call("sp_batch_run", self._h, n, stages, ctypes.c_void_p(stream or 0))

    def graph_launch(self, n: int, stream=None) -> None:
        _lib.call("sp_batch_graph_launch", self._h, n, ctypes.c_void_p(stream or 0))

    def synchronize(self) -> None:
        _lib.call("sp_batch_synchronize", self._h)
        self._hold = []

    def poll(self) -> bool:
        """True when the last run (and any asynchronous readout queued after it)
        has finished (does not wait)."""
        d = ctypes.c_int()
        _lib.call("sp_batch_poll", self._h, ctypes.byref(d))
        if d.value:
            self._hold = []
        return bool(d.value)

    def last_run(self) -> dict:
        """Database-multiply kernel of the last run ("valu" / "mfma" / None) and
        the number of captured stage-chain graphs."""
        v, g = ctypes.c_int(), ctypes.c_int()
        _lib.call("sp_batch_last_run", self._h, ctypes.byref(v), ctypes.byref(g))
        return {"dbmul": DBMUL_NAMES[v.value], "graphs": g.value}

    def tune(self, n: int) -> dict:
        """Autotune the database multiply for n queries (after one expansion has run)."""
        a, b, c = ctypes.c_float(), ctypes.c_float(), ctypes.c_int()
        _lib.call("sp_batch_tune", self._h, n, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        m2, m4 = ctypes.c_float(), ctypes.c_float()
        _lib.call("sp_batch_tune_multi_ms", self._h, ctypes.byref(m2), ctypes.byref(m4))
        opt = lambda x: x.value if x.value >= 0 else None  # noqa: E731
        return {"valu_ms": a.value, "mfma_ms": opt(b), "valu_q2_ms": opt(m2), "valu_q4_ms": opt(m4),
                "chosen": DBMUL_NAMES[c.value]}

    def set_dbmul(self, n: int, variant) -> None:
        """Database-multiply variant for later runs of n queries (a DBMUL_NAMES name or number)."""
        v = variant if isinstance(variant, int) else {name: k for k, name in DBMUL_NAMES.items()}[variant]
        _lib.call("sp_batch_set_dbmul", self._h, n, v)

    def load_many(self, first: int, queries, keys: "list[Keys]", v_folding, asynchronous: bool = False) -> None:
        """Slots [first, first + len(keys)): queries[i] ([2][2][2048] u32), resident keys[i],
        v_folding[i] (u32, fold_words).  Arrays in pinned memory make the transfers
        asynchronous DMA.  Blocking by default; with asynchronous=True the uploads
        are queued behind the batch's previous run and the call returns at once
        (the arrays are kept until the next run has finished)."""
        n = len(keys)
        s = self.db.shape
        qs = [_sized(q, s.query_words, "load_many: query") for q in queries]
        vfs = [_sized(v, s.fold_words, "load_many: v_folding") for v in v_folding]
        if len(qs) != n or len(vfs) != n:
            raise ValueError("load_many: one query, key set and v_folding per slot")
        self._check_keys(keys, "load_many")
        P = ctypes.c_void_p
        qa = (P * n)(*[q.ctypes.data for q in qs])
        va = (P * n)(*[v.ctypes.data for v in vfs])
        ka = (P * n)(*[k.handle.value for k in keys])
        if asynchronous:
            self._hold += qs + vfs
        _lib.call("sp_batch_load_many_async" if asynchronous else "sp_batch_load_many", self._h, first, n, qa, ka, va)

    def load_packed(self, first: int, queries, keys: "list[Keys]", asynchronous: bool = False) -> None:
        """Slots [first, first + len(keys)): packed queries (client.query_packed) with
        resident keys that carry a conversion key; each slot's fold vectors are
        converted from its expanded query on the device.  asynchronous as for
        load_many."""
        n = len(keys)
        s = self.db.shape
        qs = [_sized(q, s.query_words, "load_packed: query") for q in queries]
        if len(qs) != n:
            raise ValueError("load_packed: one query and key set per slot")
        self._check_keys(keys, "load_packed")
        P = ctypes.c_void_p
        qa = (P * n)(*[q.ctypes.data for q in qs])
        ka = (P * n)(*[k.handle.value for k in keys])
        if asynchronous:
            self._hold += qs
        _lib.call("sp_batch_load_packed_async" if asynchronous else "sp_batch_load_packed", self._h, first, n, qa, ka)

    def _check_keys(self, keys, who: str) -> None:
        for k in keys:
            if not isinstance(k, Keys) or not k.handle:
                raise ValueError(f"{who}: keys must be open Keys objects")

    def read_responses_async(self, n: int, out: np.ndarray, stream=None) -> None:
        """Queue the copy of the responses of slots [0, n) of the last run into `out`
        (u32, n x n2 x 2 x 2 x 2048 words, pinned for DMA) on `stream`; complete when
        poll() returns True."""
        s = self.db.shape
        o = out.reshape(-1)
        if o.dtype != np.uint32 or not o.flags["C_CONTIGUOUS"] or o.size < n * s.response_words:
            raise ValueError("read_responses_async: out must be contiguous u32 of n x response_words")
        self._hold.append(out)
        _lib.call("sp_batch_read_responses_async", self._h, n, _ptr(o), ctypes.c_void_p(stream or 0))

    def read_packed_async(self, n: int, out: np.ndarray, stream=None) -> None:
        """Queue the copy of the compact responses of slots [0, n) of the last run
        into `out` (u8, n x packed_response_bytes, pinned for DMA) on `stream`."""
        o = out.reshape(-1)
        need = n * self.db.shape.packed_response_bytes(self.q2_bits)
        if o.dtype != np.uint8 or not o.flags["C_CONTIGUOUS"] or o.size < need:
            raise ValueError("read_packed_async: out must be contiguous u8 of n x packed_response_bytes")
        self._hold.append(out)
        _lib.call("sp_batch_read_packed_async", self._h, n, _ptr(o), ctypes.c_void_p(stream or 0))

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
        lft = _sized(left, s.left_words, "Keys: left")
        rgt = _sized(right, s.right_words, "Keys: right")
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
    arrs = _query_arrays(s, "answer", query, left, right, v_folding)
    _lib.call("sp_answer", db.handle, s.t_gsw, s.t_exp, *[_ptr(a) for a in arrs], _ptr(out))
    return out


def _query_arrays(s: Shape, who: str, query, left, right, v_folding) -> list:
    """The four host arrays of a query with host keys, size-checked against the shape."""
    return [_sized(query, s.query_words, f"{who}: query"), _sized(left, s.left_words, f"{who}: left"),
            _sized(right, s.right_words, f"{who}: right"), _sized(v_folding, s.fold_words, f"{who}: v_folding")]
