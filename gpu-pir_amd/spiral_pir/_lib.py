"""ctypes binding of libspiral_hip.so (C ABI: include/spiral_pir.h).

The library is built in place (``gpu-pir_amd/libspiral_hip.so``).  Loading
fails loudly if it is missing; there is no CPU fallback anywhere in this
package.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libspiral_hip.so")

SP_OK = 0
SP_ERR_ARG = -1
SP_ERR_HIP = -2
SP_ERR_NOMEM = -3
SP_ERR_STATE = -4


class SpiralError(RuntimeError):
    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} failed ({code}): {msg}")
        self.code = code


P = ctypes.c_void_p
I = ctypes.c_int
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
SZ = ctypes.c_size_t
PP = ctypes.POINTER(ctypes.c_void_p)
PSZ = ctypes.POINTER(ctypes.c_size_t)
PI = ctypes.POINTER(ctypes.c_int)

# name -> argtypes (all return int status unless listed in _RET)
SIGNATURES = {
    "sp_ctx_create": [I, PP],
    "sp_ctx_destroy": [P],
    "sp_ctx_synchronize": [P],
    "sp_ctx_set_variant": [P, I, I],
    "sp_ntt_fwd": [P, P, SZ, I, P],
    "sp_ntt_inv": [P, P, SZ, I, I, P],
    "sp_from_ntt": [P, P, P, SZ, P],
    "sp_to_ntt": [P, P, P, SZ, P],
    "sp_automorph": [P, P, P, SZ, U32, P],
    "sp_gadget_extprod": [P, P, I, I, P, P, SZ, SZ, SZ, SZ, P],
    "sp_db_mul": [P, P, P, P, I, I, P],
    "sp_fold": [P, P, I, I, P, P, P],
    "sp_fold_noalloc": [P, P, I, I, P, P, P],
    "sp_stream_read": [P, P, SZ, P, P],
    "sp_expansion_depth": [I, I, I, PI, PI],
    "sp_expand": [P, P, I, I, I, I, P, P, I, P],
    "sp_regev_to_gsw": [P, P, SZ, I, I, I, P, P, P],
    "sp_packed_response_bytes": [I, I, PSZ],
    "sp_pack_response": [P, P, I, P, I, I, P, P],
    "sp_db_create": [P, I, I, I, PP],
    "sp_db_load_bytes": [P, P],
    "sp_db_load_ntt": [P, P],
    "sp_db_destroy": [P],
    "sp_db_build_limbs": [P, PSZ],
    "sp_query_create": [P, I, I, PP],
    "sp_query_load": [P, P, P, P, P, P],
    "sp_query_load_packed": [P, P, P],
    "sp_query_set_packing": [P, I],
    "sp_query_read_packed": [P, P],
    "sp_query_run": [P, I, P],
    "sp_query_read": [P, I, P],
    "sp_query_destroy": [P],
    "sp_answer": [P, I, I, P, P, P, P, P],
    "sp_batch_create": [P, I, I, I, PP],
    "sp_batch_load": [P, I, P, P, P, P],
    "sp_batch_run": [P, I, I, P],
    "sp_batch_graph_launch": [P, I, P],
    "sp_batch_read": [P, I, I, P],
    "sp_batch_read_responses": [P, I, P],
    "sp_batch_synchronize": [P],
    "sp_batch_poll": [P, PI],
    "sp_batch_tune": [P, I, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float), PI],
    "sp_batch_tune_multi_ms": [P, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)],
    "sp_batch_set_dbmul": [P, I, I],
    "sp_batch_destroy": [P],
    "sp_keys_create": [P, I, I, P, P, PP],
    "sp_keys_destroy": [P],
    "sp_keys_set_conversion": [P, I, P],
    "sp_keys_set_packing": [P, I, P],
    "sp_batch_set_packing": [P, I],
    "sp_batch_read_packed": [P, I, P],
    "sp_batch_load_many": [P, I, I, P, P, P],
    "sp_batch_load_packed": [P, I, I, P, P],
    "sp_batch_load_many_async": [P, I, I, P, P, P],
    "sp_batch_load_packed_async": [P, I, I, P, P],
    "sp_batch_read_responses_async": [P, I, P, P],
    "sp_batch_read_packed_async": [P, I, P, P],
    "sp_batch_last_run": [P, PI, PI],
    "sp_group_create": [P, I, PP],
    "sp_group_launch": [P, PI, P],
    "sp_group_synchronize": [P],
    "sp_group_destroy": [P],
    "sp_client_create": [I, I, I, I, I, U64, PP],
    "sp_client_destroy": [P],
    "sp_client_sizes": [P, PSZ, PSZ, PSZ, PSZ, PSZ],
    "sp_client_keygen": [P, P, P],
    "sp_client_query": [P, U64, P, P],
    "sp_client_decode": [P, P, P],
    "sp_client_conversion_key": [P, I, P],
    "sp_client_query_packed": [P, U64, P],
    "sp_client_phase": [P, P, P],
    "sp_client_secret": [P, P],
    "sp_client_packing_keys": [P, I, P],
    "sp_client_decode_packed": [P, P, SZ, I, P],
}
_RET = {
    "sp_last_error": (ctypes.c_char_p, []),
    "sp_version": (ctypes.c_char_p, []),
    "sp_client_last_error": (ctypes.c_char_p, []),
    "sp_db_device_bytes": (SZ, [P]),
    "sp_stream_sink_words": (SZ, []),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libspiral_hip.so not built at {LIB_PATH}: run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    for name, (res, args) in _RET.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def call(name: str, *args) -> None:
    L = load()
    rc = getattr(L, name)(*args)
    if rc != SP_OK:
        err = L.sp_client_last_error() if name.startswith("sp_client") else L.sp_last_error()
        raise SpiralError(name, rc, (err or b"").decode(errors="replace"))


def version() -> str:
    return load().sp_version().decode()
