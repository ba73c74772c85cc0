"""Python face of the native Spiral client (sp_client_* in libspiral_hip.so).

The reference has no client (its inputs are spiral-rs dumps); this one
produces keys and queries for the server path and decodes responses.  It
runs on the host only.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .params import T_CONV, Shape


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    assert a.flags["C_CONTIGUOUS"]
    return ctypes.c_void_p(a.ctypes.data)


class Client:
    def __init__(self, shape: Shape, seed: int):
        self.shape = shape
        h = ctypes.c_void_p()
        _lib.call("sp_client_create", shape.nu1, shape.nu2, shape.n2, shape.t_gsw, shape.t_exp,
                  ctypes.c_uint64(seed), ctypes.byref(h))
        self._h = h

    def close(self) -> None:
        if self._h:
            _lib.call("sp_client_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def keygen(self) -> tuple[np.ndarray, np.ndarray]:
        """Expansion keys: (left [g][2][t_exp][2][N], right [stop+1][2][t_exp][2][N]) u32."""
        s = self.shape
        left = np.zeros(s.left_words, np.uint32)
        right = np.zeros(s.right_words, np.uint32)
        _lib.call("sp_client_keygen", self._h, _ptr(left), _ptr(right))
        return (left.reshape(s.g, 2, s.t_exp, 2, -1), right.reshape(s.stop_round + 1, 2, s.t_exp, 2, -1))

    def query(self, index: int) -> tuple[np.ndarray, np.ndarray]:
        """(query NTT Ciphertext [2][2][N], v_folding [nu2][2][2 t_gsw][2][N]) for record `index`."""
        s = self.shape
        q = np.zeros(s.query_words, np.uint32)
        vf = np.zeros(s.fold_words, np.uint32)
        _lib.call("sp_client_query", self._h, ctypes.c_uint64(index), _ptr(q), _ptr(vf))
        return q.reshape(2, 2, -1), vf.reshape(s.nu2, 2, 2 * s.t_gsw, 2, -1)

    def conversion_key(self) -> np.ndarray:
        """Regev -> GSW conversion key CiphertextHalf[2][2 t_conv] = u32 [2][8][2][N]."""
        s = self.shape
        out = np.zeros(s.conv_words, np.uint32)
        _lib.call("sp_client_conversion_key", self._h, T_CONV, _ptr(out))
        return out.reshape(2, 2 * T_CONV, 2, -1)

    def query_packed(self, index: int) -> np.ndarray:
        """Packed query NTT Ciphertext [2][2][N] for record `index`: the server derives
        the fold vectors from it (Regev -> GSW conversion)."""
        q = np.zeros(self.shape.query_words, np.uint32)
        _lib.call("sp_client_query_packed", self._h, ctypes.c_uint64(index), _ptr(q))
        return q.reshape(2, 2, -1)

    def packing_keys(self) -> np.ndarray:
        """Response packing keys u32 [n][2][t_conv][n+1][N] (n * n = n2)."""
        s = self.shape
        out = np.zeros(s.packing_words, np.uint32)
        _lib.call("sp_client_packing_keys", self._h, T_CONV, _ptr(out))
        return out.reshape(s.pack_n, 2, T_CONV, s.pack_n + 1, -1)

    def decode_packed(self, response: np.ndarray, q2_bits: int) -> np.ndarray:
        """Compact (packed, modulus-switched) response bytes -> record bytes [n2][N]."""
        s = self.shape
        resp = np.ascontiguousarray(response, dtype=np.uint8)
        out = np.zeros((s.n2, 2048), np.uint8)
        _lib.call("sp_client_decode_packed", self._h, _ptr(resp), ctypes.c_size_t(resp.size), q2_bits, _ptr(out))
        return out

    def phase(self, ct: np.ndarray) -> np.ndarray:
        """b - a*s mod Q (u64 [N], coefficient order) of one NTT Ciphertext [2][2][N]."""
        c = np.ascontiguousarray(ct, dtype=np.uint32)
        assert c.size == 4 * 2048
        out = np.zeros(2048, np.uint64)
        _lib.call("sp_client_phase", self._h, _ptr(c), ctypes.c_void_p(out.ctypes.data))
        return out

    def secret(self) -> np.ndarray:
        out = np.zeros(2048, np.int32)
        _lib.call("sp_client_secret", self._h, ctypes.c_void_p(out.ctypes.data))
        return out

    def decode(self, response: np.ndarray) -> np.ndarray:
        """n2 raw response ciphertexts -> record bytes [n2][N]."""
        s = self.shape
        resp = np.ascontiguousarray(response, dtype=np.uint32)
        assert resp.size == s.response_words
        out = np.zeros((s.n2, 2048), np.uint8)
        _lib.call("sp_client_decode", self._h, _ptr(resp), _ptr(out))
        return out
