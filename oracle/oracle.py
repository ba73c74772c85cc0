"""ctypes binding of the CPU oracle (oracle/spiral_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg, never by the product package.

Arrays are numpy ``uint32`` in the reference's struct layouts (see
spiral_oracle.h).  Also holds the readers for the reference's ``.dat`` fixture
formats (gpu-pir/host.cu:155-173,216-260; gpu-pir/cpu_version/testing.py:355-408).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

N = 2048
Q1 = 268369921
Q2 = 249561089
Q = Q1 * Q2
QS = (Q1, Q2)

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libspiral_oracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        I = ctypes.c_int
        U = ctypes.c_uint32
        sigs = {
            "or_modulus": ([], ctypes.c_uint64),
            "or_q": ([I], U),
            "or_twiddle": ([I, I, I], U),
            "or_ntt_fwd": ([P, I], None),
            "or_ntt_inv": ([P, I], None),
            "or_ntt_brute": ([P, P, I], None),
            "or_from_ntt": ([P, P], None),
            "or_to_ntt": ([P, P], None),
            "or_automorph": ([P, P, U], None),
            "or_gadget_bits": ([I], I),
            "or_gadget_inv": ([P, I, I, P], None),
            "or_extprod_acc": ([P, I, I, P, P], None),
            "or_db_mul": ([P, P, P, I, I], None),
            "or_db_mul_omp": ([P, P, P, I, I, I], None),
            "or_fold": ([P, I, I, P, P], None),
            "or_set_threads": ([I], None),
            "or_expansion_g": ([I, I, I], I),
            "or_expansion_stop_round": ([I, I], I),
            "or_v_neg1": ([I, P], None),
            "or_expand": ([P, I, I, I, I, P, P, I], None),
            "or_folding_neg": ([P, P, I, I], None),
            "or_regev_to_gsw": ([P, I, I, I, P, P], None),
            "or_switch_modulus": ([ctypes.c_uint64, I], U),
            "or_packed_response_bytes": ([I, I, I], ctypes.c_size_t),
            "or_pack_response": ([P, I, P, I, I, I, P], None),
            "or_answer": ([P, I, I, I, I, I, P, P, P, P, P, P], None),
            "or_db_encode": ([P, P, I, I, I], None),
        }
        for name, (args, res) in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return ctypes.c_void_p(a.ctypes.data)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.uint32)


# --------------------------------------------------------------------------
# Single-function wrappers (copies in, returns new arrays)
# --------------------------------------------------------------------------

def ntt_fwd(a, m: int) -> np.ndarray:
    a = _u32(a).copy()
    lib().or_ntt_fwd(_p(a), m)
    return a


def ntt_inv(a, m: int) -> np.ndarray:
    a = _u32(a).copy()
    lib().or_ntt_inv(_p(a), m)
    return a


def ntt_brute(a, m: int) -> np.ndarray:
    a = _u32(a)
    out = np.zeros(N, np.uint32)
    lib().or_ntt_brute(_p(a), _p(out), m)
    return out


def from_ntt(ntt_poly) -> np.ndarray:
    ntt_poly = _u32(ntt_poly).reshape(2, N)
    raw = np.zeros((2, N), np.uint32)
    lib().or_from_ntt(_p(ntt_poly), _p(raw))
    return raw


def to_ntt(raw_poly) -> np.ndarray:
    raw_poly = _u32(raw_poly).reshape(2, N)
    out = np.zeros((2, N), np.uint32)
    lib().or_to_ntt(_p(raw_poly), _p(out))
    return out


def automorph(raw_poly, t: int) -> np.ndarray:
    raw_poly = _u32(raw_poly).reshape(2, N)
    out = np.zeros((2, N), np.uint32)
    lib().or_automorph(_p(raw_poly), _p(out), t)
    return out


def gadget_bits(num_elems: int) -> int:
    return lib().or_gadget_bits(num_elems)


def gadget_inv(raw_polys, rdim: int, mx: int) -> np.ndarray:
    raw_polys = _u32(raw_polys).reshape(rdim, 2, N)
    out = np.zeros((mx, N), np.uint32)
    lib().or_gadget_inv(_p(raw_polys), rdim, mx, _p(out))
    return out


def extprod_acc(raw_in, rdim: int, mx: int, keys, acc=None) -> np.ndarray:
    raw_in = _u32(raw_in)
    keys = _u32(keys)
    assert keys.size == 2 * mx * 2 * N
    acc = np.zeros((2, 2, N), np.uint32) if acc is None else _u32(acc).reshape(2, 2, N).copy()
    lib().or_extprod_acc(_p(raw_in), rdim, mx, _p(keys), _p(acc))
    return acc


def db_mul(db, first_dim, nu1: int, nu2: int, nthreads: int | None = None) -> np.ndarray:
    db = _u32(db)
    first_dim = _u32(first_dim)
    assert db.size == (1 << (nu1 + nu2)) * 2 * N
    assert first_dim.size == (1 << nu1) * 4 * N
    out = np.zeros(((1 << nu2), 2, 2, N), np.uint32)
    if nthreads is None:
        lib().or_db_mul(_p(db), _p(first_dim), _p(out), nu1, nu2)
    else:
        lib().or_db_mul_omp(_p(db), _p(first_dim), _p(out), nu1, nu2, nthreads)
    return out


def set_threads(n: int) -> None:
    """Threads of expand() and fold() (default 1); results do not depend on it."""
    lib().or_set_threads(int(n))


def fold(accs, nu2: int, t_gsw: int, v_folding, v_folding_neg) -> np.ndarray:
    accs = _u32(accs).copy()
    v_folding = _u32(v_folding)
    v_folding_neg = _u32(v_folding_neg)
    assert accs.size == (1 << nu2) * 4 * N
    assert v_folding.size == nu2 * 2 * 2 * t_gsw * 2 * N
    lib().or_fold(_p(accs), nu2, t_gsw, _p(v_folding), _p(v_folding_neg))
    return accs.reshape((1 << nu2), 2, 2, N)


def expansion_g(nu1: int, nu2: int, t_gsw: int) -> int:
    return lib().or_expansion_g(nu1, nu2, t_gsw)


def expansion_stop_round(nu2: int, t_gsw: int) -> int:
    return lib().or_expansion_stop_round(nu2, t_gsw)


def v_neg1(r: int) -> np.ndarray:
    out = np.zeros((2, N), np.uint32)
    lib().or_v_neg1(r, _p(out))
    return out


def expand(query, nu1: int, nu2: int, t_gsw: int, t_exp: int, left, right,
           rounds: int = -1) -> np.ndarray:
    g = expansion_g(nu1, nu2, t_gsw)
    v = np.zeros(((1 << g), 2, 2, N), np.uint32)
    v[0] = _u32(query).reshape(2, 2, N)
    left = _u32(left)
    right = _u32(right)
    lib().or_expand(_p(v), nu1, nu2, t_gsw, t_exp, _p(left), _p(right), rounds)
    return v


def folding_neg(v_folding, nu2: int, t_gsw: int) -> np.ndarray:
    v_folding = _u32(v_folding)
    out = np.zeros_like(v_folding)
    lib().or_folding_neg(_p(v_folding), _p(out), nu2, t_gsw)
    return out


def regev_to_gsw(v, nu2: int, t_gsw: int, t_conv: int, conv) -> np.ndarray:
    """Expanded NTT Ciphertexts v[>= 2 nu2 t_gsw] + conversion key
    CiphertextHalf[2][2 t_conv] -> v_folding CiphertextHalf[nu2][2][2 t_gsw]."""
    v = _u32(v)
    conv = _u32(conv)
    assert v.size >= 2 * nu2 * t_gsw * 4 * N and conv.size == 2 * 2 * t_conv * 2 * N
    out = np.zeros((nu2, 2, 2 * t_gsw, 2, N), np.uint32)
    lib().or_regev_to_gsw(_p(v), nu2, t_gsw, t_conv, _p(conv), _p(out))
    return out


def switch_modulus(x: int, bits: int) -> int:
    return lib().or_switch_modulus(x, bits)


def pack_response(raw_cts, n: int, w, t_conv: int, q2_bits: int, q1_bits: int = 10) -> np.ndarray:
    """n^2 raw Ciphertexts + packing keys [n][2][t_conv][n+1][N] -> compact response bytes."""
    raw_cts = _u32(raw_cts)
    w = _u32(w)
    assert raw_cts.size == n * n * 4 * N and w.size == n * 2 * t_conv * (n + 1) * N
    out = np.zeros(lib().or_packed_response_bytes(n, q2_bits, q1_bits), np.uint8)
    lib().or_pack_response(_p(raw_cts), n, _p(w), t_conv, q2_bits, q1_bits, _p(out))
    return out


def answer(db, n2: int, nu1: int, nu2: int, t_gsw: int, t_exp: int, query, left, right,
           v_folding, v_folding_neg) -> np.ndarray:
    db = _u32(db)
    out = np.zeros((n2, 2, 2, N), np.uint32)
    lib().or_answer(_p(db), n2, nu1, nu2, t_gsw, t_exp, _p(_u32(query)), _p(_u32(left)),
                    _p(_u32(right)), _p(_u32(v_folding)), _p(_u32(v_folding_neg)), _p(out))
    return out


def db_encode(records: np.ndarray, n2: int, nu1: int, nu2: int) -> np.ndarray:
    records = np.ascontiguousarray(records, dtype=np.uint8)
    assert records.size == (1 << (nu1 + nu2)) * n2 * N
    db = np.zeros((n2, 1 << nu2, 1 << nu1, 2, N), np.uint32)
    lib().or_db_encode(_p(records), _p(db), n2, nu1, nu2)
    return db


# --------------------------------------------------------------------------
# numpy helpers
# --------------------------------------------------------------------------

def raw_to_u64(raw_poly) -> np.ndarray:
    raw_poly = np.asarray(raw_poly, dtype=np.uint32).reshape(-1, 2, N)
    out = (raw_poly[:, 0].astype(np.uint64) << np.uint64(32)) | raw_poly[:, 1].astype(np.uint64)
    return out


def u64_to_raw(x) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64).reshape(-1, N)
    out = np.empty((x.shape[0], 2, N), np.uint32)
    out[:, 0] = (x >> np.uint64(32)).astype(np.uint32)
    out[:, 1] = (x & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return out


# --------------------------------------------------------------------------
# Reference fixture readers (formats: gpu-pir/cpu_version/testing.py:355-408)
# --------------------------------------------------------------------------

def read_ntt_polys(path: str) -> np.ndarray:
    """Spiral NTT-form dump: per poly u64[2048] residue q1, then u64[2048] residue q2."""
    a = np.fromfile(path, dtype=np.uint64).reshape(-1, 2, N)
    assert (a[:, 0] < Q1).all() and (a[:, 1] < Q2).all()
    return a.astype(np.uint32)


def read_raw_polys(path: str) -> np.ndarray:
    """Spiral raw dump: per poly u64[2048] (values < Q) -> raw Poly hi/lo u32."""
    return u64_to_raw(np.fromfile(path, dtype=np.uint64))


def matrices_to_keys(polys: np.ndarray, rows: int, cols: int) -> np.ndarray:
    """Row-major 2 x cols NTT matrices (one per level) -> CiphertextHalf[level][crt][col][row][z]."""
    polys = polys.reshape(-1, rows, cols, 2, N)
    return np.ascontiguousarray(polys.transpose(0, 3, 2, 1, 4))
