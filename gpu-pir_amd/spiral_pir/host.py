"""Host helpers mirroring gpu-pir/host.h: fixture loading and grouping.

``load_polys(fn, crt_form, montgomery)`` reads the spiral-rs dump formats the
reference's tests use (gpu-pir/host.cu:216-249): an NTT-form poly is
u64[2048] residue mod Q0 then u64[2048] residue mod Q1; a raw poly is
u64[2048] coefficients, split into (hi, lo) u32 planes.  Unlike the
reference, ``montgomery=True`` is not offered: this library takes keys in
standard form and natural NTT order and converts internally.
"""
from __future__ import annotations

import numpy as np

from .params import POLY_LEN, MODULI


def load_polys(fn: str, crt_form: bool) -> np.ndarray:
    """-> u32 [n_polys][2][2048] (NTT residues, or raw hi/lo planes)."""
    a = np.fromfile(fn, dtype=np.uint64)
    if crt_form:
        a = a.reshape(-1, 2, POLY_LEN)
        if not ((a[:, 0] < MODULI[0]).all() and (a[:, 1] < MODULI[1]).all()):
            raise ValueError(f"{fn}: residues out of range")
        return a.astype(np.uint32)
    a = a.reshape(-1, POLY_LEN)
    out = np.empty((a.shape[0], 2, POLY_LEN), np.uint32)
    out[:, 0] = (a >> np.uint64(32)).astype(np.uint32)
    out[:, 1] = (a & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return out


def group_to_ciphertext(polys: np.ndarray, grouping: int) -> np.ndarray:
    """gpu-pir/host.cu:251-260: groups of 2*grouping polys -> [groups][grouping] Ciphertexts,
    ciphertext c of a group = (poly c, poly c + grouping)."""
    g2 = 2 * grouping
    p = polys.reshape(-1, g2, 2, POLY_LEN)
    return np.ascontiguousarray(p.reshape(-1, 2, grouping, 2, POLY_LEN).transpose(0, 2, 1, 3, 4))


def flatten(cts: np.ndarray) -> np.ndarray:
    """gpu-pir/host.cu:262-268: first ciphertext of every group."""
    return np.ascontiguousarray(cts[:, 0])


def ciphertexts_to_key_layout(cts: np.ndarray) -> np.ndarray:
    """[level][col] Ciphertexts (as group_to_ciphertext gives for a key file) ->
    CiphertextHalf[level][crt][col][row][N], the v_folding / expansion-key layout:
    gpu-pir/fold_test.cu:56-66 copies ciphertext j of level i row by row into
    v_folding[i][crt][j]."""
    c = np.asarray(cts).reshape(cts.shape[0], cts.shape[1], 2, 2, POLY_LEN)
    return np.ascontiguousarray(c.transpose(0, 3, 1, 2, 4))


def matrices_to_key_layout(polys: np.ndarray, cols: int) -> np.ndarray:
    """Row-major 2 x cols NTT matrices (one per level) -> CiphertextHalf[level][crt][col][row][N]
    (the v_folding / expansion-key layout, structures.cuh:173-183)."""
    p = polys.reshape(-1, 2, cols, 2, POLY_LEN)
    return np.ascontiguousarray(p.transpose(0, 3, 2, 1, 4))
