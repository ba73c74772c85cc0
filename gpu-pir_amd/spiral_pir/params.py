"""Spiral parameters of the answer path (gpu-pir/params.cuh:16-84).

The reference fixes these at compile time; here they are runtime values with
the reference's defaults.  Moduli naming: the reference C++ calls them Q0/Q1
(params.cuh:25-30), its Python Q1/Q2 (cpu_version/params.py:22-27); this
package uses Q0/Q1 like the C++.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

POLY_LEN = 2048                     # params.cuh:16
LEN_LOG2 = 11                       # params.cuh:17
Q0 = 268369921                      # params.cuh:25
Q0_ROOT = 66687                     # params.cuh:26
Q1 = 249561089                      # params.cuh:28
Q1_ROOT = 158221                    # params.cuh:29
MODULI = (Q0, Q1)                   # params.cuh:31
MODULUS = Q0 * Q1                   # params.cuh:40
MOD_BITS = 56                       # params.cuh:41
PT_MODULUS = 256                    # params.cuh:42
T_CONV = 4                          # params.cuh:65
PACK_Q1_BITS = 10                   # rows 1..n of a packed response: modulus 4 p = 2^10
PACK_Q2_BITS = 20                   # row 0 of a packed response (default)
T_GSW = 8                           # params.cuh:66
T_EXP_LEFT = 8                      # params.cuh:67
T_EXP_RIGHT = 8                     # params.cuh:68
N_MAT = 2                           # plaintext matrices n x n (the reference fixes n = 1, params.cuh:49)
SIGMA = 6.4


def get_bits_per(dim: int) -> int:
    """gpu-pir/gadget.cuh:37-39: ceil(MOD_BITS / dim)."""
    return (MOD_BITS + dim - 1) // dim


def ceil_log2(x: int) -> int:
    return max(0, math.ceil(math.log2(x))) if x > 1 else 0


@dataclass(frozen=True)
class Shape:
    """A database shape: 2^nu1 rows (first dimension) x 2^nu2 columns (folded)."""

    nu1: int
    nu2: int
    n2: int = N_MAT * N_MAT
    t_gsw: int = T_GSW
    t_exp: int = T_EXP_LEFT

    @property
    def rows(self) -> int:
        return 1 << self.nu1

    @property
    def cols(self) -> int:
        return 1 << self.nu2

    @property
    def records(self) -> int:
        return 1 << (self.nu1 + self.nu2)

    @property
    def g(self) -> int:
        """Expansion depth (gpu-pir/expansion.cuh:83)."""
        return ceil_log2(self.t_gsw * self.nu2 + (1 << self.nu1))

    @property
    def stop_round(self) -> int:
        """gpu-pir/expansion.cuh:84."""
        return ceil_log2(self.t_gsw * self.nu2)

    @property
    def record_bytes(self) -> int:
        return self.n2 * POLY_LEN

    @property
    def db_ntt_bytes(self) -> int:
        """NTT form as u32 words (the reference's Database): 2 residues x 4 bytes
        per plaintext coefficient."""
        return self.records * self.n2 * 2 * POLY_LEN * 4

    @property
    def db_device_bytes(self) -> int:
        """Resident form in HBM: the 28-bit residues of groups of min(2^nu2, 8)
        columns packed into ceil(28 x group / 32) words (7/8 of the u32 form)."""
        group = min(self.cols, 8)
        words = (28 * group + 31) // 32
        return self.db_ntt_bytes // (4 * group) * 4 * words

    # words (u32) of the client/server buffers
    @property
    def query_words(self) -> int:
        return 4 * POLY_LEN

    @property
    def left_words(self) -> int:
        return self.g * 2 * self.t_exp * 2 * POLY_LEN

    @property
    def right_words(self) -> int:
        return (self.stop_round + 1) * 2 * self.t_exp * 2 * POLY_LEN

    @property
    def fold_words(self) -> int:
        return self.nu2 * 2 * (2 * self.t_gsw) * 2 * POLY_LEN

    @property
    def conv_words(self) -> int:
        """Conversion key CiphertextHalf[2][2 t_conv] (gpu-pir/structures.cuh:176)."""
        return 2 * (2 * T_CONV) * 2 * POLY_LEN

    @property
    def pack_n(self) -> int:
        """Plaintext matrix dimension n (n * n = n2) of a packed response."""
        n = int(round(self.n2 ** 0.5))
        return n if n * n == self.n2 else 0

    @property
    def packing_words(self) -> int:
        """Response packing keys u32 [n][2][t_conv][n+1][N]."""
        n = self.pack_n
        return n * 2 * T_CONV * (n + 1) * POLY_LEN

    def packed_response_bytes(self, q2_bits: int) -> int:
        """Compact response: n polys at q2_bits, n^2 polys at log2(4 p) bits per coefficient."""
        n = self.pack_n
        return n * POLY_LEN * q2_bits // 8 + n * n * POLY_LEN * PACK_Q1_BITS // 8

    @property
    def response_words(self) -> int:
        return self.n2 * 4 * POLY_LEN
