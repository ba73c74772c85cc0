"""Database files: the library's own container and the spiral-rs NTT dump.

Container (``.spdb``, little-endian), a 32-byte header then the payload:

    magic     8 bytes  b"SPDB" 01 00 00 00
    kind      u32      KIND_RECORDS: plaintext records, u8 [records][n2][2048]
                       KIND_NTT:     NTT form in the reference's Database layout,
                                     u32 [n2][2^nu2 cols][2^nu1 rows][2][2048]
                                     (gpu-pir/structures.cuh:200-210), residues
                                     canonical, standard form
    nu1, nu2, n2       u32 each
    reserved  u32 x 2  zero

Files are memory-mapped: ``Database.load_file`` streams the payload from disk
into the device in the library's own chunks, so a file larger than host memory
loads too.

spiral-rs dump (the reference's fixture format for NTT polynomials,
gpu-pir/host.cu:216-249, cpu_version/testing.py:61-114): per polynomial
u64[2048] residues mod Q0, then u64[2048] residues mod Q1.  A database dump
holds the polynomials in the reference Database order [n2][cols][rows];
``read_dat_database`` converts it to the u32 layout ``Database.load_ntt`` takes.
"""
from __future__ import annotations

import os
import struct

import numpy as np

from .params import MODULI, POLY_LEN, Shape

MAGIC = b"SPDB\x01\x00\x00\x00"
HEADER = struct.Struct("<8s6I")  # magic, kind, nu1, nu2, n2, reserved x 2
KIND_RECORDS = 0
KIND_NTT = 1


def payload_shape(shape: Shape, kind: int) -> tuple[tuple[int, ...], np.dtype]:
    if kind == KIND_RECORDS:
        return (shape.records, shape.n2, POLY_LEN), np.dtype(np.uint8)
    if kind == KIND_NTT:
        return (shape.n2, shape.cols, shape.rows, 2, POLY_LEN), np.dtype(np.uint32)
    raise ValueError(f"unknown database file kind {kind}")


def write(path: str, shape: Shape, kind: int, payload: np.ndarray) -> None:
    """Writes a container; payload in the layout of `kind` (checked)."""
    dims, dtype = payload_shape(shape, kind)
    a = np.ascontiguousarray(payload)
    if a.dtype != dtype or a.size != int(np.prod(dims)):
        raise ValueError(f"payload must be {dtype} {dims}, got {a.dtype} {a.shape}")
    if kind == KIND_NTT:
        v = a.reshape(dims)
        if not ((v[:, :, :, 0] < MODULI[0]).all() and (v[:, :, :, 1] < MODULI[1]).all()):
            raise ValueError("NTT payload: residues out of range")
    with open(path, "wb") as f:
        f.write(HEADER.pack(MAGIC, kind, shape.nu1, shape.nu2, shape.n2, 0, 0))
        a.reshape(-1).tofile(f)


def read(path: str) -> tuple[Shape, int, np.ndarray]:
    """-> (shape, kind, payload memory-mapped read-only in the layout of kind)."""
    with open(path, "rb") as f:
        head = f.read(HEADER.size)
    if len(head) != HEADER.size:
        raise ValueError(f"{path}: not a database file (short header)")
    magic, kind, nu1, nu2, n2, r0, r1 = HEADER.unpack(head)
    if magic != MAGIC or r0 or r1:
        raise ValueError(f"{path}: not a database file (bad magic)")
    if not (1 <= n2 <= 16 and 0 < nu1 < 32 and 0 < nu2 < 32):
        raise ValueError(f"{path}: shape ({nu1}, {nu2}, n2 = {n2}) out of range")
    shape = Shape(nu1, nu2, n2)
    dims, dtype = payload_shape(shape, kind)
    want = HEADER.size + int(np.prod(dims)) * dtype.itemsize
    if os.path.getsize(path) != want:
        raise ValueError(f"{path}: size {os.path.getsize(path)}, expected {want} for ({nu1}, {nu2}) kind {kind}")
    return shape, kind, np.memmap(path, dtype=dtype, mode="r", offset=HEADER.size, shape=dims)


def read_dat_database(path: str, shape: Shape, chunk_polys: int = 1 << 14) -> np.ndarray:
    """spiral-rs NTT dump of a database -> u32 [n2][cols][rows][2][2048] (residues checked)."""
    n_polys = shape.n2 * shape.cols * shape.rows
    src = np.memmap(path, dtype=np.uint64, mode="r")
    if src.size != n_polys * 2 * POLY_LEN:
        raise ValueError(f"{path}: {src.size} words, expected {n_polys * 2 * POLY_LEN} for {shape}")
    src = src.reshape(n_polys, 2, POLY_LEN)
    out = np.empty((n_polys, 2, POLY_LEN), np.uint32)
    for p0 in range(0, n_polys, chunk_polys):
        blk = src[p0:p0 + chunk_polys]
        if not ((blk[:, 0] < MODULI[0]).all() and (blk[:, 1] < MODULI[1]).all()):
            raise ValueError(f"{path}: residues out of range in polynomials {p0}..")
        out[p0:p0 + chunk_polys] = blk
    return out.reshape(shape.n2, shape.cols, shape.rows, 2, POLY_LEN)
