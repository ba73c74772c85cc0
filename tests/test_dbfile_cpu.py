"""Database files (spiral_pir.dbfile): the container round trip, its checks, and
the spiral-rs NTT dump reader (u64 residues, gpu-pir/host.cu:216-249)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng

N = 2048


def _records(shape, seed):
    return rng(seed).integers(0, 256, size=(shape.records, shape.n2, N), dtype=np.uint8)


def test_records_container_round_trip(tmp_path):
    from spiral_pir import Shape, dbfile

    s = Shape(4, 2)
    recs = _records(s, 11)
    p = str(tmp_path / "db.spdb")
    dbfile.write(p, s, dbfile.KIND_RECORDS, recs)
    shape, kind, payload = dbfile.read(p)
    assert (shape, kind) == (s, dbfile.KIND_RECORDS)
    assert isinstance(payload, np.memmap)
    np.testing.assert_array_equal(payload, recs)


def test_ntt_container_round_trip(tmp_path, oracle):
    from spiral_pir import Shape, dbfile

    s = Shape(4, 2)
    db = oracle.db_encode(_records(s, 12), s.n2, s.nu1, s.nu2)
    p = str(tmp_path / "db.spdb")
    dbfile.write(p, s, dbfile.KIND_NTT, db)
    shape, kind, payload = dbfile.read(p)
    assert (shape, kind) == (s, dbfile.KIND_NTT)
    np.testing.assert_array_equal(payload, db.reshape(payload.shape))


def test_container_checks(tmp_path, oracle):
    from spiral_pir import Shape, dbfile

    s = Shape(4, 2)
    recs = _records(s, 13)
    with pytest.raises(ValueError):  # wrong payload size
        dbfile.write(str(tmp_path / "x.spdb"), s, dbfile.KIND_RECORDS, recs[:-1])
    bad = oracle.db_encode(recs, s.n2, s.nu1, s.nu2).copy()
    bad.reshape(-1)[5] = 268369921  # = Q0: not a canonical residue
    with pytest.raises(ValueError, match="out of range"):
        dbfile.write(str(tmp_path / "y.spdb"), s, dbfile.KIND_NTT, bad)
    p = str(tmp_path / "db.spdb")
    dbfile.write(p, s, dbfile.KIND_RECORDS, recs)
    raw = bytearray(open(p, "rb").read())
    open(str(tmp_path / "trunc.spdb"), "wb").write(raw[:-1])
    with pytest.raises(ValueError, match="size"):
        dbfile.read(str(tmp_path / "trunc.spdb"))
    raw[0] = ord("X")
    open(str(tmp_path / "magic.spdb"), "wb").write(raw)
    with pytest.raises(ValueError, match="magic"):
        dbfile.read(str(tmp_path / "magic.spdb"))


def test_dat_database_reader(tmp_path, oracle):
    """A database written in the spiral-rs dump format (u64 per residue, residue Q0
    polynomial then residue Q1 per polynomial, Database order) reads back as the
    u32 layout Database.load_ntt takes; an out-of-range residue is rejected."""
    from spiral_pir import Shape, dbfile

    s = Shape(3, 2)
    db = oracle.db_encode(_records(s, 14), s.n2, s.nu1, s.nu2)
    p = str(tmp_path / "db.dat")
    db.astype(np.uint64).tofile(p)
    got = dbfile.read_dat_database(p, s, chunk_polys=5)  # chunks that do not divide the count
    assert got.shape == (s.n2, s.cols, s.rows, 2, N) and got.dtype == np.uint32
    np.testing.assert_array_equal(got, db.reshape(got.shape))
    w = db.astype(np.uint64)
    w.reshape(-1)[-1] = 249561089  # = Q1 in the last residue-Q1 polynomial
    w.tofile(p)
    with pytest.raises(ValueError, match="out of range"):
        dbfile.read_dat_database(p, s)


def test_out_of_range_ntt_container_is_refused_before_upload(tmp_path, oracle):
    """A container whose NTT payload holds a residue >= q_m (a corrupt or foreign
    file; dbfile.write refuses to make one, so it is patched in place) is rejected
    by Database.load_file before anything reaches the library, whose packed
    28-bit form would otherwise spill the word into a neighbouring column."""
    from spiral_pir import Shape, dbfile
    from spiral_pir.server import Database

    s = Shape(4, 2)
    db = oracle.db_encode(_records(s, 15), s.n2, s.nu1, s.nu2)
    p = str(tmp_path / "db.spdb")
    dbfile.write(p, s, dbfile.KIND_NTT, db)
    raw = bytearray(open(p, "rb").read())
    word = dbfile.HEADER.size + 4 * (3 * 2 * N + N + 17)  # residue q2 of polynomial 3, coefficient 17
    raw[word:word + 4] = (1 << 28).to_bytes(4, "little")
    open(p, "wb").write(raw)
    shape, kind, payload = dbfile.read(p)  # the container itself is well formed

    class Target:
        calls = []
        shape = s

        def load_ntt(self, payload, check=True):
            self.calls.append("ntt")

    t = Target()
    with pytest.raises(ValueError, match="out of range"):
        Database.load_file(t, p)
    assert t.calls == []
    with pytest.raises(ValueError, match="out of range"):
        Database.load_ntt(t, payload)
