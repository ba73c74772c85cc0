"""Databases loaded from files on the GPU: a records container, an NTT-form
container and a spiral-rs NTT dump give the same resident database (the same
answers) as the records loaded from memory."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

N = 2048


def test_database_from_files(gpu_ctx, oracle, tmp_path):
    from spiral_pir import Shape, dbfile
    from spiral_pir.client import Client
    from spiral_pir.server import Database, Query

    s = Shape(6, 3)
    recs = rng(5100).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
    db_ntt = oracle.db_encode(recs, s.n2, s.nu1, s.nu2)
    p_rec, p_ntt, p_dat = (str(tmp_path / n) for n in ("r.spdb", "n.spdb", "d.dat"))
    dbfile.write(p_rec, s, dbfile.KIND_RECORDS, recs)
    dbfile.write(p_ntt, s, dbfile.KIND_NTT, db_ntt)
    db_ntt.astype(np.uint64).tofile(p_dat)
    c = Client(s, 5200)
    left, right = c.keygen()
    idx = 300
    q, vf = c.query(idx)

    def answer(load):
        db = Database(gpu_ctx, s)
        load(db)
        one = Query(db)
        one.load(q, left, right, vf)
        one.run()
        r = one.read()
        one.close()
        db.close()
        return r

    ref = answer(lambda db: db.load_bytes(recs))
    np.testing.assert_array_equal(c.decode(ref), recs[idx])
    for name, load in (("records file", lambda db: db.load_file(p_rec)),
                       ("ntt file", lambda db: db.load_file(p_ntt)),
                       ("spiral-rs dump", lambda db: db.load_ntt(dbfile.read_dat_database(p_dat, s)))):
        np.testing.assert_array_equal(answer(load), ref, err_msg=name)
    other = Database(gpu_ctx, Shape(6, 2))
    with pytest.raises(ValueError, match="not this database"):
        other.load_file(p_rec)
    other.close()
