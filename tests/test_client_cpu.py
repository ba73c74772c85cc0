"""The client (product, host code) with the oracle as the server: the decoded
record must be the stored one.  This pins the client (not in the reference)
and, with it, the oracle's expansion rounds >= 1, database multiply and fold,
which no reference fixture covers beyond round 0."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng


def _case(oracle, nu1, nu2, index, seed_db=1000, seed_client=2000):
    from spiral_pir import Shape
    from spiral_pir.client import Client

    s = Shape(nu1, nu2)
    recs = rng(seed_db).integers(0, 256, size=(s.records, s.n2, 2048), dtype=np.uint8)
    db = oracle.db_encode(recs, s.n2, nu1, nu2)
    c = Client(s, seed_client)
    left, right = c.keygen()
    q, vf = c.query(index)
    vneg = oracle.folding_neg(vf, nu2, s.t_gsw)
    out = oracle.answer(db, s.n2, nu1, nu2, s.t_gsw, s.t_exp, q, left, right, vf, vneg)
    return c, recs, out, (q, left, right, vf)


def test_end_to_end_config0_shape(oracle):
    """(nu1, nu2) = (6, 2), n = 2 (4 sub-databases), record 37."""
    c, recs, out, _ = _case(oracle, 6, 2, 37)
    np.testing.assert_array_equal(c.decode(out), recs[37])


@pytest.mark.parametrize("index", [0, 255, 128 + 3])
def test_end_to_end_indices(oracle, index):
    c, recs, out, _ = _case(oracle, 6, 2, index, seed_client=2100 + index)
    np.testing.assert_array_equal(c.decode(out), recs[index])


def test_expanded_selectors_decrypt_to_one_hot(oracle):
    """After expansion, selector i1 encrypts floor(Q/p) and every other first-dimension
    selector encrypts 0 (decoded with p = 256)."""
    nu1, nu2, idx = 6, 2, 37
    c, recs, out, (q, left, right, vf) = _case(oracle, nu1, nu2, idx)
    from spiral_pir import Shape

    s = Shape(nu1, nu2)
    v = oracle.expand(q, nu1, nu2, s.t_gsw, s.t_exp, left, right)
    i1 = idx >> nu2
    for j in (0, i1, 5, 63):
        raw = np.stack([oracle.from_ntt(v[2 * j, 0]), oracle.from_ntt(v[2 * j, 1])])
        resp = np.zeros((s.n2, 2, 2, 2048), np.uint32)
        resp[0] = raw
        dec = c.decode(resp)[0]
        assert dec[0] == (1 if j == i1 else 0), j
        assert not dec[1:].any()


def test_client_is_deterministic():
    from spiral_pir import Shape
    from spiral_pir.client import Client

    s = Shape(6, 2)
    a, b = Client(s, 5), Client(s, 5)
    np.testing.assert_array_equal(a.keygen()[0], b.keygen()[0])
    np.testing.assert_array_equal(a.query(3)[0], b.query(3)[0])
    assert not np.array_equal(Client(s, 6).query(3)[0], Client(s, 5).query(3)[0])


def test_client_rejects_bad_index():
    from spiral_pir import Shape, SpiralError
    from spiral_pir.client import Client

    c = Client(Shape(6, 2), 1)
    with pytest.raises(SpiralError):
        c.query(1 << 8)
