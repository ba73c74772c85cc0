"""Batched answers on the GPU: every slot of a batch must equal the single-query
answer (and the oracle), alone or in the mix, eagerly or replayed from a
captured hipGraph."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

N = 2048


def _setup(gpu_ctx, nu1, nu2, seed):
    from spiral_pir import Shape
    from spiral_pir.server import Database

    s = Shape(nu1, nu2)
    recs = rng(seed).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
    db = Database(gpu_ctx, s)
    db.load_bytes(recs)
    return s, recs, db


def _queries(s, n, seed0):
    from spiral_pir.client import Client

    out = []
    for k in range(n):
        c = Client(s, seed0 + k)
        left, right = c.keygen()
        idx = int(rng(seed0 + 1000 + k).integers(0, s.records))
        q, vf = c.query(idx)
        out.append((c, idx, q, left, right, vf))
    return out


def test_batch_slots_match_oracle(gpu_ctx, oracle):
    from spiral_pir.server import Batch

    s, recs, db = _setup(gpu_ctx, 6, 2, 900)
    qs = _queries(s, 5, 3000)
    b = Batch(db, 6)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    b.run(5)
    resp = b.responses(5)
    db_ntt = oracle.db_encode(recs, s.n2, s.nu1, s.nu2)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        exp = oracle.answer(db_ntt, s.n2, s.nu1, s.nu2, s.t_gsw, s.t_exp, q, left, right, vf,
                            oracle.folding_neg(vf, s.nu2, s.t_gsw))
        np.testing.assert_array_equal(resp[slot], exp, err_msg=f"slot {slot}")
        np.testing.assert_array_equal(b.read(slot), exp)
        np.testing.assert_array_equal(c.decode(resp[slot]), recs[idx])
    b.close()
    db.close()


def test_same_result_alone_and_in_the_mix(gpu_ctx):
    """A query answered alone (batch of 1) and inside batches of different sizes and
    positions gives the same response."""
    from spiral_pir.server import Batch, Query

    s, recs, db = _setup(gpu_ctx, 5, 3, 901)
    qs = _queries(s, 7, 3100)
    c, idx, q, left, right, vf = qs[3]
    alone = Query(db)
    alone.load(q, left, right, vf)
    alone.run()
    ref = alone.read()
    alone.close()
    for order in ([3, 0, 1, 2, 4, 5, 6], [6, 5, 4, 3], [0, 1, 2, 3, 4, 5, 6]):
        b = Batch(db, len(order))
        for slot, k in enumerate(order):
            cc, ii, qq, ll, rr, vv = qs[k]
            b.load(slot, qq, ll, rr, vv)
        b.run(len(order))
        got = b.read(order.index(3))
        np.testing.assert_array_equal(got, ref)
        b.close()
    np.testing.assert_array_equal(c.decode(ref), recs[idx])
    db.close()


def test_graph_replay_matches_eager(gpu_ctx):
    """Replays of the captured stage chain, with eager work and synchronizations in
    between, give the eager result every time."""
    import torch
    from spiral_pir.server import Batch

    s, recs, db = _setup(gpu_ctx, 5, 2, 902)
    qs = _queries(s, 4, 3200)
    b = Batch(db, 4)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    b.run(4)
    eager = b.responses(4)
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(3):
        x = torch.randn(1 << 20, device="cuda")
        (x * 2).sum().item()  # unrelated eager work between replays
        b.graph_launch(4, stream)
        b.synchronize()
        np.testing.assert_array_equal(b.responses(4), eager, err_msg=f"replay {rep}")
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        np.testing.assert_array_equal(c.decode(eager[slot]), recs[idx])
    b.close()
    db.close()


def test_batch_partial_run_uses_first_slots(gpu_ctx):
    from spiral_pir.server import Batch

    s, recs, db = _setup(gpu_ctx, 4, 2, 903)
    qs = _queries(s, 3, 3300)
    b = Batch(db, 8)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    b.run(3)
    resp = b.responses(3)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        np.testing.assert_array_equal(c.decode(resp[slot]), recs[idx])
    b.close()
    db.close()


def test_batch_config2_shape(gpu_ctx):
    """(9, 6) with 8 queries in one batch: every slot equals the single-query answer
    (itself checked against the oracle at this shape) and decodes."""
    from spiral_pir.server import Batch, Query

    s, recs, db = _setup(gpu_ctx, 9, 6, 1001)
    qs = _queries(s, 8, 3400)
    b = Batch(db, 8)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    b.graph_launch(8)
    resp = b.responses(8)
    single = Query(db)
    for slot in (0, 5):
        c, idx, q, left, right, vf = qs[slot]
        single.load(q, left, right, vf)
        single.run()
        np.testing.assert_array_equal(resp[slot], single.read())
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        np.testing.assert_array_equal(c.decode(resp[slot]), recs[idx])
    single.close()
    b.close()
    db.close()
