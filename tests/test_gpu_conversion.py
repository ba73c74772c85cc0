"""Regev -> GSW conversion on the GPU (packed queries): the device conversion
against the oracle's, bit for bit, and packed queries answered end to end.

The reference leaves the conversion a stub (gpu-pir/cpu_version/conversion.py:3);
the oracle's restatement is pinned by decryption (tests/test_client_cpu.py).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import live_elements, rand_residues, rng

pytestmark = pytest.mark.gpu

N = 2048
Q0, Q1 = 268369921, 249561089


def dev(a: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32).copy()).cuda()


def host(t) -> np.ndarray:
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy().view(np.uint32)


def _rand_key(g, cols: int) -> np.ndarray:
    """Key matrix CiphertextHalf[2 crt][cols] with canonical residues."""
    k = np.empty((2, cols, 2, N), np.uint32)
    k[0] = g.integers(0, Q0, size=(cols, 2, N), dtype=np.uint32)
    k[1] = g.integers(0, Q1, size=(cols, 2, N), dtype=np.uint32)
    return k


@pytest.mark.parametrize("nu2", [2, 6])
def test_regev_to_gsw_matches_oracle(gpu_ctx, oracle, nu2):
    """Random ciphertexts (plus all-zero and all-(q-1) ones, whose raw values are
    0 and Q - 1) and a random key: every output word equals the oracle's."""
    from spiral_pir import ops

    g = rng(7000 + nu2)
    n_ct = 2 * nu2 * 8 + 3
    v = rand_residues(g, 2 * n_ct).reshape(n_ct, 2, 2, N)
    v[1] = 0
    v[3, :, 0] = Q0 - 1
    v[3, :, 1] = Q1 - 1
    conv = _rand_key(g, 8)
    out = dev(np.zeros((nu2, 2, 16, 2, N), np.uint32))
    ops.regev_to_gsw(gpu_ctx, dev(v), dev(conv), out, nu2)
    got = host(out).reshape(nu2, 2, 16, 2, N)
    want = oracle.regev_to_gsw(v, nu2, 8, 4, conv)
    for lvl in range(nu2):
        np.testing.assert_array_equal(got[lvl], want[lvl], err_msg=f"level {lvl}")


def test_regev_to_gsw_rejects_bad_arguments(gpu_ctx):
    from spiral_pir import SpiralError, ops

    v = dev(np.zeros((8, 2, 2, N), np.uint32))  # 8 ciphertexts < 2 * 2 * 8
    with pytest.raises(SpiralError):
        ops.regev_to_gsw(gpu_ctx, v, dev(np.zeros((2, 8, 2, N), np.uint32)),
                         dev(np.zeros((2, 2, 16, 2, N), np.uint32)), 2)


def _db(gpu_ctx, nu1, nu2, seed):
    from spiral_pir import Shape
    from spiral_pir.server import Database

    s = Shape(nu1, nu2)
    recs = rng(seed).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
    db = Database(gpu_ctx, s)
    db.load_bytes(recs)
    return s, recs, db


def test_batch_packed_matches_oracle(gpu_ctx, oracle):
    """(6, 2): three packed queries from two clients in one batch; expanded
    ciphertexts, converted fold vectors and responses all equal the oracle's;
    the responses decode to the records.  Graph replay and eager run agree."""
    from spiral_pir.client import Client
    from spiral_pir.server import READ_EXPANDED, READ_V_FOLDING, Batch, Keys

    s, recs, db = _db(gpu_ctx, 6, 2, 7100)
    db_ntt = oracle.db_encode(recs, s.n2, 6, 2)
    clients = [Client(s, 7200 + i) for i in range(2)]
    mats = [(c,) + c.keygen() + (c.conversion_key(),) for c in clients]
    keys = [Keys(db, left, right, conv) for (c, left, right, conv) in mats]
    entries = [(0, 37), (1, 200), (0, 3)]
    qs = [mats[ci][0].query_packed(idx) for ci, idx in entries]
    b = Batch(db, 3)
    b.load_packed(0, qs, [keys[ci] for ci, _ in entries])
    b.graph_launch(3)
    b.synchronize()
    resp_graph = b.responses(3)
    live = live_elements(oracle, 6, 2, s.t_gsw)  # a whole answer forms only what it reads
    for slot, (ci, idx) in enumerate(entries):
        c, left, right, conv = mats[ci]
        v = oracle.expand(qs[slot], 6, 2, s.t_gsw, s.t_exp, left, right)
        np.testing.assert_array_equal(b.read(slot, READ_EXPANDED)[live], v[live], err_msg=f"slot {slot} expansion")
        vf = oracle.regev_to_gsw(v, 2, s.t_gsw, 4, conv)
        np.testing.assert_array_equal(b.read(slot, READ_V_FOLDING), vf, err_msg=f"slot {slot} conversion")
        want = oracle.answer(db_ntt, s.n2, 6, 2, s.t_gsw, s.t_exp, qs[slot], left, right, vf,
                             oracle.folding_neg(vf, 2, s.t_gsw))
        np.testing.assert_array_equal(resp_graph[slot], want, err_msg=f"slot {slot} response")
        np.testing.assert_array_equal(c.decode(resp_graph[slot]), recs[idx])
    b.run(3)
    np.testing.assert_array_equal(b.responses(3), resp_graph)
    b.close()
    for k in keys:
        k.close()
    db.close()


def test_mixed_packed_and_vector_slots(gpu_ctx):
    """One run holding a packed slot and a slot loaded with fold vectors: both
    answers equal the same queries answered alone; the graph is recaptured when
    the form of the slots changes."""
    from spiral_pir.client import Client
    from spiral_pir.server import Batch, Keys, Query

    s, recs, db = _db(gpu_ctx, 5, 3, 7300)
    c = Client(s, 7400)
    left, right = c.keygen()
    conv = c.conversion_key()
    keys = Keys(db, left, right, conv)
    qp = c.query_packed(100)
    qv, vf = c.query(17)
    alone = []
    for load in (lambda q: q.load_packed(qp, keys), lambda q: q.load(qv, left, right, vf)):
        one = Query(db)
        load(one)
        one.run()
        alone.append(one.read())
        one.close()
    b = Batch(db, 2)
    b.load_packed(0, [qp], [keys])
    b.load_many(1, [qv], [keys], [vf])
    for rep in range(2):
        b.graph_launch(2)
        got = b.responses(2)
        np.testing.assert_array_equal(got[0], alone[0], err_msg=f"rep {rep} packed slot")
        np.testing.assert_array_equal(got[1], alone[1], err_msg=f"rep {rep} vector slot")
    np.testing.assert_array_equal(c.decode(got[0]), recs[100])
    np.testing.assert_array_equal(c.decode(got[1]), recs[17])
    # slot 0 switches to fold vectors: no slot is packed any more
    b.load_many(0, [qv], [keys], [vf])
    b.graph_launch(2)
    got = b.responses(2)
    np.testing.assert_array_equal(got[0], alone[1])
    np.testing.assert_array_equal(got[1], alone[1])
    b.close()
    keys.close()
    db.close()


def test_packed_query_nu9_6(gpu_ctx, oracle):
    """(9, 6): a packed query decodes; its converted fold vectors equal the
    oracle's conversion of the device's expansion, and the answer equals the
    one computed from those fold vectors given explicitly."""
    from spiral_pir.client import Client
    from spiral_pir.server import READ_EXPANDED, READ_V_FOLDING, Keys, Query

    s, recs, db = _db(gpu_ctx, 9, 6, 7500)
    c = Client(s, 7600)
    left, right = c.keygen()
    conv = c.conversion_key()
    keys = Keys(db, left, right, conv)
    idx = (300 << 6) | 45
    qp = c.query_packed(idx)
    one = Query(db)
    one.load_packed(qp, keys)
    one.run()
    resp = one.read()
    np.testing.assert_array_equal(c.decode(resp), recs[idx])
    v = one.read(READ_EXPANDED)
    vf = one.read(READ_V_FOLDING)
    np.testing.assert_array_equal(vf, oracle.regev_to_gsw(v, 6, s.t_gsw, 4, conv))
    two = Query(db)
    two.load(qp, left, right, vf)
    two.run()
    np.testing.assert_array_equal(two.read(), resp)
    one.close()
    two.close()
    keys.close()
    db.close()


def test_packed_keys_required(gpu_ctx):
    from spiral_pir import SpiralError
    from spiral_pir.client import Client
    from spiral_pir.server import Batch, Keys

    s, recs, db = _db(gpu_ctx, 5, 2, 7700)
    c = Client(s, 7800)
    left, right = c.keygen()
    keys = Keys(db, left, right)  # no conversion key
    b = Batch(db, 1)
    with pytest.raises(SpiralError, match="conversion key"):
        b.load_packed(0, [c.query_packed(1)], [keys])
    b.close()
    keys.close()
    db.close()
