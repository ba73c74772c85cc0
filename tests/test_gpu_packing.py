"""Response packing + modulus switching on the GPU: the device's compact
responses against the oracle's, byte for byte, and decoded by the client.

Not in the reference (spiral-rs code outside it); the oracle's restatement is
pinned by decryption and exact rounding checks (tests/test_client_cpu.py).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

N = 2048
Q0, Q1 = 268369921, 249561089
QQ = Q0 * Q1


def dev(a: np.ndarray):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.uint32).view(np.int32).copy()).cuda()


def _raw(values: np.ndarray) -> np.ndarray:
    """u64 coefficients [k][N] -> raw polys [k][2][N] (hi, lo)."""
    v = values.astype(np.uint64)
    return np.stack([(v >> np.uint64(32)).astype(np.uint32), (v & np.uint64(0xFFFFFFFF)).astype(np.uint32)], axis=1)


@pytest.mark.parametrize("q2_bits", [11, 20, 32])
def test_pack_response_matches_oracle(gpu_ctx, oracle, q2_bits):
    """Random raw ciphertexts (with coefficients 0, 1, Q/2 and Q - 1 planted) and
    random packing keys: every byte equals the oracle's."""
    import torch
    from spiral_pir import ops

    g = rng(8000 + q2_bits)
    vals = g.integers(0, QQ, size=(8, N), dtype=np.uint64)
    vals[:, :4] = np.array([0, 1, QQ // 2, QQ - 1], np.uint64)
    raw = _raw(vals).reshape(4, 2, 2, N)
    w = np.empty((2, 2, 4, 3, N), np.uint32)
    w[:, 0] = g.integers(0, Q0, size=(2, 4, 3, N), dtype=np.uint32)
    w[:, 1] = g.integers(0, Q1, size=(2, 4, 3, N), dtype=np.uint32)
    want = oracle.pack_response(raw, 2, w, 4, q2_bits)
    out = torch.zeros(want.size, dtype=torch.uint8, device="cuda")
    ops.pack_response(gpu_ctx, dev(raw), dev(w), out, q2_bits)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def _db(gpu_ctx, nu1, nu2, seed):
    from spiral_pir import Shape
    from spiral_pir.server import Database

    s = Shape(nu1, nu2)
    recs = rng(seed).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
    db = Database(gpu_ctx, s)
    db.load_bytes(recs)
    return s, recs, db


def test_batch_packed_responses(gpu_ctx, oracle):
    """(6, 2): packed queries with conversion and response packing in one graph:
    compact responses equal the oracle's packing of the device's own raw
    responses and decode to the records; a slot loaded without packing keys has
    no compact response."""
    from spiral_pir import SpiralError
    from spiral_pir.client import Client
    from spiral_pir.server import Batch, Keys

    s, recs, db = _db(gpu_ctx, 6, 2, 8100)
    clients = [Client(s, 8200 + i) for i in range(2)]
    mats = []
    for c in clients:
        left, right = c.keygen()
        w = c.packing_keys()
        mats.append((c, Keys(db, left, right, c.conversion_key(), w), w))
    entries = [(0, 37), (1, 200), (0, 3)]
    qs = [mats[ci][0].query_packed(idx) for ci, idx in entries]
    b = Batch(db, 3)
    b.set_packing(20)
    b.load_packed(0, qs, [mats[ci][1] for ci, _ in entries])
    for rep in range(2):
        b.graph_launch(3)
        packed = b.packed_responses(3)
        raw = b.responses(3)
        for slot, (ci, idx) in enumerate(entries):
            c, _, w = mats[ci]
            np.testing.assert_array_equal(packed[slot], oracle.pack_response(raw[slot], 2, w, 4, 20),
                                          err_msg=f"rep {rep} slot {slot}")
            np.testing.assert_array_equal(c.decode_packed(packed[slot], 20), recs[idx])
    c0 = clients[0]
    left, right = c0.keygen()
    plain = Keys(db, left, right, c0.conversion_key())  # no packing keys
    b.load_packed(1, [c0.query_packed(5)], [plain])
    b.graph_launch(3)
    with pytest.raises(SpiralError, match="packing keys"):
        b.packed_responses(3)
    b.close()
    plain.close()
    for _, k, _ in mats:
        k.close()
    db.close()


def test_packing_width_changes_between_graph_launches(gpu_ctx, oracle):
    """q2_bits cycled 11 -> 20 -> 11 -> 32 -> 20 between graph launches: each width
    has its own captured chain and the device buffer of the compact responses
    never moves, so every launch's compact responses equal those of an eager run
    and the oracle's packing of the raw responses, and decode (q2_bits >= 16).  A readout after
    the width changed without a new run is refused."""
    import torch
    from spiral_pir import SpiralError
    from spiral_pir.client import Client
    from spiral_pir.server import Batch, Keys

    s, recs, db = _db(gpu_ctx, 6, 2, 8150)
    c = Client(s, 8250)
    left, right = c.keygen()
    w = c.packing_keys()
    keys = Keys(db, left, right, c.conversion_key(), w)
    idxs = (7, 190)
    b = Batch(db, 2)
    b.load_packed(0, [c.query_packed(i) for i in idxs], [keys, keys])
    stream = torch.cuda.current_stream().cuda_stream
    for q2 in (11, 20, 11, 32, 20):
        b.set_packing(q2)
        (torch.randn(1 << 16, device="cuda") * 2).sum().item()
        b.graph_launch(2, stream)
        got = b.packed_responses(2)
        raw = b.responses(2)
        b.run(2)
        np.testing.assert_array_equal(b.packed_responses(2), got, err_msg=f"q2 {q2}: graph vs eager")
        for slot, idx in enumerate(idxs):
            np.testing.assert_array_equal(got[slot], oracle.pack_response(raw[slot], 2, w, 4, q2),
                                          err_msg=f"q2 {q2} slot {slot}")
            if q2 >= 16:  # 11 bits of row 0 keep too little of a * s for the record to decode
                np.testing.assert_array_equal(c.decode_packed(got[slot], q2), recs[idx])
    assert b.last_run()["graphs"] == 3  # one chain per width
    b.set_packing(11)
    with pytest.raises(SpiralError, match="packing changed"):
        b.packed_responses(2)
    b.close()
    keys.close()
    db.close()


def test_packed_query_and_response_nu9_6(gpu_ctx, oracle):
    """(9, 6): the whole Spiral server path for one query (expansion, Regev -> GSW
    conversion, database multiply, fold, packing, modulus switching) decodes."""
    from spiral_pir.client import Client
    from spiral_pir.server import Keys, Query

    s, recs, db = _db(gpu_ctx, 9, 6, 8300)
    c = Client(s, 8400)
    left, right = c.keygen()
    w = c.packing_keys()
    keys = Keys(db, left, right, c.conversion_key(), w)
    idx = (111 << 6) | 22
    one = Query(db)
    one.set_packing(20)
    one.load_packed(c.query_packed(idx), keys)
    one.run()
    packed = one.read_packed()
    np.testing.assert_array_equal(packed, oracle.pack_response(one.read(), 2, w, 4, 20))
    np.testing.assert_array_equal(c.decode_packed(packed, 20), recs[idx])
    one.close()
    keys.close()
    db.close()


def test_serving_packed_responses(gpu_ctx):
    """A server that packs responses: packed queries complete with compact bytes
    that decode to the records."""
    from spiral_pir.client import Client
    from spiral_pir.serving import Server

    s, recs, db = _db(gpu_ctx, 6, 2, 8500)
    srv = Server({"x": db}, max_batch=4, buffers=2, pack_q2_bits=20)
    c = Client(s, 8600)
    left, right = c.keygen()
    srv.register("x", "c", left, right, c.conversion_key(), c.packing_keys())
    srv.start()
    idxs = [int(i) for i in rng(86).integers(0, s.records, size=9)]
    reqs = [srv.submit("x", "c", c.query_packed(i), None, tag=i) for i in idxs]
    for r in reqs:
        r.wait(timeout=60)
    srv.stop()
    for r in reqs:
        assert r.response.size == s.packed_response_bytes(20)
        np.testing.assert_array_equal(c.decode_packed(r.response, 20), recs[r.tag])
    srv.close()
    db.close()
