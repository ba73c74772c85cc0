"""Serving and multi-database launches on the GPU: every response served in a
mixed stream equals the same query answered alone (bit for bit), and decodes
to the stored record."""
from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import rng

pytestmark = pytest.mark.gpu

N = 2048


def _db(gpu_ctx, nu1, nu2, seed):
    from spiral_pir import Shape
    from spiral_pir.server import Database

    s = Shape(nu1, nu2)
    recs = rng(seed).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
    db = Database(gpu_ctx, s)
    db.load_bytes(recs)
    return s, recs, db


def _pool(s, n_clients, per_client, seed):
    """Clients with keys and a few queries each: (client_id, client, keys, [(index, query, v_folding)])."""
    from spiral_pir.client import Client

    out = []
    for c in range(n_clients):
        cl = Client(s, seed + c)
        left, right = cl.keygen()
        qs = []
        for k in range(per_client):
            idx = int(rng(seed + 100 * c + k).integers(0, s.records))
            q, vf = cl.query(idx)
            qs.append((idx, q, vf))
        out.append((seed + c, cl, (left, right), qs))
    return out


def _alone(db, left, right, q, vf):
    from spiral_pir.server import Query

    one = Query(db)
    one.load(q, left, right, vf)
    one.run()
    r = one.read()
    one.close()
    return r


def test_batch_load_many_matches_per_slot_load(gpu_ctx):
    """Resident keys + bulk load give the same responses as per-slot uploads."""
    from spiral_pir.server import Batch, Keys

    s, recs, db = _db(gpu_ctx, 6, 3, 1200)
    pool = _pool(s, 2, 2, 5200)
    entries = [(cid, cl, kk, e) for cid, cl, kk, qs in pool for e in qs]
    keys = {cid: Keys(db, *kk) for cid, cl, kk, qs in pool}
    a, b = Batch(db, 4), Batch(db, 4)
    for slot, (cid, cl, kk, (idx, q, vf)) in enumerate(entries):
        a.load(slot, q, kk[0], kk[1], vf)
    b.load_many(0, [e[3][1] for e in entries], [keys[e[0]] for e in entries], [e[3][2] for e in entries])
    a.run(4)
    b.run(4)
    np.testing.assert_array_equal(b.responses(4), a.responses(4))
    for slot, (cid, cl, kk, (idx, q, vf)) in enumerate(entries):
        np.testing.assert_array_equal(cl.decode(b.read(slot)), recs[idx])
    # a partial bulk load into the middle slots
    b.load_many(1, [entries[0][3][1], entries[3][3][1]], [keys[entries[0][0]], keys[entries[3][0]]],
                [entries[0][3][2], entries[3][3][2]])
    b.run(4)
    rb = b.responses(4)
    np.testing.assert_array_equal(rb[1], a.responses(4)[0])
    np.testing.assert_array_equal(rb[2], a.responses(4)[3])
    for x in (a, b):
        x.close()
    for k in keys.values():
        k.close()
    db.close()


def test_async_loads_and_readouts_match_blocking(gpu_ctx):
    """Asynchronous bulk loads (queued behind the previous run on the device) and
    asynchronous readouts into pinned memory give the blocking path's responses:
    packed queries; fold-vector queries in two loads before one run; and a
    second load queued right after a launch, before that launch has finished,
    which must not disturb the first run's results."""
    import torch
    from spiral_pir.client import Client
    from spiral_pir.server import Batch, Keys
    from spiral_pir.serving import pinned_u32

    s, recs, db = _db(gpu_ctx, 6, 4, 1250)
    clients = [Client(s, 5250 + k) for k in range(2)]
    keys, mats = [], []
    for c in clients:
        left, right = c.keygen()
        keys.append(Keys(db, left, right, c.conversion_key()))
        mats.append((left, right))
    idxs = [int(i) for i in rng(125).integers(0, s.records, size=8)]
    packed = [clients[j % 2].query_packed(idxs[j]) for j in range(8)]
    vec = [clients[j % 2].query(idxs[j]) for j in range(8)]
    ks = [keys[j % 2] for j in range(8)]
    ref = Batch(db, 4)
    ref.load_packed(0, packed[:4], ks[:4])
    ref.run(4)
    ref_p1 = ref.responses(4)
    ref.load_packed(0, packed[4:], ks[4:])
    ref.run(4)
    ref_p2 = ref.responses(4)
    ref.load_many(0, [q for q, _ in vec[:4]], ks[:4], [v for _, v in vec[:4]])
    ref.run(4)
    ref_v = ref.responses(4)
    ref.close()
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    b = Batch(db, 4)
    out1, out2, out3 = (pinned_u32(4 * s.response_words) for _ in range(3))
    b.load_packed(0, packed[:4], ks[:4], asynchronous=True)
    b.graph_launch(4, sp)
    b.read_responses_async(4, out1, sp)
    b.load_packed(0, packed[4:], ks[4:], asynchronous=True)  # queued while the first run may still be going
    b.graph_launch(4, sp)
    b.read_responses_async(4, out2, sp)
    b.load_many(0, [q for q, _ in vec[:2]], ks[:2], [v for _, v in vec[:2]], asynchronous=True)
    b.load_many(2, [q for q, _ in vec[2:4]], ks[2:4], [v for _, v in vec[2:4]], asynchronous=True)
    b.run(4, stream=sp)
    b.read_responses_async(4, out3, sp)
    while not b.poll():
        pass
    shape = (4, s.n2, 2, 2, N)
    np.testing.assert_array_equal(out1.reshape(shape), ref_p1)
    np.testing.assert_array_equal(out2.reshape(shape), ref_p2)
    np.testing.assert_array_equal(out3.reshape(shape), ref_v)
    for j in range(4):
        np.testing.assert_array_equal(clients[j % 2].decode(out2.reshape(shape)[j]), recs[idxs[4 + j]])
    b.close()
    for k in keys:
        k.close()
    db.close()


def test_group_launch_three_databases(gpu_ctx):
    """Batches on three databases of different shapes in one graph launch, replayed
    twice with eager work between: every slot equals its query answered alone."""
    import torch
    from spiral_pir.server import Batch, Group

    lanes = []
    for nu1, nu2, nq, seed in ((6, 5, 3, 1300), (5, 6, 2, 1301), (4, 2, 2, 1302)):
        s, recs, db = _db(gpu_ctx, nu1, nu2, seed)
        pool = _pool(s, nq, 1, 5300 + 10 * nu1)
        b = Batch(db, nq)
        exp = []
        for slot, (cid, cl, (left, right), [(idx, q, vf)]) in enumerate(pool):
            b.load(slot, q, left, right, vf)
            exp.append((cl, idx, _alone(db, left, right, q, vf)))
        lanes.append((s, recs, db, b, exp))
    g = Group([ln[3] for ln in lanes])
    stream = torch.cuda.current_stream().cuda_stream
    ns = [len(ln[4]) for ln in lanes]
    for rep in range(2):
        (torch.randn(1 << 18, device="cuda") * 2).sum().item()
        g.launch(ns, stream)
        g.synchronize()
        for s, recs, db, b, exp in lanes:
            got = b.responses(len(exp))
            for slot, (cl, idx, ref) in enumerate(exp):
                np.testing.assert_array_equal(got[slot], ref, err_msg=f"rep {rep} shape {s.nu1},{s.nu2} slot {slot}")
                np.testing.assert_array_equal(cl.decode(got[slot]), recs[idx])
    # one lane idle in this launch
    g.launch([ns[0], 0, ns[2]], stream)
    g.synchronize()
    g.close()
    for s, recs, db, b, exp in lanes:
        b.close()
        db.close()


def test_serving_mixed_stream_equals_alone(gpu_ctx):
    """Continuous batching over two resident databases: requests arrive while others
    run; each response equals the same query answered alone."""
    from spiral_pir.serving import Server, latency_stats

    dbs, refs, recs_of = {}, {}, {}
    for name, (nu1, nu2, seed) in {"a": (6, 5, 1400), "b": (5, 6, 1401)}.items():
        s, recs, db = _db(gpu_ctx, nu1, nu2, seed)
        dbs[name] = db
        recs_of[name] = recs
        refs[name] = _pool(s, 3, 2, 5400 + 50 * nu1)
    srv = Server(dbs, max_batch=8, buffers=2)
    expected = {}
    for name, pool in refs.items():
        for cid, cl, (left, right), qs in pool:
            srv.register(name, cid, left, right)
            for k, (idx, q, vf) in enumerate(qs):
                expected[(name, cid, k)] = (cl, idx, _alone(dbs[name], left, right, q, vf))
    assert all(v["capacity"] == 8 for v in srv.admission["lanes"].values())
    srv.start()
    g = rng(77)
    reqs = []
    for it in range(48):
        name = "a" if g.random() < 0.7 else "b"
        cid, cl, kk, qs = refs[name][int(g.integers(0, 3))]
        k = int(g.integers(0, 2))
        reqs.append(srv.submit(name, cid, qs[k][1], qs[k][2], tag=(name, cid, k)))
        if it % 7 == 0:
            time.sleep(0.002)
    for r in reqs:
        r.wait(timeout=60)
    srv.stop()
    for r in reqs:
        cl, idx, ref = expected[r.tag]
        np.testing.assert_array_equal(r.response, ref, err_msg=str(r.tag))
    for (name, cid, k), (cl, idx, ref) in expected.items():
        np.testing.assert_array_equal(cl.decode(ref), recs_of[name][idx])
    st = latency_stats(reqs)
    assert st["completed"] == 48 and st["p99_ms"] >= st["p50_ms"] > 0
    lanes = srv.lanes.values()
    assert sum(ln.launched_queries for ln in lanes) == 48
    assert sum(ln.launched_batches for ln in lanes) < 48  # queries were batched together
    srv.close()
    for db in dbs.values():
        db.close()


def test_serving_warmup_and_size_classes(gpu_ctx):
    """Warm-up captures every size class; later batches of every size answer correctly."""
    from spiral_pir.serving import Server

    s, recs, db = _db(gpu_ctx, 5, 5, 1500)
    pool = _pool(s, 2, 3, 5500)
    srv = Server({"x": db}, max_batch=6, buffers=1)
    exp = {}
    for cid, cl, (left, right), qs in pool:
        srv.register("x", cid, left, right)
        for k, (idx, q, vf) in enumerate(qs):
            exp[(cid, k)] = (cl, idx, _alone(db, left, right, q, vf))
    cid0, _, _, qs0 = pool[0]
    srv.warmup("x", cid0, qs0[0][1], qs0[0][2])
    assert {c for (_, c) in srv.lanes["x"].ready} == {1, 2, 3, 4, 5, 6}
    srv.start()
    for n in (1, 3, 6, 5, 2):
        batch = []
        for j in range(n):
            cid, cl, kk, qs = pool[j % 2]
            k = j % 3
            batch.append(srv.submit("x", cid, qs[k][1], qs[k][2], tag=(cid, k)))
        for r in batch:
            r.wait(timeout=60)
            np.testing.assert_array_equal(r.response, exp[r.tag][2])
    srv.stop()
    srv.close()
    db.close()


def test_serving_packed_and_vector_requests(gpu_ctx):
    """Packed queries (fold vectors converted on the device) and queries with
    fold vectors, interleaved in one lane: each response equals the query
    answered alone and decodes to the record."""
    from spiral_pir.client import Client
    from spiral_pir.serving import Server

    s, recs, db = _db(gpu_ctx, 6, 4, 1600)
    srv = Server({"x": db}, max_batch=4, buffers=2)
    clients = []
    for k in range(2):
        cl = Client(s, 5600 + k)
        left, right = cl.keygen()
        conv = cl.conversion_key()
        srv.register("x", k, left, right, conv)
        clients.append((cl, left, right, conv))
    from spiral_pir.server import Keys, Query

    expected = {}
    reqs_in = []
    g = rng(88)
    for it in range(14):
        k = int(g.integers(0, 2))
        cl, left, right, conv = clients[k]
        idx = int(g.integers(0, s.records))
        one = Query(db)
        if it % 3 == 0:
            q, vf = cl.query(idx)
            one.load(q, left, right, vf)
        else:
            q, vf = cl.query_packed(idx), None
            kk = Keys(db, left, right, conv)
            one.load_packed(q, kk)
        one.run()
        expected[it] = (cl, idx, one.read())
        one.close()
        if vf is None:
            kk.close()
        reqs_in.append((k, q, vf, it))
    srv.start()
    reqs = [srv.submit("x", k, q, vf, tag=it) for (k, q, vf, it) in reqs_in]
    for r in reqs:
        r.wait(timeout=60)
    srv.stop()
    for r in reqs:
        cl, idx, ref = expected[r.tag]
        np.testing.assert_array_equal(r.response, ref, err_msg=str(r.tag))
        np.testing.assert_array_equal(cl.decode(r.response), recs[idx])
    srv.close()
    db.close()
