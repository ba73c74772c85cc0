"""The benchmark's configurations at their full sizes, bit for bit against the
oracle on the GPU:

* (10, 7): the single-query answer stage by stage (expansion g = 11, a
  1024-row database multiply, a 7-level fold in the latency form) and the
  split fold form; the expansion with random keys, element by element;
* 2^nu1 = 1024 rows through the int8-limb MFMA multiply (the persistent
  LDS-staged kernel with 32 row blocks and the streaming kernel) and the VALU
  kernels, random and all-(q-1) databases; the LDS-staged kernels for 1, 2, 4
  and 8 row blocks;
* (8, 4), 16 queries: the VALU batched path (no MFMA at 2^nu2 = 16);
* (9, 6), 64 queries: the autotuned graph batch, slots against the oracle;
* 48/32/16 queries over (9, 6)/(10, 7)/(8, 4) in one group launch;
* Poisson serving, 70/30 over (9, 6)/(10, 7), every response equal to the
  query answered alone and, for one query per database, to the oracle.

The databases are shared by the tests of this module (module-scoped) and the
oracle's NTT form of each is computed once.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rand_residues, rng

pytestmark = pytest.mark.gpu

N = 2048
Q0, Q1 = 268369921, 249561089
SEEDS = {(9, 6): 1001, (10, 7): 1002, (8, 4): 1003, (10, 5): 1005}


class _Dbs:
    """Resident databases of uniform random records and their oracle NTT forms, built on first use."""

    def __init__(self, ctx, oracle):
        self.ctx, self.oracle = ctx, oracle
        self.items: dict = {}

    def get(self, nu1: int, nu2: int):
        from spiral_pir import Shape
        from spiral_pir.server import Database

        key = (nu1, nu2)
        if key not in self.items:
            s = Shape(nu1, nu2)
            recs = rng(SEEDS[key]).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
            db = Database(self.ctx, s)
            db.load_bytes(recs)
            self.items[key] = (s, recs, db, self.oracle.db_encode(recs, s.n2, nu1, nu2))
        return self.items[key]

    def drop(self, nu1: int, nu2: int) -> None:
        item = self.items.pop((nu1, nu2), None)
        if item:
            item[2].close()

    def close(self) -> None:
        for key in list(self.items):
            self.drop(*key)


@pytest.fixture(scope="module")
def dbs(gpu_ctx, oracle):
    oracle.set_threads(16)
    d = _Dbs(gpu_ctx, oracle)
    yield d
    d.close()
    oracle.set_threads(1)


def _clients_queries(s, n, n_clients, seed):
    from spiral_pir.client import Client

    clients = [Client(s, seed + k) for k in range(n_clients)]
    keys = [c.keygen() for c in clients]
    idx = [int(x) for x in rng(seed + 100).integers(0, s.records, size=n)]
    out = []
    for j in range(n):
        c = j % n_clients
        q, vf = clients[c].query(idx[j])
        out.append((clients[c], idx[j], q, keys[c][0], keys[c][1], vf))
    return out


def _oracle_answer(oracle, s, db_ntt, q, left, right, vf):
    return oracle.answer(db_ntt, s.n2, s.nu1, s.nu2, s.t_gsw, s.t_exp, q, left, right, vf,
                         oracle.folding_neg(vf, s.nu2, s.t_gsw))


def _alone(db, entries):
    """Responses of the queries answered one at a time (single-query objects)."""
    from spiral_pir.server import Query

    one = Query(db)
    out = []
    for (cl, idx, q, left, right, vf) in entries:
        one.load(q, left, right, vf)
        one.run()
        out.append(one.read())
    one.close()
    return out


def _rand_keys(g, levels, cols):
    """Key matrices [level][crt][col][row][z] of canonical random residues."""
    k = np.empty((levels, 2, cols, 2, N), np.uint32)
    k[:, 0] = g.integers(0, Q0, size=(levels, cols, 2, N), dtype=np.uint32)
    k[:, 1] = g.integers(0, Q1, size=(levels, cols, 2, N), dtype=np.uint32)
    return k


# ---- (10, 7): single query, expansion, fold -----------------------------------

def test_answer_nu10_7_stage_by_stage(gpu_ctx, oracle, dbs):
    """(10, 7), one query: expansion (g = 11, latency and split rounds), the VALU
    multiply over 1024 rows, from_ntt and the 7-level fold in the latency form,
    each stage equal to the oracle; then the split fold form gives the same
    response; the record decodes."""
    from spiral_pir.server import Query, READ_DB_MUL, READ_EXPANDED, STAGE_DB_MUL, STAGE_EXPAND, STAGE_FOLD

    s, recs, db, db_ntt = dbs.get(10, 7)
    (cl, idx, q, left, right, vf), = _clients_queries(s, 1, 1, 2002)
    one = Query(db)
    one.load(q, left, right, vf)
    one.run(STAGE_EXPAND)
    v_gpu = one.read(READ_EXPANDED)
    v_exp = oracle.expand(q, s.nu1, s.nu2, s.t_gsw, s.t_exp, left, right)
    np.testing.assert_array_equal(v_gpu, v_exp)
    one.run(STAGE_DB_MUL)
    acc = one.read(READ_DB_MUL)
    sel = np.ascontiguousarray(v_exp[0:2 * s.rows:2])
    for t in range(s.n2):
        np.testing.assert_array_equal(acc[t], oracle.db_mul(db_ntt[t], sel, s.nu1, s.nu2, nthreads=0), err_msg=f"trial {t}")
    one.run(STAGE_FOLD)
    resp = one.read()
    exp = _oracle_answer(oracle, s, db_ntt, q, left, right, vf)
    np.testing.assert_array_equal(resp, exp)
    gpu_ctx.set_variant(1, 4)  # the split fold form (partial products, then finish, per level)
    try:
        one.run()
        np.testing.assert_array_equal(one.read(), exp)
    finally:
        gpu_ctx.set_variant(1, 0)
    np.testing.assert_array_equal(cl.decode(resp), recs[idx])
    one.close()


@pytest.mark.parametrize("variant", [0, 4, 21], ids=["auto", "split", "halves"])
def test_expansion_nu10_7_random_keys(gpu_ctx, oracle, variant):
    """g = 11 rounds with random query and keys (not client keys: every key word
    arbitrary), element by element against the oracle."""
    from spiral_pir import ops

    from test_gpu_parity import dev, host

    nu1, nu2, t = 10, 7, 8
    g = rng(1107 + variant)
    gg, st = oracle.expansion_g(nu1, nu2, t), oracle.expansion_stop_round(nu2, t)
    assert (gg, st) == (11, 6)
    q = rand_residues(g, 2).reshape(2, 2, N)
    left, right = _rand_keys(g, gg, t), _rand_keys(g, st + 1, t)
    v = np.zeros((1 << gg, 2, 2, N), np.uint32)
    v[0] = q
    d = dev(v)
    gpu_ctx.set_variant(0, variant)
    try:
        ops.coefficient_expansion(gpu_ctx, d, dev(left), dev(right), nu1, nu2, t, t)
    finally:
        gpu_ctx.set_variant(0, 0)
    got = host(d).reshape(1 << gg, 2, 2, N)
    exp = oracle.expand(q, nu1, nu2, t, t, left, right)
    bad = [i for i in range(1 << gg) if not np.array_equal(got[i], exp[i])]
    assert not bad, f"elements differ: {bad[:10]} (of {len(bad)})"


# ---- 1024 rows through every multiply kernel ----------------------------------------

@pytest.fixture
def force_dbmul(gpu_ctx):
    yield lambda v: gpu_ctx.set_variant(2, v)
    gpu_ctx.set_variant(2, 0)


def _batch(db, entries):
    from spiral_pir.server import Batch

    b = Batch(db, len(entries))
    for slot, (cl, idx, q, left, right, vf) in enumerate(entries):
        b.load(slot, q, left, right, vf)
    return b


def _dbmul_check(oracle, b, s, db_ntt, slots, what):
    from spiral_pir.server import READ_DB_MUL, READ_EXPANDED

    for slot in slots:
        v = b.read(slot, READ_EXPANDED)
        sel = np.ascontiguousarray(v[0:2 * s.rows:2])
        acc = b.read(slot, READ_DB_MUL)
        for t in range(s.n2):
            np.testing.assert_array_equal(acc[t], oracle.db_mul(db_ntt[t], sel, s.nu1, s.nu2, nthreads=0),
                                          err_msg=f"{what}: slot {slot} trial {t}")


def test_dbmul_nu1_10_mfma_ring_and_valu(gpu_ctx, oracle, dbs, force_dbmul):
    """(10, 5), 40 queries = 80 selector rows: the LDS-staged MFMA kernel with 32 row
    blocks, then 2 queries through the streaming MFMA kernel (I <= 64), and the
    VALU kernel; accumulators of the first and last slot against the oracle."""
    from spiral_pir.server import STAGE_DB_MUL, STAGE_EXPAND

    s, recs, db, db_ntt = dbs.get(10, 5)
    entries = _clients_queries(s, 40, 4, 4100)
    b = _batch(db, entries)
    b.run(40, STAGE_EXPAND)
    for var, name in ((2, "mfma persistent, 32 row blocks"), (1, "valu")):
        force_dbmul(var)
        b.run(40, STAGE_DB_MUL)
        assert b.last_run()["dbmul"] == ("mfma" if var == 2 else "valu")
        _dbmul_check(oracle, b, s, db_ntt, (0, 39), name)
    b.close()
    b = _batch(db, entries[:2])
    b.run(2, STAGE_EXPAND)
    force_dbmul(2)
    b.run(2, STAGE_DB_MUL)
    _dbmul_check(oracle, b, s, db_ntt, (0, 1), "mfma streaming")
    b.close()
    dbs.drop(10, 5)


def test_dbmul_nu1_10_extreme_residues(gpu_ctx, oracle, force_dbmul):
    """A (10, 5) database of all-(q-1) residues (every limb 127 over 1024 rows: the
    int32 limb sums reach 4 x 1024 x 127^2) through the LDS-staged and streaming MFMA
    kernels and the VALU kernels (one and four queries per read), against the oracle."""
    from spiral_pir import Shape
    from spiral_pir.server import Database, STAGE_DB_MUL, STAGE_EXPAND

    s = Shape(10, 5)
    db_ntt = np.empty((s.n2, s.cols, s.rows, 2, N), np.uint32)
    db_ntt[:, :, :, 0] = Q0 - 1
    db_ntt[:, :, :, 1] = Q1 - 1
    db = Database(gpu_ctx, s)
    db.load_ntt(db_ntt)
    for B, variants in ((34, (2, 1, 4)), (2, (2, 3))):
        entries = _clients_queries(s, B, 2, 4200 + B)
        b = _batch(db, entries)
        b.run(B, STAGE_EXPAND)
        for var in variants:
            force_dbmul(var)
            b.run(B, STAGE_DB_MUL)
            _dbmul_check(oracle, b, s, db_ntt, (B - 1,), f"B={B} variant {var}")
        b.close()
    db.close()


@pytest.mark.parametrize("nu1", [5, 6, 7, 8])
def test_dbmul_mfma_ring_row_blocks(gpu_ctx, oracle, force_dbmul, nu1):
    """33 queries (66 selector rows: the LDS-staged kernels) at 2^nu1 = 32 .. 256
    rows, i.e. 1, 2, 4 and 8 row blocks (one tile per block for 1 and 2, the
    persistent kernel for 4 and 8): the MFMA accumulators equal the VALU kernel's
    for every slot read and the oracle's for one."""
    from spiral_pir import Shape
    from spiral_pir.server import Database, READ_DB_MUL, STAGE_DB_MUL, STAGE_EXPAND

    s = Shape(nu1, 5)
    recs = rng(4300 + nu1).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
    db = Database(gpu_ctx, s)
    db.load_bytes(recs)
    B = 33
    entries = _clients_queries(s, B, 3, 4400 + nu1)
    b = _batch(db, entries)
    b.run(B, STAGE_EXPAND)
    slots = (0, 16, 32)
    force_dbmul(1)
    b.run(B, STAGE_DB_MUL)
    ref = {k: b.read(k, READ_DB_MUL) for k in slots}
    force_dbmul(2)
    b.run(B, STAGE_DB_MUL)
    assert b.last_run()["dbmul"] == "mfma"
    for k in slots:
        np.testing.assert_array_equal(b.read(k, READ_DB_MUL), ref[k], err_msg=f"slot {k}")
    _dbmul_check(oracle, b, s, oracle.db_encode(recs, s.n2, nu1, 5), (32,), "ring")
    b.close()
    db.close()


@pytest.mark.parametrize("shape,B", [((9, 6), 5), ((10, 7), 3), ((3, 3), 6)])
def test_dbmul_multi_query_valu(gpu_ctx, oracle, dbs, force_dbmul, shape, B):
    """The VALU multiply that reads the database once for 2 (variant 3) or 4
    (variant 4) queries, with the last query group partial: accumulators equal
    the one-query kernel's for every slot and the oracle's for the first and
    last slot; at (3, 3) one packed group of 8 columns and 8 rows (one per wave)."""
    from spiral_pir import Shape
    from spiral_pir.server import Database, READ_DB_MUL, STAGE_DB_MUL, STAGE_EXPAND

    if shape == (3, 3):
        s = Shape(3, 3)
        recs = rng(4800).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
        db = Database(gpu_ctx, s)
        db.load_bytes(recs)
        db_ntt = oracle.db_encode(recs, s.n2, 3, 3)
    else:
        s, recs, db, db_ntt = dbs.get(*shape)
    entries = _clients_queries(s, B, 3, 4900 + B)
    b = _batch(db, entries)
    b.run(B, STAGE_EXPAND)
    force_dbmul(1)
    b.run(B, STAGE_DB_MUL)
    ref = [b.read(k, READ_DB_MUL) for k in range(B)]
    for var, name in ((3, "valu_q2"), (4, "valu_q4")):
        force_dbmul(var)
        b.run(B, STAGE_DB_MUL)
        assert b.last_run()["dbmul"] == name
        for k in range(B):
            np.testing.assert_array_equal(b.read(k, READ_DB_MUL), ref[k], err_msg=f"{name} slot {k}")
    _dbmul_check(oracle, b, s, db_ntt, (0, B - 1), "valu_q4")
    b.close()
    if shape == (3, 3):
        db.close()


# ---- (8, 4): the VALU batched path ------------------------------------------------

def test_batch_nu8_4_valu_16_queries(gpu_ctx, oracle, dbs):
    """(8, 4), 16 queries by graph replay: 2^nu2 = 16 leaves the VALU multiply as the
    only batched form (here the one reading the database once per 4 queries);
    every slot equals the oracle's answer and decodes."""
    s, recs, db, db_ntt = dbs.get(8, 4)
    entries = _clients_queries(s, 16, 4, 4500)
    b = _batch(db, entries)
    b.graph_launch(16)
    b.synchronize()
    assert b.last_run()["dbmul"] == "valu_q4"
    got = b.responses(16)
    for slot, (cl, idx, q, left, right, vf) in enumerate(entries):
        np.testing.assert_array_equal(got[slot], _oracle_answer(oracle, s, db_ntt, q, left, right, vf),
                                      err_msg=f"slot {slot}")
        np.testing.assert_array_equal(cl.decode(got[slot]), recs[idx])
    b.close()


# ---- (9, 6), 64 queries ----------------------------------------------------------

def test_batch64_nu9_6_autotuned_graph_vs_oracle(gpu_ctx, oracle, dbs):
    """The benchmark's batched configuration: (9, 6), 64 queries from 8 clients, the
    autotuned multiply, graph replays.  Four slots equal the oracle's answer, every
    slot decodes, the bulk readout equals the per-slot readout, and the VALU-forced
    eager run gives the same responses."""
    import torch
    from spiral_pir.server import READ_RESPONSE, STAGE_EXPAND

    s, recs, db, db_ntt = dbs.get(9, 6)
    B = 64
    entries = _clients_queries(s, B, 8, 3900)
    b = _batch(db, entries)
    b.run(B, STAGE_EXPAND)
    info = b.tune(B)
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(2):
        (torch.randn(1 << 18, device="cuda") * 2).sum().item()  # unrelated eager work between replays
        b.graph_launch(B, stream)
        b.synchronize()
    got = b.responses(B)
    for slot in (0, 17, 40, 63):
        cl, idx, q, left, right, vf = entries[slot]
        np.testing.assert_array_equal(got[slot], _oracle_answer(oracle, s, db_ntt, q, left, right, vf),
                                      err_msg=f"slot {slot} ({info})")
        np.testing.assert_array_equal(b.read(slot, READ_RESPONSE), got[slot])
    for slot, (cl, idx, q, left, right, vf) in enumerate(entries):
        np.testing.assert_array_equal(cl.decode(got[slot]), recs[idx], err_msg=f"slot {slot}")
    gpu_ctx.set_variant(2, 1)
    try:
        b.run(B)
        np.testing.assert_array_equal(b.responses(B), got)
    finally:
        gpu_ctx.set_variant(2, 0)
    b.close()


def test_graph_cache_follows_variant_changes(gpu_ctx, dbs):
    """A graph captured with one multiply kernel is not replayed after the variant
    changes: the next launch captures a new chain with the new kernel, and the
    responses stay equal."""
    import torch

    s, recs, db, db_ntt = dbs.get(9, 6)
    B = 40
    entries = _clients_queries(s, B, 4, 4600)
    b = _batch(db, entries)
    stream = torch.cuda.current_stream().cuda_stream
    gpu_ctx.set_variant(2, 1)
    try:
        b.graph_launch(B, stream)
        b.synchronize()
        first = b.last_run()
        ref = b.responses(B)
        gpu_ctx.set_variant(2, 2)
        b.graph_launch(B, stream)
        b.synchronize()
        second = b.last_run()
        np.testing.assert_array_equal(b.responses(B), ref)
        gpu_ctx.set_variant(2, 1)
        b.graph_launch(B, stream)
        b.synchronize()
        third = b.last_run()
    finally:
        gpu_ctx.set_variant(2, 0)
    assert first == {"dbmul": "valu", "graphs": 1}
    assert second == {"dbmul": "mfma", "graphs": 2}
    assert third == {"dbmul": "valu", "graphs": 2}  # the first graph again
    np.testing.assert_array_equal(b.responses(B), ref)
    b.close()


# ---- 48/32/16 over three databases --------------------------------------------

def test_group_48_32_16_three_shapes(gpu_ctx, oracle, dbs):
    """The exact uneven launch: 48 queries on (9, 6), 32 on (10, 7), 16 on (8, 4),
    autotuned, one graph with one branch per database, replayed twice with eager
    work in between.  Every slot equals its query answered alone; the first and
    last slot of every shape equal the oracle's answer; every slot decodes."""
    import torch
    from spiral_pir.server import Group, STAGE_EXPAND

    plan = (((9, 6), 48, 3100), ((10, 7), 32, 3200), ((8, 4), 16, 3300))
    lanes = []
    for shape, n, seed in plan:
        s, recs, db, db_ntt = dbs.get(*shape)
        entries = _clients_queries(s, n, 4, seed)
        b = _batch(db, entries)
        b.run(n, STAGE_EXPAND)
        b.tune(n)
        lanes.append((s, recs, db, db_ntt, b, entries, _alone(db, entries)))
    g = Group([ln[4] for ln in lanes])
    ns = [len(ln[5]) for ln in lanes]
    stream = torch.cuda.current_stream().cuda_stream
    for rep in range(2):
        (torch.randn(1 << 18, device="cuda") * 2).sum().item()
        g.launch(ns, stream)
        g.synchronize()
        for s, recs, db, db_ntt, b, entries, alone in lanes:
            got = b.responses(len(entries))
            for slot, ref in enumerate(alone):
                np.testing.assert_array_equal(got[slot], ref, err_msg=f"rep {rep} ({s.nu1},{s.nu2}) slot {slot}")
    for s, recs, db, db_ntt, b, entries, alone in lanes:
        for slot in (0, len(entries) - 1):
            cl, idx, q, left, right, vf = entries[slot]
            np.testing.assert_array_equal(alone[slot], _oracle_answer(oracle, s, db_ntt, q, left, right, vf),
                                          err_msg=f"({s.nu1},{s.nu2}) slot {slot}")
        for slot, (cl, idx, q, left, right, vf) in enumerate(entries):
            np.testing.assert_array_equal(cl.decode(alone[slot]), recs[idx])
    g.close()
    for ln in lanes:
        ln[4].close()


# ---- Poisson serving over (9, 6) / (10, 7) --------------------------------------

def test_serving_poisson_nu9_6_and_nu10_7(gpu_ctx, oracle, dbs):
    """About 1 s of Poisson arrivals at 1000 queries/s, 70/30 over the resident
    (9, 6) and (10, 7) databases, continuous batching with asynchronous uploads
    and readouts: every served response equals its query answered alone, and the
    first query of each pool equals the oracle's answer."""
    from spiral_pir.serving import Server, drive, latency_stats, pinned_u32, poisson_times

    pools = {}
    for name in ((9, 6), (10, 7)):
        s, recs, db, db_ntt = dbs.get(*name)
        entries = _clients_queries(s, 6, 3, 4700 + name[0])
        alone = _alone(db, entries)
        cl, idx, q, left, right, vf = entries[0]
        np.testing.assert_array_equal(alone[0], _oracle_answer(oracle, s, db_ntt, q, left, right, vf))
        pools[name] = (db, entries, alone)
    srv = Server({k: v[0] for k, v in pools.items()}, max_batch=64, buffers=2)
    pinned = {}
    for name, (db, entries, alone) in pools.items():
        for j, (cl, idx, q, left, right, vf) in enumerate(entries):
            cid = (name, j % 3)
            if cid not in srv.lanes[name].keys:
                srv.register(name, cid, left, right)
            pq, pv = pinned_u32(q.size), pinned_u32(vf.size)
            pq[:] = q.reshape(-1)
            pv[:] = vf.reshape(-1)
            pinned[(name, j)] = (cid, pq, pv)
        cid, pq, pv = pinned[(name, 0)]
        srv.warmup(name, cid, pq, pv)
    g = rng(3000)
    arrivals = []
    for t in poisson_times(1000.0, 1.0, g):
        name = (9, 6) if g.random() < 0.7 else (10, 7)
        j = int(g.integers(0, len(pools[name][1])))
        cid, pq, pv = pinned[(name, j)]
        arrivals.append((float(t), name, cid, pq, pv, (name, j)))
    srv.start()
    reqs = drive(srv, arrivals)
    for r in reqs:
        r.wait(timeout=60)
    srv.stop()
    for r in reqs:
        name, j = r.tag
        np.testing.assert_array_equal(r.response, pools[name][2][j], err_msg=str(r.tag))
    st = latency_stats(reqs)
    assert st["completed"] == len(arrivals) > 500
    assert sum(ln.launched_batches for ln in srv.lanes.values()) < len(arrivals)
    print("serving (1000 q/s offered, 1 s):", st)
    srv.close()
