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


@pytest.mark.parametrize("nu1,nu2", [(6, 1), (5, 2), (4, 5)])
def test_fold_latency_form_matches_oracle_and_split(gpu_ctx, oracle, nu1, nu2):
    """Runs of at most two queries fold in the latency form (one launch per level
    past the first, finish of each level inside the next, alternating buffers);
    with nu2 = 1, 2, 5 the chain ends in place or out of place and uses both
    launch shapes.  The responses equal the oracle's and those of the split
    form (fold variant 4), in one- and two-query runs."""
    from spiral_pir.server import Batch

    s, recs, db = _setup(gpu_ctx, nu1, nu2, 950 + nu2)
    qs = _queries(s, 2, 3500 + 10 * nu2)
    b = Batch(db, 3)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    db_ntt = oracle.db_encode(recs, s.n2, s.nu1, s.nu2)
    exp = [oracle.answer(db_ntt, s.n2, s.nu1, s.nu2, s.t_gsw, s.t_exp, q, left, right, vf,
                         oracle.folding_neg(vf, s.nu2, s.t_gsw)) for (c, idx, q, left, right, vf) in qs]
    for n in (1, 2):
        b.run(n)
        fused = b.responses(n)
        gpu_ctx.set_variant(1, 4)
        try:
            b.run(n)
            split = b.responses(n)
        finally:
            gpu_ctx.set_variant(1, 0)
        for slot in range(n):
            np.testing.assert_array_equal(fused[slot], exp[slot], err_msg=f"n={n} slot {slot}")
            np.testing.assert_array_equal(split[slot], exp[slot], err_msg=f"n={n} slot {slot} split")
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        np.testing.assert_array_equal(c.decode(fused[slot]), recs[idx])
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


@pytest.fixture
def force_dbmul(gpu_ctx):
    def setv(v):
        gpu_ctx.set_variant(2, v)
    yield setv
    gpu_ctx.set_variant(2, 0)


@pytest.mark.parametrize("nu1,nu2,B", [(6, 5, 5), (5, 6, 16), (7, 5, 1)])
def test_dbmul_mfma_matches_oracle(gpu_ctx, oracle, force_dbmul, nu1, nu2, B):
    """The int8-limb MFMA database multiply gives the same accumulators as the oracle
    (and so the same responses), including batches that do not fill the 32-row tiles."""
    from spiral_pir.server import Batch, READ_DB_MUL, READ_EXPANDED, STAGE_DB_MUL, STAGE_EXPAND, STAGE_FOLD

    s, recs, db = _setup(gpu_ctx, nu1, nu2, 910 + nu1)
    qs = _queries(s, B, 3500 + nu2)
    b = Batch(db, B)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    force_dbmul(2)
    b.run(B, STAGE_EXPAND)
    b.run(B, STAGE_DB_MUL)
    db_ntt = oracle.db_encode(recs, s.n2, nu1, nu2)
    for slot in sorted({0, B - 1}):
        v = b.read(slot, READ_EXPANDED)
        sel = np.ascontiguousarray(v[0:2 * s.rows:2])
        acc = b.read(slot, READ_DB_MUL)
        for t in range(s.n2):
            np.testing.assert_array_equal(acc[t], oracle.db_mul(db_ntt[t], sel, nu1, nu2, nthreads=0),
                                          err_msg=f"slot {slot} trial {t}")
    b.run(B, STAGE_FOLD)
    resp = b.responses(B)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        np.testing.assert_array_equal(c.decode(resp[slot]), recs[idx])
    b.close()
    db.close()


@pytest.mark.parametrize("B", [2, 40])
def test_dbmul_mfma_extreme_residues(gpu_ctx, oracle, force_dbmul, B):
    """A database of all-(q-1) residues (every limb at its maximum, 512 rows) through the
    MFMA multiply (B = 2: the streaming kernel; B = 40, 80 selector rows: the
    LDS-staged kernel with a partial row tile): the int32 limb sums stay exact, against
    the oracle."""
    from spiral_pir import Shape
    from spiral_pir.server import Batch, Database, READ_DB_MUL, READ_EXPANDED, STAGE_DB_MUL, STAGE_EXPAND

    s = Shape(9, 5)
    R, C = s.rows, s.cols
    Q0, Q1 = 268369921, 249561089
    db_ntt = np.empty((s.n2, C, R, 2, N), np.uint32)
    db_ntt[:, :, :, 0] = Q0 - 1
    db_ntt[:, :, :, 1] = Q1 - 1
    db = Database(gpu_ctx, s)
    db.load_ntt(db_ntt)
    qs = _clients_queries(s, B, 2, 3600)
    b = Batch(db, B)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    force_dbmul(2)
    b.run(B, STAGE_EXPAND)
    b.run(B, STAGE_DB_MUL)
    v = b.read(B - 1, READ_EXPANDED)
    sel = np.ascontiguousarray(v[0:2 * R:2])
    acc = b.read(B - 1, READ_DB_MUL)
    np.testing.assert_array_equal(acc[0], oracle.db_mul(db_ntt[0], sel, s.nu1, s.nu2, nthreads=0))
    b.close()
    db.close()


def test_autotuner_picks_and_keeps_results(gpu_ctx):
    from spiral_pir.server import Batch, STAGE_EXPAND

    s, recs, db = _setup(gpu_ctx, 6, 5, 920)
    qs = _queries(s, 8, 3700)
    b = Batch(db, 8)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    b.run(8)
    ref = b.responses(8)
    b.run(8, STAGE_EXPAND)
    info = b.tune(8)
    assert info["chosen"] in ("valu", "mfma", "valu_q2", "valu_q4")
    assert info["valu_ms"] > 0 and info["mfma_ms"] > 0 and info["valu_q2_ms"] > 0 and info["valu_q4_ms"] > 0
    b.run(8)
    np.testing.assert_array_equal(b.responses(8), ref)
    b.graph_launch(8)
    np.testing.assert_array_equal(b.responses(8), ref)
    b.close()
    db.close()


def _clients_queries(s, B, n_clients, seed):
    from spiral_pir.client import Client

    clients = [Client(s, seed + k) for k in range(n_clients)]
    keys = [c.keygen() for c in clients]
    idx = [int(x) for x in rng(seed + 100).integers(0, s.records, size=B)]
    qs = []
    for slot in range(B):
        c = slot % n_clients
        q, vf = clients[c].query(idx[slot])
        qs.append((clients[c], idx[slot], q, keys[c][0], keys[c][1], vf))
    return qs


def test_dbmul_mfma_partial_row_tile(gpu_ctx, force_dbmul):
    """48 queries = 96 selector rows: the last 32-row wave tile of the 128-row block is
    empty and returns early; accumulators equal the VALU multiply's for every read slot."""
    from spiral_pir.server import Batch, READ_DB_MUL, STAGE_DB_MUL, STAGE_EXPAND

    s, recs, db = _setup(gpu_ctx, 9, 5, 930)
    B = 48
    qs = _clients_queries(s, B, 4, 3800)
    b = Batch(db, B)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    b.run(B, STAGE_EXPAND)
    force_dbmul(1)
    b.run(B, STAGE_DB_MUL)
    slots = (0, 31, 47)
    ref = {k: b.read(k, READ_DB_MUL) for k in slots}
    force_dbmul(2)
    b.run(B, STAGE_DB_MUL)
    for k in slots:
        np.testing.assert_array_equal(b.read(k, READ_DB_MUL), ref[k], err_msg=f"slot {k}")
    b.close()
    db.close()


@pytest.mark.parametrize("nu1,nu2,B", [(9, 5, 48), (6, 5, 5)])
def test_mfma_products_to_fold_directly(gpu_ctx, force_dbmul, nu1, nu2, B):
    """After the MFMA multiply the fold stage takes its products out of the NTT domain
    straight from the multiply's p-major output (no readout in between, so nothing
    transposes them first): the raw accumulators and the responses equal those of the
    VALU multiply (B = 48: LDS-staged kernel, partial row tile; B = 5: streaming kernel)."""
    from spiral_pir.server import Batch, READ_RAW_ACCS, STAGE_DB_MUL, STAGE_EXPAND, STAGE_FOLD

    s, recs, db = _setup(gpu_ctx, nu1, nu2, 940 + B)
    qs = _clients_queries(s, B, 3, 3900 + B)
    b = Batch(db, B)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    slots = sorted({0, B // 2, B - 1})
    out = {}
    for var in (1, 2):
        force_dbmul(var)
        b.run(B, STAGE_EXPAND)
        b.run(B, STAGE_DB_MUL)
        assert b.last_run()["dbmul"] == ("mfma" if var == 2 else "valu")
        b.run(B, STAGE_FOLD)
        out[var] = (b.responses(B), {k: b.read(k, READ_RAW_ACCS) for k in slots})
    np.testing.assert_array_equal(out[2][0], out[1][0])
    for k in slots:
        np.testing.assert_array_equal(out[2][1][k], out[1][1][k], err_msg=f"slot {k}")
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        np.testing.assert_array_equal(c.decode(out[2][0][slot]), recs[idx])
    b.close()
    db.close()


def test_expansion_pair_digits_below_and_at_stop_round(gpu_ctx, oracle):
    """80 queries at (5, 5), t_GSW = 8 (g = 7, stop_round 6): rounds 5 and 6 have more
    one-step digit blocks than 16 per CU (64 x 80 and 104 x 80 against 16 x 256), so
    their digits come from the parent-and-child kernel, below stop_round (every
    element works) and at it (even parents, then the odd ones that still work), the
    children's coefficients by rotation of the parents'; the earlier rounds take one
    step per block.  The expansion of the first, a middle and the last slot equals
    the oracle's, element by element, and every record decodes."""
    from spiral_pir.server import Batch, READ_EXPANDED, STAGE_DB_MUL, STAGE_EXPAND, STAGE_FOLD

    s, recs, db = _setup(gpu_ctx, 5, 5, 970)
    B = 80
    qs = _clients_queries(s, B, 5, 4700)
    b = Batch(db, B)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    b.run(B, STAGE_EXPAND)
    for slot in (0, 41, B - 1):
        c, idx, q, left, right, vf = qs[slot]
        np.testing.assert_array_equal(b.read(slot, READ_EXPANDED),
                                      oracle.expand(q, s.nu1, s.nu2, s.t_gsw, s.t_exp, left, right),
                                      err_msg=f"slot {slot}")
    b.run(B, STAGE_DB_MUL)
    b.run(B, STAGE_FOLD)
    resp = b.responses(B)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        np.testing.assert_array_equal(c.decode(resp[slot]), recs[idx], err_msg=f"slot {slot}")
    b.close()
    db.close()


def test_dbmul_mfma_two_row_tiles_against_oracle(gpu_ctx, oracle, force_dbmul):
    """100 queries at (7, 5) = 200 selector rows (I = 224): the persistent MFMA kernel's
    row tiles of 128 are a whole one and a partial one (rows 128-199 hold queries,
    200-223 zero selectors, 224-255 are masked), 4 row blocks of 32; its
    accumulators equal the VALU multiply's for slots in both
    tiles and the oracle's for one slot of the second; the expansion of rounds 5-7
    takes the parent-and-child digits kernel; every record decodes."""
    from spiral_pir.server import Batch, READ_DB_MUL, READ_EXPANDED, STAGE_DB_MUL, STAGE_EXPAND, STAGE_FOLD

    s, recs, db = _setup(gpu_ctx, 7, 5, 980)
    B = 100
    qs = _clients_queries(s, B, 5, 4800)
    b = Batch(db, B)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        b.load(slot, q, left, right, vf)
    b.run(B, STAGE_EXPAND)
    force_dbmul(1)
    b.run(B, STAGE_DB_MUL)
    slots = (0, 63, 64, 99)
    ref = {k: b.read(k, READ_DB_MUL) for k in slots}
    force_dbmul(2)
    b.run(B, STAGE_DB_MUL)
    assert b.last_run()["dbmul"] == "mfma"
    for k in slots:
        np.testing.assert_array_equal(b.read(k, READ_DB_MUL), ref[k], err_msg=f"slot {k}")
    db_ntt = oracle.db_encode(recs, s.n2, s.nu1, s.nu2)
    v = b.read(99, READ_EXPANDED)
    c, idx, q, left, right, vf = qs[99]
    np.testing.assert_array_equal(v, oracle.expand(q, s.nu1, s.nu2, s.t_gsw, s.t_exp, left, right))
    sel = np.ascontiguousarray(v[0:2 * s.rows:2])
    for t in range(s.n2):
        np.testing.assert_array_equal(ref[99][t], oracle.db_mul(db_ntt[t], sel, s.nu1, s.nu2, nthreads=0),
                                      err_msg=f"trial {t}")
    b.run(B, STAGE_FOLD)
    resp = b.responses(B)
    for slot, (c, idx, q, left, right, vf) in enumerate(qs):
        np.testing.assert_array_equal(c.decode(resp[slot]), recs[idx], err_msg=f"slot {slot}")
    b.close()
    db.close()


@pytest.mark.parametrize("nu1,nu2,B", [(6, 2, 2), (6, 2, 20), (7, 5, 40)])
def test_full_answer_forms_only_live_elements(gpu_ctx, oracle, nu1, nu2, B):
    """A whole answer (all stages in one run) of queries with fold vectors forms
    every even expanded element (the selectors and their ancestors) exactly as the
    oracle does, leaves the odd i >= 2^(stop_round + 1) (v_neg1 products of elements
    skipped past stop_round) as they were, and does not run the odd steps (only the
    conversion of packed queries reads odd elements): after an expansion-only run
    of other queries, a full run of new ones holds the oracle's expansion of the new
    queries on the even elements and the earlier values on those odd ones, and every
    response decodes.  B = 2 takes the latency forms for the leading rounds, B = 20
    the split form throughout, B = 40 at (7, 5) the parent-and-child digits kernel
    in the last round."""
    from spiral_pir.server import Batch, READ_EXPANDED, STAGE_EXPAND, STAGES_ALL

    s, recs, db = _setup(gpu_ctx, nu1, nu2, 1500 + nu1 + B)
    old = _clients_queries(s, B, 3, 5100 + B)
    new = _clients_queries(s, B, 3, 5300 + B)
    b = Batch(db, B)
    for slot, (c, idx, q, left, right, vf) in enumerate(old):
        b.load(slot, q, left, right, vf)
    b.run(B, STAGE_EXPAND)
    slots = sorted({0, B - 1})
    prev = {slot: b.read(slot, READ_EXPANDED) for slot in slots}
    for slot, (c, idx, q, left, right, vf) in enumerate(new):
        b.load(slot, q, left, right, vf)
    b.run(B, STAGES_ALL)
    i = np.arange(1 << s.g)
    even = i % 2 == 0
    untouched = ~even & (i >= (2 << oracle.expansion_stop_round(nu2, s.t_gsw)))
    assert untouched.any()
    for slot in slots:
        c, idx, q, left, right, vf = new[slot]
        v = b.read(slot, READ_EXPANDED)
        exp = oracle.expand(q, nu1, nu2, s.t_gsw, s.t_exp, left, right)
        np.testing.assert_array_equal(v[even], exp[even], err_msg=f"slot {slot} even elements")
        np.testing.assert_array_equal(v[untouched], prev[slot][untouched], err_msg=f"slot {slot} elements left as they were")
    resp = b.responses(B)
    for slot, (c, idx, q, left, right, vf) in enumerate(new):
        np.testing.assert_array_equal(c.decode(resp[slot]), recs[idx], err_msg=f"slot {slot}")
    b.close()
    db.close()
