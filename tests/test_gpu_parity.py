"""GPU parity: every HIP operator against the CPU oracle, bit for bit.

Inputs are seeded; expected values come from the oracle (oracle/), which is
itself pinned to the reference's fixtures by tests/test_oracle_kat.py.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rand_raw, rand_residues, rng

pytestmark = pytest.mark.gpu

N = 2048
Q0, Q1 = 268369921, 249561089
QQ = Q0 * Q1


def dev(a: np.ndarray):
    import torch

    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    elif a.dtype == np.uint64:
        a = a.view(np.int64)
    return torch.from_numpy(a.copy()).cuda()


def host(t, dtype=np.uint32) -> np.ndarray:
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy().view(dtype)


def special_polys(m: int) -> np.ndarray:
    q = (Q0, Q1)[m]
    out = np.zeros((4, N), np.uint32)
    out[1, :] = q - 1
    out[2, 0] = 1
    out[3, ::2] = q - 1
    return out


@pytest.mark.parametrize("m", [0, 1])
def test_ntt_fwd_matches_oracle(gpu_ctx, oracle, m):
    from spiral_pir import ops

    g = rng(100 + m)
    q = (Q0, Q1)[m]
    polys = np.concatenate([g.integers(0, q, size=(61, N), dtype=np.uint32), special_polys(m)])
    t = dev(polys)
    ops.NTT_R(gpu_ctx, t, m)
    got = host(t).reshape(-1, N)
    for i in range(polys.shape[0]):
        np.testing.assert_array_equal(got[i], oracle.ntt_fwd(polys[i], m), err_msg=f"poly {i}")


@pytest.mark.parametrize("m", [0, 1])
def test_ntt_inv_matches_oracle(gpu_ctx, oracle, m):
    from spiral_pir import ops

    g = rng(200 + m)
    q = (Q0, Q1)[m]
    polys = np.concatenate([g.integers(0, q, size=(29, N), dtype=np.uint32), special_polys(m)])
    t = dev(polys)
    ops.INTT_R(gpu_ctx, t, m, scale=True)
    got = host(t).reshape(-1, N)
    t2 = dev(polys)
    ops.INTT_R(gpu_ctx, t2, m, scale=False)  # reference convention: no N^-1
    got_unscaled = host(t2).reshape(-1, N)
    for i in range(polys.shape[0]):
        exp = oracle.ntt_inv(polys[i], m)
        np.testing.assert_array_equal(got[i], exp, err_msg=f"poly {i}")
        np.testing.assert_array_equal(got_unscaled[i], (exp.astype(np.uint64) * N % q).astype(np.uint32))


def test_ntt_round_trip_many(gpu_ctx):
    """gpu-pir/ntt_test.cu:84-99 at scale: INTT(NTT(a)) * N^-1 = a over 4096 polys."""
    from spiral_pir import ops

    g = rng(7)
    for m, q in enumerate((Q0, Q1)):
        a = g.integers(0, q, size=(4096, N), dtype=np.uint32)
        t = dev(a)
        ops.NTT_R(gpu_ctx, t, m)
        ops.INTT_R(gpu_ctx, t, m, scale=True)
        np.testing.assert_array_equal(host(t).reshape(-1, N), a)


def test_from_to_ntt(gpu_ctx, oracle):
    from spiral_pir import ops
    import torch

    g = rng(300)
    ntt_in = rand_residues(g, 16)
    raw_out = torch.zeros(16 * 2 * N, dtype=torch.int32, device="cuda")
    ops.from_ntt(gpu_ctx, dev(ntt_in), raw_out)
    got = host(raw_out).reshape(16, 2, N)
    for i in range(16):
        np.testing.assert_array_equal(got[i], oracle.from_ntt(ntt_in[i]))
    raw_in = rand_raw(g, 16, include_q=True)
    ntt_out = torch.zeros(16 * 2 * N, dtype=torch.int32, device="cuda")
    ops.to_ntt(gpu_ctx, dev(raw_in), ntt_out)
    got = host(ntt_out).reshape(16, 2, N)
    for i in range(16):
        np.testing.assert_array_equal(got[i], oracle.to_ntt(raw_in[i]))


@pytest.mark.parametrize("t", [3, 2049, 1025, 513, 65, 5, 4095])
def test_automorph(gpu_ctx, oracle, t):
    from spiral_pir import ops
    import torch

    g = rng(400 + t)
    raw = rand_raw(g, 4)
    raw[:, :, 5] = 0  # a zero that gets negated must become Q
    x = oracle.raw_to_u64(raw)
    out = torch.zeros(4 * N, dtype=torch.int64, device="cuda")
    ops.automorph(gpu_ctx, dev(x), out, t)
    got = host(out, np.uint64).reshape(4, N)
    for i in range(4):
        np.testing.assert_array_equal(got[i], oracle.raw_to_u64(oracle.automorph(raw[i], t))[0])


@pytest.mark.parametrize("rdim,mx", [(1, 8), (2, 16), (1, 4), (2, 8)])
def test_gadget_extprod(gpu_ctx, oracle, rdim, mx):
    from spiral_pir import ops

    g = rng(500 + mx * rdim)
    items = 3
    inp = rand_raw(g, items * rdim, include_q=True).reshape(items, rdim, 2, N)
    keys = np.empty((items, 2, mx, 2, N), np.uint32)
    keys[:, 0] = g.integers(0, Q0, size=(items, mx, 2, N), dtype=np.uint32)
    keys[:, 1] = g.integers(0, Q1, size=(items, mx, 2, N), dtype=np.uint32)
    acc0 = rand_residues(g, items * 2).reshape(items, 2, 2, N)
    acc = dev(acc0)
    ops.inv_gadget_ntt_mult(gpu_ctx, dev(inp), dev(keys), acc, rdim, mx, n_items=items,
                            in_stride=rdim * 2 * N, key_stride=2 * mx * 2 * N, acc_stride=4 * N)
    got = host(acc).reshape(items, 2, 2, N)
    for i in range(items):
        exp = oracle.extprod_acc(inp[i], rdim, mx, keys[i], acc0[i])
        np.testing.assert_array_equal(got[i], exp, err_msg=f"item {i}")


@pytest.mark.parametrize("nu1,nu2", [(6, 4), (2, 1), (9, 3)])
def test_db_mul_reference_layout(gpu_ctx, oracle, nu1, nu2):
    from spiral_pir import ops
    import torch

    g = rng(600 + nu1 * 16 + nu2)
    R, C = 1 << nu1, 1 << nu2
    db = rand_residues(g, R * C).reshape(C, R, 2, N)
    sel = rand_residues(g, R * 2).reshape(R, 2, 2, N)
    out = torch.zeros(C * 4 * N, dtype=torch.int32, device="cuda")
    ops.db_mul(gpu_ctx, dev(db), dev(sel), out, nu1, nu2)
    got = host(out).reshape(C, 2, 2, N)
    np.testing.assert_array_equal(got, oracle.db_mul(db, sel, nu1, nu2))


def test_db_mul_extreme_values(gpu_ctx, oracle):
    """All operands q-1: the lazy u64 accumulation at its largest (512 rows)."""
    from spiral_pir import ops
    import torch

    nu1, nu2 = 9, 1
    R, C = 1 << nu1, 1 << nu2
    db = np.empty((C, R, 2, N), np.uint32)
    db[:, :, 0] = Q0 - 1
    db[:, :, 1] = Q1 - 1
    sel = np.empty((R, 2, 2, N), np.uint32)
    sel[:, :, 0] = Q0 - 1
    sel[:, :, 1] = Q1 - 1
    out = torch.zeros(C * 4 * N, dtype=torch.int32, device="cuda")
    ops.db_mul(gpu_ctx, dev(db), dev(sel), out, nu1, nu2)
    np.testing.assert_array_equal(host(out).reshape(C, 2, 2, N), oracle.db_mul(db, sel, nu1, nu2))


def _rand_keys(g, levels, mx):
    k = np.empty((levels, 2, mx, 2, N), np.uint32)
    k[:, 0] = g.integers(0, Q0, size=(levels, mx, 2, N), dtype=np.uint32)
    k[:, 1] = g.integers(0, Q1, size=(levels, mx, 2, N), dtype=np.uint32)
    return k


@pytest.fixture
def variant(gpu_ctx, request):
    """(stage, variant) forced for one test, automatic choice restored afterwards."""
    stage, var = request.param
    gpu_ctx.set_variant(stage, var)
    yield var
    gpu_ctx.set_variant(stage, 0)


@pytest.mark.parametrize("variant", [(1, 1), (1, 2), (1, 3), (1, 4)], indirect=True,
                         ids=["fold-4wave", "fold-8wave", "fold-16wave", "fold-split"])
@pytest.mark.parametrize("nu2", [1, 2, 4])
@pytest.mark.parametrize("general", [False, True])
def test_fold(gpu_ctx, oracle, nu2, general, variant):
    from spiral_pir import ops

    g = rng(700 + nu2 + 10 * general)
    t_gsw = 8
    accs = rand_raw(g, (1 << nu2) * 2).reshape(1 << nu2, 2, 2, N)
    vf = _rand_keys(g, nu2, 2 * t_gsw)
    vneg = _rand_keys(g, nu2, 2 * t_gsw) if general else oracle.folding_neg(vf, nu2, t_gsw)
    t = dev(accs)
    ops.FOLD_CIPHERTEXTS(gpu_ctx, t, dev(vf), dev(vneg) if general else None, nu2=nu2, t_gsw=t_gsw)
    got = host(t).reshape(1 << nu2, 2, 2, N)
    exp = oracle.fold(accs, nu2, t_gsw, vf, vneg)
    np.testing.assert_array_equal(got[0], exp[0])


@pytest.mark.parametrize("variant", [(0, 1), (0, 2), (0, 3), (0, 4), (0, 21)], indirect=True,
                         ids=["expand-4wave", "expand-8wave", "expand-16wave", "expand-split", "expand-halves"])
@pytest.mark.parametrize("nu1,nu2", [(6, 2), (3, 1)])
def test_expansion_random_keys(gpu_ctx, oracle, nu1, nu2, variant):
    from spiral_pir import ops

    g = rng(800 + nu1)
    t_gsw = t_exp = 8
    gg = oracle.expansion_g(nu1, nu2, t_gsw)
    st = oracle.expansion_stop_round(nu2, t_gsw)
    q = rand_residues(g, 2).reshape(2, 2, N)
    left = _rand_keys(g, gg, t_exp)
    right = _rand_keys(g, st + 1, t_exp)
    v = np.zeros((1 << gg, 2, 2, N), np.uint32)
    v[0] = q
    t = dev(v)
    ops.coefficient_expansion(gpu_ctx, t, dev(left), dev(right), nu1, nu2, t_gsw, t_exp)
    got = host(t).reshape(1 << gg, 2, 2, N)
    exp = oracle.expand(q, nu1, nu2, t_gsw, t_exp, left, right)
    for i in range(1 << gg):
        np.testing.assert_array_equal(got[i], exp[i], err_msg=f"element {i}")


def _answer_case(gpu_ctx, oracle, nu1, nu2, seed_db, seed_client, index):
    from spiral_pir import Shape
    from spiral_pir.client import Client
    from spiral_pir.server import Database, Query, READ_DB_MUL, READ_EXPANDED, STAGE_DB_MUL, STAGE_EXPAND, STAGE_FOLD

    s = Shape(nu1, nu2)
    recs = rng(seed_db).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
    db = Database(gpu_ctx, s)
    db.load_bytes(recs)
    c = Client(s, seed_client)
    left, right = c.keygen()
    qct, vf = c.query(index)
    q = Query(db)
    q.load(qct, left, right, vf)
    q.run(STAGE_EXPAND)
    v_gpu = q.read(READ_EXPANDED)
    q.run(STAGE_DB_MUL)
    acc_gpu = q.read(READ_DB_MUL)
    q.run(STAGE_FOLD)
    resp = q.read()
    # oracle, stage by stage
    v_exp = oracle.expand(qct, nu1, nu2, s.t_gsw, s.t_exp, left, right)
    np.testing.assert_array_equal(v_gpu, v_exp)
    db_ntt = oracle.db_encode(recs, s.n2, nu1, nu2)
    sel = np.ascontiguousarray(v_exp[0:2 * s.rows:2])
    for t in range(s.n2):
        np.testing.assert_array_equal(acc_gpu[t], oracle.db_mul(db_ntt[t], sel, nu1, nu2, nthreads=0),
                                      err_msg=f"trial {t}")
    vneg = oracle.folding_neg(vf, nu2, s.t_gsw)
    exp = oracle.answer(db_ntt, s.n2, nu1, nu2, s.t_gsw, s.t_exp, qct, left, right, vf, vneg)
    np.testing.assert_array_equal(resp, exp)
    np.testing.assert_array_equal(c.decode(resp), recs[index])
    q.close()
    db.close()


def test_answer_small_stage_by_stage(gpu_ctx, oracle):
    """Config-0 shape (nu1, nu2) = (6, 2), n = 2: every stage bit-exact, record decoded."""
    _answer_case(gpu_ctx, oracle, 6, 2, 1000, 2000, 37)


def test_answer_small_legacy_kernels(gpu_ctx, oracle):
    """The same answer through the 4-wave expansion and fold kernels."""
    gpu_ctx.set_variant(0, 1)
    gpu_ctx.set_variant(1, 1)
    try:
        _answer_case(gpu_ctx, oracle, 6, 2, 1000, 2000, 37)
    finally:
        gpu_ctx.set_variant(0, 0)
        gpu_ctx.set_variant(1, 0)


def test_answer_other_shape(gpu_ctx, oracle):
    _answer_case(gpu_ctx, oracle, 4, 3, 1004, 2004, 101)


def test_answer_config1_shape(gpu_ctx, oracle):
    """(nu1, nu2) = (9, 6), n = 2: the single-query benchmark shape, every stage bit-exact."""
    _answer_case(gpu_ctx, oracle, 9, 6, 1001, 2001, 12345)


def test_sp_answer_matches_oracle(gpu_ctx, oracle):
    """sp_answer, the one-call host entry point (host query, keys and fold vectors
    in, n2 raw response ciphertexts out), at (6, 2) and (9, 6): equal to the
    oracle's answer; wrong-size inputs are refused before the library call."""
    from spiral_pir import Shape
    from spiral_pir.client import Client
    from spiral_pir.server import Database, answer

    for (nu1, nu2, seed, index) in ((6, 2, 1000, 37), (9, 6, 1001, 12345)):
        s = Shape(nu1, nu2)
        recs = rng(seed).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
        db = Database(gpu_ctx, s)
        db.load_bytes(recs)
        c = Client(s, 2000 + seed)
        left, right = c.keygen()
        qct, vf = c.query(index)
        got = answer(db, qct, left, right, vf)
        exp = oracle.answer(oracle.db_encode(recs, s.n2, nu1, nu2), s.n2, nu1, nu2, s.t_gsw, s.t_exp, qct, left,
                            right, vf, oracle.folding_neg(vf, nu2, s.t_gsw))
        np.testing.assert_array_equal(got, exp, err_msg=f"({nu1},{nu2})")
        np.testing.assert_array_equal(c.decode(got), recs[index])
        with pytest.raises(ValueError, match="v_folding"):
            answer(db, qct, left, right, vf.reshape(-1)[:-1])
        db.close()


def test_query_rerun_is_identical(gpu_ctx):
    """Running the same resident query twice gives the same response (stage 1 restarts
    from the stored query)."""
    from spiral_pir import Shape
    from spiral_pir.client import Client
    from spiral_pir.server import Database, Query

    s = Shape(5, 2)
    recs = rng(31).integers(0, 256, size=(s.records, s.n2, N), dtype=np.uint8)
    db = Database(gpu_ctx, s)
    db.load_bytes(recs)
    c = Client(s, 32)
    left, right = c.keygen()
    qct, vf = c.query(77)
    q = Query(db)
    q.load(qct, left, right, vf)
    q.run()
    r1 = q.read()
    q.run()
    r2 = q.read()
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_array_equal(c.decode(r1), recs[77])
    q.close()
    db.close()


@pytest.mark.parametrize("words", [4, 1 << 20, (1 << 22) + 12])
def test_stream_read_reads_every_word(gpu_ctx, words):
    """The streaming-read ceiling kernel (bench measurement) touches every word:
    the xor of its sink equals the xor of the buffer, for sizes below one pass
    of the grid, a whole number of passes and a ragged tail."""
    import torch

    a = rng(990 + words % 97).integers(0, 2**32, size=words, dtype=np.uint32)
    d = dev(a)
    sink = torch.full((gpu_ctx.stream_sink_words(),), -1, dtype=torch.int32, device="cuda")
    gpu_ctx.stream_read(d.data_ptr(), 4 * words, sink.data_ptr())
    got = int(np.bitwise_xor.reduce(host(sink)))
    assert got == int(np.bitwise_xor.reduce(a))
