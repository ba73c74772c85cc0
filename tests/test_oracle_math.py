"""The oracle against what the mathematics fixes, independent of the
reference's fixtures: closed forms, brute force on small inputs, algebraic
laws.  Each check is chosen so that a plausible slip (a wrong twiddle index,
sign, operand order or bit width) fails it."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import rand_raw, rand_residues, rng

N = 2048
Q0, Q1 = 268369921, 249561089
QQ = Q0 * Q1
QS = (Q0, Q1)


def negacyclic_mul(a, b, q):
    """Schoolbook product in Z_q[x]/(x^N + 1) with Python ints."""
    a = [int(x) for x in a]
    b = [int(x) for x in b]
    out = [0] * N
    nz = [j for j in range(N) if b[j]]
    for i in range(N):
        if not a[i]:
            continue
        for j in nz:
            k = i + j
            if k < N:
                out[k] += a[i] * b[j]
            else:
                out[k - N] -= a[i] * b[j]
    return np.array([x % q for x in out], dtype=np.uint32)


def test_constants(oracle):
    assert oracle.lib().or_q(0) == Q0 and oracle.lib().or_q(1) == Q1
    for m, psi in enumerate((66687, 158221)):
        q = QS[m]
        assert pow(psi, 2048, q) == q - 1  # primitive 4096-th root: psi^N = -1
        assert pow(psi, 4096, q) == 1
    # gpu-pir/params.cuh:34-35: CRT constants = Bezout coefficient * N^-1 * 2^32.
    assert (pow(Q1, -1, Q0) * pow(N, -1, Q0) * 2**32) % Q0 == 88833249
    assert (pow(Q0, -1, Q1) * pow(N, -1, Q1) * 2**32) % Q1 == 166953760
    # gpu-pir/ntt_test.cu:442: 268238881 = 2048^-1 mod Q0.
    assert pow(N, -1, Q0) == 268238881
    # gpu-pir/params.cuh:36-39: Montgomery constants.
    assert (4026597377 * Q0) % 2**32 == 1 and (4045406209 * Q1) % 2**32 == 1
    assert 2**64 % Q0 == 234877184 and 2**64 % Q1 == 148369395


@pytest.mark.parametrize("m", [0, 1])
def test_ntt_closed_form(oracle, m):
    g = rng(10 + m)
    a = g.integers(0, QS[m], size=N, dtype=np.uint32)
    np.testing.assert_array_equal(oracle.ntt_fwd(a, m), oracle.ntt_brute(a, m))


@pytest.mark.parametrize("m", [0, 1])
def test_ntt_convolution(oracle, m):
    """NTT turns negacyclic products into pointwise products (sparse b keeps it fast)."""
    g = rng(20 + m)
    q = QS[m]
    a = g.integers(0, q, size=N, dtype=np.uint32)
    b = np.zeros(N, np.uint32)
    b[g.choice(N, 12, replace=False)] = g.integers(1, q, size=12, dtype=np.uint32)
    b[N - 1] = 7  # exercises the wrap-around sign
    prod = (oracle.ntt_fwd(a, m).astype(np.uint64) * oracle.ntt_fwd(b, m) % q).astype(np.uint32)
    np.testing.assert_array_equal(oracle.ntt_inv(prod, m), negacyclic_mul(a, b, q))


@pytest.mark.parametrize("m", [0, 1])
def test_ntt_round_trip(oracle, m):
    g = rng(30 + m)
    for _ in range(8):
        a = g.integers(0, QS[m], size=N, dtype=np.uint32)
        np.testing.assert_array_equal(oracle.ntt_inv(oracle.ntt_fwd(a, m), m), a)


def test_crt_from_ntt(oracle):
    """from_ntt(to_ntt(x)) = x for x in [0, Q), and the CRT output is the unique lift."""
    g = rng(40)
    raw = rand_raw(g, 3)
    for i in range(3):
        np.testing.assert_array_equal(oracle.from_ntt(oracle.to_ntt(raw[i])), raw[i])
    # Q itself reduces to 0 in both residues, so from_ntt gives 0, not Q.
    r = np.zeros((2, N), np.uint32)
    r[0, :] = QQ >> 32
    r[1, :] = QQ & 0xFFFFFFFF
    assert not oracle.to_ntt(r).any()


@pytest.mark.parametrize("t", [3, 5, 2049, 1025, 129, 4095])
def test_automorph_brute(oracle, t):
    """gpu-pir/automorph_test.cu:14-22 written out with Python ints, plus the Q convention."""
    g = rng(50 + t)
    raw = rand_raw(g, 1)[0]
    raw[:, 3] = 0
    x = oracle.raw_to_u64(raw)[0]
    exp = [0] * N
    for i in range(N):
        idx, flip = (i * t) % N, ((i * t) // N) % 2
        exp[idx] = (QQ - int(x[i])) if flip else int(x[i])
    got = oracle.raw_to_u64(oracle.automorph(raw, t))[0]
    assert [int(v) for v in got] == exp
    if ((3 * t) // N) % 2:
        assert int(got[(3 * t) % N]) == QQ


def test_automorph_composition(oracle):
    """tau_s(tau_t(x)) = tau_{st mod 2N}(x) on values reduced mod Q."""
    g = rng(60)
    raw = rand_raw(g, 1)[0]
    lhs = oracle.raw_to_u64(oracle.automorph(oracle.automorph(raw, 5), 2049))[0] % QQ
    rhs = oracle.raw_to_u64(oracle.automorph(raw, (5 * 2049) % 4096))[0] % QQ
    np.testing.assert_array_equal(lhs, rhs)


def test_automorph_in_ntt_domain_is_a_permutation(oracle):
    """NTT(tau_t f)[k] = NTT(f)[k'] with 2 brv(k') + 1 = t (2 brv(k) + 1) mod 2N —
    the identity the GPU expansion uses to skip an NTT."""
    def brv(x):
        return int(f"{x:011b}"[::-1], 2)

    g = rng(70)
    raw = rand_raw(g, 1)[0]
    for t in (2049, 1025, 65):
        a = oracle.to_ntt(raw)
        b = oracle.to_ntt(oracle.automorph(raw, t))
        for k in (0, 1, 77, 2047):
            e = (t * (2 * brv(k) + 1)) % 4096
            k2 = brv((e - 1) // 2)
            assert b[0, k] == a[0, k2] and b[1, k] == a[1, k2]


@pytest.mark.parametrize("rdim,mx", [(1, 8), (2, 16), (2, 8), (1, 4)])
def test_gadget_recombines(oracle, rdim, mx):
    g = rng(80 + mx)
    raw = rand_raw(g, rdim, include_q=True)
    d = oracle.gadget_inv(raw, rdim, mx).astype(object)
    k = mx // rdim
    bits = oracle.gadget_bits(k)
    assert bits == -(-56 // k)
    assert int(d.max()) < (1 << bits)
    for j in range(rdim):
        x = oracle.raw_to_u64(raw[j])[0]
        rec = sum(d[j + i * rdim] << (bits * i) for i in range(k))
        assert [int(v) for v in rec] == [int(v) for v in x]


def test_extprod_is_linear_and_matches_direct(oracle):
    """M * NTT(g^-1(x)): against a direct per-digit computation with Python ints at a few slots."""
    g = rng(90)
    mx = 8
    raw = rand_raw(g, 1)
    keys = np.empty((2, mx, 2, N), np.uint32)
    keys[0] = g.integers(0, Q0, size=(mx, 2, N), dtype=np.uint32)
    keys[1] = g.integers(0, Q1, size=(mx, 2, N), dtype=np.uint32)
    acc = oracle.extprod_acc(raw, 1, mx, keys)
    digits = oracle.gadget_inv(raw, 1, mx)
    for m, q in enumerate(QS):
        dn = [oracle.ntt_fwd(digits[c] % q, m) for c in range(mx)]
        for row in range(2):
            for k in (0, 5, 2047):
                exp = sum(int(dn[c][k]) * int(keys[m, c, row, k]) for c in range(mx)) % q
                assert acc[row, m, k] == exp


def test_db_mul_brute(oracle):
    """out[c][r][m][z] = sum_j sel[j][r][m][z] * db[c][j][m][z] mod q, checked with Python ints."""
    g = rng(100)
    nu1, nu2 = 4, 2
    R, C = 1 << nu1, 1 << nu2
    db = rand_residues(g, R * C).reshape(C, R, 2, N)
    sel = rand_residues(g, 2 * R).reshape(R, 2, 2, N)
    out = oracle.db_mul(db, sel, nu1, nu2)
    for c in range(C):
        for r in range(2):
            for m, q in enumerate(QS):
                for z in (0, 1, 1000, 2047):
                    exp = sum(int(sel[j, r, m, z]) * int(db[c, j, m, z]) for j in range(R)) % q
                    assert out[c, r, m, z] == exp
    np.testing.assert_array_equal(oracle.db_mul(db, sel, nu1, nu2, nthreads=4), out)


def test_db_mul_linearity(oracle):
    g = rng(110)
    nu1, nu2 = 3, 1
    R, C = 1 << nu1, 1 << nu2
    db = rand_residues(g, R * C).reshape(C, R, 2, N)
    s1 = rand_residues(g, 2 * R).reshape(R, 2, 2, N)
    s2 = rand_residues(g, 2 * R).reshape(R, 2, 2, N)
    s12 = s1.copy()
    for m, q in enumerate(QS):
        s12[:, :, m] = ((s1[:, :, m].astype(np.uint64) + s2[:, :, m]) % q).astype(np.uint32)
    o1, o2, o12 = (oracle.db_mul(db, s, nu1, nu2) for s in (s1, s2, s12))
    for m, q in enumerate(QS):
        np.testing.assert_array_equal(o12[:, :, m], ((o1[:, :, m].astype(np.uint64) + o2[:, :, m]) % q))


def test_v_neg1_closed_form(oracle):
    """v_neg1[r] = NTT(-x^(N - 2^r)) equals x^(-2^r) at every evaluation point."""
    def brv(x):
        return int(f"{x:011b}"[::-1], 2)

    for r in (0, 3, 10):
        v = oracle.v_neg1(r)
        for m, (q, psi) in enumerate(zip(QS, (66687, 158221))):
            for k in (0, 9, 2047):
                e = (-(1 << r) * (2 * brv(k) + 1)) % 4096
                assert v[m, k] == pow(psi, e, q)


def test_fold_picks_by_bit(oracle):
    """With C = GSW(1) built from G alone (no noise), fold returns acc[num_per + i]; with
    C = 0 it returns acc[i] — the selection law behind the reference's fold."""
    g = rng(120)
    nu2, t_gsw = 2, 8
    accs = rand_raw(g, 8).reshape(4, 2, 2, N)
    zero = np.zeros((nu2, 2, 16, 2, N), np.uint32)
    gad = oracle.folding_neg(zero, nu2, t_gsw)  # G - 0 = G
    # bits (1, 0): level 1 selects the upper half, level 0 the lower.
    vf = np.stack([zero[0], gad[1]])
    out = oracle.fold(accs, nu2, t_gsw, vf, oracle.folding_neg(vf, nu2, t_gsw))
    np.testing.assert_array_equal(out[0], accs[2])


def test_threads_do_not_change_results(oracle):
    """The oracle's expansion and fold give the same words on 1 and 4 threads
    (the CPU baseline times them on all host cores)."""
    g = rng(130)
    nu1, nu2, t_gsw, t_exp = 6, 2, 8, 8
    gexp = oracle.expansion_g(nu1, nu2, t_gsw)
    stop = oracle.expansion_stop_round(nu2, t_gsw)
    query = rand_residues(g, 2).reshape(2, 2, N)
    left = rand_residues(g, gexp * 16).reshape(gexp, 2, 8, 2, N)
    right = rand_residues(g, (stop + 1) * 16).reshape(stop + 1, 2, 8, 2, N)
    accs = rand_raw(g, 8).reshape(4, 2, 2, N)
    vf = rand_residues(g, nu2 * 32).reshape(nu2, 2, 16, 2, N)
    vneg = oracle.folding_neg(vf, nu2, t_gsw)
    out = {}
    try:
        for n in (1, 4):
            oracle.set_threads(n)
            out[n] = (oracle.expand(query, nu1, nu2, t_gsw, t_exp, left, right),
                      oracle.fold(accs, nu2, t_gsw, vf, vneg))
    finally:
        oracle.set_threads(1)
    np.testing.assert_array_equal(out[1][0], out[4][0])
    np.testing.assert_array_equal(out[1][1], out[4][1])


def test_expansion_key_set_choice_planted(oracle):
    """Which key set each expansion step uses (gpu-pir/expansion.cuh:121-130): in
    round r, element i takes left[r] when r != 0 and i is even, else right[r].
    Planted mistakes in the oracle's inputs: corrupting right[r] (1 <= r <=
    stop_round) changes exactly the odd elements of round r and no even one,
    corrupting left[r] exactly the even ones; in round 0 right[0] serves both
    elements and left[0] neither; past stop_round left[r] changes the even
    elements only (odd ones are not updated there).  A swap of the key sets, a
    wrong parity test or a wrong round index fails one of these."""
    nu1, nu2, t = 6, 2, 8
    g_, st = oracle.expansion_g(nu1, nu2, t), oracle.expansion_stop_round(nu2, t)
    assert (g_, st) == (7, 4)
    g = rng(4242)
    q = rand_residues(g, 2).reshape(2, 2, N)
    def keys(gen, levels):  # key matrices [level][crt][col][row][z], canonical residues
        k = np.empty((levels, 2, t, 2, N), np.uint32)
        k[:, 0] = gen.integers(0, Q0, size=(levels, t, 2, N), dtype=np.uint32)
        k[:, 1] = gen.integers(0, Q1, size=(levels, t, 2, N), dtype=np.uint32)
        return k

    left, right = keys(g, g_), keys(g, st + 1)

    def changed(r, which):
        base = oracle.expand(q, nu1, nu2, t, t, left, right, rounds=r + 1)
        lft, rgt = left.copy(), right.copy()
        k = lft if which == "left" else rgt
        k[r] = keys(rng(100 + r), 1)[0]
        bad = oracle.expand(q, nu1, nu2, t, t, lft, rgt, rounds=r + 1)
        diff = {i for i in range(1 << g_) if not np.array_equal(base[i], bad[i])}
        assert all(np.array_equal(base[i], bad[i]) for i in range(1 << (r + 1), 1 << g_))
        return diff

    live = lambda r: set(range(1 << (r + 1)))  # noqa: E731
    assert changed(0, "right") == {0, 1}
    assert changed(0, "left") == set()
    for r in range(1, st + 1):
        assert changed(r, "right") == {i for i in live(r) if i % 2}, f"right[{r}]"
        assert changed(r, "left") == {i for i in live(r) if i % 2 == 0}, f"left[{r}]"
    for r in range(st + 1, g_):
        assert changed(r, "left") == {i for i in live(r) if i % 2 == 0}, f"left[{r}] past stop_round"
