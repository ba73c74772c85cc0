"""The oracle against the reference's own known-answer fixtures.

The fixtures are spiral-rs dumps kept by the reference under
gpu-pir/test_data (formats: gpu-pir/cpu_version/testing.py:355-408).  They are
read where the reference is mounted (this container) and the tests skip
elsewhere; tests/golden/ carries SHA-256 digests of the expected outputs and
oracle-made vectors that travel with the repository.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def _td(ref_data, name):
    return os.path.join(ref_data, "test_data", name)


def test_ntt_kat(oracle, ref_data):
    """gpu-pir/cpu_version/ntt.py:379-403: raw_poly.dat -> ntt_poly.dat (both residues)."""
    raw = oracle.read_raw_polys(_td(ref_data, "raw_poly.dat"))[0]
    ntt = oracle.read_ntt_polys(_td(ref_data, "ntt_poly.dat"))[0]
    np.testing.assert_array_equal(oracle.to_ntt(raw), ntt)


def test_from_ntt_kat(oracle, ref_data):
    """INTT + inverse CRT takes ntt_poly.dat back to raw_poly.dat."""
    raw = oracle.read_raw_polys(_td(ref_data, "raw_poly.dat"))[0]
    ntt = oracle.read_ntt_polys(_td(ref_data, "ntt_poly.dat"))[0]
    np.testing.assert_array_equal(oracle.from_ntt(ntt), raw)


def test_fold_kat(oracle, ref_data):
    """gpu-pir/fold_test.cu:30-94: client_fold_input + keys -> client_fold_output (nu2 = 2, T_GSW = 8)."""
    inp = oracle.read_raw_polys(_td(ref_data, "client_fold_input.dat")).reshape(4, 2, 2, 2048)
    keys = oracle.matrices_to_keys(oracle.read_ntt_polys(_td(ref_data, "client_folding_keys.dat")), 2, 16)
    kneg = oracle.matrices_to_keys(oracle.read_ntt_polys(_td(ref_data, "client_folding_keys_neg.dat")), 2, 16)
    out = oracle.read_raw_polys(_td(ref_data, "client_fold_output.dat")).reshape(2, 2, 2048)
    got = oracle.fold(inp, 2, 8, keys, kneg)
    np.testing.assert_array_equal(got[0], out)


def test_folding_neg_is_gadget_minus_keys(oracle, ref_data):
    """client_folding_keys_neg.dat = G - client_folding_keys.dat exactly."""
    keys = oracle.matrices_to_keys(oracle.read_ntt_polys(_td(ref_data, "client_folding_keys.dat")), 2, 16)
    kneg = oracle.matrices_to_keys(oracle.read_ntt_polys(_td(ref_data, "client_folding_keys_neg.dat")), 2, 16)
    np.testing.assert_array_equal(oracle.folding_neg(keys, 2, 8), kneg)


def test_expansion_round0_kat(oracle, ref_data):
    """gpu-pir/cpu_version/expansion.py:262-294: one expansion round of client_expand_input.dat
    with client_expand_right_keys.dat gives client_expand_output_1round.dat (v[0] and v[1]);
    this also fixes v_neg1 = -x^(N-1)."""
    q = oracle.read_ntt_polys(_td(ref_data, "client_expand_input.dat")).reshape(2, 2, 2048)
    right = oracle.matrices_to_keys(oracle.read_ntt_polys(_td(ref_data, "client_expand_right_keys.dat")), 2, 8)
    left = oracle.matrices_to_keys(oracle.read_ntt_polys(_td(ref_data, "client_expand_left_keys.dat")), 2, 8)
    assert left.shape[0] == oracle.expansion_g(6, 2, 8) == 7
    assert right.shape[0] == oracle.expansion_stop_round(2, 8) + 1 == 5
    exp = oracle.read_ntt_polys(_td(ref_data, "client_expand_output_1round.dat")).reshape(2, 2, 2, 2048)
    v = oracle.expand(q, 6, 2, 8, 8, left, right, rounds=1)
    np.testing.assert_array_equal(v[0], exp[0])
    np.testing.assert_array_equal(v[1], exp[1])


def test_fixture_digests(ref_data):
    """The reference's expected outputs are the files whose digests tests/golden records."""
    with open(os.path.join(HERE, "golden", "reference_digests.json")) as f:
        digests = json.load(f)
    for name, digest in digests.items():
        with open(os.path.join(ref_data, name), "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == digest, name


def test_db_mul_kat_inputs_are_unusable_without_db(ref_data):
    """eric_first_dim.dat / eric_folding_inputs.dat exist but eric_db.dat does not
    (.MISSING_LARGE_BLOBS), so the reference's DB-multiply KAT cannot run; the DB
    multiply is pinned by the closed form and brute force instead."""
    assert os.path.exists(os.path.join(ref_data, "data", "eric_first_dim.dat"))
    assert not os.path.exists(os.path.join(ref_data, "data", "eric_db.dat"))


# ---- the same vectors from tests/golden/reference_kat.npz (run everywhere) ----

def test_golden_vectors_are_the_reference_files(oracle, ref_data, kat):
    """reference_kat.npz holds exactly what the reference's files hold."""
    np.testing.assert_array_equal(kat["raw_poly"], oracle.read_raw_polys(_td(ref_data, "raw_poly.dat"))[0])
    np.testing.assert_array_equal(kat["ntt_poly"], oracle.read_ntt_polys(_td(ref_data, "ntt_poly.dat"))[0])
    np.testing.assert_array_equal(
        kat["folding_keys"],
        oracle.matrices_to_keys(oracle.read_ntt_polys(_td(ref_data, "client_folding_keys.dat")), 2, 16))
    np.testing.assert_array_equal(
        kat["expand_right_keys0"],
        oracle.matrices_to_keys(oracle.read_ntt_polys(_td(ref_data, "client_expand_right_keys.dat")), 2, 8)[:1])
    np.testing.assert_array_equal(
        kat["fold_output"], oracle.read_raw_polys(_td(ref_data, "client_fold_output.dat")).reshape(2, 2, 2048))
    np.testing.assert_array_equal(
        kat["expand_output_1round"],
        oracle.read_ntt_polys(_td(ref_data, "client_expand_output_1round.dat")).reshape(2, 2, 2, 2048))


def test_golden_ntt(oracle, kat):
    np.testing.assert_array_equal(oracle.to_ntt(kat["raw_poly"]), kat["ntt_poly"])
    np.testing.assert_array_equal(oracle.from_ntt(kat["ntt_poly"]), kat["raw_poly"])


def test_golden_fold(oracle, kat):
    keys = kat["folding_keys"]
    got = oracle.fold(kat["fold_input"], 2, 8, keys, oracle.folding_neg(keys, 2, 8))
    np.testing.assert_array_equal(got[0], kat["fold_output"])


def test_golden_expansion_round0(oracle, kat):
    from conftest import kat_expansion_keys

    left, right = kat_expansion_keys(kat)
    v = oracle.expand(kat["expand_input"], 6, 2, 8, 8, left, right, rounds=1)
    np.testing.assert_array_equal(v[0], kat["expand_output_1round"][0])
    np.testing.assert_array_equal(v[1], kat["expand_output_1round"][1])
