"""spiral_pir.host, the product-side mirror of gpu-pir/host.h (load_polys,
group_to_ciphertext, flatten; host.cu:216-268): read the spiral-rs dump
formats the reference's tests use into this library's layouts.

The expected arrays are the reference's known-answer vectors kept in
tests/golden/reference_kat.npz (written from the reference's fixture files by
the oracle's readers).  Where the reference is mounted, the fixture files
themselves are read; elsewhere dump files are written from the vectors in the
format the reference documents (gpu-pir/cpu_version/testing.py:61-114)."""
from __future__ import annotations

import os

import numpy as np
import pytest

N = 2048


def _write_raw(path, raw):
    """raw Polys (hi/lo u32 planes) -> u64[2048] per polynomial."""
    r = np.asarray(raw, np.uint32).reshape(-1, 2, N)
    ((r[:, 0].astype(np.uint64) << np.uint64(32)) | r[:, 1].astype(np.uint64)).tofile(path)


def _write_ntt(path, polys):
    """NTT Polys u32 [n][2][2048] -> u64 residues, residue q1 then q2 per polynomial."""
    np.asarray(polys, np.uint32).reshape(-1, 2, N).astype(np.uint64).tofile(path)


def _key_file_polys(keys):
    """Key layout [level][crt][col][row][z] -> the dump's polynomial order: row-major
    2 x cols matrices per level, each polynomial [crt][z]."""
    return np.ascontiguousarray(keys.transpose(0, 3, 2, 1, 4))


@pytest.fixture
def fixture_files(tmp_path, kat, request):
    """Paths of client_fold_input / client_folding_keys / client_fold_output /
    ntt_poly / raw_poly: the reference's files if mounted, else written from the KAT."""
    ref = "/root/reference/gpu-pir/test_data"
    names = ("client_fold_input.dat", "client_folding_keys.dat", "client_fold_output.dat", "ntt_poly.dat",
             "raw_poly.dat")
    if request.param == "reference":
        if not os.path.isdir(ref):
            pytest.skip("reference fixtures not mounted")
        return {n: os.path.join(ref, n) for n in names}
    p = {n: str(tmp_path / n) for n in names}
    _write_raw(p["client_fold_input.dat"], kat["fold_input"])
    _write_ntt(p["client_folding_keys.dat"], _key_file_polys(kat["folding_keys"]))
    _write_raw(p["client_fold_output.dat"], kat["fold_output"])
    _write_ntt(p["ntt_poly.dat"], kat["ntt_poly"])
    _write_raw(p["raw_poly.dat"], kat["raw_poly"])
    return p


@pytest.mark.parametrize("fixture_files", ["reference", "written"], indirect=True)
def test_fold_fixtures_through_host_helpers(fixture_files, kat):
    """The fold known-answer files read the way gpu-pir/fold_test.cu:33-36 reads
    them (raw input grouped by 1 and flattened; keys in NTT form grouped by
    2 T_GSW and copied into v_folding) give the KAT's fold input, keys and output."""
    from spiral_pir import host

    fin = host.flatten(host.group_to_ciphertext(host.load_polys(fixture_files["client_fold_input.dat"], False), 1))
    np.testing.assert_array_equal(fin, kat["fold_input"])
    cts = host.group_to_ciphertext(host.load_polys(fixture_files["client_folding_keys.dat"], True), 16)
    assert cts.shape == (2, 16, 2, 2, N)  # V2 levels x 2 T_GSW ciphertexts (fold_test.cu:37-38)
    np.testing.assert_array_equal(host.ciphertexts_to_key_layout(cts), kat["folding_keys"])
    np.testing.assert_array_equal(
        host.matrices_to_key_layout(host.load_polys(fixture_files["client_folding_keys.dat"], True), 16),
        kat["folding_keys"])
    fout = host.flatten(host.group_to_ciphertext(host.load_polys(fixture_files["client_fold_output.dat"], False), 1))
    np.testing.assert_array_equal(fout[0], kat["fold_output"])


@pytest.mark.parametrize("fixture_files", ["reference", "written"], indirect=True)
def test_single_polys_through_host_helpers(fixture_files, kat, oracle):
    from spiral_pir import host

    np.testing.assert_array_equal(host.load_polys(fixture_files["ntt_poly.dat"], True)[0], kat["ntt_poly"])
    np.testing.assert_array_equal(host.load_polys(fixture_files["raw_poly.dat"], False)[0], kat["raw_poly"])
    # the same files through the oracle's readers
    np.testing.assert_array_equal(host.load_polys(fixture_files["ntt_poly.dat"], True),
                                  oracle.read_ntt_polys(fixture_files["ntt_poly.dat"]))
    np.testing.assert_array_equal(host.load_polys(fixture_files["raw_poly.dat"], False),
                                  oracle.read_raw_polys(fixture_files["raw_poly.dat"]))


def test_load_polys_rejects_out_of_range_residues(tmp_path, kat):
    from spiral_pir import host

    p = str(tmp_path / "bad.dat")
    polys = kat["ntt_poly"].reshape(1, 2, N).astype(np.uint64)
    polys[0, 1, 9] = 249561089  # = Q1
    polys.tofile(p)
    with pytest.raises(ValueError, match="out of range"):
        host.load_polys(p, True)


def test_group_to_ciphertext_pairs_halves():
    """host.cu:186-196: ciphertext c of a group of 2g polynomials is (poly c, poly c + g)."""
    from spiral_pir import host

    polys = np.arange(6 * 2 * N, dtype=np.uint32).reshape(6, 2, N)
    cts = host.group_to_ciphertext(polys, 3)
    assert cts.shape == (1, 3, 2, 2, N)
    for c in range(3):
        np.testing.assert_array_equal(cts[0, c, 0], polys[c])
        np.testing.assert_array_equal(cts[0, c, 1], polys[c + 3])
    np.testing.assert_array_equal(host.flatten(cts), cts[:, 0])
