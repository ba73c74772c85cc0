#!/usr/bin/env python3
"""Writes tests/golden/reference_kat.npz: the reference's known-answer vectors for
this path, as data, so that tests can check the oracle and the GPU against them
where the reference is not mounted.

Source files (spiral-rs dumps kept by the reference, formats of
gpu-pir/cpu_version/testing.py:61-114), read from /root/reference/gpu-pir/test_data:

  raw_poly.dat, ntt_poly.dat            NTT / from_ntt KAT (gpu-pir/cpu_version/ntt.py:379-403)
  client_fold_input.dat, client_folding_keys.dat, client_fold_output.dat
                                        fold KAT, nu2 = 2, T_GSW = 8 (gpu-pir/fold_test.cu:30-94)
  client_expand_input.dat, client_expand_right_keys.dat (level 0 only),
  client_expand_output_1round.dat       expansion round 0 (gpu-pir/cpu_version/expansion.py:262-294)

Stored as u32 in the oracle's layouts (natural Spiral NTT order, raw polys as
hi/lo planes, key matrices as [level][crt][col][row][z]).  client_folding_keys_neg.dat
is not stored: it equals G - client_folding_keys.dat, which a test checks where
the reference is mounted.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402  (readers only; no oracle computation is stored)

SRC = "/root/reference/gpu-pir/test_data"


def main() -> None:
    td = lambda n: os.path.join(SRC, n)  # noqa: E731
    raw = O.read_raw_polys(td("raw_poly.dat"))[0]
    ntt = O.read_ntt_polys(td("ntt_poly.dat"))[0]
    fold_in = O.read_raw_polys(td("client_fold_input.dat")).reshape(4, 2, 2, 2048)
    fold_keys = O.matrices_to_keys(O.read_ntt_polys(td("client_folding_keys.dat")), 2, 16)
    fold_out = O.read_raw_polys(td("client_fold_output.dat")).reshape(2, 2, 2048)
    exp_in = O.read_ntt_polys(td("client_expand_input.dat")).reshape(2, 2, 2048)
    right0 = O.matrices_to_keys(O.read_ntt_polys(td("client_expand_right_keys.dat")), 2, 8)[:1]
    exp_out = O.read_ntt_polys(td("client_expand_output_1round.dat")).reshape(2, 2, 2, 2048)
    out = os.path.join(HERE, "reference_kat.npz")
    np.savez_compressed(out, raw_poly=raw, ntt_poly=ntt, fold_input=fold_in, folding_keys=fold_keys,
                        fold_output=fold_out, expand_input=exp_in, expand_right_keys0=right0,
                        expand_output_1round=exp_out)
    print(out, os.path.getsize(out))


if __name__ == "__main__":
    main()
