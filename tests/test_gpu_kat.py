"""The GPU operators against the reference's own known-answer vectors
(tests/golden/reference_kat.npz), independent of the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import kat_expansion_keys
from test_gpu_parity import dev, host

pytestmark = pytest.mark.gpu

N = 2048


def test_gpu_ntt_kat(gpu_ctx, kat):
    """raw_poly.dat -> ntt_poly.dat (gpu-pir/cpu_version/ntt.py:379-403) through to_ntt and back."""
    import torch
    from spiral_pir import ops

    out = torch.zeros(2 * N, dtype=torch.int32, device="cuda")
    ops.to_ntt(gpu_ctx, dev(kat["raw_poly"]), out)
    np.testing.assert_array_equal(host(out).reshape(2, N), kat["ntt_poly"])
    back = torch.zeros(2 * N, dtype=torch.int32, device="cuda")
    ops.from_ntt(gpu_ctx, dev(kat["ntt_poly"]), back)
    np.testing.assert_array_equal(host(back).reshape(2, N), kat["raw_poly"])
    for m in (0, 1):  # the bare residue transform on one row
        t = dev(np.ascontiguousarray(kat["ntt_poly"][m]))
        ops.INTT_R(gpu_ctx, t, m, scale=True)
        ops.NTT_R(gpu_ctx, t, m)
        np.testing.assert_array_equal(host(t), kat["ntt_poly"][m])


@pytest.mark.parametrize("given_neg", [False, True], ids=["neg=G-C", "neg given"])
def test_gpu_fold_kat(gpu_ctx, oracle, kat, given_neg):
    """client_fold_input + client_folding_keys -> client_fold_output (gpu-pir/fold_test.cu:30-94)."""
    from spiral_pir import ops

    keys = kat["folding_keys"]
    neg = dev(oracle.folding_neg(keys, 2, 8)) if given_neg else None
    t = dev(kat["fold_input"])
    ops.FOLD_CIPHERTEXTS(gpu_ctx, t, dev(keys), neg, nu2=2, t_gsw=8)
    np.testing.assert_array_equal(host(t).reshape(4, 2, 2, N)[0], kat["fold_output"])


def test_gpu_expansion_round0_kat(gpu_ctx, kat):
    """client_expand_input + right keys -> client_expand_output_1round (gpu-pir/cpu_version/expansion.py:262-294)."""
    from spiral_pir import ops

    left, right = kat_expansion_keys(kat)
    v = np.zeros((1 << 7, 2, 2, N), np.uint32)
    v[0] = kat["expand_input"]
    t = dev(v)
    ops.coefficient_expansion(gpu_ctx, t, dev(left), dev(right), 6, 2, 8, 8, rounds=1)
    got = host(t).reshape(1 << 7, 2, 2, N)
    np.testing.assert_array_equal(got[0], kat["expand_output_1round"][0])
    np.testing.assert_array_equal(got[1], kat["expand_output_1round"][1])
