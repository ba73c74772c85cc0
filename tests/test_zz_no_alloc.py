"""sp_fold_noalloc, the fold entry point that allocates nothing (its temporaries
live in a static device array): called through the C ABI, its result equals
sp_fold's on the same inputs, bit for bit, over two calls in a row (the second
smaller, reusing the array with the first call's contents still in it), in the
fast (G - C) and the general form; sizes beyond the array are refused."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from conftest import rand_raw, rng

pytestmark = pytest.mark.gpu

N = 2048
Q0, Q1 = 268369921, 249561089


def _keys(g, levels, mx):
    k = np.empty((levels, 2, mx, 2, N), np.uint32)
    k[:, 0] = g.integers(0, Q0, size=(levels, mx, 2, N), dtype=np.uint32)
    k[:, 1] = g.integers(0, Q1, size=(levels, mx, 2, N), dtype=np.uint32)
    return k


def _dev(a):
    import torch

    return torch.from_numpy(np.ascontiguousarray(a).view(np.int32).copy()).cuda()


def _fold(gpu_ctx, name, accs, vf, vneg, nu2, t_gsw, stream):
    from spiral_pir import _lib

    d = _dev(accs)
    kv = _dev(vf)
    kn = None if vneg is None else _dev(vneg)
    _lib.call(name, gpu_ctx.handle, ctypes.c_void_p(d.data_ptr()), nu2, t_gsw, ctypes.c_void_p(kv.data_ptr()),
              ctypes.c_void_p(0 if kn is None else kn.data_ptr()), ctypes.c_void_p(stream))
    return d


@pytest.mark.parametrize("general", [False, True], ids=["neg=G-C", "neg given"])
def test_fold_noalloc_equals_fold(gpu_ctx, oracle, general):
    import torch

    stream = torch.cuda.current_stream().cuda_stream
    t_gsw = 8
    for call, nu2 in enumerate((7, 3) if not general else (6, 2)):
        g = rng(9100 + 10 * nu2 + general)
        accs = rand_raw(g, (1 << nu2) * 2).reshape(1 << nu2, 2, 2, N)
        vf = _keys(g, nu2, 2 * t_gsw)
        vneg = _keys(g, nu2, 2 * t_gsw) if general else None
        ref = _fold(gpu_ctx, "sp_fold", accs, vf, vneg, nu2, t_gsw, stream)
        got = _fold(gpu_ctx, "sp_fold_noalloc", accs, vf, vneg, nu2, t_gsw, stream)
        torch.cuda.synchronize()
        r = ref.cpu().numpy().view(np.uint32).reshape(1 << nu2, 2, 2, N)
        a = got.cpu().numpy().view(np.uint32).reshape(1 << nu2, 2, 2, N)
        np.testing.assert_array_equal(a[0], r[0], err_msg=f"call {call}: nu2 {nu2}")
        exp = oracle.fold(accs, nu2, t_gsw, vf, vneg if general else oracle.folding_neg(vf, nu2, t_gsw))
        np.testing.assert_array_equal(a[0], exp[0], err_msg=f"call {call}: nu2 {nu2} vs oracle")


def test_fold_noalloc_refuses_what_does_not_fit(gpu_ctx):
    from spiral_pir import SpiralError, _lib

    import torch

    t = torch.zeros(4, dtype=torch.int32, device="cuda")
    p = ctypes.c_void_p(t.data_ptr())
    with pytest.raises(SpiralError, match="scratch"):
        _lib.call("sp_fold_noalloc", gpu_ctx.handle, p, 8, 8, p, ctypes.c_void_p(0), ctypes.c_void_p(0))
    with pytest.raises(SpiralError, match="scratch"):
        _lib.call("sp_fold_noalloc", gpu_ctx.handle, p, 7, 16, p, p, ctypes.c_void_p(0))
