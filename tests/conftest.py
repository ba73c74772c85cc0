"""Shared test setup.

Markers: ``gpu`` tests need an MI355X (run with ``-m gpu``); everything else
runs on the CPU.  The oracle (oracle/) is the checker; the product package
is gpu-pir_amd/spiral_pir.
"""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "gpu-pir_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

REF_DATA = "/root/reference/gpu-pir"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X); run with -m gpu")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O

    O.lib()
    return O


@pytest.fixture(scope="session")
def kat():
    """The reference's known-answer vectors for this path, committed as data
    (tests/golden/reference_kat.npz, written by tests/golden/make_reference_kat.py)."""
    with np.load(os.path.join(ROOT, "tests", "golden", "reference_kat.npz")) as d:
        return {k: d[k] for k in d.files}


def kat_expansion_keys(kat):
    """Left/right key arrays for the (6, 2) round-0 vector: only right level 0 is read
    by round 0; the other levels are zero."""
    right = np.zeros((5, 2, 8, 2, 2048), np.uint32)
    right[0] = kat["expand_right_keys0"][0]
    left = np.zeros((7, 2, 8, 2, 2048), np.uint32)
    return left, right


@pytest.fixture(scope="session")
def ref_data():
    """Directory of the reference's fixture files (present only where the reference is mounted)."""
    if not os.path.isdir(os.path.join(REF_DATA, "test_data")):
        pytest.skip("reference fixtures not mounted")
    return REF_DATA


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("gpu test selected but no GPU is visible")
    from spiral_pir.server import Context

    torch.cuda.init()
    ctx = Context(0)
    yield ctx
    ctx.close()


def rng(seed: int) -> np.random.Generator:
    return np.random.default_rng(seed)


def rand_residues(g: np.random.Generator, n: int) -> np.ndarray:
    """n NTT Polys with canonical residues: u32 [n][2][2048]."""
    out = np.empty((n, 2, 2048), np.uint32)
    out[:, 0] = g.integers(0, 268369921, size=(n, 2048), dtype=np.uint32)
    out[:, 1] = g.integers(0, 249561089, size=(n, 2048), dtype=np.uint32)
    return out


def rand_raw(g: np.random.Generator, n: int, include_q: bool = False) -> np.ndarray:
    """n raw Polys with coefficients in [0, Q) (optionally a few equal to Q)."""
    Q = 268369921 * 249561089
    x = g.integers(0, Q, size=(n, 2048), dtype=np.uint64)
    if include_q:
        x[:, ::97] = Q
    out = np.empty((n, 2, 2048), np.uint32)
    out[:, 0] = (x >> np.uint64(32)).astype(np.uint32)
    out[:, 1] = (x & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    return out


def live_elements(oracle, nu1: int, nu2: int, t_gsw: int) -> np.ndarray:
    """Mask of the expanded elements a whole answer forms: every element but the odd
    i >= 2^(stop_round + 1), the v_neg1 products of elements skipped past stop_round
    (gpu-pir/expansion.cuh:92-97), which no step, selector or GSW slot reads."""
    g, st = oracle.expansion_g(nu1, nu2, t_gsw), oracle.expansion_stop_round(nu2, t_gsw)
    i = np.arange(1 << g)
    dead = (i % 2 == 1) & (i >= (2 << st))
    assert np.all((i[dead] >> 1) >= nu2 * t_gsw)  # no GSW slot among them
    return ~dead
