"""Reference operator surface on device tensors.

Each function wraps one reference kernel (names kept): the arguments are
torch tensors on the GPU holding u32 data (dtype int32 or uint32, same bits)
in the reference's struct layouts, natural Spiral NTT order, standard form.
Work runs on torch's current stream.  There is no CPU path: these raise if
the HIP library is missing or a tensor is not on the GPU.
"""
from __future__ import annotations

import ctypes

from . import _lib
from .params import T_CONV, T_GSW, T_EXP_LEFT, T_EXP_RIGHT
from .server import Context


def _dev(t, name: str):
    if not t.is_cuda:
        raise ValueError(f"{name} must be a GPU tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.element_size() not in (4, 8):
        raise ValueError(f"{name} must hold 32- or 64-bit words")
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def NTT_R(ctx: Context, polys, residue: int) -> None:
    """In-place forward NTT of polys (n x 2048) mod MODULI[residue] (ntt_28bit.cuh:251)."""
    _lib.call("sp_ntt_fwd", ctx.handle, _dev(polys, "polys"), polys.numel() // 2048, residue, _stream())


def INTT_R(ctx: Context, polys, residue: int, scale: bool = False) -> None:
    """In-place inverse NTT (ntt_28bit.cuh:292); scale=False leaves out N^-1 as the reference does."""
    _lib.call("sp_ntt_inv", ctx.handle, _dev(polys, "polys"), polys.numel() // 2048, residue, int(scale), _stream())


def from_ntt(ctx: Context, ntt_polys, raw_out) -> None:
    """NTT Polys -> raw Polys (gadget.cuh:8-24)."""
    _lib.call("sp_from_ntt", ctx.handle, _dev(ntt_polys, "ntt"), _dev(raw_out, "raw"), ntt_polys.numel() // 4096,
              _stream())


def to_ntt(ctx: Context, raw_polys, ntt_out) -> None:
    """raw Polys -> NTT Polys (classes.cuh:140-173)."""
    _lib.call("sp_to_ntt", ctx.handle, _dev(raw_polys, "raw"), _dev(ntt_out, "ntt"), raw_polys.numel() // 4096,
              _stream())


def automorph(ctx: Context, poly_coeffs, dest_poly_coeffs, t: int) -> None:
    """x -> x^t on u64 coefficients (automorph.cuh:54-62); tensors of int64/uint64."""
    if poly_coeffs.element_size() != 8:
        raise ValueError("automorph works on 64-bit coefficients")
    _lib.call("sp_automorph", ctx.handle, _dev(poly_coeffs, "in"), _dev(dest_poly_coeffs, "out"),
              poly_coeffs.numel() // 2048, t, _stream())


def inv_gadget_ntt_mult(ctx: Context, inp, keys, acc, rdim: int, mx: int, n_items: int = 1,
                        in_stride: int = 0, key_stride: int = 0, acc_stride: int = 0) -> None:
    """acc += M * NTT(g^-1(inp)) (gadget.cuh:84-101, rdim 1 or 2)."""
    _lib.call("sp_gadget_extprod", ctx.handle, _dev(inp, "inp"), rdim, mx, _dev(keys, "keys"), _dev(acc, "acc"),
              n_items, in_stride, key_stride, acc_stride, _stream())


def db_mul(ctx: Context, db, first_dim, folding_inputs, nu1: int, nu2: int) -> None:
    """db_mul_kernel (db_mul.cuh:26-70) on the reference Database/FirstDim/FoldingInputs layouts."""
    _lib.call("sp_db_mul", ctx.handle, _dev(db, "db"), _dev(first_dim, "first_dim"),
              _dev(folding_inputs, "folding_inputs"), nu1, nu2, _stream())


def FOLD_CIPHERTEXTS(ctx: Context, accumulators, v_folding, v_folding_neg=None, nu2: int = 2,
                     t_gsw: int = T_GSW) -> None:
    """fold_all_ciphertexts (folding.cuh:45-65), in place; result in accumulators[0].

    v_folding_neg=None takes G - v_folding (the server-derived form)."""
    neg = ctypes.c_void_p(0) if v_folding_neg is None else _dev(v_folding_neg, "v_folding_neg")
    _lib.call("sp_fold", ctx.handle, _dev(accumulators, "accumulators"), nu2, t_gsw, _dev(v_folding, "v_folding"),
              neg, _stream())


def coefficient_expansion(ctx: Context, v, left, right, nu1: int, nu2: int, t_gsw: int = T_GSW,
                          t_exp: int = T_EXP_LEFT, rounds: int = -1) -> None:
    """All rounds of coefficient_expansion_r (expansion.cuh:222-260), in place on v."""
    assert T_EXP_LEFT == T_EXP_RIGHT
    _lib.call("sp_expand", ctx.handle, _dev(v, "v"), nu1, nu2, t_gsw, t_exp, _dev(left, "left"),
              _dev(right, "right"), rounds, _stream())


def regev_to_gsw(ctx: Context, v, conversion, v_folding_out, nu2: int, t_gsw: int = T_GSW) -> None:
    """Regev -> GSW conversion of the expanded query v (NTT Ciphertexts, as
    coefficient_expansion leaves them) with the conversion key
    CiphertextHalf[2][2 T_CONV] (PublicParameters.conversion, structures.cuh:176)
    into v_folding_out CiphertextHalf[nu2][2][2 t_gsw]; the reference leaves
    regev_to_gsw a stub (cpu_version/conversion.py:3)."""
    _lib.call("sp_regev_to_gsw", ctx.handle, _dev(v, "v"), v.numel() // (4 * 2048), nu2, t_gsw, T_CONV,
              _dev(conversion, "conversion"), _dev(v_folding_out, "v_folding_out"), _stream())


def pack_response(ctx: Context, raw_cts, packing_keys, out, q2_bits: int) -> None:
    """Response packing + modulus switching of n^2 = 4 raw Ciphertexts (contiguous)
    with packing keys u32 [n][2][T_CONV][n+1][N] into the compact response `out`
    (uint8 tensor of sp_packed_response_bytes bytes).  Not in the reference."""
    if not out.is_cuda or not out.is_contiguous():
        raise ValueError("out must be a contiguous GPU tensor")
    _lib.call("sp_pack_response", ctx.handle, _dev(raw_cts, "raw_cts"), raw_cts.numel() // (4 * 2048),
              _dev(packing_keys, "packing_keys"), T_CONV, q2_bits, ctypes.c_void_p(out.data_ptr()), _stream())
