// sp_device.hpp — device-side building blocks for the Spiral PIR answer path
// on CDNA4 (gfx950): modular arithmetic on the two 28-bit CRT primes, the
// wave-level 2048-point negacyclic NTT/INTT, and the library's internal
// NTT-domain layout.
//
// Mathematics follows EricArcherman/GPU-PIR (gpu-pir/ntt_28bit.cuh,
// arith.cuh, gadget.cuh); the schedule is this library's own:
//
//  * One wavefront (64 lanes) owns one 2048-point transform; each lane holds
//    32 coefficients in VGPRs.  No workgroup barrier is needed inside a
//    transform: the single index transpose goes through a wave-private LDS
//    tile (rows of 66 words keep both access patterns free of bank conflicts
//    with one base address per lane, see tile_at) and the one cross-lane
//    butterfly stage exchanges register halves between lanes l and l + 32
//    with v_permlane32_swap (one instruction per register pair, no selects).
//  * Butterflies are Harvey's lazy Cooley-Tukey / Gentleman-Sande forms with
//    Shoup multiplication (w, w' = floor(w * 2^32 / q)): 3 integer multiplies
//    per butterfly.  Forward values live in [0, 4q), inverse in [0, 2q).
//
// Index bookkeeping (N = 2048, index bits b10..b0):
//   phase A   lane l, register i  <->  coefficient 64*i + l
//   phase B   lane l = L + 32*b, register i <-> index 64*L + 2*i + b
//   NTT side  lane l, register r  <->  NTT position 32*vlane(l) + r,
//             vlane(l) = 2*(l mod 32) + l / 32
// Internal NTT-domain layout "P" stores position z = 32*v + 4*c + e at word
// 256*c + 4*v + e, so that chunk c of a lane's 32 positions is one coalesced
// 16-byte load across the wave (lane l takes v = vlane(l)).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <utility>

namespace sp {

constexpr int N = 2048;
constexpr int LOGN = 11;
constexpr uint32_t Q0 = 268369921u;  // gpu-pir/params.cuh:25
constexpr uint32_t Q1 = 249561089u;  // gpu-pir/params.cuh:28
constexpr uint64_t QQ = (uint64_t)Q0 * (uint64_t)Q1;
constexpr int LDS_TILE = 2112;  // words of wave-private transpose scratch

// Twiddle table per (residue, direction), u32 words:
//   [0, 64)          phase-A twiddles: (w, w') for index 1..31, uniform per wave
//   [64, 64+47*128)  phase-B slots: slot s -> w at 64 + (2s)*64 + lane,
//                    w' at 64 + (2s+1)*64 + lane
// w' is the Shoup companion floor(w 2^32 / q); the forward tables hold
// -w mod 2^32 in place of w (the forward butterflies' multiply-add form).
constexpr int TW_A_WORDS = 64;
constexpr int TW_B_SLOTS = 47;
constexpr int TW_WORDS = TW_A_WORDS + TW_B_SLOTS * 2 * 64;  // 6080

// Per-residue constants (device copy lives in the context's constant block).
struct ModConst {
  uint32_t q;
  uint32_t qinv_neg;    // -q^-1 mod 2^32  (Montgomery REDC)
  uint32_t r2;          // 2^64 mod q
  uint32_t crt_c;       // N^-1 * (q_other^-1 mod q) mod q
  uint32_t crt_cp;      // Shoup companion of crt_c
  uint32_t ninv;        // N^-1 mod q
  uint32_t ninv_p;      // Shoup companion
  uint32_t two32;       // 2^32 mod q
  uint32_t two32_p;     // Shoup companion of two32
  uint32_t mu1;         // floor(2^32 / q)  (Shoup companion of 1)
  uint32_t crt_c32;     // crt_c * 2^32 mod q (CRT constant for Montgomery-form inputs)
  uint32_t crt_c32p;    // Shoup companion of crt_c32
};

__device__ __forceinline__ uint32_t shoup_mul(uint32_t y, uint32_t w, uint32_t wp, uint32_t q) {
  // y * w mod q in [0, 2q) for any 32-bit y and w < q (Shoup / Harvey).
  const uint32_t hi = __umulhi(y, wp);
  return y * w - hi * q;
}

__device__ __forceinline__ uint32_t reduce_2q(uint32_t x, uint32_t two_q) {
  // [0, 4q) -> [0, 2q): min(x, x - 2q) as unsigned.
  return min(x, x - two_q);
}

__device__ __forceinline__ uint32_t canon(uint32_t x, uint32_t q) {
  // [0, 2q) -> [0, q)
  return min(x, x - q);
}

__device__ __forceinline__ uint32_t canon4(uint32_t x, uint32_t q) {
  // [0, 4q) -> [0, q)
  x = min(x, x - 2 * q);
  return min(x, x - q);
}

// Forward butterflies take the twiddle negated: nw = -w mod 2^32 (the
// forward tables hold nw in place of w, with the Shoup companion wp of w).
// Then nt = y * nw + floor(y * wp / 2^32) * q is -t mod 2^32 for the Shoup
// product t = y * w mod q in [0, 2q), which costs one multiply and one
// multiply-add (v_mad_u64_u32), and y' = x - t + 2q is one three-input add:
// 5 VALU instructions for a lazy butterfly instead of 7.
__device__ __forceinline__ uint32_t shoup_mul_neg(uint32_t y, uint32_t nw, uint32_t wp, uint32_t q) {
  return y * nw + __umulhi(y, wp) * q;
}

__device__ __forceinline__ void ct_bf(uint32_t &x, uint32_t &y, uint32_t nw, uint32_t wp, uint32_t q) {
  // Harvey forward butterfly: inputs [0, 4q), outputs [0, 4q).
  const uint32_t tq = 2 * q;
  const uint32_t xr = reduce_2q(x, tq);
  const uint32_t nt = shoup_mul_neg(y, nw, wp, q);
  y = xr + tq + nt;
  x = xr - nt;
}

// Forward butterfly without the reduction of x: x' = x + t, y' = x - t + 2q
// with t = y * w mod q in [0, 2q).  If x, y < B q then x', y' < (B + 2) q.
__device__ __forceinline__ void ct_bf_lazy(uint32_t &x, uint32_t &y, uint32_t nw, uint32_t wp, uint32_t q) {
  const uint32_t nt = shoup_mul_neg(y, nw, wp, q);
  y = x + 2 * q + nt;
  x = x - nt;
}

// Forward butterfly for any x < 2^32: x is first brought to [0, 2q) by a
// Shoup multiplication by 1 (mu1 = floor(2^32 / q)); outputs in [0, 4q).
__device__ __forceinline__ void ct_bf_full(uint32_t &x, uint32_t &y, uint32_t nw, uint32_t wp, uint32_t q,
                                           uint32_t mu1) {
  const uint32_t xr = x - __umulhi(x, mu1) * q;
  const uint32_t nt = shoup_mul_neg(y, nw, wp, q);
  y = xr + 2 * q + nt;
  x = xr - nt;
}

// Stage st (1..11) of a forward transform whose first LAZY stages skip the
// reduction of x; stage LAZY + 1 reduces x fully; later stages are lazy again
// when TAIL (values then grow from 4q by 2q per stage), else Harvey's.
template <int LAZY, bool TAIL>
__device__ __forceinline__ void ct_stage_bf(int st, uint32_t &x, uint32_t &y, uint32_t w, uint32_t wp, uint32_t q,
                                            uint32_t mu1) {
  if (st <= LAZY)
    ct_bf_lazy(x, y, w, wp, q);
  else if (LAZY > 0 && st == LAZY + 1)
    ct_bf_full(x, y, w, wp, q, mu1);
  else if (TAIL)
    ct_bf_lazy(x, y, w, wp, q);
  else
    ct_bf(x, y, w, wp, q);
}

// -q mod 2^32, kept opaque so that the compiler does not fold y + h * (-q)
// back into y - h * q (a multiply and a subtract instead of one multiply-add).
__device__ __forceinline__ uint32_t neg_q(uint32_t q) {
  uint32_t nq = 0u - q;
  asm("" : "+v"(nq));
  return nq;
}

// Harvey inverse butterfly: inputs [0, 2q), outputs [0, 2q).
__device__ __forceinline__ void gs_bf(uint32_t &x, uint32_t &y, uint32_t w, uint32_t wp, uint32_t q) {
  const uint32_t tq = 2 * q;
  const uint32_t u = reduce_2q(x + y, tq);
  const uint32_t v = x - y + tq;
  y = shoup_mul(v, w, wp, q);
  x = u;
}

// Montgomery reduction of a 64-bit accumulator: returns acc * 2^-32 mod q,
// canonical.  Valid for acc < 2^62 (result < 2^30 + q < 5q before the final
// corrections).
__device__ __forceinline__ uint32_t redc64(uint64_t acc, uint32_t q, uint32_t qinv_neg) {
  const uint32_t m = (uint32_t)acc * qinv_neg;
  const uint64_t t = (acc + (uint64_t)m * q) >> 32;
  uint32_t r = (uint32_t)t;
  r = min(r, r - 4 * q);
  r = min(r, r - 2 * q);
  return min(r, r - q);
}

__device__ __forceinline__ void wave_lds_sync() {
  // Orders this wave's LDS traffic (LDS instructions of one wave execute in
  // issue order; the fences stop the compiler from moving accesses across).
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The NTT-side lane index: lane l of a wave transform holds NTT positions
// 32 vlane(l) .. 32 vlane(l) + 31.
__device__ __forceinline__ int vlane(int lane) { return 2 * (lane & 31) + (lane >> 5); }

// x of lanes 0-31 and y of lanes 32-63 stay; lanes 32-63 of x receive y of
// lanes 0-31 and lanes 0-31 of y receive x of lanes 32-63 (v_permlane32_swap).
__device__ __forceinline__ void swap_halves(uint32_t &x, uint32_t &y) {
  const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
  x = r[0];
  y = r[1];
}

// Transpose tile slot of coefficient 64*i + l (phase-A register i, lane l):
// rows of 66 words.  A phase-A access (fixed i: 66 i + l) and a phase-B access
// (lane L + 32b reading row L, column 2i + b: 66 L + 2 i + b) both touch 64
// distinct banks (2 L + b distinct mod 64; a 65-word row gives L + b, and the
// counters then show 1.3 conflict cycles per LDS instruction), and both are
// one per-lane base plus a constant offset per register.
__device__ __forceinline__ int tile_at(int i, int l) { return 66 * i + l; }

__device__ __forceinline__ uint32_t tw_a(const uint32_t *__restrict__ tw, int idx, int k) {
  return tw[idx * 2 + k];
}
__device__ __forceinline__ uint32_t tw_b(const uint32_t *__restrict__ tw, int slot, int k, int lane) {
  return tw[TW_A_WORDS + (slot * 2 + k) * 64 + lane];
}

// Forward negacyclic NTT of one polynomial held by one wave.
// In:  r[i] = coefficient 64*i + lane, values in [0, 4q).
// Out: r[j] = NTT position 32*lane + j (natural Spiral order), in [0, 4q).
// LAZY > 0: the first LAZY stages skip the reduction of x, which is valid
// while the values stay below 2^32: inputs below B q grow to (B + 2 LAZY) q,
// and both moduli satisfy 16 q + 128 < 2^32 (so 8 stages for digits below
// 128, 7 for inputs below 2q).  Stage LAZY + 1 reduces x fully (outputs
// < 4q); with TAIL the remaining 11 - LAZY - 1 stages are lazy too and the
// outputs are below (4 + 2 (10 - LAZY)) q, else Harvey's bound 4q holds.
// mu1 = floor(2^32 / q).
//
// HALF: the two half-branches of a "virtual" array.  After stage 1 (stride
// 1024) the outputs of positions [0, 1024) (branch 0: x + t) and [1024, 2048)
// (branch 1: x - t) depend only on their own half of the array, through the
// same butterfly network with the twiddles of that half.  With HALF the
// caller has done stage 1 of TWO polynomials A and B for one branch h, A's
// branch in r[0..15] (indices 64 i + lane) and B's in r[16..31]; stages
// 2-11 then run with branch h's twiddles in both halves (phase A: block blk
// of stage s takes the twiddle of block (blk mod 2^(s-1)) + h 2^(s-1); phase
// B: lane l takes the per-lane twiddles of lane l with bit 4 set to h, the
// lane whose L = l mod 32 lies in half h).
// Out: lane l = L + 32 b holds virtual positions 64 L + 32 b + j: A's
// positions 1024 h + 64 L + 32 b + j for L < 16, B's 1024 h + 64 (L - 16) +
// 32 b + j for L >= 16.  One wave thus transforms two polynomials' halves
// for the cost of one transform.
template <int LAZY, bool TAIL = false, bool HALF = false>
__device__ __forceinline__ void ntt_fwd_wave_t(uint32_t (&r)[32], const uint32_t *__restrict__ tw,
                                               uint32_t *__restrict__ lds, uint32_t q, uint32_t mu1, int lane,
                                               int h = 0) {
  static_assert(LAZY >= 0 && LAZY <= 8, "at most 8 lazy stages");
  static_assert(!TAIL || (LAZY > 0 && 4 + 2 * (10 - LAZY) <= 15), "lazy tail bound");
  static_assert(!HALF || LAZY > 0, "the caller's stage 1 is lazy");
  // Phase A: strides 1024 .. 64 (index bits 10..6 live in the register index).
#pragma unroll
  for (int s = HALF ? 1 : 0; s < 5; s++) {
    const int half = 16 >> s;
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++) {
      const int ti = HALF ? (1 << s) + (blk & ((1 << (s - 1)) - 1)) + (h << (s - 1)) : (1 << s) + blk;
      const uint32_t w = tw_a(tw, ti, 0), wp = tw_a(tw, ti, 1);
#pragma unroll
      for (int i = 0; i < half; i++)
        ct_stage_bf<LAZY, TAIL>(1 + s, r[blk * 2 * half + i], r[blk * 2 * half + i + half], w, wp, q, mu1);
    }
  }
  // Transpose to phase B through the wave's LDS tile.
#pragma unroll
  for (int i = 0; i < 32; i++) lds[tile_at(i, lane)] = r[i];
  wave_lds_sync();
  const int L = lane & 31, b = lane >> 5;
#pragma unroll
  for (int i = 0; i < 32; i++) r[i] = lds[tile_at(L, 2 * i + b)];
  wave_lds_sync();
  // Phase B: strides 32 .. 2 in registers.
  const int tl = HALF ? ((lane & ~16) | (h << 4)) : lane;  // lane whose per-lane twiddles apply
  int slot = 0;
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const int half = 16 >> s;
#pragma unroll
    for (int blk = 0; blk < (1 << s); blk++) {
      const uint32_t w = tw_b(tw, slot + blk, 0, tl), wp = tw_b(tw, slot + blk, 1, tl);
#pragma unroll
      for (int i = 0; i < half; i++)
        ct_stage_bf<LAZY, TAIL>(6 + s, r[blk * 2 * half + i], r[blk * 2 * half + i + half], w, wp, q, mu1);
    }
    slot += (1 << s);
  }
  // Stride 1 pairs lanes L and L + 32: after the swap of r[k] (lanes 32-63)
  // with r[16 + k] (lanes 0-31), lane L + 32 b holds the pair of index
  // 64 L + 2 (16 b + k) in (r[k], r[16 + k]) and ends holding the 32
  // consecutive outputs 64 L + 32 b + j.
  uint32_t o[32];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t x = r[k], y = r[16 + k];
    swap_halves(x, y);
    ct_stage_bf<LAZY, TAIL>(11, x, y, tw_b(tw, 31 + k, 0, tl), tw_b(tw, 31 + k, 1, tl), q, mu1);
    o[2 * k] = x;
    o[2 * k + 1] = y;
  }
#pragma unroll
  for (int j = 0; j < 32; j++) r[j] = o[j];
}

__device__ __forceinline__ void ntt_fwd_wave(uint32_t (&r)[32], const uint32_t *__restrict__ tw,
                                             uint32_t *__restrict__ lds, uint32_t q, int lane) {
  ntt_fwd_wave_t<0>(r, tw, lds, q, 0u, lane);
}

// Stage 1 of two byte-packed digit polynomials A and B (byte b of word i4 is
// coefficient 64 (4 i4 + b) + lane, values < 128), kept for branch H in the
// form ntt_fwd_wave_t<LAZY, TAIL, true> takes: r[i] = A's branch value at
// index 64 i + lane, r[16 + i] = B's (lazy: x + t for H = 0, x - t + 2q for
// H = 1, with t = y w mod q in [0, 2q) for y the coefficient 1024 further on;
// outputs < 128 + 2q).
template <int H>
__device__ __forceinline__ void digits_stage1_half(uint32_t (&r)[32], const uint32_t (&da)[8], const uint32_t (&db)[8],
                                                   const uint32_t *__restrict__ tw, uint32_t q) {
  const uint32_t nw = tw_a(tw, 1, 0), wp = tw_a(tw, 1, 1);
#pragma unroll
  for (int p = 0; p < 2; p++) {
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t wlo = p ? db[i >> 2] : da[i >> 2], whi = p ? db[4 + (i >> 2)] : da[4 + (i >> 2)];
      const uint32_t x = (wlo >> (8 * (i & 3))) & 0xFFu, y = (whi >> (8 * (i & 3))) & 0xFFu;
      const uint32_t nt = shoup_mul_neg(y, nw, wp, q);
      r[16 * p + i] = H ? x + 2 * q + nt : x - nt;
    }
  }
}

// Inverse negacyclic NTT (without the factor N^-1) of one polynomial.
// In:  r[j] = NTT position 32*lane + j, values in [0, 2q).
// Out: r[i] = N * coefficient (64*i + lane) mod q, values in [0, 2q).
__device__ __forceinline__ void ntt_inv_wave(uint32_t (&r)[32], const uint32_t *__restrict__ tw,
                                             uint32_t *__restrict__ lds, uint32_t q, int lane) {
  const int L = lane & 31, b = lane >> 5;
#pragma unroll
  for (int k = 0; k < 16; k++)
    gs_bf(r[2 * k], r[2 * k + 1], tw_b(tw, 31 + k, 0, lane), tw_b(tw, 31 + k, 1, lane), q);
  // to phase B: lane L + 32 b' takes index 64 L + 2 i + b' in n[i]
  uint32_t n[32];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t x = r[2 * k], y = r[2 * k + 1];
    swap_halves(x, y);
    n[k] = x;
    n[16 + k] = y;
  }
  // Strides 2 .. 32 in registers.
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const int half = 1 << s;
    const int nblk = 16 >> s;
#pragma unroll
    for (int blk = 0; blk < nblk; blk++) {
      const uint32_t w = tw_b(tw, nblk - 1 + blk, 0, lane), wp = tw_b(tw, nblk - 1 + blk, 1, lane);
#pragma unroll
      for (int i = 0; i < half; i++) gs_bf(n[blk * 2 * half + i], n[blk * 2 * half + i + half], w, wp, q);
    }
  }
#pragma unroll
  for (int i = 0; i < 32; i++) lds[tile_at(L, 2 * i + b)] = n[i];
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < 32; i++) r[i] = lds[tile_at(i, lane)];
  wave_lds_sync();
  // Phase A: strides 64 .. 1024.
#pragma unroll
  for (int s = 0; s < 5; s++) {
    const int half = 1 << s;
    const int nblk = 16 >> s;
#pragma unroll
    for (int blk = 0; blk < nblk; blk++) {
      const uint32_t w = tw_a(tw, nblk + blk, 0), wp = tw_a(tw, nblk + blk, 1);
#pragma unroll
      for (int i = 0; i < half; i++) gs_bf(r[blk * 2 * half + i], r[blk * 2 * half + i + half], w, wp, q);
    }
  }
}

template <int... I, class F>
__device__ __forceinline__ void static_for_impl(std::integer_sequence<int, I...>, F &&f) {
  (f(std::integral_constant<int, I>{}), ...);
}
// f(integral_constant<int, 0>), ..., f(integral_constant<int, n - 1>).
template <int n, class F>
__device__ __forceinline__ void static_for(F &&f) {
  static_for_impl(std::make_integer_sequence<int, n>{}, f);
}

// Inverse transform with lazy sums, for callers that multiply every output by
// a Shoup constant next (any 32-bit input): inputs in [0, 2q), outputs < 8q.
// Stage g (1..11) takes inputs below B = 2q, 4q or 8q (g mod 3 = 1, 2, 0):
// v = x - y + B and the Shoup product (multiply-add form) is below 2q; the sum
// x + y (below 2B) is reduced by a Barrett step only when g is a multiple of 3,
// so two stages in three skip the reduction (2 of their 8 instructions).
// MAD = false: the products as plain multiplies and subtracts (no 64-bit
// multiply-add destinations: for kernels whose register allocation they upset).
template <bool REDUCE, bool MAD>
__device__ __forceinline__ void gs_bf_lazy(uint32_t &x, uint32_t &y, uint32_t w, uint32_t wp, uint32_t q,
                                           uint32_t nq, uint32_t mu1, uint32_t B) {
  const uint32_t s = x + y;
  const uint32_t v = x - y + B;
  if (MAD) {
    y = v * w + __umulhi(v, wp) * nq;
    x = REDUCE ? s + __umulhi(s, mu1) * nq : s;
  } else {
    y = shoup_mul(v, w, wp, q);
    x = REDUCE ? s - __umulhi(s, mu1) * q : s;
  }
}
template <int G, bool MAD>
__device__ __forceinline__ void gs_stage_lazy(uint32_t &x, uint32_t &y, uint32_t w, uint32_t wp, uint32_t q,
                                              uint32_t nq, uint32_t mu1) {
  constexpr int PH = G % 3;  // 1: inputs < 2q, 2: < 4q, 0: < 8q (reduce)
  gs_bf_lazy<PH == 0, MAD>(x, y, w, wp, q, nq, mu1, (PH == 1 ? 2u : PH == 2 ? 4u : 8u) * q);
}
template <bool MAD = true>
__device__ __forceinline__ void ntt_inv_wave_lazy(uint32_t (&r)[32], const uint32_t *__restrict__ tw,
                                                  uint32_t *__restrict__ lds, uint32_t q, uint32_t mu1, int lane) {
  const int L = lane & 31, b = lane >> 5;
  const uint32_t nq = MAD ? neg_q(q) : 0u;
#pragma unroll
  for (int k = 0; k < 16; k++)
    gs_stage_lazy<1, MAD>(r[2 * k], r[2 * k + 1], tw_b(tw, 31 + k, 0, lane), tw_b(tw, 31 + k, 1, lane), q, nq, mu1);
  uint32_t n[32];
#pragma unroll
  for (int k = 0; k < 16; k++) {
    uint32_t x = r[2 * k], y = r[2 * k + 1];
    swap_halves(x, y);
    n[k] = x;
    n[16 + k] = y;
  }
  // Strides 2 .. 32 in registers (stages 2-6).
  static_for<5>([&](auto ss) {
    constexpr int s = decltype(ss)::value;
    constexpr int half = 1 << s, nblk = 16 >> s;
#pragma unroll
    for (int blk = 0; blk < nblk; blk++) {
      const uint32_t w = tw_b(tw, nblk - 1 + blk, 0, lane), wp = tw_b(tw, nblk - 1 + blk, 1, lane);
#pragma unroll
      for (int i = 0; i < half; i++)
        gs_stage_lazy<2 + s, MAD>(n[blk * 2 * half + i], n[blk * 2 * half + i + half], w, wp, q, nq, mu1);
    }
  });
#pragma unroll
  for (int i = 0; i < 32; i++) lds[tile_at(L, 2 * i + b)] = n[i];
  wave_lds_sync();
#pragma unroll
  for (int i = 0; i < 32; i++) r[i] = lds[tile_at(i, lane)];
  wave_lds_sync();
  // Phase A: strides 64 .. 1024 (stages 7-11).
  static_for<5>([&](auto ss) {
    constexpr int s = decltype(ss)::value;
    constexpr int half = 1 << s, nblk = 16 >> s;
#pragma unroll
    for (int blk = 0; blk < nblk; blk++) {
      const uint32_t w = tw_a(tw, nblk + blk, 0), wp = tw_a(tw, nblk + blk, 1);
#pragma unroll
      for (int i = 0; i < half; i++)
        gs_stage_lazy<7 + s, MAD>(r[blk * 2 * half + i], r[blk * 2 * half + i + half], w, wp, q, nq, mu1);
    }
  });
}

// ---- internal NTT-domain layout P --------------------------------------
__device__ __forceinline__ void load_p(uint32_t (&r)[32], const uint32_t *__restrict__ base, int lane) {
  const uint4 *b4 = reinterpret_cast<const uint4 *>(base);
  const int vl = vlane(lane);
#pragma unroll
  for (int c = 0; c < 8; c++) {
    const uint4 v = b4[c * 64 + vl];
    r[4 * c] = v.x; r[4 * c + 1] = v.y; r[4 * c + 2] = v.z; r[4 * c + 3] = v.w;
  }
}
__device__ __forceinline__ void store_p(uint32_t *__restrict__ base, const uint32_t (&r)[32], int lane) {
  uint4 *b4 = reinterpret_cast<uint4 *>(base);
  const int vl = vlane(lane);
#pragma unroll
  for (int c = 0; c < 8; c++) b4[c * 64 + vl] = make_uint4(r[4 * c], r[4 * c + 1], r[4 * c + 2], r[4 * c + 3]);
}
__host__ __device__ __forceinline__ int p_of_z(int z) {
  // natural NTT position -> internal word
  const int l = z >> 5, c = (z >> 2) & 7, e = z & 3;
  return 256 * c + 4 * l + e;
}
__host__ __device__ __forceinline__ int z_of_p(int p) {
  const int c = p >> 8, l = (p >> 2) & 63, e = p & 3;
  return 32 * l + 4 * c + e;
}

// phase-A coefficient load/store (natural coefficient order, coalesced).
__device__ __forceinline__ void load_coeffs(uint32_t (&r)[32], const uint32_t *__restrict__ base, int lane) {
#pragma unroll
  for (int i = 0; i < 32; i++) r[i] = base[64 * i + lane];
}

}  // namespace sp
