// sp_kernels.hip — CDNA4 kernels of the Spiral PIR answer path.
//
// All kernels compute exact modular integer arithmetic; every output is a
// canonical residue (or a canonical raw value in [0, Q]) so results compare
// bit for bit with the reference semantics.  See sp_device.hpp for the NTT
// schedule and the internal NTT-domain layout P.
#include <cmath>
#include <utility>

#include "sp_kernels.hpp"

namespace sp {

// ---------------------------------------------------------------------------
// Boundary transforms (natural order in and out)
// ---------------------------------------------------------------------------
template <bool INV>
__global__ __launch_bounds__(256) void k_ntt_natural(uint32_t *__restrict__ polys, size_t n,
                                                     const uint32_t *__restrict__ tw, ModConst mc,
                                                     int scale) {
  __shared__ uint32_t lds_all[4 * LDS_TILE];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  uint32_t *lds = lds_all + wave * LDS_TILE;
  const uint32_t q = mc.q;
  for (size_t pi = (size_t)blockIdx.x * 4 + wave; pi < n; pi += (size_t)gridDim.x * 4) {
    uint32_t *p = polys + pi * N;
    uint32_t r[32];
    if (!INV) {
      load_coeffs(r, p, lane);
      ntt_fwd_wave(r, tw, lds, q, lane);
      uint4 *o = reinterpret_cast<uint4 *>(p + 32 * vlane(lane));
#pragma unroll
      for (int c = 0; c < 8; c++)
        o[c] = make_uint4(canon4(r[4 * c], q), canon4(r[4 * c + 1], q), canon4(r[4 * c + 2], q),
                          canon4(r[4 * c + 3], q));
    } else {
      const uint4 *in = reinterpret_cast<const uint4 *>(p + 32 * vlane(lane));
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const uint4 v = in[c];
        r[4 * c] = v.x; r[4 * c + 1] = v.y; r[4 * c + 2] = v.z; r[4 * c + 3] = v.w;
      }
      ntt_inv_wave(r, tw, lds, q, lane);
#pragma unroll
      for (int i = 0; i < 32; i++) {
        uint32_t x = r[i];
        if (scale) x = shoup_mul(x, mc.ninv, mc.ninv_p, q);
        p[64 * i + lane] = canon(x, q);
      }
    }
  }
}

// NTT Poly (natural positions, or layout P) -> raw Poly.  One poly per
// 128-thread block; wave m inverts residue m, the CRT pass is split.
template <bool LP>
__global__ __launch_bounds__(128) void k_from_ntt(const uint32_t *__restrict__ ntt,
                                                  uint32_t *__restrict__ raw, size_t n,
                                                  size_t in_stride, size_t out_stride,
                                                  const uint32_t *__restrict__ tw_base, ModConst mc0,
                                                  ModConst mc1) {
  __shared__ uint32_t lds_all[2 * LDS_TILE];
  __shared__ uint32_t xm[2][N];
  const int m = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  uint32_t *lds = lds_all + m * LDS_TILE;
  const ModConst mc = m ? mc1 : mc0;
  const uint32_t *tw = tw_base + (m * 2 + 1) * TW_WORDS;
  for (size_t pi = blockIdx.x; pi < n; pi += gridDim.x) {
    const uint32_t *src = ntt + pi * in_stride + m * N;
    uint32_t r[32];
    if (LP) {
      load_p(r, src, lane);
    } else {
      const uint4 *in = reinterpret_cast<const uint4 *>(src + 32 * vlane(lane));
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const uint4 v = in[c];
        r[4 * c] = v.x; r[4 * c + 1] = v.y; r[4 * c + 2] = v.z; r[4 * c + 3] = v.w;
      }
    }
    ntt_inv_wave_lazy<false>(r, tw, lds, mc.q, mc.mu1, lane);
#pragma unroll
    for (int i = 0; i < 32; i++) xm[m][64 * i + lane] = canon(shoup_mul(r[i], mc.crt_c, mc.crt_cp, mc.q), mc.q);
    __syncthreads();
    uint32_t *dst = raw + pi * out_stride;
#pragma unroll
    for (int ii = 0; ii < 16; ii++) {
      const int z = 64 * (16 * m + ii) + lane;
      uint64_t x = (uint64_t)xm[0][z] * Q1 + (uint64_t)xm[1][z] * Q0;
      x = (x >= QQ) ? x - QQ : x;
      dst[z] = (uint32_t)(x >> 32);
      dst[N + z] = (uint32_t)x;
    }
    __syncthreads();
  }
}

// x (raw, given as hi/lo) mod q, canonical.  hi < 2^32, lo < 2^32.
__device__ __forceinline__ uint32_t raw_mod(uint32_t hi, uint32_t lo, const ModConst &mc) {
  const uint32_t t1 = shoup_mul(hi, mc.two32, mc.two32_p, mc.q);  // hi * 2^32 mod q, [0, 2q)
  const uint32_t t2 = lo - __umulhi(lo, mc.mu1) * mc.q;           // lo mod q, [0, 2q)
  return canon4(t1 + t2, mc.q);
}

// raw Poly -> NTT Poly (natural positions).
__global__ __launch_bounds__(128) void k_to_ntt(const uint32_t *__restrict__ raw, uint32_t *__restrict__ ntt,
                                                size_t n, const uint32_t *__restrict__ tw_base,
                                                ModConst mc0, ModConst mc1) {
  __shared__ uint32_t lds_all[2 * LDS_TILE];
  const int m = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  uint32_t *lds = lds_all + m * LDS_TILE;
  const ModConst mc = m ? mc1 : mc0;
  const uint32_t *tw = tw_base + (m * 2 + 0) * TW_WORDS;
  for (size_t pi = blockIdx.x; pi < n; pi += gridDim.x) {
    const uint32_t *src = raw + pi * 2 * N;
    uint32_t r[32];
#pragma unroll
    for (int i = 0; i < 32; i++) r[i] = raw_mod(src[64 * i + lane], src[N + 64 * i + lane], mc);
    __syncthreads();  // both waves read both planes before either overwrites one
    ntt_fwd_wave(r, tw, lds, mc.q, lane);
    uint4 *o = reinterpret_cast<uint4 *>(ntt + pi * 2 * N + m * N + 32 * vlane(lane));
#pragma unroll
    for (int c = 0; c < 8; c++)
      o[c] = make_uint4(canon4(r[4 * c], mc.q), canon4(r[4 * c + 1], mc.q), canon4(r[4 * c + 2], mc.q),
                        canon4(r[4 * c + 3], mc.q));
    __syncthreads();
  }
}

// automorph kernel (gpu-pir/automorph.cuh:54-62 semantics) on u64 coefficients.
__global__ __launch_bounds__(256) void k_automorph_u64(const uint64_t *__restrict__ in,
                                                       uint64_t *__restrict__ out, size_t n, uint32_t t) {
  const size_t total = n * N;
  for (size_t g = (size_t)blockIdx.x * 256 + threadIdx.x; g < total; g += (size_t)gridDim.x * 256) {
    const size_t poly = g / N;
    const uint64_t i = g % N;
    const uint64_t it = i * t;
    const uint64_t x = in[g];
    const uint64_t idx = it & (N - 1);
    out[poly * N + idx] = ((it >> LOGN) & 1) ? QQ - x : x;
  }
}

// ---------------------------------------------------------------------------
// Gadget external product (boundary form: natural keys in standard form)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t digit_of(uint32_t hi, uint32_t lo, int shift, uint32_t mask) {
  const uint64_t x = ((uint64_t)hi << 32) | lo;
  return shift >= 64 ? 0u : (uint32_t)((x >> shift) & mask);
}

__global__ __launch_bounds__(256) void k_extprod(const uint32_t *__restrict__ in, int rdim, int mx, int bits,
                                                 const uint32_t *__restrict__ keys, uint32_t *__restrict__ acc,
                                                 size_t n, size_t in_stride, size_t key_stride, size_t acc_stride,
                                                 const uint32_t *__restrict__ tw_base, ModConst mc0, ModConst mc1) {
  __shared__ uint32_t lds_all[4 * LDS_TILE];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int m = wave >> 1, h = wave & 1;
  uint32_t *lds = lds_all + wave * LDS_TILE;
  const ModConst mc = m ? mc1 : mc0;
  const uint32_t q = mc.q;
  const uint32_t *tw = tw_base + (m * 2 + 0) * TW_WORDS;
  const uint32_t mask = (bits >= 32) ? 0xFFFFFFFFu : ((1u << bits) - 1u);
  for (size_t it = blockIdx.x; it < n; it += gridDim.x) {
    uint64_t a0[32], a1[32];
#pragma unroll
    for (int j = 0; j < 32; j++) { a0[j] = 0; a1[j] = 0; }
    for (int col = h; col < mx; col += 2) {
      const int j = col % rdim, k = col / rdim;
      const uint32_t *hi = in + it * in_stride + (size_t)j * 2 * N;
      const uint32_t *lo = hi + N;
      uint32_t r[32];
#pragma unroll
      for (int i = 0; i < 32; i++) r[i] = digit_of(hi[64 * i + lane], lo[64 * i + lane], k * bits, mask);
      ntt_fwd_wave(r, tw, lds, q, lane);
      const uint32_t *k0 = keys + it * key_stride + ((size_t)(m * mx + col) * 2 + 0) * N + 32 * vlane(lane);
      const uint32_t *k1 = k0 + N;
#pragma unroll
      for (int jj = 0; jj < 32; jj++) {
        const uint32_t d = reduce_2q(r[jj], 2 * q);
        a0[jj] += (uint64_t)d * k0[jj];
        a1[jj] += (uint64_t)d * k1[jj];
      }
    }
    uint32_t own[32];
    __syncthreads();  // every wave is done with its transpose tile
#pragma unroll
    for (int jj = 0; jj < 32; jj++) {
      const uint32_t v0 = shoup_mul(redc64(a0[jj], q, mc.qinv_neg), mc.two32, mc.two32_p, q);
      const uint32_t v1 = shoup_mul(redc64(a1[jj], q, mc.qinv_neg), mc.two32, mc.two32_p, q);
      own[jj] = canon(h ? v1 : v0, q);
      lds_all[wave * N + 32 * lane + jj] = canon(h ? v0 : v1, q);
    }
    __syncthreads();
    uint32_t *dst = acc + it * acc_stride + (size_t)(h * 2 + m) * N + 32 * vlane(lane);
    const uint32_t *other = lds_all + (wave ^ 1) * N + 32 * lane;
#pragma unroll
    for (int jj = 0; jj < 32; jj++) {
      const uint32_t s = canon(own[jj] + other[jj], q);
      dst[jj] = canon(s + dst[jj], q);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Layout conversion natural <-> P (optionally scaling residues by 2^32)
// ---------------------------------------------------------------------------
// Polys are indexed k = 0..n-1; poly k is residue (k / m_div) & 1.  Any
// 32-bit input word is accepted: the output is canonical, x mod q (mont = 0)
// or x 2^32 mod q (mont = 1), so a non-canonical residue from a caller cannot
// break the lazy-reduction bounds of the kernels that read the result.
__global__ __launch_bounds__(256) void k_nat_to_p(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                  size_t n, int m_div, int mont, ModConst mc0, ModConst mc1) {
  const size_t total = n * N;
  for (size_t g = (size_t)blockIdx.x * 256 + threadIdx.x; g < total; g += (size_t)gridDim.x * 256) {
    const size_t k = g / N;
    const int p = (int)(g % N);
    const uint32_t x = in[k * N + z_of_p(p)];
    const ModConst &mc = ((k / m_div) & 1) ? mc1 : mc0;
    // Shoup multiplication by 2^32 mod q, or by 1 (companion floor(2^32 / q)): [0, 2q), then canonical
    out[g] = mont ? canon(shoup_mul(x, mc.two32, mc.two32_p, mc.q), mc.q) : canon(shoup_mul(x, 1u, mc.mu1, mc.q), mc.q);
  }
}

__global__ __launch_bounds__(256) void k_p_to_nat(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                  size_t n) {
  const size_t total = n * N;
  for (size_t g = (size_t)blockIdx.x * 256 + threadIdx.x; g < total; g += (size_t)gridDim.x * 256) {
    const size_t k = g / N;
    const int z = (int)(g % N);
    out[g] = in[k * N + p_of_z(z)];
  }
}

// P -> natural for 2^32-scaled residues: out = in * 2^-32 mod q (canonical).
__global__ __launch_bounds__(256) void k_p_to_nat_unscale(const uint32_t *__restrict__ in,
                                                          uint32_t *__restrict__ out, size_t n, int m_div,
                                                          ModConst mc0, ModConst mc1) {
  const size_t total = n * N;
  for (size_t g = (size_t)blockIdx.x * 256 + threadIdx.x; g < total; g += (size_t)gridDim.x * 256) {
    const size_t k = g / N;
    const int z = (int)(g % N);
    const ModConst &mc = ((k / m_div) & 1) ? mc1 : mc0;
    out[g] = redc64((uint64_t)in[k * N + p_of_z(z)], mc.q, mc.qinv_neg);
  }
}

// Reference Database layout [c][j][m][z] -> internal [m][j][c][p] (one trial),
// residues reduced mod q_m: the packed resident form holds 28 bits per
// residue, so a word >= 2^28 from a caller must not reach k_db_pack.
__global__ __launch_bounds__(256) void k_db_from_ref(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                     int R, int C, ModConst mc0, ModConst mc1) {
  const size_t total = (size_t)2 * R * C * N;
  for (size_t g = (size_t)blockIdx.x * 256 + threadIdx.x; g < total; g += (size_t)gridDim.x * 256) {
    const int p = (int)(g % N);
    size_t rest = g / N;
    const int c = (int)(rest % C); rest /= C;
    const int j = (int)(rest % R); rest /= R;
    const int m = (int)rest;
    const ModConst &mc = m ? mc1 : mc0;
    out[g] = canon(shoup_mul(in[(((size_t)c * R + j) * 2 + m) * N + z_of_p(p)], 1u, mc.mu1, mc.q), mc.q);
  }
}

// ---------------------------------------------------------------------------
// Database ingest: plaintext bytes -> NTT residues, u32 staging layout
//   out[t][m][j - j0][c][p], record index = j * C + c, j0 = rec0 / C (rec0 is a
// multiple of C, so a chunk holds whole rows).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_db_encode(const uint8_t *__restrict__ bytes, size_t rec0, size_t nrec,
                                                   uint32_t *__restrict__ db, int n2, int rows, int C,
                                                   const uint32_t *__restrict__ tw_base, ModConst mc0,
                                                   ModConst mc1) {
  __shared__ uint32_t lds_all[4 * LDS_TILE];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  uint32_t *lds = lds_all + wave * LDS_TILE;
  const size_t total = nrec * n2 * 2;
  for (size_t w = (size_t)blockIdx.x * 4 + wave; w < total; w += (size_t)gridDim.x * 4) {
    const int m = (int)(w & 1);
    const size_t pt = w >> 1;  // (record, trial) within the chunk
    const int t = (int)(pt % n2);
    const size_t rl = pt / n2;
    const size_t rec = rec0 + rl;
    const int j = (int)(rec / C - rec0 / C), c = (int)(rec % C);
    const uint8_t *src = bytes + pt * N;
    const ModConst mc = m ? mc1 : mc0;
    uint32_t r[32];
#pragma unroll
    for (int i = 0; i < 32; i++) r[i] = src[64 * i + lane];
    ntt_fwd_wave(r, tw_base + (m * 2) * TW_WORDS, lds, mc.q, lane);
#pragma unroll
    for (int i = 0; i < 32; i++) r[i] = canon4(r[i], mc.q);
    store_p(db + ((((size_t)t * 2 + m) * rows + j) * C + c) * N, r, lane);
  }
}

// ---------------------------------------------------------------------------
// Resident database format.  Residues are below 2^28, so the CG = min(C, 8)
// columns of one group (trial t, residue m, row j, position p) are packed into
// W = ceil(28 CG / 32) words: bits 28 c .. 28 c + 27 of the group hold column
// CG g + c.  Word k of all positions forms one plane:
//   db[t][m][j][g][k][P],   g < C / CG,  k < W.
// At C >= 8 the multiply streams 7/8 of the bytes of the plain u32 form.
// ---------------------------------------------------------------------------

template <int c, int W>
__device__ __forceinline__ uint32_t unpack28(const uint32_t (&w)[W]) {
  constexpr int pos = 28 * c, k = pos >> 5, sh = pos & 31;
  static_assert(k < W && (sh + 28 <= 32 || k + 1 < W), "column outside its group");
  if constexpr (sh + 28 <= 32)
    return (w[k] >> sh) & 0x0FFFFFFFu;
  else
    return __builtin_amdgcn_alignbit(w[k + 1], w[k], sh) & 0x0FFFFFFFu;
}

// Column c of the group whose word 0 is at w0 (planes `plane` words apart).
__device__ __forceinline__ uint32_t unpack28_at(const uint32_t *__restrict__ w0, size_t plane, int c) {
  const int pos = 28 * c, k = pos >> 5, sh = pos & 31;
  const uint32_t lo = w0[(size_t)k * plane];
  const uint32_t hi = sh + 28 > 32 ? w0[(size_t)(k + 1) * plane] : 0u;
  return __builtin_amdgcn_alignbit(hi, lo, sh) & 0x0FFFFFFFu;
}

// u32 staging rows [t][m][rows][C][P] -> rows j0 .. j0 + rows - 1 of the
// resident format (R rows per residue).
__global__ __launch_bounds__(256) void k_db_pack(const uint32_t *__restrict__ in, uint32_t *__restrict__ out, int n2,
                                                 int rows, int j0, int R, int C) {
  const int CG = db_group_cols(C), W = db_group_words(C), G = C / CG;
  const size_t total = (size_t)n2 * 2 * rows * G * N;
  for (size_t gid = (size_t)blockIdx.x * 256 + threadIdx.x; gid < total; gid += (size_t)gridDim.x * 256) {
    const int p = (int)(gid % N);
    size_t rest = gid / N;
    const int g = (int)(rest % G); rest /= G;
    const int jl = (int)(rest % rows);
    const size_t tm = rest / rows;  // t * 2 + m
    uint32_t w[7] = {0, 0, 0, 0, 0, 0, 0};
    const uint32_t *src = in + ((tm * rows + jl) * C + (size_t)g * CG) * N + p;
    for (int c = 0; c < CG; c++) {
      const uint32_t x = src[(size_t)c * N];
      const int pos = 28 * c, k = pos >> 5, sh = pos & 31;
      w[k] |= x << sh;
      if (sh + 28 > 32) w[k + 1] |= x >> (32 - sh);
    }
    uint32_t *o = out + ((tm * R + j0 + jl) * G + g) * W * N + p;
    for (int k = 0; k < W; k++) o[(size_t)k * N] = w[k];
  }
}

// ---------------------------------------------------------------------------
// First-dimension DB multiply (single query), internal layouts:
//   db  [T][m][R][C/CG][W][P] (resident format)   sel j at sel + j*sel_stride: [row][m][P]
//   out [T][C][row][m][P]
// Block = WPB waves splitting the R rows; a lane owns 2 positions and the CG
// columns of one group, accumulating sel*db lazily in u64 (folded every 128 rows).  The
// database is streamed once with non-temporal loads (no reuse to cache);
// the selectors are reused by every column group and trial and stay cached.
// Blocks are numbered so that the 8 XCDs each own a quarter of the
// (residue, position-block) selector groups, keeping selector re-reads in
// one XCD's L2.
// ---------------------------------------------------------------------------
// Exact x mod q for any 64-bit x (canonical): (hi * 2^32 + lo) mod q.
__device__ __forceinline__ uint32_t mod_u64(uint64_t x, const ModConst &mc) {
  return raw_mod((uint32_t)(x >> 32), (uint32_t)x, mc);
}

typedef unsigned int v2u32 __attribute__((ext_vector_type(2)));

template <int CG, int WPB>
__global__ __launch_bounds__(WPB * 64) void k_db_mul(const uint32_t *__restrict__ db,
                                                     const uint32_t *__restrict__ sel_all, size_t sel_stride,
                                                     size_t sel_qstride, uint32_t *__restrict__ out_all,
                                                     size_t out_qstride, int nq, int T, int R, int C, ModConst mc0,
                                                     ModConst mc1) {
  constexpr int W = (28 * CG + 31) / 32;
  __shared__ uint32_t red[WPB][CG][2][128];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ncg = C / CG;
  const int bid = blockIdx.x;
  // Consecutive blocks of one XCD take the same database tile for different
  // queries, so a batch reads each tile from HBM about once.
  const int xcd = bid & 7, loc = bid >> 3;
  const int qi = loc % nq;
  const int rest0 = loc / nq;
  const int sg = xcd * 4 + (rest0 & 3);
  const int rest = rest0 >> 2;
  const int tt = rest % T, cg = rest / T;
  if (cg >= ncg) return;
  const uint32_t *__restrict__ sel = sel_all + (size_t)qi * sel_qstride;
  uint32_t *__restrict__ out = out_all + (size_t)qi * out_qstride;
  const int m = sg >> 4, pb = sg & 15;
  const ModConst mc = m ? mc1 : mc0;
  const int p0 = pb * 128 + lane * 2;
  const int rows_per_wave = R / WPB;
  const int j0 = wave * rows_per_wave;

  uint64_t acc[CG][2][2];
#pragma unroll
  for (int c = 0; c < CG; c++)
#pragma unroll
    for (int pp = 0; pp < 2; pp++) { acc[c][pp][0] = 0; acc[c][pp][1] = 0; }

  // Products are < q^2 < 2^56; after each 128-row block the accumulators are
  // reduced mod q, so a block starts below q and ends below q + 128 q^2 < 2^63.
  const uint32_t *dbase = db + ((((size_t)tt * 2 + m) * R) * ncg + cg) * W * N + p0;
  for (int jb = 0; jb < rows_per_wave; jb += 128) {
    const int jend = min(rows_per_wave, jb + 128);
#pragma unroll 2
    for (int jj = jb; jj < jend; jj++) {
      const int j = j0 + jj;
      const uint32_t *sp = sel + (size_t)j * sel_stride + (size_t)m * N + p0;
      const uint2 s0 = *reinterpret_cast<const uint2 *>(sp);
      const uint2 s1 = *reinterpret_cast<const uint2 *>(sp + 2 * N);
      const uint32_t *dp = dbase + (size_t)j * ncg * W * N;
      uint32_t wx[W], wy[W];  // the packed groups of positions p0 and p0 + 1
#pragma unroll
      for (int k = 0; k < W; k++) {
        const v2u32 d = __builtin_nontemporal_load(reinterpret_cast<const v2u32 *>(dp + (size_t)k * N));
        wx[k] = d.x;
        wy[k] = d.y;
      }
      static_for<CG>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
        const uint32_t dx = unpack28<c, W>(wx), dy = unpack28<c, W>(wy);
        acc[c][0][0] += (uint64_t)s0.x * dx;
        acc[c][1][0] += (uint64_t)s0.y * dy;
        acc[c][0][1] += (uint64_t)s1.x * dx;
        acc[c][1][1] += (uint64_t)s1.y * dy;
      });
    }
    if (jend < rows_per_wave) {
#pragma unroll
      for (int c = 0; c < CG; c++)
#pragma unroll
        for (int pp = 0; pp < 2; pp++) {
          acc[c][pp][0] = mod_u64(acc[c][pp][0], mc);
          acc[c][pp][1] = mod_u64(acc[c][pp][1], mc);
        }
    }
  }
#pragma unroll
  for (int c = 0; c < CG; c++)
#pragma unroll
    for (int row = 0; row < 2; row++) {
      red[wave][c][row][lane * 2] = mod_u64(acc[c][0][row], mc);
      red[wave][c][row][lane * 2 + 1] = mod_u64(acc[c][1][row], mc);
    }
  __syncthreads();
  for (int idx = threadIdx.x; idx < CG * 256; idx += WPB * 64) {
    const int c = idx >> 8, row = (idx >> 7) & 1, pos = idx & 127;
    uint32_t s = red[0][c][row][pos];
#pragma unroll
    for (int w = 1; w < WPB; w++) s = canon(s + red[w][c][row][pos], mc.q);
    out[((((size_t)tt * C + cg * CG + c) * 2 + row) * 2 + m) * N + pb * 128 + pos] = s;
  }
}

// ---------------------------------------------------------------------------
// Fold: one level, one block per (pair i, trial, query).
//   accs: raw Ciphertexts, set (query, trial) at accs + set*set_stride
//   keys: level key matrix, layout [m][col][row][P], 2^32-scaled
// FAST evaluates  Y + C * NTT(g^-1(X) - g^-1(Y))  (valid when
// C_neg = G - C); otherwise  C * NTT(g^-1 X) + C_neg * NTT(g^-1 Y).
// ---------------------------------------------------------------------------
template <bool FAST>
__global__ __launch_bounds__(256) void k_fold_level(uint32_t *__restrict__ accs, size_t set_stride, int n_trials,
                                                    int num_per, const uint32_t *__restrict__ cpos,
                                                    const uint32_t *__restrict__ cneg, size_t key_query_stride,
                                                    int t_gsw, int bits, const uint32_t *__restrict__ tw_base,
                                                    ModConst mc0, ModConst mc1) {
  __shared__ uint32_t lds_all[4 * LDS_TILE];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int m = wave >> 1, h = wave & 1;
  uint32_t *lds = lds_all + wave * LDS_TILE;
  const ModConst mc = m ? mc1 : mc0;
  const uint32_t q = mc.q;
  const int mx = 2 * t_gsw;
  const uint32_t mask = (1u << bits) - 1u;
  const int i = blockIdx.x;
  const int set = blockIdx.z * n_trials + blockIdx.y;
  uint32_t *base = accs + (size_t)set * set_stride;
  const uint32_t *X = base + (size_t)(num_per + i) * 4 * N;
  uint32_t *Y = base + (size_t)i * 4 * N;
  const uint32_t *kp = cpos + (size_t)blockIdx.z * key_query_stride + (size_t)m * mx * 2 * N;
  const uint32_t *kn = FAST ? nullptr : cneg + (size_t)blockIdx.z * key_query_stride + (size_t)m * mx * 2 * N;
  const uint32_t *tw = tw_base + (m * 2 + 0) * TW_WORDS;

  uint64_t a0[32], a1[32];
#pragma unroll
  for (int j = 0; j < 32; j++) { a0[j] = 0; a1[j] = 0; }

  if (FAST) {
    const uint32_t *xh = X + (size_t)h * 2 * N, *xl = xh + N;
    const uint32_t *yh = Y + (size_t)h * 2 * N, *yl = yh + N;
    for (int k = 0; k < t_gsw; k++) {
      uint32_t r[32];
#pragma unroll
      for (int ii = 0; ii < 32; ii++) {
        const int z = 64 * ii + lane;
        r[ii] = digit_of(xh[z], xl[z], k * bits, mask) + q - digit_of(yh[z], yl[z], k * bits, mask);
      }
      ntt_fwd_wave(r, tw, lds, q, lane);
      const uint32_t *k0 = kp + ((size_t)(h + 2 * k) * 2) * N;
      uint32_t kv0[32], kv1[32];
      load_p(kv0, k0, lane);
      load_p(kv1, k0 + N, lane);
#pragma unroll
      for (int jj = 0; jj < 32; jj++) {
        const uint32_t d = reduce_2q(r[jj], 2 * q);
        a0[jj] += (uint64_t)d * kv0[jj];
        a1[jj] += (uint64_t)d * kv1[jj];
      }
    }
  } else {
    const uint32_t *in = h ? Y : X;
    const uint32_t *kk = h ? kn : kp;
    for (int col = 0; col < mx; col++) {
      const int j = col & 1, k = col >> 1;
      const uint32_t *ih = in + (size_t)j * 2 * N, *il = ih + N;
      uint32_t r[32];
#pragma unroll
      for (int ii = 0; ii < 32; ii++) {
        const int z = 64 * ii + lane;
        r[ii] = digit_of(ih[z], il[z], k * bits, mask);
      }
      ntt_fwd_wave(r, tw, lds, q, lane);
      uint32_t kv0[32], kv1[32];
      load_p(kv0, kk + ((size_t)col * 2) * N, lane);
      load_p(kv1, kk + ((size_t)col * 2 + 1) * N, lane);
#pragma unroll
      for (int jj = 0; jj < 32; jj++) {
        const uint32_t d = reduce_2q(r[jj], 2 * q);
        a0[jj] += (uint64_t)d * kv0[jj];
        a1[jj] += (uint64_t)d * kv1[jj];
      }
    }
  }
  // Row sums across the two waves of this residue.
  uint32_t s[32];
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < 32; jj++) {
    const uint32_t v0 = redc64(a0[jj], q, mc.qinv_neg);
    const uint32_t v1 = redc64(a1[jj], q, mc.qinv_neg);
    s[jj] = h ? v1 : v0;
    lds_all[wave * N + 32 * lane + jj] = h ? v0 : v1;
  }
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < 32; jj++) s[jj] = canon(s[jj] + lds_all[(wave ^ 1) * N + 32 * lane + jj], q);
  __syncthreads();
  ntt_inv_wave(s, tw_base + (m * 2 + 1) * TW_WORDS, lds, q, lane);
#pragma unroll
  for (int ii = 0; ii < 32; ii++) s[ii] = canon(shoup_mul(s[ii], mc.crt_c, mc.crt_cp, q), q);
  __syncthreads();
  // CRT: wave (m, h) finishes coefficients 64*ii + lane, ii in [16m, 16m+16).
  uint32_t *xm = lds_all + wave * 1024;
#pragma unroll
  for (int ii = 0; ii < 16; ii++) xm[ii * 64 + lane] = s[16 * (1 - m) + ii];
  __syncthreads();
  const uint32_t *xo = lds_all + (wave ^ 2) * 1024;
  uint32_t *yh = Y + (size_t)h * 2 * N, *yl = yh + N;
#pragma unroll
  for (int ii = 0; ii < 16; ii++) {
    const int z = 64 * (16 * m + ii) + lane;
    const uint32_t own = s[16 * m + ii], oth = xo[ii * 64 + lane];
    const uint32_t x0 = m ? oth : own, x1 = m ? own : oth;
    uint64_t x = (uint64_t)x0 * Q1 + (uint64_t)x1 * Q0;
    x = (x >= QQ) ? x - QQ : x;
    if (FAST) {
      x += ((uint64_t)yh[z] << 32) | yl[z];
      x = (x >= QQ) ? x - QQ : x;
    }
    yh[z] = (uint32_t)(x >> 32);
    yl[z] = (uint32_t)x;
  }
}

// ---------------------------------------------------------------------------
// Expansion
// ---------------------------------------------------------------------------
// v[num_in + i] = v_neg1[r] (.) v[i] for i < num_in; neg1: [m][2 (w, w')][P].
__global__ __launch_bounds__(256) void k_expand_neg1(uint32_t *__restrict__ v, size_t q_stride, int num_in,
                                                     const uint32_t *__restrict__ neg1, ModConst mc0,
                                                     ModConst mc1) {
  uint32_t *vq = v + (size_t)blockIdx.y * q_stride;
  const size_t total = (size_t)num_in * (4 * N / 4);
  for (size_t g = (size_t)blockIdx.x * 256 + threadIdx.x; g < total; g += (size_t)gridDim.x * 256) {
    const size_t i = g / (N);  // N uint4 per ciphertext
    const int u = (int)(g % N);
    const int word = 4 * u;
    const int m = (word / N) & 1;
    const int p = word % N;
    const ModConst &mc = m ? mc1 : mc0;
    const uint4 x = reinterpret_cast<const uint4 *>(vq + i * 4 * N)[u];
    const uint4 w = reinterpret_cast<const uint4 *>(neg1 + (size_t)m * 2 * N + p)[0];
    const uint4 wp = reinterpret_cast<const uint4 *>(neg1 + (size_t)m * 2 * N + N + p)[0];
    uint4 y;
    y.x = canon(shoup_mul(x.x, w.x, wp.x, mc.q), mc.q);
    y.y = canon(shoup_mul(x.y, w.y, wp.y, mc.q), mc.q);
    y.z = canon(shoup_mul(x.z, w.z, wp.z, mc.q), mc.q);
    y.w = canon(shoup_mul(x.w, w.w, wp.w, mc.q), mc.q);
    reinterpret_cast<uint4 *>(vq + ((size_t)num_in + i) * 4 * N)[u] = y;
  }
}

// v[b][0] = query[b] for b < nq (restarts the in-place expansion).
__global__ __launch_bounds__(256) void k_copy_queries(const uint32_t *__restrict__ query, size_t query_qstride,
                                                      uint32_t *__restrict__ v, size_t v_qstride) {
  const uint4 *src = reinterpret_cast<const uint4 *>(query + (size_t)blockIdx.y * query_qstride);
  uint4 *dst = reinterpret_cast<uint4 *>(v + (size_t)blockIdx.y * v_qstride);
  for (int u = blockIdx.x * 256 + threadIdx.x; u < N; u += gridDim.x * 256) dst[u] = src[u];
}

__device__ __forceinline__ uint32_t brv11(uint32_t x) { return __brev(x) >> 21; }

// One action_expand step (gpu-pir/expansion.cuh:87-166) for element i of
// round r, in place on v[i] (layout P).  Block = 4 waves.
__global__ __launch_bounds__(256) void k_expand_action(uint32_t *__restrict__ v, size_t q_stride, int r,
                                                       int stop, int max_bits, const uint32_t *__restrict__ left,
                                                       const uint32_t *__restrict__ right, size_t left_qstride,
                                                       size_t right_qstride, int t_exp, int bits,
                                                       const uint32_t *__restrict__ tw_base, ModConst mc0,
                                                       ModConst mc1) {
  __shared__ uint32_t scratch[4 * LDS_TILE];
  __shared__ uint64_t ta[N];
  __shared__ uint32_t as_e[2][N];
  const int i = blockIdx.x;
  if (stop > 0 && r > stop && (i & 1)) return;
  if (stop > 0 && r == stop && (i & 1) && (i / 2) >= max_bits) return;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  uint32_t *vi = v + (size_t)blockIdx.y * q_stride + (size_t)i * 4 * N;
  const uint32_t t = (uint32_t)(N >> r) + 1u;
  const bool use_left = (r != 0) && ((i & 1) == 0);
  const uint32_t *W = use_left ? left + (size_t)blockIdx.y * left_qstride : right + (size_t)blockIdx.y * right_qstride;
  uint32_t *xm = scratch + 2 * LDS_TILE;  // tiles of waves 2 and 3, unused in phase 1

  // Phase 1: waves 0/1 take v_i.a to raw residues; waves 2/3 stage v_i.as_e.
  if (wave < 2) {
    const int m = wave;
    const ModConst mc = m ? mc1 : mc0;
    uint32_t rr[32];
    load_p(rr, vi + (size_t)m * N, lane);
    ntt_inv_wave(rr, tw_base + (m * 2 + 1) * TW_WORDS, scratch + wave * LDS_TILE, mc.q, lane);
#pragma unroll
    for (int ii = 0; ii < 32; ii++)
      xm[m * N + 64 * ii + lane] = canon(shoup_mul(rr[ii], mc.crt_c, mc.crt_cp, mc.q), mc.q);
  } else {
    const int m = wave - 2;
    const uint4 *src = reinterpret_cast<const uint4 *>(vi + (size_t)(2 + m) * N);
    uint4 *dst = reinterpret_cast<uint4 *>(as_e[m]);
#pragma unroll
    for (int c = 0; c < 8; c++) dst[c * 64 + lane] = src[c * 64 + lane];
  }
  __syncthreads();
  // Phase 2: CRT and the raw automorphism tau_t of a (zero negates to Q).
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t z = threadIdx.x + 256 * k;
    uint64_t x = (uint64_t)xm[z] * Q1 + (uint64_t)xm[N + z] * Q0;
    x = (x >= QQ) ? x - QQ : x;
    const uint32_t zt = z * t;
    ta[zt & (N - 1)] = ((zt >> LOGN) & 1) ? QQ - x : x;
  }
  __syncthreads();
  // Phase 3: gadget digits of tau(a), NTT, multiply-accumulate with W.
  const int m = wave >> 1, h = wave & 1;
  const ModConst mc = m ? mc1 : mc0;
  const uint32_t q = mc.q;
  const uint32_t mask = (1u << bits) - 1u;
  uint64_t a0[32], a1[32];
#pragma unroll
  for (int j = 0; j < 32; j++) { a0[j] = 0; a1[j] = 0; }
  const int per = t_exp / 2;
  for (int kk = 0; kk < per + (h ? (t_exp & 1) : 0); kk++) {
    const int k = h * per + kk;
    uint32_t rr[32];
    const int shift = k * bits;
#pragma unroll
    for (int ii = 0; ii < 32; ii++) {
      const uint64_t x = ta[64 * ii + lane];
      rr[ii] = shift >= 64 ? 0u : (uint32_t)((x >> shift) & mask);
    }
    ntt_fwd_wave(rr, tw_base + (m * 2) * TW_WORDS, scratch + wave * LDS_TILE, q, lane);
    uint32_t kv0[32], kv1[32];
    const uint32_t *wk = W + ((size_t)(m * t_exp + k) * 2) * N;
    load_p(kv0, wk, lane);
    load_p(kv1, wk + N, lane);
#pragma unroll
    for (int jj = 0; jj < 32; jj++) {
      const uint32_t d = reduce_2q(rr[jj], 2 * q);
      a0[jj] += (uint64_t)d * kv0[jj];
      a1[jj] += (uint64_t)d * kv1[jj];
    }
  }
  uint32_t s[32];
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < 32; jj++) {
    const uint32_t v0 = redc64(a0[jj], q, mc.qinv_neg);
    const uint32_t v1 = redc64(a1[jj], q, mc.qinv_neg);
    s[jj] = h ? v1 : v0;
    scratch[wave * N + 32 * lane + jj] = h ? v0 : v1;
  }
  __syncthreads();
#pragma unroll
  for (int jj = 0; jj < 32; jj++) s[jj] = canon(s[jj] + scratch[(wave ^ 1) * N + 32 * lane + jj], q);
  // Phase 5: v_i(row h, residue m) += W g^-1(tau a)  [+ NTT(tau as_e) on row 1].
  uint32_t vv[32];
  uint32_t *dst = vi + (size_t)(h * 2 + m) * N;
  load_p(vv, dst, lane);
#pragma unroll
  for (int jj = 0; jj < 32; jj++) {
    uint32_t x = vv[jj] + s[jj];
    if (h) {
      // NTT-domain automorphism: position z of tau(f) is position z' of f,
      // 2*brv(z')+1 = t*(2*brv(z)+1) mod 2N.
      const uint32_t z = 32 * vlane(lane) + jj;
      const uint32_t e = ((2 * brv11(z) + 1) * t) & (2 * N - 1);
      const uint32_t zs = brv11((e - 1) >> 1);
      x += as_e[m][p_of_z((int)zs)];
    }
    x = min(x, x - 2 * q);
    vv[jj] = canon(x, q);
  }
  store_p(dst, vv, lane);
}

}  // namespace sp

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
namespace sp {

static inline unsigned grid_for(size_t items, size_t per_block, unsigned cap = 65535u * 4) {
  size_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

hipError_t launch_ntt_natural(bool inv, uint32_t *polys, size_t n, const DevTables &t, int m, int scale,
                              hipStream_t s) {
  const unsigned g = grid_for(n, 4);
  if (inv)
    k_ntt_natural<true><<<g, 256, 0, s>>>(polys, n, tw_of(t, m, 1), t.mc[m], scale);
  else
    k_ntt_natural<false><<<g, 256, 0, s>>>(polys, n, tw_of(t, m, 0), t.mc[m], scale);
  return hipGetLastError();
}

hipError_t launch_from_ntt(bool layout_p, const uint32_t *ntt, uint32_t *raw, size_t n, size_t in_stride,
                           size_t out_stride, const DevTables &t, hipStream_t s) {
  const unsigned g = grid_for(n, 1);
  if (layout_p)
    k_from_ntt<true><<<g, 128, 0, s>>>(ntt, raw, n, in_stride, out_stride, t.tw, t.mc[0], t.mc[1]);
  else
    k_from_ntt<false><<<g, 128, 0, s>>>(ntt, raw, n, in_stride, out_stride, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_to_ntt(const uint32_t *raw, uint32_t *ntt, size_t n, const DevTables &t, hipStream_t s) {
  k_to_ntt<<<grid_for(n, 1), 128, 0, s>>>(raw, ntt, n, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_automorph_u64(const uint64_t *in, uint64_t *out, size_t n, uint32_t tt, hipStream_t s) {
  k_automorph_u64<<<grid_for(n * N, 256), 256, 0, s>>>(in, out, n, tt);
  return hipGetLastError();
}

hipError_t launch_extprod(const uint32_t *in, int rdim, int mx, int bits, const uint32_t *keys, uint32_t *acc,
                          size_t n, size_t in_stride, size_t key_stride, size_t acc_stride, const DevTables &t,
                          hipStream_t s) {
  k_extprod<<<grid_for(n, 1), 256, 0, s>>>(in, rdim, mx, bits, keys, acc, n, in_stride, key_stride, acc_stride,
                                           t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_nat_to_p(const uint32_t *in, uint32_t *out, size_t n, int m_div, int mont, const DevTables &t,
                           hipStream_t s) {
  k_nat_to_p<<<grid_for(n * N, 256), 256, 0, s>>>(in, out, n, m_div, mont, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_p_to_nat(const uint32_t *in, uint32_t *out, size_t n, hipStream_t s) {
  k_p_to_nat<<<grid_for(n * N, 256), 256, 0, s>>>(in, out, n);
  return hipGetLastError();
}

hipError_t launch_p_to_nat_unscale(const uint32_t *in, uint32_t *out, size_t n, int m_div, const DevTables &t,
                                   hipStream_t s) {
  k_p_to_nat_unscale<<<grid_for(n * N, 256), 256, 0, s>>>(in, out, n, m_div, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_db_from_ref(const uint32_t *in, uint32_t *out, int R, int C, const DevTables &t, hipStream_t s) {
  k_db_from_ref<<<grid_for((size_t)2 * R * C * N, 256), 256, 0, s>>>(in, out, R, C, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_db_encode(const uint8_t *bytes, size_t rec0, size_t nrec, uint32_t *db, int n2, int C,
                            const DevTables &t, hipStream_t s) {
  const int rows = (int)(nrec / C);
  k_db_encode<<<grid_for(nrec * n2 * 2, 4), 256, 0, s>>>(bytes, rec0, nrec, db, n2, rows, C, t.tw, t.mc[0],
                                                         t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_db_pack(const uint32_t *in, uint32_t *out, int n2, int rows, int j0, int R, int C, hipStream_t s) {
  const size_t items = (size_t)n2 * 2 * rows * (C / db_group_cols(C)) * N;
  k_db_pack<<<grid_for(items, 256), 256, 0, s>>>(in, out, n2, rows, j0, R, C);
  return hipGetLastError();
}

// Database multiply of several queries per database read (VALU).  k_db_mul
// reads the whole database once per query; this kernel reads it once per Q
// queries: a lane owns PPL positions and the 8 columns of one packed group
// (CG = 8 in W = 7 words) and keeps Q x 8 x PPL x 2 lazy u64 sums; the WPB
// waves of a block split the rows (UNR rows per loop pass) and reduce through
// LDS, one query at a time.  Block = (residue, block of 64 PPL positions,
// trial, column group, query group); each XCD owns a contiguous range of
// (residue, position block) groups so that the re-reads of the selectors stay
// in its L2, and the query groups of one database tile are consecutive on an
// XCD.  Needs C >= 8 and R >= WPB; queries of the last group past nq repeat
// query nq - 1 and are not written.
template <int Q, int PPL, int WPB, int UNR>
__global__ __launch_bounds__(WPB * 64) void k_db_mul_q(const uint32_t *__restrict__ db,
                                                       const uint32_t *__restrict__ sel_all, size_t sel_stride,
                                                       size_t sel_qstride, uint32_t *__restrict__ out_all,
                                                       size_t out_qstride, int nq, int T, int R, int C, ModConst mc0,
                                                       ModConst mc1) {
  constexpr int CG = 8, W = 7, PB = 64 * PPL, NPB = N / PB, SGX = 2 * NPB / 8;
  typedef unsigned int vu __attribute__((ext_vector_type(PPL)));
  __shared__ uint32_t red[WPB][CG][2][PB];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int ncg = C / CG;
  const int nqg = (nq + Q - 1) / Q;
  const int xcd = blockIdx.x & 7, loc = blockIdx.x >> 3;
  const int qg = loc % nqg;
  const int rest0 = loc / nqg;
  const int sg = xcd * SGX + rest0 % SGX;
  const int rest = rest0 / SGX;
  const int tt = rest % T, cg = rest / T;
  if (cg >= ncg) return;
  const int m = sg / NPB, pb = sg % NPB;
  const ModConst mc = m ? mc1 : mc0;
  const int p0 = pb * PB + lane * PPL;
  const uint32_t *sel[Q];
#pragma unroll
  for (int qi = 0; qi < Q; qi++)
    sel[qi] = sel_all + (size_t)min(qg * Q + qi, nq - 1) * sel_qstride + (size_t)m * N + p0;
  const int rows_per_wave = R / WPB;
  const int j0 = wave * rows_per_wave;
  uint64_t acc[Q][CG][PPL][2];
#pragma unroll
  for (int qi = 0; qi < Q; qi++)
#pragma unroll
    for (int c = 0; c < CG; c++)
#pragma unroll
      for (int pp = 0; pp < PPL; pp++) { acc[qi][c][pp][0] = 0; acc[qi][c][pp][1] = 0; }
  // Products are < q^2 < 2^56; every 128 rows the sums are reduced mod q, so a
  // block of rows starts below q and ends below q + 128 q^2 < 2^63.
  const uint32_t *dbase = db + ((((size_t)tt * 2 + m) * R) * ncg + cg) * W * N + p0;
  for (int jb = 0; jb < rows_per_wave; jb += 128) {
    const int jend = min(rows_per_wave, jb + 128);
#pragma unroll UNR
    for (int jj = jb; jj < jend; jj++) {
      const int j = j0 + jj;
      const uint32_t *dp = dbase + (size_t)j * ncg * W * N;
      uint32_t w[PPL][W];
#pragma unroll
      for (int k = 0; k < W; k++) {
        const vu d = __builtin_nontemporal_load(reinterpret_cast<const vu *>(dp + (size_t)k * N));
#pragma unroll
        for (int pp = 0; pp < PPL; pp++) w[pp][k] = d[pp];
      }
      uint32_t sv[Q][2][PPL];
#pragma unroll
      for (int qi = 0; qi < Q; qi++)
#pragma unroll
        for (int row = 0; row < 2; row++) {
          const vu x = *reinterpret_cast<const vu *>(sel[qi] + (size_t)j * sel_stride + (size_t)row * 2 * N);
#pragma unroll
          for (int pp = 0; pp < PPL; pp++) sv[qi][row][pp] = x[pp];
        }
      static_for<CG>([&](auto cc) {
        constexpr int c = decltype(cc)::value;
#pragma unroll
        for (int pp = 0; pp < PPL; pp++) {
          const uint32_t d = unpack28<c, W>(w[pp]);
#pragma unroll
          for (int qi = 0; qi < Q; qi++) {
            acc[qi][c][pp][0] += (uint64_t)sv[qi][0][pp] * d;
            acc[qi][c][pp][1] += (uint64_t)sv[qi][1][pp] * d;
          }
        }
      });
    }
    if (jend < rows_per_wave) {
#pragma unroll
      for (int qi = 0; qi < Q; qi++)
#pragma unroll
        for (int c = 0; c < CG; c++)
#pragma unroll
          for (int pp = 0; pp < PPL; pp++) {
            acc[qi][c][pp][0] = mod_u64(acc[qi][c][pp][0], mc);
            acc[qi][c][pp][1] = mod_u64(acc[qi][c][pp][1], mc);
          }
    }
  }
#pragma unroll
  for (int qi = 0; qi < Q; qi++) {
    const int q = qg * Q + qi;
    if (q >= nq) break;  // uniform over the block
#pragma unroll
    for (int c = 0; c < CG; c++)
#pragma unroll
      for (int row = 0; row < 2; row++)
#pragma unroll
        for (int pp = 0; pp < PPL; pp++) red[wave][c][row][lane * PPL + pp] = mod_u64(acc[qi][c][pp][row], mc);
    __syncthreads();
    uint32_t *out = out_all + (size_t)q * out_qstride;
    for (int idx = threadIdx.x; idx < CG * 2 * PB; idx += WPB * 64) {
      const int c = idx / (2 * PB), row = (idx / PB) & 1, pos = idx % PB;
      uint32_t v = red[0][c][row][pos];
#pragma unroll
      for (int w = 1; w < WPB; w++) v = canon(v + red[w][c][row][pos], mc.q);
      out[((((size_t)tt * C + cg * CG + c) * 2 + row) * 2 + m) * N + pb * PB + pos] = v;
    }
    __syncthreads();
  }
}

// Launch of the multi-query multiply.  form 1: 2 queries per read, 2 positions
// per lane, 4-wave blocks (3 per CU at 161-166 VGPRs): at (10, 7), 1.36 ms for
// 2 queries (one database read at HBM speed is 1.17 ms), against 1.43 ms in
// 8-wave blocks and 2.25 ms for one read per query.  form 2: 4 queries per
// read, 1 position per lane, 2-wave blocks without unrolling: at (9, 6), 8
// queries 0.85 ms against 0.96 ms in 4-wave blocks unrolled twice (at (10, 7)
// the 4-wave form took 1.85 ms for 4 queries, 8-wave blocks 2.56 ms).  Deeper
// unrolling and half-group lanes (4 columns x 2 positions) were slower.
hipError_t launch_db_mul_multi(const uint32_t *db, const uint32_t *sel, size_t sel_stride, size_t sel_qstride,
                               uint32_t *out, size_t out_qstride, int nq, int T, int R, int C, int form,
                               const DevTables &t, hipStream_t s) {
  const int ncg = C / 8;
#define SP_DBQ(QV, PPLV)                                                                                        \
  k_db_mul_q<QV, PPLV, 4, 2><<<(unsigned)(2 * (N / (64 * PPLV)) * T * ncg * ((nq + QV - 1) / QV)), 256, 0, s>>>( \
      db, sel, sel_stride, sel_qstride, out, out_qstride, nq, T, R, C, t.mc[0], t.mc[1])
  if (form == 2)
    k_db_mul_q<4, 1, 2, 1><<<(unsigned)(2 * (N / 64) * T * ncg * ((nq + 3) / 4)), 128, 0, s>>>(
        db, sel, sel_stride, sel_qstride, out, out_qstride, nq, T, R, C, t.mc[0], t.mc[1]);
  else
    SP_DBQ(2, 2);
#undef SP_DBQ
  return hipGetLastError();
}

hipError_t launch_db_mul(const uint32_t *db, const uint32_t *sel, size_t sel_stride, size_t sel_qstride,
                         uint32_t *out, size_t out_qstride, int nq, int T, int R, int C, const DevTables &t,
                         hipStream_t s) {
  // One query, whole packed groups: k_db_mul_q in 2-wave blocks, 2 positions
  // per lane, rows not unrolled (1024 blocks at (9, 6), one per 128 positions,
  // trial, residue and column group; each wave streams 256 rows).  Measured at
  // (9, 6) back to back: 271-273 us against 298 us for k_db_mul<8, 8>; 4-wave
  // blocks 279 us, 1-wave blocks 406 us, 4 positions per lane 275-279 us,
  // unrolled rows 278-346 us.  For several queries k_db_mul, whose consecutive
  // blocks on an XCD take one database tile for different queries, stays
  // ahead (8 queries: 2.09 against 2.15-2.18 ms).
  if (nq == 1 && C >= 8 && R >= 2) {
    k_db_mul_q<1, 2, 2, 1><<<(unsigned)(2 * (N / 128) * T * (C / 8)), 128, 0, s>>>(
        db, sel, sel_stride, sel_qstride, out, out_qstride, 1, T, R, C, t.mc[0], t.mc[1]);
    return hipGetLastError();
  }
  const int cg = db_group_cols(C);
  const unsigned blocks = (unsigned)(32 * T * (C / cg) * nq);
  // 8 waves split the rows (R >= 8); the smallest databases (R = 4) use 4.
#define SP_DBMUL(CGV, W)                                                                                      \
  k_db_mul<CGV, W><<<blocks, W * 64, 0, s>>>(db, sel, sel_stride, sel_qstride, out, out_qstride, nq, T, R, C, \
                                             t.mc[0], t.mc[1])
  if (R >= 8) {
    switch (cg) {
      case 8: SP_DBMUL(8, 8); break;
      case 4: SP_DBMUL(4, 8); break;
      case 2: SP_DBMUL(2, 8); break;
      default: SP_DBMUL(1, 8); break;
    }
  } else {
    switch (cg) {
      case 8: SP_DBMUL(8, 4); break;
      case 4: SP_DBMUL(4, 4); break;
      case 2: SP_DBMUL(2, 4); break;
      default: SP_DBMUL(1, 4); break;
    }
  }
#undef SP_DBMUL
  return hipGetLastError();
}

hipError_t launch_fold_level(bool fast, uint32_t *accs, size_t set_stride, int n_trials, int n_queries,
                             int num_per, const uint32_t *cpos, const uint32_t *cneg, size_t key_query_stride,
                             int t_gsw, int bits, const DevTables &t, hipStream_t s) {
  dim3 grid(num_per, n_trials, n_queries);
  if (fast)
    k_fold_level<true><<<grid, 256, 0, s>>>(accs, set_stride, n_trials, num_per, cpos, cneg, key_query_stride,
                                            t_gsw, bits, t.tw, t.mc[0], t.mc[1]);
  else
    k_fold_level<false><<<grid, 256, 0, s>>>(accs, set_stride, n_trials, num_per, cpos, cneg, key_query_stride,
                                             t_gsw, bits, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_expand_neg1(uint32_t *v, size_t q_stride, int n_queries, int num_in, const uint32_t *neg1_r,
                              const DevTables &t, hipStream_t s) {
  dim3 grid(grid_for((size_t)num_in * N, 256, 65535u), n_queries);
  k_expand_neg1<<<grid, 256, 0, s>>>(v, q_stride, num_in, neg1_r, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_expand_action(uint32_t *v, size_t q_stride, int n_queries, int r, int num_out, int stop,
                                int max_bits, const uint32_t *left, const uint32_t *right, size_t left_qstride,
                                size_t right_qstride, int t_exp, int bits, const DevTables &t, hipStream_t s) {
  dim3 grid(num_out, n_queries);
  k_expand_action<<<grid, 256, 0, s>>>(v, q_stride, r, stop, max_bits, left, right, left_qstride, right_qstride,
                                       t_exp, bits, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

}  // namespace sp

namespace sp {
hipError_t launch_copy_queries(const uint32_t *query, size_t query_qstride, uint32_t *v, size_t v_qstride, int nq,
                               hipStream_t s) {
  k_copy_queries<<<dim3(2, nq), 256, 0, s>>>(query, query_qstride, v, v_qstride);
  return hipGetLastError();
}

// One ciphertext (4N words) per blockIdx.y, gathered into consecutive rows.
// Per-slot flags written by a kernel whose arguments carry the values, so that
// an asynchronous load needs no host buffer that outlives the call.
struct SlotFlags {
  int v[SLOT_FLAGS_PER_LAUNCH];
};

__global__ __launch_bounds__(64) void k_set_slot_flags(int *__restrict__ dst, int n, SlotFlags f) {
  const int i = threadIdx.x;
  if (i < n) dst[i] = f.v[i];
}

hipError_t launch_set_slot_flags(int *dst, int n, const int *vals, hipStream_t s) {
  for (int i0 = 0; i0 < n; i0 += SLOT_FLAGS_PER_LAUNCH) {
    const int k = n - i0 < SLOT_FLAGS_PER_LAUNCH ? n - i0 : SLOT_FLAGS_PER_LAUNCH;
    SlotFlags f{};
    for (int i = 0; i < k; i++) f.v[i] = vals[i0 + i];
    k_set_slot_flags<<<1, 64, 0, s>>>(dst + i0, k, f);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

__global__ __launch_bounds__(256) void k_gather_rows(const uint32_t *__restrict__ src, size_t src_stride,
                                                     uint32_t *__restrict__ dst) {
  const uint4 *from = reinterpret_cast<const uint4 *>(src + (size_t)blockIdx.y * src_stride);
  uint4 *to = reinterpret_cast<uint4 *>(dst + (size_t)blockIdx.y * 4 * N);
  for (int u = blockIdx.x * 256 + threadIdx.x; u < N; u += gridDim.x * 256) to[u] = from[u];
}

hipError_t launch_gather_rows(const uint32_t *src, size_t src_stride, int rows, uint32_t *dst, hipStream_t s) {
  if (rows < 1) return hipSuccess;
  k_gather_rows<<<dim3(2, rows), 256, 0, s>>>(src, src_stride, dst);
  return hipGetLastError();
}
}  // namespace sp

// ===========================================================================
// Round-structured external-product kernels (gadget base 2^7 x 8 digits).
//
// All gadget digits of a block's inputs are extracted once into LDS as
// packed bytes.  Each round, every wave transforms one digit polynomial into
// its own LDS tile (layout P); then every thread multiply-accumulates the
// round's tiles at its own PPT positions against the key matrix.  The 64-bit
// accumulators live only in the multiply-accumulate phase, so the NTT keeps
// its registers and the block runs NW = 8 transforms at a time.
// Digit bytes: coefficient z = 64 i + l of digit polynomial k sits in byte
// (i & 3) of word (i >> 2) * 64 + l, so lane l of a wave reads its phase-A
// registers with 8 conflict-free 4-byte LDS reads.
// ===========================================================================
namespace sp {

constexpr int DIG_WORDS = N / 4;  // 512 words per byte-packed digit polynomial

// Tile store for transform outputs below 2^32: each value is reduced to
// [0, 2q) by a Shoup multiplication by 1.
__device__ __forceinline__ void store_tile_p_full(uint32_t *__restrict__ tile, const uint32_t (&r)[32], int lane,
                                                  uint32_t q, uint32_t mu1) {
  uint4 *t4 = reinterpret_cast<uint4 *>(tile);
#pragma unroll
  for (int c = 0; c < 8; c++) {
    uint32_t v[4];
#pragma unroll
    for (int e = 0; e < 4; e++) v[e] = r[4 * c + e] - __umulhi(r[4 * c + e], mu1) * q;
    t4[c * 64 + vlane(lane)] = make_uint4(v[0], v[1], v[2], v[3]);
  }
}

// Fold and expansion kernels come in two shapes:
//  * NW = 8 waves (512 threads, 8 transform tiles = 67.6 KB of LDS, at most
//    128 VGPRs): two blocks per CU, for grids that fill the chip;
//  * NW = 16 waves (1024 threads, 16 tiles = 135 KB): one block per CU with
//    half the passes, for grids smaller than the CU count (latency).
// Every wave transforms one digit polynomial per pass into its own tile; after
// a barrier every thread multiply-accumulates the pass's tiles at its own
// positions against the key.

// Fold digit differences g^-1(X) - g^-1(Y) as signed bytes into dig
// ([j][k - K0][DIG_WORDS] for digits K0 .. K0 + NK - 1 of both polys j).
template <int NT, int K0 = 0, int NK = 8>
__device__ __forceinline__ void fold_digits(const uint32_t *__restrict__ X, const uint32_t *__restrict__ Y,
                                            uint32_t *__restrict__ dig, int tid) {
  for (int it = tid; it < 2 * DIG_WORDS; it += NT) {
    const int j = it / DIG_WORDS, w = it % DIG_WORDS, l = w & 63, i4 = w >> 6;
    const uint32_t *xh = X + (size_t)j * 2 * N, *xl = xh + N;
    const uint32_t *yh = Y + (size_t)j * 2 * N, *yl = yh + N;
    uint32_t packed[NK];
#pragma unroll
    for (int k = 0; k < NK; k++) packed[k] = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
      const int z = 64 * (4 * i4 + b) + l;
      const uint64_t x = ((uint64_t)xh[z] << 32) | xl[z];
      const uint64_t y = ((uint64_t)yh[z] << 32) | yl[z];
#pragma unroll
      for (int k = 0; k < NK; k++) {
        const int d = (int)((x >> (7 * (K0 + k))) & 127u) - (int)((y >> (7 * (K0 + k))) & 127u);
        packed[k] |= ((uint32_t)d & 0xFFu) << (8 * b);
      }
    }
#pragma unroll
    for (int k = 0; k < NK; k++) dig[(j * NK + k) * DIG_WORDS + w] = packed[k];
  }
}

// One residue of a fold level: 16 / NW passes over the key columns owned by
// this wave (column c = wave + NW h), one forward transform per wave and
// pass, then a multiply-accumulate of the pass's NW tiles.
// Returns the Montgomery-reduced sums of rows 0/1 at this thread's positions.
template <int NW, int M>
__device__ __forceinline__ void fold_residue(const uint32_t (&dw)[16 / NW][8], uint32_t (&sum)[2][N / (NW * 64)],
                                             uint32_t (*tiles)[LDS_TILE], const uint32_t *__restrict__ kp,
                                             const uint32_t *__restrict__ tw_base, const ModConst &mc, int wave,
                                             int lane, int tid) {
  constexpr int PPT = N / (NW * 64);
  uint64_t acc[2][PPT];
#pragma unroll
  for (int row = 0; row < 2; row++)
#pragma unroll
    for (int e = 0; e < PPT; e++) acc[row][e] = 0;
#pragma unroll
  for (int h = 0; h < 16 / NW; h++) {
    uint32_t r[32];
#pragma unroll
    for (int i4 = 0; i4 < 8; i4++)
#pragma unroll
      for (int b = 0; b < 4; b++) r[4 * i4 + b] = (uint32_t)((int32_t)(int8_t)(dw[h][i4] >> (8 * b)) + (int32_t)mc.q);
    const uint32_t *twm = tw_base + (M * 2) * TW_WORDS;
    asm volatile("" : "+s"(twm));  // re-read per pass, not held in registers
    // inputs in (q - 128, q + 128) < 2q: stages 1-7 and 9-11 skip the reduction (outputs < 10q)
    ntt_fwd_wave_t<7, true>(r, twm, tiles[wave], mc.q, mc.mu1, lane);
    store_tile_p_full(tiles[wave], r, lane, mc.q, mc.mu1);
    __syncthreads();
#pragma unroll
    for (int w2 = 0; w2 < NW; w2++) {
      const int c = w2 + NW * h;
      const uint32_t *k0 = kp + ((size_t)(M * 16 + c) * 2) * N + PPT * tid;
      const uint32_t *k1 = k0 + N;
      const uint32_t *tl = &tiles[w2][PPT * tid];
#pragma unroll
      for (int e = 0; e < PPT; e++) {
        const uint32_t d = tl[e];
        acc[0][e] += (uint64_t)d * k0[e];
        acc[1][e] += (uint64_t)d * k1[e];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int row = 0; row < 2; row++)
#pragma unroll
    for (int e = 0; e < PPT; e++) sum[row][e] = redc64(acc[row][e], mc.q, mc.qinv_neg);
}

// Fold epilogue: tiles 2m + row hold the NTT-domain sums (P layout, < q) of
// row `row` mod q_m; waves 0-3 take them out of the NTT domain, combine the
// residues by CRT, add Ysrc (raw hi/lo planes) mod Q and write the result to Y
// (Ysrc may be Y).
__device__ __forceinline__ void fold_tail(uint32_t (*tiles)[LDS_TILE], uint32_t *Y, const uint32_t *Ysrc,
                                          const uint32_t *__restrict__ tw_base, const ModConst &mc0,
                                          const ModConst &mc1, int wave, int lane) {
  uint32_t s[32];
  if (wave < 4) {
    const int m = wave >> 1;
    const ModConst mc = m ? mc1 : mc0;
    load_p(s, tiles[wave], lane);
    wave_lds_sync();
    ntt_inv_wave_lazy(s, tw_base + (m * 2 + 1) * TW_WORDS, tiles[wave], mc.q, mc.mu1, lane);
#pragma unroll
    for (int ii = 0; ii < 32; ii++) s[ii] = canon(shoup_mul(s[ii], mc.crt_c, mc.crt_cp, mc.q), mc.q);
  }
  __syncthreads();
  if (wave < 4) {
    const int m = wave >> 1;
    uint32_t *xm = tiles[wave];
#pragma unroll
    for (int ii = 0; ii < 16; ii++) xm[ii * 64 + lane] = s[16 * (1 - m) + ii];
  }
  __syncthreads();
  if (wave < 4) {
    const int m = wave >> 1, row = wave & 1;
    const uint32_t *xo = tiles[wave ^ 2];
    uint32_t *yh = Y + (size_t)row * 2 * N, *yl = yh + N;
    const uint32_t *sh = Ysrc + (size_t)row * 2 * N, *sl = sh + N;
#pragma unroll
    for (int ii = 0; ii < 16; ii++) {
      const int z = 64 * (16 * m + ii) + lane;
      const uint32_t own = s[16 * m + ii], oth = xo[ii * 64 + lane];
      const uint32_t x0 = m ? oth : own, x1 = m ? own : oth;
      uint64_t x = (uint64_t)x0 * Q1 + (uint64_t)x1 * Q0;
      x = (x >= QQ) ? x - QQ : x;
      x += ((uint64_t)sh[z] << 32) | sl[z];
      x = (x >= QQ) ? x - QQ : x;
      yh[z] = (uint32_t)(x >> 32);
      yl[z] = (uint32_t)x;
    }
  }
}

// Fold level, server form C_neg = G - C:  acc[i] <- Y + from_ntt(C * NTT(g^-1 X - g^-1 Y)),
// X = acc[num_per + i], Y = acc[i].  The 16 digit polynomials (2 polys x 8
// digits, key column c = j + 2k) are extracted once into tiles 0-3; wave w
// keeps columns w + NW h packed in registers (4 coefficients per word) and
// transforms each once per residue.
template <int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(4)))
void k_fold8(uint32_t *__restrict__ accs, size_t set_stride, int n_trials, int num_per,
             const uint32_t *__restrict__ cpos, size_t key_qstride, const uint32_t *__restrict__ tw_base,
             ModConst mc0, ModConst mc1) {
  constexpr int NT = NW * 64;
  constexpr int PPT = N / NT;
  constexpr int H = 16 / NW;
  static_assert(NW == 8 || NW == 16, "8 or 16 waves");
  __shared__ uint32_t tiles[NW][LDS_TILE];
  uint32_t *dig = &tiles[0][0];  // digit bytes, 8192 words over tiles 0-3
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int i = blockIdx.x;
  const int set = blockIdx.z * n_trials + blockIdx.y;
  uint32_t *base = accs + (size_t)set * set_stride;
  const uint32_t *X = base + (size_t)(num_per + i) * 4 * N;
  uint32_t *Y = base + (size_t)i * 4 * N;
  const uint32_t *kp = cpos + (size_t)blockIdx.z * key_qstride;

  fold_digits<NT>(X, Y, dig, tid);
  __syncthreads();
  uint32_t dw[H][8];
#pragma unroll
  for (int h = 0; h < H; h++) {
    const int c = wave + NW * h, j = c & 1, k = c >> 1;
#pragma unroll
    for (int i4 = 0; i4 < 8; i4++) dw[h][i4] = dig[(j * 8 + k) * DIG_WORDS + i4 * 64 + lane];
  }
  __syncthreads();  // the digit bytes are dead from here on

  uint32_t sum[2][2][PPT];  // [m][row][e]
  fold_residue<NW, 0>(dw, sum[0], tiles, kp, tw_base, mc0, wave, lane, tid);
  fold_residue<NW, 1>(dw, sum[1], tiles, kp, tw_base, mc1, wave, lane, tid);
  // Sums of polynomial (row, m) into tile 2m + row, then INTT + CRT + Y.
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int row = 0; row < 2; row++)
#pragma unroll
      for (int e = 0; e < PPT; e++) tiles[2 * m + row][PPT * tid + e] = sum[m][row][e];
  __syncthreads();
  fold_tail(tiles, Y, Y, tw_base, mc0, mc1, wave, lane);
}

// ---- split fold level: partial products + finish ----------------------------
// k_fold_part<NG>: one block of NW = 16 / NG waves per (pair i, trial, query,
// residue m, key-column group g < NG): key columns c = NW g .. NW g + NW - 1
// (c = j + 2k: digits NW g / 2 .. of both polys), one transform per wave, a
// multiply-accumulate of the NW tiles, and the Montgomery-reduced partial sums
// of both rows (NTT domain, layout P, < q) to part[set][i][m][g][row][P]
// (set = query * n_trials + trial).  NG = 2 (8-wave blocks) for grids that fill
// the chip; NG = 4 (4-wave blocks: one transform per SIMD, twice the CUs) for
// small grids, where the blocks' latency is the cost.
template <int NG>
__device__ __forceinline__ void fold_part_digits(const uint32_t *__restrict__ X, const uint32_t *__restrict__ Y,
                                                 uint32_t *__restrict__ dig, int g, int tid) {
  constexpr int NK = 8 / NG, NT = 16 / NG * 64;
  if (NG == 2) {
    if (g == 0) fold_digits<NT, 0, NK>(X, Y, dig, tid);
    else fold_digits<NT, NK, NK>(X, Y, dig, tid);
  } else {
    if (g == 0) fold_digits<NT, 0, NK>(X, Y, dig, tid);
    else if (g == 1) fold_digits<NT, NK, NK>(X, Y, dig, tid);
    else if (g == 2) fold_digits<NT, 2 * NK, NK>(X, Y, dig, tid);
    else fold_digits<NT, 3 * NK, NK>(X, Y, dig, tid);
  }
}

template <int NG>
__global__ __launch_bounds__(16 / NG * 64) __attribute__((amdgpu_waves_per_eu(4)))
void k_fold_part(const uint32_t *__restrict__ accs, size_t set_stride, int n_trials, int num_per,
                 const uint32_t *__restrict__ cpos, size_t key_qstride, uint32_t *__restrict__ part,
                 size_t part_set_stride, const uint32_t *__restrict__ tw_base, ModConst mc0, ModConst mc1) {
  constexpr int NW = 16 / NG, NK = 8 / NG;
  static_assert(NW <= 8, "lazy digit tiles: at most 8 products per REDC");
  constexpr int PPT = N / (NW * 64);
  __shared__ uint32_t tiles[NW][LDS_TILE];
  uint32_t *dig = &tiles[0][0];  // 2 polys x NK digit polynomials, 1024 * NK words
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int i = blockIdx.x;
  const int g = blockIdx.y % NG, m = (blockIdx.y / NG) & 1, trial = blockIdx.y / (2 * NG);
  const int set = blockIdx.z * n_trials + trial;
  const uint32_t *base = accs + (size_t)set * set_stride;
  const uint32_t *X = base + (size_t)(num_per + i) * 4 * N;
  const uint32_t *Y = base + (size_t)i * 4 * N;
  const uint32_t *kp = cpos + (size_t)blockIdx.z * key_qstride;
  fold_part_digits<NG>(X, Y, dig, g, tid);
  __syncthreads();
  const int c = NW * g + wave, j = c & 1, kl = (c >> 1) - NK * g;
  uint32_t dw[8];
#pragma unroll
  for (int i4 = 0; i4 < 8; i4++) dw[i4] = dig[(j * NK + kl) * DIG_WORDS + i4 * 64 + lane];
  __syncthreads();  // the digit bytes are dead from here on
  const uint32_t q = m ? mc1.q : mc0.q, mu1 = m ? mc1.mu1 : mc0.mu1, qinv = m ? mc1.qinv_neg : mc0.qinv_neg;
  {
    uint32_t r[32];
#pragma unroll
    for (int i4 = 0; i4 < 8; i4++)
#pragma unroll
      for (int b = 0; b < 4; b++) r[4 * i4 + b] = (uint32_t)((int32_t)(int8_t)(dw[i4] >> (8 * b)) + (int32_t)q);
    // inputs in (q - 128, q + 128) < 2q: stages 1-7 and 9-11 skip the reduction (outputs < 10q)
    ntt_fwd_wave_t<7, true>(r, tw_base + (m * 2) * TW_WORDS, tiles[wave], q, mu1, lane);
    // stored lazy: a sum of at most 8 products with keys < q is below 80 q^2
    // < 2^63 and its REDC below 6q, which redc64 brings to [0, q)
    store_p(tiles[wave], r, lane);
  }
  __syncthreads();
  uint64_t acc[2][PPT];
#pragma unroll
  for (int row = 0; row < 2; row++)
#pragma unroll
    for (int e = 0; e < PPT; e++) acc[row][e] = 0;
#pragma unroll
  for (int w2 = 0; w2 < NW; w2++) {
    const uint32_t *k0 = kp + ((size_t)(m * 16 + NW * g + w2) * 2) * N + PPT * tid;
    const uint32_t *k1 = k0 + N;
    const uint32_t *tl = &tiles[w2][PPT * tid];
#pragma unroll
    for (int e = 0; e < PPT; e++) {
      const uint32_t d = tl[e];
      acc[0][e] += (uint64_t)d * k0[e];
      acc[1][e] += (uint64_t)d * k1[e];
    }
  }
  uint32_t *dst = part + (size_t)set * part_set_stride + ((((size_t)i * 2 + m) * NG + g) * 2) * N + PPT * tid;
#pragma unroll
  for (int row = 0; row < 2; row++)
#pragma unroll
    for (int e = 0; e < PPT; e++) dst[(size_t)row * N + e] = redc64(acc[row][e], q, qinv);
}

// Tiles 0-3 (256 threads): tile w = sum over g < NG of the polynomials at
// src(w, g) (layout P, canonical residues mod q_(w >> qshift & 1)).  The loads
// of four rounds are issued together: a round trip to memory is the cost here.
template <int NG, int QSHIFT, class Src>
__device__ __forceinline__ void sum_partials4(uint32_t (*tiles)[LDS_TILE], const Src &src, const ModConst &mc0,
                                              const ModConst &mc1, int tid) {
#pragma unroll
  for (int half = 0; half < 2; half++) {
    uint4 v[4][NG];
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int w = (4 * half + it) >> 1, k4 = tid + ((4 * half + it) & 1) * 256;
#pragma unroll
      for (int g = 0; g < NG; g++) v[it][g] = reinterpret_cast<const uint4 *>(src(w, g))[k4];
    }
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int w = (4 * half + it) >> 1, k4 = tid + ((4 * half + it) & 1) * 256;
      const uint32_t q = ((w >> QSHIFT) & 1) ? mc1.q : mc0.q;
      uint4 a = v[it][0];
#pragma unroll
      for (int g = 1; g < NG; g++)  // partial sums are < q: add and reduce once per term
        a = make_uint4(canon(a.x + v[it][g].x, q), canon(a.y + v[it][g].y, q), canon(a.z + v[it][g].z, q),
                       canon(a.w + v[it][g].w, q));
      reinterpret_cast<uint4 *>(tiles[w])[k4] = a;
    }
  }
}

// k_fold_finish<NG>: one 4-wave block per (pair i, trial, query): tile 2m + row =
// sum over g of part[m][g][row] mod q_m, then the fold epilogue: Y = Ysrc +
// from_ntt(sums), Y = accs[set][i], Ysrc = ysrc[set][i] (ysrc may be accs).
template <int NG>
__global__ __launch_bounds__(256) void k_fold_finish(uint32_t *accs, size_t set_stride, int n_trials,
                                                     const uint32_t *ysrc, size_t ysrc_set_stride,
                                                     const uint32_t *__restrict__ part, size_t part_set_stride,
                                                     const uint32_t *__restrict__ tw_base, ModConst mc0,
                                                     ModConst mc1) {
  __shared__ uint32_t tiles[4][LDS_TILE];
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int i = blockIdx.x;
  const int set = blockIdx.z * n_trials + blockIdx.y;
  uint32_t *Y = accs + (size_t)set * set_stride + (size_t)i * 4 * N;
  const uint32_t *Ysrc = ysrc + (size_t)set * ysrc_set_stride + (size_t)i * 4 * N;
  const uint32_t *pp = part + (size_t)set * part_set_stride + (size_t)i * 2 * NG * 2 * N;  // [m][g][row][P]
  // tile w = 2 m + row
  sum_partials4<NG, 1>(
      tiles, [&](int w, int g) { return pp + (size_t)(((w >> 1) * NG + g) * 2 + (w & 1)) * N; }, mc0, mc1, tid);
  __syncthreads();
  fold_tail(tiles, Y, Ysrc, tw_base, mc0, mc1, wave, lane);
}

// ---- fold, latency form -------------------------------------------------------
// For batches whose fold grids leave most CUs idle: level 0 runs k_fold_part<2>
// alone, every later level L runs one k_fold_fused launch, and one
// k_fold_finish closes the last level: nu2 + 1 launches instead of 2 nu2.
// A k_fold_fused block (pair i of level L, residue m, row j, digit half h,
// trial, query; 4 waves, one transform per SIMD in each phase) owns the key
// columns c = j + 2 k, k = 4 h .. 4 h + 3.  It first finishes level L - 1 for
// row j of the two elements the pair reads, X = element num_per + i and Y =
// element i of level L - 1:
//   raw(e)_j = src_acc[e]_j + from_ntt(sum over groups of src_part[e][m'][group][j])
// (wave 2 e' + m': one inverse transform each, CRT with the wave of the other
// residue), forms its 4 digit differences of X_j - Y_j from those raw values
// and runs the partial products of level L for its columns into dst_part
// [pair][m][group = 2 j + h][row][P] (16N words per pair).  The (m = 0, h = 0)
// blocks of a pair also write raw(Y) (row j each) to dst_acc[i], the Y that
// level L + 1 (or the finish) adds.  The blocks of a pair repeat the inverse
// transforms of their row (4 times), which costs nothing in a launch that
// leaves CUs idle.  src and dst buffers alternate between levels, so no block
// reads what another block of the same launch writes.
// NGP: groups of partial sums per pair at level L - 1 (2 after k_fold_part<2>).
template <int NGP>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
void k_fold_fused(const uint32_t *__restrict__ src_acc, size_t src_acc_stride, uint32_t *__restrict__ dst_acc,
                  size_t dst_acc_stride, const uint32_t *__restrict__ src_part, size_t src_part_stride,
                  uint32_t *__restrict__ dst_part, size_t dst_part_stride, int n_trials, int num_per,
                  const uint32_t *__restrict__ cpos, size_t key_qstride, const uint32_t *__restrict__ tw_base,
                  ModConst mc0, ModConst mc1) {
  constexpr int NW = 4, NK = 4, PPT = N / (NW * 64);
  __shared__ uint32_t tiles[NW][LDS_TILE];
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int i = blockIdx.x;
  const int h = blockIdx.y & 1, j = (blockIdx.y >> 1) & 1, m = (blockIdx.y >> 2) & 1, trial = blockIdx.y >> 3;
  const int set = blockIdx.z * n_trials + trial;
  // ---- finish of level L - 1, row j: tile 2 e + mm = sum of the partials of element e, residue mm
  const int e = wave >> 1, mm = wave & 1;
  // this wave's Y_old values (row j of element e), loaded ahead of the transforms
  uint32_t yo[2][16];
  {
    const uint32_t *sh = src_acc + (size_t)set * src_acc_stride + (size_t)(e ? i : num_per + i) * 4 * N +
                         (size_t)j * 2 * N;
#pragma unroll
    for (int ii = 0; ii < 16; ii++) {
      const int z = 64 * (16 * mm + ii) + lane;
      yo[0][ii] = sh[z];
      yo[1][ii] = sh[N + z];
    }
  }
  {
    const uint32_t *pb = src_part + (size_t)set * src_part_stride;
    sum_partials4<NGP, 0>(
        tiles,
        [&](int w, int gp) {
          return pb + (size_t)((w >> 1) ? i : num_per + i) * 4 * NGP * N + (size_t)(((w & 1) * NGP + gp) * 2 + j) * N;
        },
        mc0, mc1, tid);
  }
  __syncthreads();
  const ModConst mcm = mm ? mc1 : mc0;
  uint32_t s[32];
  load_p(s, tiles[wave], lane);
  wave_lds_sync();
  ntt_inv_wave_lazy(s, tw_base + (mm * 2 + 1) * TW_WORDS, tiles[wave], mcm.q, mcm.mu1, lane);
#pragma unroll
  for (int ii = 0; ii < 32; ii++) s[ii] = canon(shoup_mul(s[ii], mcm.crt_c, mcm.crt_cp, mcm.q), mcm.q);
  __syncthreads();
#pragma unroll
  for (int ii = 0; ii < 16; ii++) tiles[wave][ii * 64 + lane] = s[16 * (1 - mm) + ii];
  __syncthreads();
  // raw value of coefficient z = 64 (16 mm + ii) + lane of row j of element e
  uint64_t xr[16];
  {
    const uint32_t *xo = tiles[wave ^ 1];
#pragma unroll
    for (int ii = 0; ii < 16; ii++) {
      const uint32_t own = s[16 * mm + ii], oth = xo[ii * 64 + lane];
      const uint32_t x0 = mm ? oth : own, x1 = mm ? own : oth;
      uint64_t x = (uint64_t)x0 * Q1 + (uint64_t)x1 * Q0;
      x = (x >= QQ) ? x - QQ : x;
      x += ((uint64_t)yo[0][ii] << 32) | yo[1][ii];
      xr[ii] = (x >= QQ) ? x - QQ : x;
    }
  }
  __syncthreads();  // the exchange tiles are dead
  if (e == 1) {  // Y: raw values to LDS for the X waves; the (m = 0, h = 0) blocks store them
    uint32_t *ty = tiles[wave];
#pragma unroll
    for (int ii = 0; ii < 16; ii++) {
      ty[ii * 64 + lane] = (uint32_t)(xr[ii] >> 32);
      ty[1024 + ii * 64 + lane] = (uint32_t)xr[ii];
    }
    if (m == 0 && h == 0) {
      uint32_t *dh = dst_acc + (size_t)set * dst_acc_stride + (size_t)i * 4 * N + (size_t)j * 2 * N;
#pragma unroll
      for (int ii = 0; ii < 16; ii++) {
        const int z = 64 * (16 * mm + ii) + lane;
        dh[z] = (uint32_t)(xr[ii] >> 32);
        dh[N + z] = (uint32_t)xr[ii];
      }
    }
  }
  __syncthreads();
  uint32_t *dig = &tiles[0][0];  // [k - 4 h][DIG_WORDS], 2048 words in tile 0
  if (e == 0) {  // X: digits 4 h .. 4 h + 3 of X_j - Y_j, coefficient 64 (4 i4 + b) + lane at byte b of word i4
    const uint32_t *ty = tiles[wave + 2];
    uint32_t packed[NK][4];
#pragma unroll
    for (int kl = 0; kl < NK; kl++)
#pragma unroll
      for (int q4 = 0; q4 < 4; q4++) packed[kl][q4] = 0;
#pragma unroll
    for (int ii = 0; ii < 16; ii++) {
      const uint64_t x = xr[ii];
      const uint64_t y = ((uint64_t)ty[ii * 64 + lane] << 32) | ty[1024 + ii * 64 + lane];
#pragma unroll
      for (int kl = 0; kl < NK; kl++) {
        const int sh = 7 * (NK * h + kl);
        const int d = (int)((x >> sh) & 127u) - (int)((y >> sh) & 127u);
        packed[kl][ii >> 2] |= ((uint32_t)d & 0xFFu) << (8 * (ii & 3));
      }
    }
#pragma unroll
    for (int kl = 0; kl < NK; kl++)
#pragma unroll
      for (int q4 = 0; q4 < 4; q4++) dig[kl * DIG_WORDS + (4 * mm + q4) * 64 + lane] = packed[kl][q4];
  }
  __syncthreads();
  // ---- partial products of level L: wave w transforms the digit of column c = j + 2 (4 h + w)
  uint32_t dw[8];
#pragma unroll
  for (int i4 = 0; i4 < 8; i4++) dw[i4] = dig[wave * DIG_WORDS + i4 * 64 + lane];
  __syncthreads();  // the digit bytes are dead from here on
  const uint32_t q = m ? mc1.q : mc0.q, mu1 = m ? mc1.mu1 : mc0.mu1, qinv = m ? mc1.qinv_neg : mc0.qinv_neg;
  {
    uint32_t r[32];
#pragma unroll
    for (int i4 = 0; i4 < 8; i4++)
#pragma unroll
      for (int b = 0; b < 4; b++) r[4 * i4 + b] = (uint32_t)((int32_t)(int8_t)(dw[i4] >> (8 * b)) + (int32_t)q);
    // inputs in (q - 128, q + 128) < 2q: stages 1-7 and 9-11 skip the reduction (outputs < 10q)
    ntt_fwd_wave_t<7, true>(r, tw_base + (m * 2) * TW_WORDS, tiles[wave], q, mu1, lane);
    // stored lazy: a sum of at most 8 products with keys < q is below 80 q^2
    // < 2^63 and its REDC below 6q, which redc64 brings to [0, q)
    store_p(tiles[wave], r, lane);
  }
  __syncthreads();
  const uint32_t *kp = cpos + (size_t)blockIdx.z * key_qstride;
  uint64_t acc[2][PPT];
#pragma unroll
  for (int rr = 0; rr < 2; rr++)
#pragma unroll
    for (int ee = 0; ee < PPT; ee++) acc[rr][ee] = 0;
#pragma unroll
  for (int w2 = 0; w2 < NW; w2++) {
    const int c = j + 2 * (NK * h + w2);
    const uint32_t *k0 = kp + ((size_t)(m * 16 + c) * 2) * N + PPT * tid;
    const uint32_t *k1 = k0 + N;
    const uint32_t *tl = &tiles[w2][PPT * tid];
#pragma unroll
    for (int ee = 0; ee < PPT; ee++) {
      const uint32_t d = tl[ee];
      acc[0][ee] += (uint64_t)d * k0[ee];
      acc[1][ee] += (uint64_t)d * k1[ee];
    }
  }
  uint32_t *dst =
      dst_part + (size_t)set * dst_part_stride + ((((size_t)i * 2 + m) * 4 + 2 * j + h) * 2) * N + PPT * tid;
#pragma unroll
  for (int rr = 0; rr < 2; rr++)
#pragma unroll
    for (int ee = 0; ee < PPT; ee++) dst[(size_t)rr * N + ee] = redc64(acc[rr][ee], q, qinv);
}

// ---- Regev -> GSW conversion ------------------------------------------------
// Spiral's RegevToGSW (the reference declares the conversion key and the
// external-product shape "2 by 2t_conv" x "2 by 1 -> 2t_conv by 1",
// gpu-pir/structures.cuh:173-177, gadget.cuh:41-53, and leaves the function a
// stub, cpu_version/conversion.py:3).  Block (k, query) converts GSW slot
// k = l t_gsw + j: ct = v[2k + 1] (NTT form, layout P, canonical).
//   column 2j + 1 of level l = ct
//   column 2j     of level l = conv * NTT(g^-1(ct)), rdim 2, 4 digits of 14
//                              bits, key column c = row + 2 d
// Waves 0-3 take the four (row, residue) polynomials of ct out of the NTT
// domain; wave c then owns digit polynomial c, transforms it in each residue
// into its tile, and every thread multiply-accumulates the 8 tiles at its 4
// positions against the key (layout [m][c][row][P], 2^32-scaled).
// Digits are below 2^14 and 16 q + 2^14 < 2^32 for both moduli, so the first
// 8 forward stages skip the reduction (outputs < 8q, reduced when stored).
// MONT: outputs 2^32-scaled (the fold's key form), else standard form.
// only (optional): per-query flags; queries whose flag is 0 are left alone.
constexpr int CONV_COLS = 8;  // 2 t_conv, t_conv = 4

template <bool MONT, int M>
__device__ __forceinline__ void conv_residue(const uint32_t (&dp)[16], uint32_t (*tiles)[LDS_TILE],
                                             const uint32_t *__restrict__ kc, uint32_t *__restrict__ out_col,
                                             const uint32_t *__restrict__ tw_base, const ModConst &mc, int wave,
                                             int lane, int tid) {
  constexpr int PPT = N / 512;
  uint32_t r[32];
#pragma unroll
  for (int i = 0; i < 32; i++) r[i] = (dp[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
  ntt_fwd_wave_t<8, true>(r, tw_base + (M * 2) * TW_WORDS, tiles[wave], mc.q, mc.mu1, lane);
  store_tile_p_full(tiles[wave], r, lane, mc.q, mc.mu1);
  __syncthreads();
  uint64_t acc[2][PPT];
#pragma unroll
  for (int row = 0; row < 2; row++)
#pragma unroll
    for (int e = 0; e < PPT; e++) acc[row][e] = 0;
#pragma unroll
  for (int c = 0; c < CONV_COLS; c++) {
    const uint32_t *k0 = kc + ((size_t)(M * CONV_COLS + c) * 2) * N + PPT * tid;
    const uint32_t *k1 = k0 + N;
    const uint32_t *tl = &tiles[c][PPT * tid];
#pragma unroll
    for (int e = 0; e < PPT; e++) {
      const uint32_t d = tl[e];
      acc[0][e] += (uint64_t)d * k0[e];
      acc[1][e] += (uint64_t)d * k1[e];
    }
  }
#pragma unroll
  for (int row = 0; row < 2; row++)
#pragma unroll
    for (int e = 0; e < PPT; e++) {
      uint32_t x = redc64(acc[row][e], mc.q, mc.qinv_neg);
      if (MONT) x = canon(shoup_mul(x, mc.two32, mc.two32_p, mc.q), mc.q);
      out_col[(size_t)row * N + PPT * tid + e] = x;
    }
  __syncthreads();
}

template <bool MONT>
__global__ __launch_bounds__(512) void k_regev_to_gsw(const uint32_t *__restrict__ v, size_t v_q, int t_gsw,
                                                      const uint32_t *__restrict__ conv, size_t conv_q,
                                                      uint32_t *__restrict__ fk, size_t fk_q,
                                                      const int *__restrict__ only,
                                                      const uint32_t *__restrict__ tw_base, ModConst mc0,
                                                      ModConst mc1) {
  if (only && !only[blockIdx.y]) return;  // uniform over the block
  __shared__ uint32_t tiles[CONV_COLS][LDS_TILE];
  uint32_t *xm = &tiles[4][0];  // CRT-scaled residues [row][m][N] (tiles 4-7)
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int k = blockIdx.x, l = k / t_gsw, j = k % t_gsw;
  const int mx = 2 * t_gsw;
  const uint32_t *ct = v + (size_t)blockIdx.y * v_q + (size_t)(2 * k + 1) * 4 * N;
  const uint32_t *kc = conv + (size_t)blockIdx.y * conv_q;
  uint32_t *lv = fk + (size_t)blockIdx.y * fk_q + (size_t)l * 2 * mx * 2 * N;  // level l: [m][col][row][P]

  for (int u = tid; u < N; u += 512) {  // column 2j + 1 = ct: 4 polys x N/4 uint4
    const int poly = u / (N / 4), k4 = u % (N / 4), row = poly >> 1, m = poly & 1;
    uint4 x = reinterpret_cast<const uint4 *>(ct + (size_t)poly * N)[k4];
    if (MONT) {
      const ModConst mc = m ? mc1 : mc0;
      x.x = canon(shoup_mul(x.x, mc.two32, mc.two32_p, mc.q), mc.q);
      x.y = canon(shoup_mul(x.y, mc.two32, mc.two32_p, mc.q), mc.q);
      x.z = canon(shoup_mul(x.z, mc.two32, mc.two32_p, mc.q), mc.q);
      x.w = canon(shoup_mul(x.w, mc.two32, mc.two32_p, mc.q), mc.q);
    }
    reinterpret_cast<uint4 *>(lv + ((size_t)(m * mx + 2 * j + 1) * 2 + row) * N)[k4] = x;
  }
  if (wave < 4) {
    const int row = wave >> 1, m = wave & 1;
    const ModConst mc = m ? mc1 : mc0;
    uint32_t r[32];
    load_p(r, ct + (size_t)(row * 2 + m) * N, lane);
    ntt_inv_wave_lazy(r, tw_base + (m * 2 + 1) * TW_WORDS, tiles[wave], mc.q, mc.mu1, lane);
#pragma unroll
    for (int i = 0; i < 32; i++)
      xm[(row * 2 + m) * N + 64 * i + lane] = canon(shoup_mul(r[i], mc.crt_c, mc.crt_cp, mc.q), mc.q);
  }
  __syncthreads();
  // Digit polynomial c = wave: row c & 1, digit d = c >> 1 of the raw value
  // x = CRT(x0, x1) in [0, Q); coefficient 64 i + lane, two digits per word.
  const int drow = wave & 1, shift = 14 * (wave >> 1);
  uint32_t dp[16];
#pragma unroll
  for (int i2 = 0; i2 < 16; i2++) {
    uint32_t w = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int z = 64 * (2 * i2 + h) + lane;
      uint64_t x = (uint64_t)xm[(drow * 2) * N + z] * Q1 + (uint64_t)xm[(drow * 2 + 1) * N + z] * Q0;
      x = (x >= QQ) ? x - QQ : x;
      w |= (uint32_t)((x >> shift) & 0x3FFFu) << (16 * h);
    }
    dp[i2] = w;
  }
  __syncthreads();  // the residues in tiles 4-7 are dead: every wave now uses its own tile
  uint32_t *col = lv + (size_t)(2 * j) * 2 * N;
  conv_residue<MONT, 0>(dp, tiles, kc, col, tw_base, mc0, wave, lane, tid);
  conv_residue<MONT, 1>(dp, tiles, kc, col + (size_t)mx * 2 * N, tw_base, mc1, wave, lane, tid);
}

hipError_t launch_regev_to_gsw(bool mont, const uint32_t *v, size_t v_q, int n_queries, int nu2, int t_gsw,
                               const uint32_t *conv, size_t conv_q, uint32_t *fk, size_t fk_q, const int *only,
                               const DevTables &t, hipStream_t s) {
  const dim3 grid((unsigned)(nu2 * t_gsw), (unsigned)n_queries);
  if (mont)
    k_regev_to_gsw<true><<<grid, 512, 0, s>>>(v, v_q, t_gsw, conv, conv_q, fk, fk_q, only, t.tw, t.mc[0], t.mc[1]);
  else
    k_regev_to_gsw<false><<<grid, 512, 0, s>>>(v, v_q, t_gsw, conv, conv_q, fk, fk_q, only, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

// ---- response packing + modulus switching ---------------------------------
// Spiral's response path (not in the reference): the n^2 = 4 scalar Regev
// responses (trial t = r n + c, raw) of a query become one 3 x 2 matrix
// ciphertext; column c = from_ntt(sum_r W_r NTT(g^-1(a_{2r+c}))) + (0, b_c,
// b_{2+c}) with 4 digits of 14 bits (rdim 1), W_r the packing keys
// [r][m][k][row][P] (2^32-scaled).  Row 0 is rescaled to 2^q2_bits, rows 1-2
// to 2^10 = 4 p, and all three polynomials of the column are bit-packed into
// the compact response.  Block (c, query): wave 4 r + d owns digit d of
// a_{2r+c}; after the per-residue transforms and multiply-accumulates, waves
// 0-5 take the six (row, residue) sums out of the NTT domain.
constexpr int PACK_N = 2;      // plaintext matrix dimension n (n^2 = 4 sub-databases)
constexpr int PACK_ROWS = PACK_N + 1;
constexpr int PACK_Q1_BITS = 10;

// floor((x 2^bits + floor(Q/2)) / Q) mod 2^bits for x < Q, bits <= 32, exactly:
// a double-precision estimate (off by at most 2) corrected with 128-bit products.
__device__ __forceinline__ uint32_t switch_modulus(uint64_t x, int bits, double scale) {
  typedef unsigned __int128 u128;
  const u128 num = ((u128)x << bits) + (QQ >> 1);
  uint64_t est = (uint64_t)((double)x * scale);
  u128 prod = (u128)est * QQ;
  for (int i = 0; i < 3 && prod > num; i++) {
    est--;
    prod -= QQ;
  }
  for (int i = 0; i < 3 && prod + QQ <= num; i++) {
    est++;
    prod += QQ;
  }
  return (uint32_t)(est & ((1ull << bits) - 1));
}

template <int M>
__device__ __forceinline__ void pack_residue(const uint32_t (&dp)[16], uint32_t (*tiles)[LDS_TILE],
                                             const uint32_t *__restrict__ wk, uint32_t (&res)[PACK_ROWS][N / 512],
                                             const uint32_t *__restrict__ tw_base, const ModConst &mc, int wave,
                                             int lane, int tid) {
  constexpr int PPT = N / 512;
  uint32_t r[32];
#pragma unroll
  for (int i = 0; i < 32; i++) r[i] = (dp[i >> 1] >> (16 * (i & 1))) & 0xFFFFu;
  ntt_fwd_wave_t<8, true>(r, tw_base + (M * 2) * TW_WORDS, tiles[wave], mc.q, mc.mu1, lane);
  store_tile_p_full(tiles[wave], r, lane, mc.q, mc.mu1);
  __syncthreads();
  uint64_t acc[PACK_ROWS][PPT];
#pragma unroll
  for (int row = 0; row < PACK_ROWS; row++)
#pragma unroll
    for (int e = 0; e < PPT; e++) acc[row][e] = 0;
#pragma unroll
  for (int w2 = 0; w2 < 8; w2++) {
    const int r2 = w2 >> 2, d2 = w2 & 3;
    const uint32_t *tl = &tiles[w2][PPT * tid];
#pragma unroll
    for (int row = 0; row < PACK_ROWS; row++) {
      const uint32_t *kk = wk + ((((size_t)r2 * 2 + M) * 4 + d2) * PACK_ROWS + row) * N + PPT * tid;
#pragma unroll
      for (int e = 0; e < PPT; e++) acc[row][e] += (uint64_t)tl[e] * kk[e];
    }
  }
#pragma unroll
  for (int row = 0; row < PACK_ROWS; row++)
#pragma unroll
    for (int e = 0; e < PPT; e++) res[row][e] = redc64(acc[row][e], mc.q, mc.qinv_neg);
  __syncthreads();
}

__global__ __launch_bounds__(512) void k_pack_response(const uint32_t *__restrict__ raw, size_t raw_q,
                                                       size_t trial_stride, const uint32_t *__restrict__ pk,
                                                       size_t pk_q, const int *__restrict__ only,
                                                       uint8_t *__restrict__ out, size_t out_q, int q2_bits,
                                                       double scale2, double scale1,
                                                       const uint32_t *__restrict__ tw_base, ModConst mc0,
                                                       ModConst mc1) {
  constexpr int PPT = N / 512;
  const int c = blockIdx.x, b = blockIdx.y;
  if (only && !only[b]) return;  // uniform over the block
  __shared__ uint32_t tiles[8][LDS_TILE];
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const uint32_t *base = raw + (size_t)b * raw_q;
  const uint32_t *wk = pk + (size_t)b * pk_q;
  // digit d = wave & 3 of a_{r n + c}, r = wave >> 2; coefficient 64 i + lane, two per word
  const uint32_t *a = base + (size_t)((wave >> 2) * PACK_N + c) * trial_stride;
  const int shift = 14 * (wave & 3);
  uint32_t dp[16];
#pragma unroll
  for (int i2 = 0; i2 < 16; i2++) {
    uint32_t w = 0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int z = 64 * (2 * i2 + h) + lane;
      const uint64_t x = ((uint64_t)a[z] << 32) | a[N + z];
      w |= (uint32_t)((x >> shift) & 0x3FFFu) << (16 * h);
    }
    dp[i2] = w;
  }
  uint32_t res[2][PACK_ROWS][PPT];
  pack_residue<0>(dp, tiles, wk, res[0], tw_base, mc0, wave, lane, tid);
  pack_residue<1>(dp, tiles, wk, res[1], tw_base, mc1, wave, lane, tid);
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int row = 0; row < PACK_ROWS; row++)
#pragma unroll
      for (int e = 0; e < PPT; e++) tiles[row * 2 + m][PPT * tid + e] = res[m][row][e];
  __syncthreads();
  if (wave < 2 * PACK_ROWS) {  // poly (row, m) = tile wave: out of the NTT domain, CRT-scaled, coefficient order
    const int m = wave & 1;
    const ModConst mc = m ? mc1 : mc0;
    uint32_t sv[32];
    load_p(sv, tiles[wave], lane);
    wave_lds_sync();
    ntt_inv_wave_lazy(sv, tw_base + (m * 2 + 1) * TW_WORDS, tiles[wave], mc.q, mc.mu1, lane);
#pragma unroll
    for (int ii = 0; ii < 32; ii++) tiles[wave][64 * ii + lane] = canon(shoup_mul(sv[ii], mc.crt_c, mc.crt_cp, mc.q), mc.q);
  }
  __syncthreads();
  // Three polys x 256 groups of 8 coefficients: CRT, + b, rescale, bit-pack (8 * bits bits = bits bytes).
  uint8_t *ob = out + (size_t)b * out_q;
  for (int task = tid; task < PACK_ROWS * 256; task += 512) {
    const int row = task >> 8, grp = task & 255;
    const int bits = row ? PACK_Q1_BITS : q2_bits;
    uint8_t *dst = ob + (row == 0 ? (size_t)c * N * q2_bits / 8
                                  : (size_t)PACK_N * N * q2_bits / 8 +
                                        (size_t)((row - 1) * PACK_N + c) * N * PACK_Q1_BITS / 8) +
                   (size_t)grp * bits;
    const uint32_t *bb = base + (size_t)((row > 0 ? row - 1 : 0) * PACK_N + c) * trial_stride + 2 * N;
    uint64_t bitbuf = 0;
    int nbits = 0, nb = 0;
    for (int k = 0; k < 8; k++) {
      const int z = grp * 8 + k;
      uint64_t x = (uint64_t)tiles[2 * row][z] * Q1 + (uint64_t)tiles[2 * row + 1][z] * Q0;
      x = (x >= QQ) ? x - QQ : x;
      if (row) {
        x += ((uint64_t)bb[z] << 32) | bb[N + z];
        x = (x >= QQ) ? x - QQ : x;
      }
      bitbuf |= (uint64_t)switch_modulus(x, bits, row ? scale1 : scale2) << nbits;
      nbits += bits;
      while (nbits >= 8) {
        dst[nb++] = (uint8_t)bitbuf;
        bitbuf >>= 8;
        nbits -= 8;
      }
    }
  }
}

size_t packed_response_bytes(int q2_bits) {
  return (size_t)PACK_N * N * q2_bits / 8 + (size_t)PACK_N * PACK_N * N * PACK_Q1_BITS / 8;
}

hipError_t launch_pack_response(const uint32_t *raw, size_t raw_q, size_t trial_stride, int n_queries,
                                const uint32_t *pk, size_t pk_q, const int *only, uint8_t *out, size_t out_q,
                                int q2_bits, const DevTables &t, hipStream_t s) {
  const double scale2 = std::ldexp(1.0, q2_bits) / (double)QQ;
  const double scale1 = std::ldexp(1.0, PACK_Q1_BITS) / (double)QQ;
  k_pack_response<<<dim3(PACK_N, (unsigned)n_queries), 512, 0, s>>>(raw, raw_q, trial_stride, pk, pk_q, only, out,
                                                                     out_q, q2_bits, scale2, scale1, t.tw, t.mc[0],
                                                                     t.mc[1]);
  return hipGetLastError();
}

// Expansion phases 1-2 for element v_i: waves 0/1 take v_i.a out of the NTT
// domain (one residue each, CRT halves into xm = tiles 2-3); then all threads
// form the CRT value, apply tau_t (a negated zero becomes Q) and write the 8
// digit bytes of every coefficient into dig = tiles 4-5.
template <int NT>
__device__ __forceinline__ void expand_digits(const uint32_t *__restrict__ vi, uint32_t (*tiles)[LDS_TILE],
                                              uint32_t t, const uint32_t *__restrict__ tw_base, const ModConst &mc0,
                                              const ModConst &mc1, int wave, int lane, int tid) {
  uint32_t *xm = tiles[2];
  if (wave < 2) {
    const int m = wave;
    const ModConst mc = m ? mc1 : mc0;
    uint32_t rr[32];
    load_p(rr, vi + (size_t)m * N, lane);
    ntt_inv_wave_lazy(rr, tw_base + (m * 2 + 1) * TW_WORDS, tiles[wave], mc.q, mc.mu1, lane);
#pragma unroll
    for (int ii = 0; ii < 32; ii++) xm[m * N + 64 * ii + lane] = canon(shoup_mul(rr[ii], mc.crt_c, mc.crt_cp, mc.q), mc.q);
  }
  __syncthreads();
  uint8_t *db = reinterpret_cast<uint8_t *>(tiles[4]);
  for (int z = tid; z < N; z += NT) {
    uint64_t x = (uint64_t)xm[z] * Q1 + (uint64_t)xm[N + z] * Q0;
    x = (x >= QQ) ? x - QQ : x;
    const uint32_t zt = (uint32_t)z * t;
    const uint32_t dst = zt & (N - 1);
    const uint64_t y = ((zt >> LOGN) & 1) ? QQ - x : x;
    const uint32_t word = (dst >> 8) * 64 + (dst & 63), byte = (dst >> 6) & 3;
#pragma unroll
    for (int k = 0; k < 8; k++) db[((size_t)k * DIG_WORDS + word) * 4 + byte] = (uint8_t)((y >> (7 * k)) & 127u);
  }
  __syncthreads();
}

// Expansion, multiply-accumulate of residue m at this thread's PPT positions
// (P index p0 + e) over the 8 digit tiles tiles[tile0 + k], then
// v_i += W g^-1(tau a) + (0, NTT(tau as_e)) for residue m.  tau(as_e) in the
// NTT domain is a permutation of positions; it is gathered before `bar`
// (a block barrier), after which the residue is written.
// The 8 digit tiles hold the forward transforms' lazy outputs (< 8q, not
// reduced): with keys < q each sum of 8 products is below 64 q^2 < 2^62, and
// its REDC, taken without a reduction of its own, is below 5q.  The old value
// of v_i (< q) and, on row 1, the gathered e term (< q) are added to it (< 7q
// < 8q), and one reduction (conditional subtractions of 4q, 2q, q) brings the
// sum to [0, q).  Precondition: v_i, and so the gathered e term, canonical.
// The e-term gather of expand_update: tau(as_e) at this thread's positions,
// from row 1 of residue m of v_i (`row`, layout P, in global memory or LDS).
template <int PPT>
__device__ __forceinline__ void expand_gather_e(uint32_t (&ase)[PPT], const uint32_t *row, int p0, uint32_t t) {
#pragma unroll
  for (int e = 0; e < PPT; e++) {
    const uint32_t z = (uint32_t)z_of_p(p0 + e);
    const uint32_t ev = ((2 * brv11(z) + 1) * t) & (2 * N - 1);
    ase[e] = row[p_of_z((int)brv11((ev - 1) >> 1))];
  }
}

// Key words of digits k < KP (rows 0 and 1) at this thread's 4 positions,
// loaded ahead of the digit transforms by callers that have the registers,
// and the gathered e term (filled by the caller).
template <int KP>
struct ExpandKeys {
  uint4 k[KP > 0 ? KP : 1][2];
  uint32_t ase[4];
};
template <int KP>
__device__ __forceinline__ void expand_prefetch_keys(ExpandKeys<KP> &pk, const uint32_t *__restrict__ W, int m,
                                                     int p0) {
#pragma unroll
  for (int k = 0; k < KP; k++) {
    const uint32_t *k0 = W + ((size_t)(m * 8 + k) * 2) * N + p0;
    pk.k[k][0] = *reinterpret_cast<const uint4 *>(k0);
    pk.k[k][1] = *reinterpret_cast<const uint4 *>(k0 + N);
  }
}

// tile_of(k): this thread's words of digit k's transform in LDS (PPT of them);
// key_of(k, row): its key words of digit k and row `row` (global memory or LDS).
template <int PPT, int KP = 0, class TileOf, class KeyOf>
__device__ __forceinline__ void expand_update_tk(uint32_t *__restrict__ vi, TileOf tile_of, KeyOf key_of, int m, int p0,
                                                 uint32_t t, uint32_t q, uint32_t qinv,
                                                 const ExpandKeys<KP> *pk = nullptr) {
  static_assert(KP == 0 || PPT == 4, "prefetched keys cover 4 positions");
  uint32_t ase[PPT];
  if (pk) {  // gathered ahead by the caller (4 positions)
#pragma unroll
    for (int e = 0; e < PPT; e++) ase[e] = pk->ase[e & 3];
  } else {
    expand_gather_e<PPT>(ase, vi + (size_t)(2 + m) * N, p0, t);
  }
  uint64_t acc[2][PPT];
#pragma unroll
  for (int row = 0; row < 2; row++)
#pragma unroll
    for (int e = 0; e < PPT; e++) acc[row][e] = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const uint32_t *k0 = key_of(k, 0), *k1 = key_of(k, 1);
    const uint32_t *tl = tile_of(k);
    if (k < KP) {
      const uint4 a = pk->k[k < KP ? k : 0][0], b = pk->k[k < KP ? k : 0][1];
      const uint32_t ka[4] = {a.x, a.y, a.z, a.w}, kb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
      for (int e = 0; e < PPT; e++) {
        const uint32_t d = tl[e];
        acc[0][e] += (uint64_t)d * ka[e & 3];
        acc[1][e] += (uint64_t)d * kb[e & 3];
      }
    } else {
#pragma unroll
      for (int e = 0; e < PPT; e++) {
        const uint32_t d = tl[e];
        acc[0][e] += (uint64_t)d * k0[e];
        acc[1][e] += (uint64_t)d * k1[e];
      }
    }
  }
  __syncthreads();  // tiles read and row 1 of residue m gathered by every thread
#pragma unroll
  for (int row = 0; row < 2; row++) {
    uint32_t *dst = vi + (size_t)(row * 2 + m) * N + p0;
#pragma unroll
    for (int e = 0; e < PPT; e++) {
      // REDC without its own reduction: below 5q (sums below 64 q^2); with
      // the old value (< q) and, on row 1, the gathered e term (< q): below 7q.
      const uint32_t mq = (uint32_t)acc[row][e] * qinv;
      uint32_t x = dst[e] + (uint32_t)((acc[row][e] + (uint64_t)mq * q) >> 32);
      if (row) x += ase[e];
      x = min(x, x - 4 * q);
      x = min(x, x - 2 * q);
      dst[e] = canon(x, q);
    }
  }
}

template <int PPT, int KP = 0, class TileOf>
__device__ __forceinline__ void expand_update_t(uint32_t *__restrict__ vi, const uint32_t *__restrict__ W,
                                                TileOf tile_of, int m, int p0, uint32_t t, uint32_t q, uint32_t qinv,
                                                const ExpandKeys<KP> *pk = nullptr) {
  expand_update_tk<PPT, KP>(
      vi, tile_of, [&](int k, int row) { return W + ((size_t)(m * 8 + k) * 2 + row) * N + p0; }, m, p0, t, q, qinv,
      pk);
}
template <int PPT, int KP = 0>
__device__ __forceinline__ void expand_update(uint32_t *__restrict__ vi, const uint32_t *__restrict__ W,
                                              uint32_t (*tiles)[LDS_TILE], int tile0, int m, int p0, uint32_t t,
                                              uint32_t q, uint32_t qinv, const ExpandKeys<KP> *pk = nullptr) {
  expand_update_t<PPT, KP>(vi, W, [&](int k) { return &tiles[tile0 + k][p0]; }, m, p0, t, q, qinv, pk);
}

// Expansion step (gpu-pir/expansion.cuh:87-166) for element i of round r,
// gadget 2^7 x 8.  The grid holds only working steps: below stop_round block b
// is step i = b; from stop_round on, blocks [0, 2^r) are the even steps 2b and
// the rest the odd steps 2(b - 2^r) + 1 that stop_round still computes (odd
// steps past it are skipped, expansion.cuh:92-97).  After phases 1-2, wave w
// keeps digit k = w mod 8 packed in registers.  NW = 8: two passes (residue 0, then 1), each one transform per
// wave and an update of that residue by all threads.  NW = 16: waves 8-15
// take residue 1 in the same pass; threads 0-511 update residue 0 and
// 512-1023 residue 1.
template <int NW>
__global__ __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(4)))
void k_expand8(uint32_t *__restrict__ v, size_t q_stride, int r, int stop, int max_bits,
               const uint32_t *__restrict__ left, const uint32_t *__restrict__ right, size_t left_qstride,
               size_t right_qstride, const uint32_t *__restrict__ tw_base, ModConst mc0, ModConst mc1) {
  constexpr int NT = NW * 64;
  static_assert(NW == 8 || NW == 16, "8 or 16 waves");
  __shared__ uint32_t tiles[NW][LDS_TILE];
  const int b = blockIdx.x, half = 1 << r;
  const int i = (stop > 0 && r >= stop) ? (b < half ? 2 * b : 2 * (b - half) + 1) : b;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  uint32_t *vi = v + (size_t)blockIdx.y * q_stride + (size_t)i * 4 * N;
  const uint32_t t = (uint32_t)(N >> r) + 1u;
  const bool use_left = (r != 0) && ((i & 1) == 0);
  const uint32_t *W = use_left ? left + (size_t)blockIdx.y * left_qstride : right + (size_t)blockIdx.y * right_qstride;

  expand_digits<NT>(vi, tiles, t, tw_base, mc0, mc1, wave, lane, tid);
  uint32_t dw[8];  // digit wave mod 8: byte b of word i4 is coefficient 64 (4 i4 + b) + lane
  const uint32_t *dig = tiles[4];
#pragma unroll
  for (int i4 = 0; i4 < 8; i4++) dw[i4] = dig[(wave & 7) * DIG_WORDS + i4 * 64 + lane];
  __syncthreads();  // the digit bytes are dead from here on

#pragma unroll
  for (int pass = 0; pass < 16 / NW; pass++) {
    const int m = NW == 16 ? (wave >> 3) : pass;  // residue transformed by this wave
    {
      const uint32_t q = m ? mc1.q : mc0.q;
      uint32_t rr[32];
#pragma unroll
      for (int i4 = 0; i4 < 8; i4++)
#pragma unroll
        for (int b = 0; b < 4; b++) rr[4 * i4 + b] = (dw[i4] >> (8 * b)) & 0xFFu;
      const uint32_t *twm = tw_base + (m * 2) * TW_WORDS;
      asm volatile("" : "+s"(twm));  // re-read per pass, not held in registers
      // digits < 128: stages 1-8 and 10-11 skip the reduction (outputs < 8q)
      const uint32_t mu1 = m ? mc1.mu1 : mc0.mu1;
      ntt_fwd_wave_t<8, true>(rr, twm, tiles[wave], q, mu1, lane);
      store_p(tiles[wave], rr, lane);  // lazy (< 8q): expand_update reduces after the MAC
    }
    __syncthreads();
    // residue updated by this thread, and its positions
    const int mu = NW == 16 ? (tid >> 9) : pass;
    const int p0 = 4 * (tid & 511);
    expand_update<4>(vi, W, tiles, NW == 16 ? 8 * mu : 0, mu, p0, t, mu ? mc1.q : mc0.q,
                     mu ? mc1.qinv_neg : mc0.qinv_neg);
  }
}

// ---- split expansion round: digits kernel + update kernel -------------------
// Round r, gadget 2^7 x 8, in two launches so that one step spreads over
// three blocks (one for its digits, one per residue for its transforms).
//
// Element i is skipped in round r: odd steps past stop_round (expansion.cuh:92-97).
__device__ __forceinline__ bool expand_skipped(int i, int r, int stop, int max_bits) {
  return stop > 0 && (i & 1) && (r > stop || (r == stop && (i >> 1) >= max_bits));
}

// dst (residue m, rows 0-1) = v_neg1 (.) src at 4 consecutive words from p0; the
// row-0 result also goes to snap (if given).
__device__ __forceinline__ void neg1_rows4(const uint32_t *__restrict__ src, uint32_t *__restrict__ dst,
                                           uint32_t *__restrict__ snap, const uint32_t *__restrict__ neg1,
                                           int m, int p0, const ModConst &mc) {
  const uint4 w = *reinterpret_cast<const uint4 *>(neg1 + (size_t)m * 2 * N + p0);
  const uint4 wp = *reinterpret_cast<const uint4 *>(neg1 + (size_t)m * 2 * N + N + p0);
#pragma unroll
  for (int row = 0; row < 2; row++) {
    const uint4 x = *reinterpret_cast<const uint4 *>(src + (size_t)(row * 2 + m) * N + p0);
    uint4 y;
    y.x = canon(shoup_mul(x.x, w.x, wp.x, mc.q), mc.q);
    y.y = canon(shoup_mul(x.y, w.y, wp.y, mc.q), mc.q);
    y.z = canon(shoup_mul(x.z, w.z, wp.z, mc.q), mc.q);
    y.w = canon(shoup_mul(x.w, w.w, wp.w, mc.q), mc.q);
    *reinterpret_cast<uint4 *>(dst + (size_t)(row * 2 + m) * N + p0) = y;
    if (row == 0 && snap) *reinterpret_cast<uint4 *>(snap + (size_t)m * N + p0) = y;
  }
}

// k_expand_digits: one 2-wave block per element i of the round.  For
// i >= 2^r the element is formed first as v[i] = v_neg1[r] * v[i - 2^r] (all
// four polynomials, written back; the reference computes the product for
// every i, expansion.cuh:171-219), unless children_formed says that the
// previous round's launch already formed the working children (then only the
// extra blocks form theirs).  For a working step, wave m then takes
// residue m of v_i.a out of the NTT domain, and the block forms the CRT
// value, applies tau_t (a negated zero becomes Q) and writes the 8 digit
// bytes of every coefficient to dig[b] (b = the step's index in the working
// order of k_expand8, layout as the digit bytes there).  Grid x: [0, work)
// the working steps, then `extra` elements that only take the v_neg1
// product: the odd i >= 2^r skipped at this round, i = odd_start + 2e + 1.
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4)))
void k_expand_digits(uint32_t *__restrict__ v, size_t q_stride, int r, int stop,
                                                       int work, int odd_start, int children_formed,
                                                       const int *__restrict__ odd_live,
                                                       const uint32_t *__restrict__ neg1_r,
                                                       uint32_t *__restrict__ dig, size_t dig_qstride,
                                                       const uint32_t *__restrict__ tw_base, ModConst mc0,
                                                       ModConst mc1) {
  __shared__ uint32_t tiles[2][LDS_TILE];
  const int b = blockIdx.x, half = 1 << r;
  const bool step = b < work;
  const int i = !step ? odd_start + 2 * (b - work) + 1
                      : ((stop > 0 && r >= stop) ? (b < half ? 2 * b : 2 * (b - half) + 1) : b);
  if (odd_live && (i & 1) && !odd_live[blockIdx.y]) return;  // uniform per block: an odd step nothing reads
  const int tid = threadIdx.x;
  const int m = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const ModConst mc = m ? mc1 : mc0;
  uint32_t *vq = v + (size_t)blockIdx.y * q_stride;
  uint32_t *vi = vq + (size_t)i * 4 * N;
  uint32_t a[32];
  if (i >= half && !(step && children_formed)) {
    // wave m: rows 0 and 1 of residue m of v_neg1 * v[i - 2^r]
    const uint32_t *src = vq + (size_t)(i - half) * 4 * N;
    const uint4 *w4 = reinterpret_cast<const uint4 *>(neg1_r + (size_t)m * 2 * N);
    const uint4 *wp4 = reinterpret_cast<const uint4 *>(neg1_r + (size_t)m * 2 * N + N);
#pragma unroll
    for (int row = 0; row < 2; row++) {
      const uint4 *s4 = reinterpret_cast<const uint4 *>(src + (size_t)(row * 2 + m) * N);
      uint4 *d4 = reinterpret_cast<uint4 *>(vi + (size_t)(row * 2 + m) * N);
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const uint4 x = s4[c * 64 + lane], w = w4[c * 64 + lane], wp = wp4[c * 64 + lane];
        uint4 y;
        y.x = canon(shoup_mul(x.x, w.x, wp.x, mc.q), mc.q);
        y.y = canon(shoup_mul(x.y, w.y, wp.y, mc.q), mc.q);
        y.z = canon(shoup_mul(x.z, w.z, wp.z, mc.q), mc.q);
        y.w = canon(shoup_mul(x.w, w.w, wp.w, mc.q), mc.q);
        d4[c * 64 + lane] = y;
      }
    }
  }
  if (!step) return;  // uniform per block
  load_p(a, vi + (size_t)m * N, lane);  // (a child just formed above is this lane's own stores)
  ntt_inv_wave_lazy(a, tw_base + (m * 2 + 1) * TW_WORDS, tiles[m], mc.q, mc.mu1, lane);
#pragma unroll
  for (int ii = 0; ii < 32; ii++) a[ii] = canon(shoup_mul(a[ii], mc.crt_c, mc.crt_cp, mc.q), mc.q);
  // Wave m combines the coefficients 64 (16 m + jj) + lane: it hands the
  // other half to the other wave through its own tile and takes the other
  // residue of its half from the other wave's tile (tiles of 2112 words: the
  // block needs no separate CRT buffer, so more blocks fit per CU).
  uint32_t own[16];
#pragma unroll
  for (int jj = 0; jj < 16; jj++) {
    own[jj] = m ? a[16 + jj] : a[jj];
    tiles[m][jj * 64 + lane] = m ? a[jj] : a[16 + jj];
  }
  __syncthreads();
  uint32_t oth[16];
#pragma unroll
  for (int jj = 0; jj < 16; jj++) oth[jj] = tiles[1 - m][jj * 64 + lane];
  __syncthreads();  // the tiles become the digit bytes
  const uint32_t t = (uint32_t)(N >> r) + 1u;
  uint8_t *db = reinterpret_cast<uint8_t *>(&tiles[0][0]);  // 4096 words over tiles 0-1
#pragma unroll
  for (int jj = 0; jj < 16; jj++) {
    const int z = 64 * (16 * m + jj) + lane;
    const uint32_t x0 = m ? oth[jj] : own[jj], x1 = m ? own[jj] : oth[jj];
    uint64_t x = (uint64_t)x0 * Q1 + (uint64_t)x1 * Q0;
    x = (x >= QQ) ? x - QQ : x;
    const uint32_t zt = (uint32_t)z * t;
    const uint32_t dst = zt & (N - 1);
    const uint64_t y = ((zt >> LOGN) & 1) ? QQ - x : x;
    const uint32_t word = (dst >> 8) * 64 + (dst & 63), byte = (dst >> 6) & 3;
#pragma unroll
    for (int k = 0; k < 8; k++) db[((size_t)k * DIG_WORDS + word) * 4 + byte] = (uint8_t)((y >> (7 * k)) & 127u);
  }
  __syncthreads();
  const uint4 *src4 = reinterpret_cast<const uint4 *>(&tiles[0][0]);
  uint4 *dst4 = reinterpret_cast<uint4 *>(dig + (size_t)blockIdx.y * dig_qstride + (size_t)b * 8 * DIG_WORDS);
  for (int u = tid; u < 8 * DIG_WORDS / 4; u += 128) dst4[u] = src4[u];
}

// k_expand_digits2: the digits of a working parent p < 2^r and of its child
// c = p + 2^r in one 2-wave block.  The child is formed from p before round
// r's updates, v_c = v_neg1[r] (.) v_p, so v_c.a = v_p.a * x^(-2^r): its
// coefficients are p's shifted down by 2^r, those that wrap around negated
// (Q - x; 0 stays 0, as the canonical inverse transform of v_c.a gives),
// and one pair of inverse transforms serves both steps.  Grid x: [0, npair)
// the working parents (below stop_round all p < 2^r; from stop_round on the
// even ones, then the odd ones that work in the round), then `extra` blocks
// that only form the v_neg1 children of skipped odd elements i = odd_start +
// 2e + 1, as in k_expand_digits.  Digit slot of element i: its step index in
// the update kernel's order (from stop_round on, even steps first).
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(4)))
void k_expand_digits2(uint32_t *__restrict__ v, size_t q_stride, int r, int stop, int max_bits, int npair,
                      int odd_start, int children_formed, const int *__restrict__ odd_live,
                      const uint32_t *__restrict__ neg1_r,
                      uint32_t *__restrict__ dig, size_t dig_qstride, const uint32_t *__restrict__ tw_base,
                      ModConst mc0, ModConst mc1) {
  __shared__ uint32_t tiles[2][LDS_TILE];
  const int b = blockIdx.x, half = 1 << r;
  const int tid = threadIdx.x;
  const int m = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const ModConst mc = m ? mc1 : mc0;
  uint32_t *vq = v + (size_t)blockIdx.y * q_stride;
  const bool split = stop > 0 && r >= stop;
  // wave m: rows 0 and 1 of residue m of element i = v_neg1 (.) v[i - 2^r]
  auto form_child = [&](int i) {
    const uint32_t *src = vq + (size_t)(i - half) * 4 * N;
    uint32_t *dstv = vq + (size_t)i * 4 * N;
    const uint4 *w4 = reinterpret_cast<const uint4 *>(neg1_r + (size_t)m * 2 * N);
    const uint4 *wp4 = reinterpret_cast<const uint4 *>(neg1_r + (size_t)m * 2 * N + N);
#pragma unroll
    for (int row = 0; row < 2; row++) {
      const uint4 *s4 = reinterpret_cast<const uint4 *>(src + (size_t)(row * 2 + m) * N);
      uint4 *d4 = reinterpret_cast<uint4 *>(dstv + (size_t)(row * 2 + m) * N);
#pragma unroll
      for (int c = 0; c < 8; c++) {
        const uint4 x = s4[c * 64 + lane], w = w4[c * 64 + lane], wp = wp4[c * 64 + lane];
        uint4 y;
        y.x = canon(shoup_mul(x.x, w.x, wp.x, mc.q), mc.q);
        y.y = canon(shoup_mul(x.y, w.y, wp.y, mc.q), mc.q);
        y.z = canon(shoup_mul(x.z, w.z, wp.z, mc.q), mc.q);
        y.w = canon(shoup_mul(x.w, w.w, wp.w, mc.q), mc.q);
        d4[c * 64 + lane] = y;
      }
    }
  };
  if (b >= npair) {  // uniform per block
    form_child(odd_start + 2 * (b - npair) + 1);
    return;
  }
  const int p = split ? (b < half / 2 ? 2 * b : 2 * (b - half / 2) + 1) : b;
  if (odd_live && (p & 1) && !odd_live[blockIdx.y]) return;  // an odd parent and child nothing reads
  const int c = p + half;
  const bool work_c = !expand_skipped(c, r, stop, max_bits);
  if (work_c && !children_formed) form_child(c);  // round 0: nothing formed v_1 yet
  auto slot = [&](int i) { return split ? ((i & 1) ? half + (i >> 1) : (i >> 1)) : i; };
  uint32_t a[32];
  load_p(a, vq + (size_t)p * 4 * N + (size_t)m * N, lane);
  ntt_inv_wave_lazy(a, tw_base + (m * 2 + 1) * TW_WORDS, tiles[m], mc.q, mc.mu1, lane);
#pragma unroll
  for (int ii = 0; ii < 32; ii++) a[ii] = canon(shoup_mul(a[ii], mc.crt_c, mc.crt_cp, mc.q), mc.q);
  // CRT exchange as in k_expand_digits: wave m combines the coefficients
  // 64 (16 m + jj) + lane.
  uint32_t own[16];
#pragma unroll
  for (int jj = 0; jj < 16; jj++) {
    own[jj] = m ? a[16 + jj] : a[jj];
    tiles[m][jj * 64 + lane] = m ? a[jj] : a[16 + jj];
  }
  __syncthreads();
  uint64_t xp[16];
#pragma unroll
  for (int jj = 0; jj < 16; jj++) {
    const uint32_t oth = tiles[1 - m][jj * 64 + lane];
    const uint32_t x0 = m ? oth : own[jj], x1 = m ? own[jj] : oth;
    uint64_t x = (uint64_t)x0 * Q1 + (uint64_t)x1 * Q0;
    xp[jj] = (x >= QQ) ? x - QQ : x;
  }
  __syncthreads();  // the exchange tiles are dead
  uint64_t xc[16];
  if (work_c) {  // uniform per block: the child's coefficients through the tiles (2048 x 8 bytes)
    uint64_t *xl = reinterpret_cast<uint64_t *>(&tiles[0][0]);
#pragma unroll
    for (int jj = 0; jj < 16; jj++) xl[64 * (16 * m + jj) + lane] = xp[jj];
    __syncthreads();
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
      const int s = 64 * (16 * m + jj) + lane + half;
      const uint64_t y = xl[s & (N - 1)];
      xc[jj] = s < N ? y : (y ? QQ - y : 0);
    }
    __syncthreads();
  }
  const uint32_t t = (uint32_t)(N >> r) + 1u;
  uint8_t *db = reinterpret_cast<uint8_t *>(&tiles[0][0]);  // 4096 words over tiles 0-1
  auto digits_out = [&](const uint64_t (&x)[16], int i) {
#pragma unroll
    for (int jj = 0; jj < 16; jj++) {
      const int z = 64 * (16 * m + jj) + lane;
      const uint32_t zt = (uint32_t)z * t;
      const uint32_t dst = zt & (N - 1);
      const uint64_t y = ((zt >> LOGN) & 1) ? QQ - x[jj] : x[jj];
      const uint32_t word = (dst >> 8) * 64 + (dst & 63), byte = (dst >> 6) & 3;
#pragma unroll
      for (int k = 0; k < 8; k++) db[((size_t)k * DIG_WORDS + word) * 4 + byte] = (uint8_t)((y >> (7 * k)) & 127u);
    }
    __syncthreads();
    const uint4 *src4 = reinterpret_cast<const uint4 *>(&tiles[0][0]);
    uint4 *dst4 =
        reinterpret_cast<uint4 *>(dig + (size_t)blockIdx.y * dig_qstride + (size_t)slot(i) * 8 * DIG_WORDS);
    for (int u = tid; u < 8 * DIG_WORDS / 4; u += 128) dst4[u] = src4[u];
    __syncthreads();  // the tiles are free again
  };
  digits_out(xp, p);
  if (work_c) digits_out(xc, c);
}

#ifndef SP_UPD_KP
#define SP_UPD_KP 3
#endif
#ifndef SP_DIGITS_PAIR
#define SP_DIGITS_PAIR 1
#endif

// k_expand_update: one 8-wave block per (working step b, query, residue m):
// wave k transforms digit k of the step for residue m; then every thread
// updates residue m of v_i at its 4 positions (expand_update).  Grid x (2 work
// blocks) holds (b, m) for 16-block groups of 8 steps, residue 0 then 1: the
// two blocks of a step are 8 apart, so they run on the same XCD at nearly the
// same time and the second reads the step's digit bytes from that XCD's L2
// instead of HBM.  Rounds of fewer than 8 steps (work 2 or 4) take the two
// residues of a step side by side, so their few blocks spread over all XCDs.
// The digit bytes (and row 1 for the e term) are requested first and the
// keys only once those are in: a wave's loads complete in issue order, so the
// transforms wait for no key load.
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
void k_expand_update(uint32_t *__restrict__ v, size_t q_stride, int r, int stop, int max_bits, int work,
                     const int *__restrict__ odd_live, const uint32_t *__restrict__ neg1_next,
                     const uint32_t *__restrict__ dig, size_t dig_qstride,
                     const uint32_t *__restrict__ left, const uint32_t *__restrict__ right, size_t left_qstride,
                     size_t right_qstride, const uint32_t *__restrict__ tw_base, ModConst mc0, ModConst mc1) {
  __shared__ __align__(16) uint32_t tiles[8][LDS_TILE];
  const int x = blockIdx.x, y = blockIdx.y;
  const int half = 1 << r;
  const bool groups = (work & 7) == 0;
  const int b = groups ? (x & 7) | ((x >> 4) << 3) : x >> 1, m = groups ? (x >> 3) & 1 : x & 1;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int i = (stop > 0 && r >= stop) ? (b < half ? 2 * b : 2 * (b - half) + 1) : b;
  uint32_t *vi = v + (size_t)y * q_stride + (size_t)i * 4 * N;
  uint32_t dw[8];
  {
    const uint32_t *dg = dig + (size_t)y * dig_qstride + ((size_t)b * 8 + wave) * DIG_WORDS;
#pragma unroll
    for (int i4 = 0; i4 < 8; i4++) dw[i4] = dg[i4 * 64 + lane];
  }
  const uint4 row1 = reinterpret_cast<const uint4 *>(vi + (size_t)(2 + m) * N)[tid];
  // The loads above go out before the flag's load below: otherwise the
  // compiler moves them behind it and its wait.
  asm volatile("" ::: "memory");
  if (odd_live && (i & 1) && !odd_live[y]) return;  // uniform per block: an odd step nothing reads
  const uint32_t t = (uint32_t)(N >> r) + 1u;
  const bool use_left = (r != 0) && ((i & 1) == 0);
  const uint32_t *W = use_left ? left + (size_t)y * left_qstride : right + (size_t)y * right_qstride;
  const uint32_t q = m ? mc1.q : mc0.q, mu1 = m ? mc1.mu1 : mc0.mu1, qinv = m ? mc1.qinv_neg : mc0.qinv_neg;
  ExpandKeys<SP_UPD_KP> pk;
  // The e term: row 1 of residue m of v_i (old values) passes through LDS, one
  // coalesced 16-byte load per thread instead of 4 scattered 4-byte ones.
  reinterpret_cast<uint4 *>(tiles[0])[tid] = row1;
  __syncthreads();
  expand_gather_e<4>(pk.ase, tiles[0], 4 * tid, t);
  __syncthreads();  // tiles[0] is wave 0's transform tile from here on
  uint32_t rr[32];
#pragma unroll
  for (int i4 = 0; i4 < 8; i4++)
#pragma unroll
    for (int bb = 0; bb < 4; bb++) rr[4 * i4 + bb] = (dw[i4] >> (8 * bb)) & 0xFFu;
  // keys of the first SP_UPD_KP digits, loaded ahead of the digit transforms
  // so that the multiply-accumulate waits for less
  if constexpr (SP_UPD_KP > 0) expand_prefetch_keys<SP_UPD_KP>(pk, W, m, 4 * tid);
  // digits < 128: stages 1-8 and 10-11 skip the reduction (outputs < 8q)
  ntt_fwd_wave_t<8, true>(rr, tw_base + (m * 2) * TW_WORDS, tiles[wave], q, mu1, lane);
  store_p(tiles[wave], rr, lane);  // lazy (< 8q): expand_update reduces after the MAC
  __syncthreads();
  expand_update<4, SP_UPD_KP>(vi, W, tiles, 0, m, 4 * tid, t, q, qinv, &pk);
  // Round r + 1's child v[i + 2^(r+1)] = v_neg1[r+1] (.) v_i (residue m, this
  // thread's positions, just written by this thread), unless i is skipped in
  // round r + 1; k_expand_digits of round r + 1 then reads it as formed.
  if (neg1_next && !expand_skipped(i, r + 1, stop, max_bits)) {
    const ModConst mcm = m ? mc1 : mc0;
    neg1_rows4(vi, vi + (size_t)(2 * half) * 4 * N, nullptr, neg1_next, m, 4 * tid, mcm);
  }
}

// ---- latency form of an expansion round (few steps) -------------------------
// The split kernels above need two launches per round, because the v_neg1
// product that creates element i >= 2^r reads v[i - 2^r] while that element's
// own blocks update it.  For rounds with few steps the whole round runs in one
// launch, one 8-wave block per (step, residue m):
//   * v_i.a of both residues comes from a snapshot S_r of the round's inputs
//     (row 0, layout [m][P] per element), which no block of the round writes;
//     each block writes only residue m of v_i;
//   * the epilogue prepares round r + 1: S_{r+1}[i] (the new v_i.a, residue m)
//     and, when element i works in round r + 1, its child
//     v[i + 2^(r+1)] = v_neg1[r+1] (.) v_i (residue m) with its snapshot, which
//     is what the reference computes at the start of round r + 1;
//   * blocks past the working steps form the v_neg1 children of skipped
//     elements (odd, past stop_round), which no block of the round writes.
// Round 0's snapshot and child come from k_expand_prep.  A run uses this form
// for a prefix of the rounds, then the split kernels, whose first round finds
// its working children formed by this epilogue.

// Round 0 of the latency form: v[0] = the query (when `query` is given; else
// v[0] is in place), S_0[0] = v[0].a, v[1] = v_neg1[0] (.) v[0], S_0[1] = v[1].a.
// Block (residue m, query); each thread owns 4 consecutive words of each poly.
__global__ __launch_bounds__(256) void k_expand_prep(const uint32_t *__restrict__ query, size_t query_qstride,
                                                     uint32_t *__restrict__ v, size_t q_stride,
                                                     uint32_t *__restrict__ snap, size_t s_qstride,
                                                     const uint32_t *__restrict__ neg1_0, ModConst mc0,
                                                     ModConst mc1) {
  uint32_t *vq = v + (size_t)blockIdx.y * q_stride;
  uint32_t *sq = snap + (size_t)blockIdx.y * s_qstride;
  const int m = blockIdx.x;
  const ModConst mc = m ? mc1 : mc0;
  for (int p0 = 4 * threadIdx.x; p0 < N; p0 += 4 * 256) {
    if (query) {
      const uint32_t *qq = query + (size_t)blockIdx.y * query_qstride;
#pragma unroll
      for (int row = 0; row < 2; row++)
        *reinterpret_cast<uint4 *>(vq + (size_t)(row * 2 + m) * N + p0) =
            *reinterpret_cast<const uint4 *>(qq + (size_t)(row * 2 + m) * N + p0);
    }
    *reinterpret_cast<uint4 *>(sq + (size_t)m * N + p0) = *reinterpret_cast<const uint4 *>(vq + (size_t)m * N + p0);
    neg1_rows4(vq, vq + 4 * N, sq + 2 * N, neg1_0, m, p0, mc);
  }
}

__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4)))
void k_expand_fused(uint32_t *__restrict__ v, size_t q_stride, int r, int stop, int max_bits, int work, int last,
                    const int *__restrict__ odd_live, const uint32_t *__restrict__ neg1_r,
                    const uint32_t *__restrict__ neg1_next,
                    const uint32_t *__restrict__ s_in, uint32_t *__restrict__ s_out, size_t s_qstride,
                    const uint32_t *__restrict__ left, const uint32_t *__restrict__ right, size_t left_qstride,
                    size_t right_qstride, const uint32_t *__restrict__ tw_base, ModConst mc0, ModConst mc1) {
  __shared__ uint32_t tiles[8][LDS_TILE];
  const int b = blockIdx.x, half = 1 << r, m = blockIdx.z;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const ModConst &mcm = m ? mc1 : mc0;
  uint32_t *vq = v + (size_t)blockIdx.y * q_stride;
  if (b >= work) {  // v_neg1 child of a skipped (odd) element: i = half + 2 (b - work) + 1
    const int i = half + 2 * (b - work) + 1;
    for (int p0 = 4 * tid; p0 < N; p0 += 4 * 512)
      neg1_rows4(vq + (size_t)(i - half) * 4 * N, vq + (size_t)i * 4 * N, nullptr, neg1_r, m, p0, mcm);
    return;
  }
  const int i = (stop > 0 && r >= stop) ? (b < half ? 2 * b : 2 * (b - half) + 1) : b;
  if (odd_live && (i & 1) && !odd_live[blockIdx.y]) return;  // uniform per block: an odd step nothing reads
  uint32_t *vi = vq + (size_t)i * 4 * N;
  const uint32_t t = (uint32_t)(N >> r) + 1u;
  const bool use_left = (r != 0) && ((i & 1) == 0);
  const uint32_t *W = use_left ? left + (size_t)blockIdx.y * left_qstride : right + (size_t)blockIdx.y * right_qstride;
  const uint32_t *si = s_in + (size_t)blockIdx.y * s_qstride + (size_t)i * 2 * N;
  expand_digits<512>(si, tiles, t, tw_base, mc0, mc1, wave, lane, tid);
  uint32_t dw[8];  // digit `wave`: byte b of word i4 is coefficient 64 (4 i4 + b) + lane
  const uint32_t *dig = tiles[4];
#pragma unroll
  for (int i4 = 0; i4 < 8; i4++) dw[i4] = dig[wave * DIG_WORDS + i4 * 64 + lane];
  __syncthreads();  // the digit bytes are dead from here on
  {
    uint32_t rr[32];
#pragma unroll
    for (int i4 = 0; i4 < 8; i4++)
#pragma unroll
      for (int bb = 0; bb < 4; bb++) rr[4 * i4 + bb] = (dw[i4] >> (8 * bb)) & 0xFFu;
    // digits < 128: stages 1-8 and 10-11 skip the reduction (outputs < 8q)
    ntt_fwd_wave_t<8, true>(rr, tw_base + (m * 2) * TW_WORDS, tiles[wave], mcm.q, mcm.mu1, lane);
    store_p(tiles[wave], rr, lane);  // lazy (< 8q): expand_update reduces after the MAC
  }
  __syncthreads();
  const int p0 = 4 * tid;
  expand_update<4>(vi, W, tiles, 0, m, p0, t, mcm.q, mcm.qinv_neg);
  if (last) return;
  // Round r + 1: this thread's 4 positions of residue m, as just written.
  uint32_t *so = s_out + (size_t)blockIdx.y * s_qstride;
  *reinterpret_cast<uint4 *>(so + (size_t)i * 2 * N + (size_t)m * N + p0) =
      *reinterpret_cast<const uint4 *>(vi + (size_t)m * N + p0);
  if (!expand_skipped(i, r + 1, stop, max_bits)) {
    const int c = i + 2 * half;
    neg1_rows4(vi, vq + (size_t)c * 4 * N, so + (size_t)c * 2 * N, neg1_next, m, p0, mcm);
  }
}

// Latency form in half-position blocks: one 4-wave block per (step, residue m,
// position half h), as k_expand_fused but each block owns the NTT positions
// 1024 h .. 1024 h + 1023 of residue m.  After stage 1 (stride 1024) a forward
// transform's two halves are independent, so wave w runs the half-h branches
// of digits w and w + 4 as one transform (ntt_fwd_wave_t<.., true>): the 8
// digit transforms of a step take one transform per SIMD instead of two, and a
// step's work spreads over twice the CUs.  The v_i.a inverse transforms, CRT
// and digits are computed by every block of the step (free in a mostly idle
// chip).  tau_t with t = 1 (mod 4) (every round with 2^r <= N / 4) maps each
// half of the NTT positions onto itself, so the e-term gather of a block reads
// only positions of its own half, which no other block writes; the caller
// keeps rounds with t = 3 (mod 4) on the other forms.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4)))
void k_expand_fused_h(uint32_t *__restrict__ v, size_t q_stride, int r, int stop, int max_bits, int work, int last,
                      const int *__restrict__ odd_live, const uint32_t *__restrict__ neg1_r,
                      const uint32_t *__restrict__ neg1_next,
                      const uint32_t *__restrict__ s_in, uint32_t *__restrict__ s_out, size_t s_qstride,
                      const uint32_t *__restrict__ left, const uint32_t *__restrict__ right, size_t left_qstride,
                      size_t right_qstride, const uint32_t *__restrict__ tw_base, ModConst mc0, ModConst mc1) {
  __shared__ uint32_t tiles[6][LDS_TILE];
  const int b = blockIdx.x, half = 1 << r, m = blockIdx.z & 1, h = blockIdx.z >> 1;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const ModConst &mcm = m ? mc1 : mc0;
  // This thread's 4 words of half h in layout P: positions 1024 h + 32 l' + 4 c
  // + e (l' = tid mod 32, c = tid / 32) at word 256 c + 4 (l' + 32 h) + e.
  const int gp = 256 * (tid >> 5) + 4 * (tid & 31) + 128 * h;
  uint32_t *vq = v + (size_t)blockIdx.y * q_stride;
  if (b >= work) {  // v_neg1 child of a skipped (odd) element: i = half + 2 (b - work) + 1
    const int i = half + 2 * (b - work) + 1;
    neg1_rows4(vq + (size_t)(i - half) * 4 * N, vq + (size_t)i * 4 * N, nullptr, neg1_r, m, gp, mcm);
    return;
  }
  const int i = (stop > 0 && r >= stop) ? (b < half ? 2 * b : 2 * (b - half) + 1) : b;
  if (odd_live && (i & 1) && !odd_live[blockIdx.y]) return;  // uniform per block: an odd step nothing reads
  uint32_t *vi = vq + (size_t)i * 4 * N;
  const uint32_t t = (uint32_t)(N >> r) + 1u;
  const bool use_left = (r != 0) && ((i & 1) == 0);
  const uint32_t *W = use_left ? left + (size_t)blockIdx.y * left_qstride : right + (size_t)blockIdx.y * right_qstride;
  const uint32_t *si = s_in + (size_t)blockIdx.y * s_qstride + (size_t)i * 2 * N;
  expand_digits<256>(si, tiles, t, tw_base, mc0, mc1, wave, lane, tid);
  uint32_t da[8], db[8];  // digits wave and wave + 4: byte b of word i4 is coefficient 64 (4 i4 + b) + lane
  const uint32_t *dig = tiles[4];
#pragma unroll
  for (int i4 = 0; i4 < 8; i4++) {
    da[i4] = dig[wave * DIG_WORDS + i4 * 64 + lane];
    db[i4] = dig[(wave + 4) * DIG_WORDS + i4 * 64 + lane];
  }
  __syncthreads();  // the digit bytes are dead from here on
  {
    const uint32_t *twm = tw_base + (m * 2) * TW_WORDS;
    uint32_t rr[32];
    if (h)
      digits_stage1_half<1>(rr, da, db, twm, mcm.q);
    else
      digits_stage1_half<0>(rr, da, db, twm, mcm.q);
    // digits < 128: stages 1-8 and 10-11 skip the reduction (outputs < 8q)
    ntt_fwd_wave_t<8, true, true>(rr, twm, tiles[wave], mcm.q, mcm.mu1, lane, h);
    // half tiles: digit `wave` at words [0, 1024) of tiles[wave], digit wave + 4
    // at [1024, 2048); position 1024 h + 32 l' + 4 c + e at word 128 c + 4 l' + e.
    // Lane L + 32 b holds virtual positions 64 L + 32 b + j: digit wave + 4 for
    // L >= 16, at l' = 2 (L mod 16) + b.
    uint4 *t4 = reinterpret_cast<uint4 *>(tiles[wave] + ((lane >> 4) & 1) * 1024) + (2 * (lane & 15) + (lane >> 5));
#pragma unroll
    for (int c = 0; c < 8; c++) t4[c * 32] = make_uint4(rr[4 * c], rr[4 * c + 1], rr[4 * c + 2], rr[4 * c + 3]);
  }
  __syncthreads();
  const int hp = 4 * tid;  // = 128 c + 4 l': this thread's words in every half tile
  expand_update_t<4>(vi, W, [&](int k) { return &tiles[k & 3][(k >> 2) * 1024 + hp]; }, m, gp, t, mcm.q,
                     mcm.qinv_neg);
  if (last) return;
  // Round r + 1: this thread's 4 positions of residue m, as just written.
  uint32_t *so = s_out + (size_t)blockIdx.y * s_qstride;
  *reinterpret_cast<uint4 *>(so + (size_t)i * 2 * N + (size_t)m * N + gp) =
      *reinterpret_cast<const uint4 *>(vi + (size_t)m * N + gp);
  if (!expand_skipped(i, r + 1, stop, max_bits)) {
    const int c = i + 2 * half;
    neg1_rows4(vi, vq + (size_t)c * 4 * N, so + (size_t)c * 2 * N, neg1_next, m, gp, mcm);
  }
}

// Working steps of round r (grid of k_expand8 / k_expand_update), and the
// skipped odd elements >= 2^r that still take the v_neg1 product.
static void round_plan(int r, int stop, int max_bits, int &work, int &odd_start, int &extra) {
  const int half = 1 << r;
  work = 2 * half;
  odd_start = 0;
  extra = 0;
  if (stop > 0 && r > stop) {
    work = half;
    odd_start = half;
    extra = half / 2;
  } else if (stop > 0 && r == stop) {
    work = half + (half < max_bits ? half : max_bits);
    odd_start = half > 2 * max_bits ? half : 2 * max_bits;
    extra = odd_start < 2 * half ? (2 * half - odd_start) / 2 : 0;
  }
}

int expand_split_work(int r, int stop, int max_bits) {
  int work, odd_start, extra;
  round_plan(r, stop, max_bits, work, odd_start, extra);
  return work;
}

hipError_t launch_expand_prep(const uint32_t *query, size_t query_qstride, uint32_t *v, size_t q_stride,
                              int n_queries, uint32_t *snap, size_t s_qstride, const DevTables &t, hipStream_t s) {
  k_expand_prep<<<dim3(2, n_queries), 256, 0, s>>>(query, query_qstride, v, q_stride, snap, s_qstride, t.neg1,
                                                   t.mc[0], t.mc[1]);
  return hipGetLastError();
}

int expand_fused_blocks(int r, int stop, int max_bits) {
  int work, odd_start, extra;
  round_plan(r, stop, max_bits, work, odd_start, extra);
  return work + (stop > 0 && r > stop ? (1 << r) / 2 : 0);
}

hipError_t launch_expand_fused(uint32_t *v, size_t q_stride, int n_queries, int r, int end, int stop, int max_bits,
                               const uint32_t *left, const uint32_t *right, size_t left_qstride, size_t right_qstride,
                               uint32_t *snap, size_t s_qstride, size_t s_round, const DevTables &t, hipStream_t s,
                               bool live_only, const int *odd_live) {
  int work, odd_start, extra;
  round_plan(r, stop, max_bits, work, odd_start, extra);
  // Children of skipped elements only: at stop_round the round plan's extra
  // elements are children of working elements, made by the previous epilogue.
  const int neg1_only = (stop > 0 && r > stop && !live_only) ? (1 << r) / 2 : 0;
  const uint32_t *neg1_r = t.neg1 + (size_t)r * 2 * 2 * N;
  const uint32_t *neg1_next = r + 1 < end ? t.neg1 + (size_t)(r + 1) * 2 * 2 * N : nullptr;
  const uint32_t *s_in = snap + (size_t)(r & 1) * s_round;
  uint32_t *s_out = snap + (size_t)((r + 1) & 1) * s_round;
  k_expand_fused<<<dim3(work + neg1_only, n_queries, 2), 512, 0, s>>>(
      v, q_stride, r, stop, max_bits, work, r + 1 >= end, live_only ? odd_live : nullptr, neg1_r, neg1_next, s_in, s_out, s_qstride, left, right,
      left_qstride, right_qstride, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_expand_fused_h(uint32_t *v, size_t q_stride, int n_queries, int r, int end, int stop, int max_bits,
                                 const uint32_t *left, const uint32_t *right, size_t left_qstride,
                                 size_t right_qstride, uint32_t *snap, size_t s_qstride, size_t s_round,
                                 const DevTables &t, hipStream_t s, bool live_only, const int *odd_live) {
  if ((N >> r) < 4) return hipErrorInvalidValue;  // tau_t must keep each position half (t = 1 mod 4)
  int work, odd_start, extra;
  round_plan(r, stop, max_bits, work, odd_start, extra);
  const int neg1_only = (stop > 0 && r > stop && !live_only) ? (1 << r) / 2 : 0;
  const uint32_t *neg1_r = t.neg1 + (size_t)r * 2 * 2 * N;
  const uint32_t *neg1_next = r + 1 < end ? t.neg1 + (size_t)(r + 1) * 2 * 2 * N : nullptr;
  const uint32_t *s_in = snap + (size_t)(r & 1) * s_round;
  uint32_t *s_out = snap + (size_t)((r + 1) & 1) * s_round;
  k_expand_fused_h<<<dim3(work + neg1_only, n_queries, 4), 256, 0, s>>>(
      v, q_stride, r, stop, max_bits, work, r + 1 >= end, live_only ? odd_live : nullptr, neg1_r, neg1_next, s_in, s_out, s_qstride, left, right,
      left_qstride, right_qstride, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}
hipError_t launch_expand_split(uint32_t *v, size_t q_stride, int n_queries, int r, int end, int stop, int max_bits,
                               const uint32_t *left, const uint32_t *right, size_t left_qstride, size_t right_qstride,
                               uint32_t *dig, size_t dig_qstride, const DevTables &t, hipStream_t s,
                               bool live_only, const int *odd_live) {
  int work, odd_start, extra;
  round_plan(r, stop, max_bits, work, odd_start, extra);
  if (live_only) extra = 0;  // the children of skipped elements feed nothing the answer reads
  if (!live_only) odd_live = nullptr;
  const uint32_t *neg1_r = t.neg1 + (size_t)r * 2 * 2 * N;
  const uint32_t *neg1_next = r + 1 < end ? t.neg1 + (size_t)(r + 1) * 2 * 2 * N : nullptr;
  // From round 1 on, the previous round's update (or latency-form) launch has
  // formed the children of its working elements; the extra blocks still form
  // the children of skipped elements (and, at stop_round, recompute children
  // already formed: same parent, same value).
  // Rounds whose one-step blocks exceed two full waves of the chip (16 two-wave
  // blocks per CU) take the pair kernel, which saves the inverse transforms of
  // every child (64 queries at (9,6): expansion 4.85 -> 4.69 ms, measured on
  // one box); smaller rounds keep one step per block, whose shorter blocks
  // finish sooner (one query: 598 against 602 us per answer with pairs).
  if (SP_DIGITS_PAIR && (long)work * n_queries > 16L * t.n_cu) {
    const int half = 1 << r;  // parents whose pair block forms the digits (see k_expand_digits2)
    const int npair = (stop > 0 && r >= stop) ? half / 2 + (r == stop ? std::min(half / 2, max_bits) : 0) : half;
    k_expand_digits2<<<dim3(npair + extra, n_queries), 128, 0, s>>>(v, q_stride, r, stop, max_bits, npair,
                                                                    odd_start, r > 0 ? 1 : 0, odd_live, neg1_r,
                                                                    dig, dig_qstride, t.tw, t.mc[0], t.mc[1]);
  } else {
    k_expand_digits<<<dim3(work + extra, n_queries), 128, 0, s>>>(v, q_stride, r, stop, work, odd_start,
                                                                  r > 0 ? 1 : 0, odd_live, neg1_r, dig, dig_qstride,
                                                                  t.tw, t.mc[0], t.mc[1]);
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_expand_update<<<dim3(2 * work, n_queries), 512, 0, s>>>(v, q_stride, r, stop, max_bits, work, odd_live, neg1_next, dig, dig_qstride,
                                       left, right, left_qstride, right_qstride, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

// Waves per block for a grid of `blocks` working blocks: 16 (one block per CU,
// half the passes) while the grid does not fill the CUs, else 8 (two per CU).
// The fold takes 16 at every size: its 16-wave block (two passes, 16 tiles
// per multiply-accumulate) measured faster than two 8-wave blocks per CU
// also on full grids (64 queries at (9,6): 3.19 vs 3.61 ms per fold stage).
static int fold_expand_waves(long blocks, const DevTables &t, int variant, bool fold) {
  if (variant == 2) return 8;
  if (variant == 3) return 16;
  if (fold) return 16;
  return blocks <= t.n_cu ? 16 : 8;
}

hipError_t launch_fold8(uint32_t *accs, size_t set_stride, int n_trials, int n_queries, int num_per,
                        const uint32_t *cpos, size_t key_qstride, const DevTables &t, int variant, uint32_t *part,
                        size_t part_set_stride, hipStream_t s) {
  dim3 grid(num_per, n_trials, n_queries);
  const long pairs = (long)num_per * n_trials * n_queries;
  // Split form: four partial blocks per pair, then a finish launch; the
  // automatic choice whenever the scratch is there (measured at (9,6): 0.16
  // vs 0.26 ms for one query, 2.75 vs 3.17 ms for 64 queries against the
  // 16-wave blocks).
  if (part && (variant == 4 || variant == 0)) {
    // Quarter blocks when even they give each CU at most one block (one
    // transform per SIMD) and the partial sums of four groups fit the scratch
    // (part_set_stride holds 8N words for each of the first level's pairs:
    // four groups need 16N per pair).  At (9,6), one query: levels 2-5,
    // 8.6-9.1 us per partial launch against 10.2-11.2 us for halves.
    const bool quarters = 8L * pairs <= (long)t.n_cu && (size_t)num_per * 16 * N <= part_set_stride;
    if (quarters) {
      k_fold_part<4><<<dim3(num_per, n_trials * 8, n_queries), 256, 0, s>>>(
          accs, set_stride, n_trials, num_per, cpos, key_qstride, part, part_set_stride, t.tw, t.mc[0], t.mc[1]);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      k_fold_finish<4><<<grid, 256, 0, s>>>(accs, set_stride, n_trials, accs, set_stride, part, part_set_stride,
                                            t.tw, t.mc[0], t.mc[1]);
      return hipGetLastError();
    }
    k_fold_part<2><<<dim3(num_per, n_trials * 4, n_queries), 512, 0, s>>>(accs, set_stride, n_trials, num_per, cpos,
                                                                           key_qstride, part, part_set_stride, t.tw,
                                                                           t.mc[0], t.mc[1]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    k_fold_finish<2><<<grid, 256, 0, s>>>(accs, set_stride, n_trials, accs, set_stride, part, part_set_stride, t.tw,
                                          t.mc[0], t.mc[1]);
    return hipGetLastError();
  }
  if (fold_expand_waves(pairs, t, variant, true) == 16)
    k_fold8<16><<<grid, 1024, 0, s>>>(accs, set_stride, n_trials, num_per, cpos, key_qstride, t.tw, t.mc[0],
                                      t.mc[1]);
  else
    k_fold8<8><<<grid, 512, 0, s>>>(accs, set_stride, n_trials, num_per, cpos, key_qstride, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_fold_latency(uint32_t *accs, size_t set_stride, uint32_t *alt_acc, size_t alt_set_stride,
                               uint32_t *part, size_t part_set_stride, uint32_t *alt_part, size_t alt_part_set_stride,
                               int n_trials, int n_queries, int nu2, const uint32_t *fkeys, size_t level_words,
                               size_t key_qstride, const DevTables &t, hipStream_t s) {
  int num_per = 1 << (nu2 - 1);
  const uint32_t *keys0 = fkeys + (size_t)(nu2 - 1) * level_words;  // level 0 uses the last key level
  k_fold_part<2><<<dim3(num_per, n_trials * 4, n_queries), 512, 0, s>>>(accs, set_stride, n_trials, num_per, keys0,
                                                                         key_qstride, part, part_set_stride, t.tw,
                                                                         t.mc[0], t.mc[1]);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint32_t *src_acc = accs, *src_part = part;
  size_t src_as = set_stride, src_ps = part_set_stride;
  int ngp = 2;  // partial-sum groups per pair of the previous level
  for (int cur = 1; cur < nu2; cur++) {
    num_per >>= 1;
    uint32_t *dst_acc = (cur & 1) ? alt_acc : accs, *dst_part = (cur & 1) ? alt_part : part;
    const size_t dst_as = (cur & 1) ? alt_set_stride : set_stride;
    const size_t dst_ps = (cur & 1) ? alt_part_set_stride : part_set_stride;
    const dim3 grid(num_per, n_trials * 8, n_queries);
    const uint32_t *keys = fkeys + (size_t)(nu2 - 1 - cur) * level_words;
    if (ngp == 2)
      k_fold_fused<2><<<grid, 256, 0, s>>>(src_acc, src_as, dst_acc, dst_as, src_part, src_ps, dst_part, dst_ps,
                                           n_trials, num_per, keys, key_qstride, t.tw, t.mc[0], t.mc[1]);
    else
      k_fold_fused<4><<<grid, 256, 0, s>>>(src_acc, src_as, dst_acc, dst_as, src_part, src_ps, dst_part, dst_ps,
                                           n_trials, num_per, keys, key_qstride, t.tw, t.mc[0], t.mc[1]);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    src_acc = dst_acc; src_part = dst_part;
    src_as = dst_as; src_ps = dst_ps;
    ngp = 4;
  }
  const dim3 fgrid(1, n_trials, n_queries);
  if (ngp == 2)
    k_fold_finish<2><<<fgrid, 256, 0, s>>>(accs, set_stride, n_trials, src_acc, src_as, src_part, src_ps, t.tw,
                                           t.mc[0], t.mc[1]);
  else
    k_fold_finish<4><<<fgrid, 256, 0, s>>>(accs, set_stride, n_trials, src_acc, src_as, src_part, src_ps, t.tw,
                                           t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_expand8(uint32_t *v, size_t q_stride, int n_queries, int r, int num_out, int stop, int max_bits,
                          const uint32_t *left, const uint32_t *right, size_t left_qstride, size_t right_qstride,
                          const DevTables &t, int variant, hipStream_t s) {
  // working steps per query (k_expand8 maps them back to i)
  int work, odd_start, extra;
  round_plan(r, stop, max_bits, work, odd_start, extra);
  (void)num_out;
  dim3 grid(work, n_queries);
  if (fold_expand_waves((long)work * n_queries, t, variant, false) == 16)
    k_expand8<16><<<grid, 1024, 0, s>>>(v, q_stride, r, stop, max_bits, left, right, left_qstride, right_qstride,
                                        t.tw, t.mc[0], t.mc[1]);
  else
    k_expand8<8><<<grid, 512, 0, s>>>(v, q_stride, r, stop, max_bits, left, right, left_qstride, right_qstride,
                                      t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

}  // namespace sp

// ===========================================================================
// Batched database multiply on int8 MFMA (gfx950 v_mfma_i32_32x32x32_i8).
//
// Per (trial t, residue m, position p) the batch computes the modular
// product O[i][c] = sum_j A[i][j] * D[j][c] mod q with i = 2*query + row.
// Residues (< 2^28) split into four 7-bit limbs, so with A = sum_a 2^7a A_a
// and D = sum_b 2^7b D_b,  O = sum_w 2^7w P_w,  P_w = sum_{a+b=w} A_a D_b:
// 16 exact int8 products accumulated in 7 int32 weights (each below
// 4 * 1024 * 127^2 < 2^31 for up to 1024 rows), combined mod q at the end
// (limb_combine, which leaves O * 2^-32 mod q).
// Operand layouts (bytes, K = j contiguous in blocks of 32):
//   A limbs  [m][p][jb][limb][I][32]      I = padded 2*batch rows
//   D limbs  [t][m][p][jb][limb][C][32]
// Output (u32)  [m][t][I][N/4][C][4] (outp_word), read by k_from_ntt_pm (or
// transposed into the per-query accumulator layout [b][t][c][row][m][P] by
// k_outp_to_acc for readouts).
// A lane of the 32x32x32 instruction holds 16 K-bytes of one row of A (and
// of one column of B); both operands are read with the same byte->K mapping,
// so the K order inside the instruction does not matter.
// ===========================================================================
namespace sp {

// Row tiles per block of the LDS-staged kernel (4: 8-wave blocks of 128 x 64
// outputs, 2: 4-wave blocks of 64 x 64) and its prefetch ring depth.
#ifndef SP_RING_RB
#define SP_RING_RB 4
#endif
#ifndef SP_RING_PD
#define SP_RING_PD 4
#endif
// Row tiles per block of the persistent kernel (2 or 4; 0: the one-tile-per-block
// kernel), its ring depth, and the limbs multiplied before its staging step.
#ifndef SP_PERS_RB
#define SP_PERS_RB 4
#endif
#ifndef SP_PERS_PD
#define SP_PERS_PD 2
#endif
#ifndef SP_PERS_SPLIT
#define SP_PERS_SPLIT 4
#endif

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

// Word of product (residue m, position p, trial t, row i, column c) in the
// multiply's output [m][t][I][N/4][C][4]: four consecutive positions of one
// column are 16 contiguous bytes, so a from_ntt wave loads its polynomial in
// layout-P chunks.  A tile writes one position, so a 128-byte line is whole
// only after the tiles of four positions (which run close together on one
// XCD) have written it; measured at 64 queries, 1.48x the product bytes reach
// HBM (lines leave L2 before they are whole).
__device__ __forceinline__ size_t outp_word(int m, int p, int t, int i, int c, int T, int I, int C) {
  return (((((size_t)m * T + t) * I + i) * (N / 4) + (p >> 2)) * C + c) * 4 + (p & 3);
}

// Product O = sum_w 2^(7w) P_w mod q from the 7 limb-weight sums (each below
// 2^31), times 2^-32: the Montgomery form, which k_from_ntt_pm undoes in its
// CRT constant and k_outp_to_acc with a multiplication by 2^32 mod q.
// 2^(7w) < q for w < 4, so those terms are exact shifts; the three others
// take their weight mod q (pwm = 2^(7w) mod q).  The sum stays below
// 2^53 + 3 * 2^59 < 2^61 and its REDC below 2^29 + q < 4q.
__device__ __forceinline__ uint32_t limb_combine(const v16i (&acc)[7], int v, const uint32_t *__restrict__ pwm,
                                                 const ModConst &mc) {
  uint64_t s = (uint64_t)(uint32_t)acc[0][v] + ((uint64_t)(uint32_t)acc[1][v] << 7) +
               ((uint64_t)(uint32_t)acc[2][v] << 14) + ((uint64_t)(uint32_t)acc[3][v] << 21);
  s += (uint64_t)(uint32_t)acc[4][v] * pwm[4] + (uint64_t)(uint32_t)acc[5][v] * pwm[5] +
       (uint64_t)(uint32_t)acc[6][v] * pwm[6];
  const uint32_t mm = (uint32_t)s * mc.qinv_neg;
  uint32_t r = (uint32_t)((s + (uint64_t)mm * mc.q) >> 32);
  r = min(r, r - 2 * mc.q);
  return min(r, r - mc.q);
}

// Selector limbs from the expanded ciphertexts of a batch.  Block: 4 rows i x
// 32 j x 64 positions; 16-byte loads of 4 positions into an LDS tile (rows of
// 68 words, 16-byte aligned), then each thread packs the 128-byte chunk of one
// (position, limb) and the chunks leave through the same LDS (rows of 33
// words) so that every store instruction writes whole 128-byte lines (8 lanes
// per chunk): 2.1x faster than each lane storing its own chunk.
__global__ __launch_bounds__(256) void k_sel_to_limbs(const uint32_t *__restrict__ v, size_t v_q, int nq, int R,
                                                      int I, uint8_t *__restrict__ A) {
  constexpr int PS = 68;
  __shared__ __align__(16) uint32_t tile[4 * 32 * PS];
  const int pb = blockIdx.x;           // position block of 64
  const int jb = blockIdx.y;           // j block of 32
  const int m = blockIdx.z & 1;
  const int i0 = (blockIdx.z >> 1) * 4;
  const int JB = R / 32;
  const int tid = threadIdx.x;
#pragma unroll
  for (int k = tid; k < 4 * 32 * 16; k += 256) {  // (row, j, 4 positions)
    const int p4 = k & 15, j = (k >> 4) & 31, ii = k >> 9;
    const int i = i0 + ii;
    uint4 x = make_uint4(0, 0, 0, 0);
    if ((i >> 1) < nq)
      x = *reinterpret_cast<const uint4 *>(v + (size_t)(i >> 1) * v_q + (size_t)(2 * (jb * 32 + j)) * 4 * N +
                                           (size_t)((i & 1) * 2 + m) * N + pb * 64 + 4 * p4);
    *reinterpret_cast<uint4 *>(&tile[(ii * 32 + j) * PS + 4 * p4]) = x;
  }
  __syncthreads();
  // 64 positions x 4 limbs x 4 rows x 32 bytes: thread -> (p, limb), 128 bytes.
  const int p = tid & 63, limb = tid >> 6;
  uint32_t words[32];
#pragma unroll
  for (int ii = 0; ii < 4; ii++)
#pragma unroll
    for (int w = 0; w < 8; w++) {
      uint32_t packed = 0;
#pragma unroll
      for (int b = 0; b < 4; b++) packed |= ((tile[(ii * 32 + w * 4 + b) * PS + p] >> (7 * limb)) & 127u) << (8 * b);
      words[ii * 8 + w] = packed;
    }
  __syncthreads();  // the tile is dead: it stages the chunks, [chunk = tid][33 words]
#pragma unroll
  for (int w = 0; w < 32; w++) tile[tid * 33 + w] = words[w];
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 8; u++) {
    const int g = u * 256 + tid, c = g >> 3, piece = g & 7;
    const int cp = c & 63, cl = c >> 6;
    const uint32_t *src = &tile[c * 33 + 4 * piece];
    uint4 *dst = reinterpret_cast<uint4 *>(A + (((((size_t)m * N + pb * 64 + cp) * JB + jb) * 4 + cl) * I + i0) * 32) +
                 piece;
    *dst = make_uint4(src[0], src[1], src[2], src[3]);
  }
}

// Database limbs from the resident database [t][m][j][g][k][P] (one-time).
__global__ __launch_bounds__(256) void k_db_to_limbs(const uint32_t *__restrict__ db, int R, int C,
                                                     uint8_t *__restrict__ D) {
  __shared__ uint32_t tile[4][32][65];
  const int pb = blockIdx.x, jb = blockIdx.y;
  const int cblocks = C / 4;
  const int c0 = (blockIdx.z % cblocks) * 4;
  const int tm = blockIdx.z / cblocks;  // t * 2 + m
  const int JB = R / 32;
  const int CG = db_group_cols(C), W = db_group_words(C), G = C / CG;
  const int tid = threadIdx.x;
  for (int k = tid; k < 4 * 32 * 64; k += 256) {
    const int p = k & 63, j = (k >> 6) & 31, cc = k >> 11;
    const int c = c0 + cc;
    const uint32_t *w0 = db + (((size_t)tm * R + jb * 32 + j) * G + c / CG) * W * N + pb * 64 + p;
    tile[cc][j][p] = unpack28_at(w0, N, c % CG);
  }
  __syncthreads();
  const int p = tid & 63, limb = tid >> 6;
  uint32_t words[32];
#pragma unroll
  for (int cc = 0; cc < 4; cc++)
#pragma unroll
    for (int w = 0; w < 8; w++) {
      uint32_t packed = 0;
#pragma unroll
      for (int b = 0; b < 4; b++) packed |= ((tile[cc][w * 4 + b][p] >> (7 * limb)) & 127u) << (8 * b);
      words[cc * 8 + w] = packed;
    }
  uint4 *dst = reinterpret_cast<uint4 *>(D + (((((size_t)tm * N + pb * 64 + p) * JB + jb) * 4 + limb) * C + c0) * 32);
#pragma unroll
  for (int u = 0; u < 8; u++) dst[u] = make_uint4(words[4 * u], words[4 * u + 1], words[4 * u + 2], words[4 * u + 3]);
}

// One block = 8 waves = a 128 (i) x 64 (c) output tile of one (t, m, p).
// Workgroups are dealt to the 8 XCDs round-robin (hardware block b runs on
// XCD b % 8); each XCD is given one contiguous range of logical tiles, so the
// tiles of one (m, p) for all trials and column groups share an L2 and the
// selector limbs are read from HBM about once.  gridDim.x is a multiple of 8.
__global__ __launch_bounds__(512) void k_dbmul_mfma(const uint8_t *__restrict__ A, const uint8_t *__restrict__ D,
                                                    uint32_t *__restrict__ outp, int T, int R, int C, int I,
                                                    int n_ig, int n_cg, ModConst mc0, ModConst mc1,
                                                    const uint32_t *__restrict__ pw) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  // logical tile -> (t fastest, then c group, i group, position, residue)
  int bid = (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
  const int t = bid % T; bid /= T;
  const int cg = bid % n_cg; bid /= n_cg;
  const int ig = bid % n_ig; bid /= n_ig;
  const int p = bid % N;
  const int m = bid / N;
  const int row0 = ig * 128 + (wave >> 1) * 32;
  const int col0 = cg * 64 + (wave & 1) * 32;
  if (row0 >= I || col0 >= C) return;
  const int JB = R / 32;
  const int h = lane >> 5, l32 = lane & 31;
  const uint8_t *a_base = A + ((((size_t)m * N + p) * JB) * 4 * I + row0 + l32) * 32 + 16 * h;
  const uint8_t *d_base = D + (((((size_t)t * 2 + m) * N + p) * JB) * 4 * C + col0 + l32) * 32 + 16 * h;
  const size_t a_limb = (size_t)I * 32, a_jb = 4 * a_limb;
  const size_t d_limb = (size_t)C * 32, d_jb = 4 * d_limb;
  v16i acc[7];
#pragma unroll
  for (int w = 0; w < 7; w++)
#pragma unroll
    for (int v = 0; v < 16; v++) acc[w][v] = 0;
  v4i fa[4], fd[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    fa[a] = *reinterpret_cast<const v4i *>(a_base + a * a_limb);
    fd[a] = *reinterpret_cast<const v4i *>(d_base + a * d_limb);
  }
  for (int jb = 0; jb < JB; jb++) {
    v4i na[4], nd[4];
    const int nj = jb + 1 < JB ? jb + 1 : jb;
#pragma unroll
    for (int a = 0; a < 4; a++) {
      na[a] = *reinterpret_cast<const v4i *>(a_base + nj * a_jb + a * a_limb);
      nd[a] = *reinterpret_cast<const v4i *>(d_base + nj * d_jb + a * d_limb);
    }
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
      for (int b = 0; b < 4; b++) acc[a + b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[a], fd[b], acc[a + b], 0, 0, 0);
#pragma unroll
    for (int a = 0; a < 4; a++) { fa[a] = na[a]; fd[a] = nd[a]; }
  }
  const ModConst &mc = m ? mc1 : mc0;
  const uint32_t *pwm = pw + m * 8;
  const int c = col0 + l32;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int i = row0 + 8 * (v >> 2) + 4 * h + (v & 3);
    outp[outp_word(m, p, t, i, c, T, I, C)] = limb_combine(acc, v, pwm, mc);
  }
}

// Same tile and block mapping as k_dbmul_mfma, with the operands staged through
// LDS and loaded PD j blocks ahead.  The database limbs stream from HBM once
// (no reuse), so their latency has to be covered by loads in flight: with one
// block per CU and loads one j block ahead the multiply ran at the latency-bound
// ~1.4 TB/s of database limbs.  Here each thread loads three 16-byte pieces per
// j block (selector limbs 0-1 and 2-3 of one row, database limbs of one
// column) into a ring of PD register slots, and the block's 8 waves read their
// fragments from LDS stages.  With one barrier per j block, the two waves of a
// SIMD used to read their fragments together and then wait for each other's
// multiplies (MFMA busy 43% at 64 queries); now the fragments of block jb + 1
// are read while the multiplies of block jb run: three LDS stages hold blocks
// jb (done), jb + 1 (being read) and jb + 2 (being written).  Rows past I and
// columns past C load row / column 0 and are zeroed.
struct MfmaPiece {
  uint4 a0, a1, d;
};

// LDS slot of operand piece idx (16-byte pieces, two per 32-byte row-limb
// chunk): bit 0 flips with bit 3 ^ bit 4.  A fragment read (lane l32 of half h
// reads piece 2 (row0 + l32) + h) then touches 8 (16) distinct 16-byte bank
// groups in every 8 (16) lanes, and the staging writes (piece = thread) stay
// contiguous per 8 lanes: without it lanes l32 and l32 + 4 of a fragment read
// hit the same banks.
__device__ __forceinline__ int mfma_slot(int idx) { return idx ^ (((idx >> 3) ^ (idx >> 4)) & 1); }

// JBC: the number of j blocks when known at compile time (the loop is then
// unrolled whole, so that the waits for the ring's loads count exactly the
// loads issued after them; 0 = R / 32 at run time).
template <int PD, int JBC, int RB>
__global__ __launch_bounds__(RB * 128) void k_dbmul_mfma_ring(const uint8_t *__restrict__ A, const uint8_t *__restrict__ D,
                                                              uint32_t *__restrict__ outp, int T, int R, int C, int I,
                                                              int n_ig, int n_cg, ModConst mc0, ModConst mc1,
                                                              const uint32_t *__restrict__ pw) {
  constexpr int NT = RB * 128, AP = RB * 32 * 4 * 2;  // threads; 16-byte A pieces per j block
  constexpr int DPT = 512 / NT;                        // D pieces (64 columns x 4 limbs x 2) per thread
  __shared__ uint4 sA[3][AP];
  __shared__ uint4 sD[3][512];
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  int bid = (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
  const int t = bid % T; bid /= T;
  const int cg = bid % n_cg; bid /= n_cg;
  const int ig = bid % n_ig; bid /= n_ig;
  const int p = bid % N;
  const int m = bid / N;
  const int JB = JBC ? JBC : R / 32;
  // A pieces tid and tid + NT: (limb al and al + 2, row ar, half hh); D pieces
  // tid + k NT: (limb dl + 2k (RB = 2) or dl, column dc, half hh).  Rows past I
  // and columns past C load row / column 0 (I and C are multiples of 32).
  const int ar = (tid >> 1) & (RB * 32 - 1), al = tid / (RB * 64), hh = tid & 1;
  const int dl = tid >> 7, dc = (tid >> 1) & 63;
  const bool a_ok = ig * RB * 32 + ar < I, d_ok = cg * 64 + dc < C;
  const size_t a_jb = (size_t)4 * I * 32, d_jb = (size_t)4 * C * 32;
  const uint8_t *a0 = A + (((size_t)m * N + p) * JB * 4 * I + (size_t)al * I + (a_ok ? ig * RB * 32 + ar : 0)) * 32 +
                      16 * hh;
  const uint8_t *a1 = a0 + (size_t)2 * I * 32;
  const uint8_t *d0 = D + ((((size_t)t * 2 + m) * N + p) * JB * 4 * C + (size_t)dl * C + (d_ok ? cg * 64 + dc : 0)) * 32 +
                      16 * hh;
  struct Piece {
    uint4 a0, a1, d[DPT];
  };
  auto load = [&](Piece &pc, int jb) {
    pc.a0 = *reinterpret_cast<const uint4 *>(a0 + jb * a_jb);
    pc.a1 = *reinterpret_cast<const uint4 *>(a1 + jb * a_jb);
#pragma unroll
    for (int k = 0; k < DPT; k++) pc.d[k] = *reinterpret_cast<const uint4 *>(d0 + jb * d_jb + (size_t)k * 2 * C * 32);
  };
  const int st_slot = mfma_slot(tid);  // tid + NT has the same bits 3-4
  // Pieces of rows past I / columns past C are zeroed as they are staged.  The
  // products would be harmless (they land in outputs that are never stored),
  // but without the masking the compiler sinks the ring's loads toward their
  // stage (about 3 loads in flight instead of 12: 3x slower, measured).
  const uint32_t ma = a_ok ? ~0u : 0u, md = d_ok ? ~0u : 0u;
  auto masked = [](uint4 x, uint32_t mk) { return make_uint4(x.x & mk, x.y & mk, x.z & mk, x.w & mk); };
  auto stage = [&](const Piece &pc, int st) {
    sA[st][st_slot] = masked(pc.a0, ma);
    sA[st][st_slot + NT] = masked(pc.a1, ma);
#pragma unroll
    for (int k = 0; k < DPT; k++) sD[st][st_slot + k * NT] = masked(pc.d[k], md);
  };
  const int rw = wave >> 1, cw = wave & 1;
  const int h = lane >> 5, l32 = lane & 31;
  // per-lane fragment slots: limb a adds a * RB * 64 (A) / a * 128 (D), which keeps bits 3-4
  const int fa_slot = mfma_slot((rw * 32 + l32) * 2 + h), fd_slot = mfma_slot((cw * 32 + l32) * 2 + h);
  auto frags = [&](v4i (&fa)[4], v4i (&fd)[4], int st) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const uint4 x = sA[st][a * RB * 64 + fa_slot];
      const uint4 y = sD[st][a * 128 + fd_slot];
      fa[a] = v4i{(int)x.x, (int)x.y, (int)x.z, (int)x.w};
      fd[a] = v4i{(int)y.x, (int)y.y, (int)y.z, (int)y.w};
    }
  };
  // Prologue: blocks 0 and 1 staged, blocks 2 .. PD + 1 in the ring (slot b % PD).
  {
    Piece p0, p1;
    load(p0, 0);
    load(p1, min(1, JB - 1));
    stage(p0, 0);
    stage(p1, 1);
  }
  Piece ring[PD];
#pragma unroll
  for (int k = 0; k < PD; k++) {
    const int b = 2 + ((k - 2 + 2 * PD) % PD);  // the block of ring slot k among 2 .. PD + 1
    load(ring[k], min(b, JB - 1));
  }
  v16i acc[7];
#pragma unroll
  for (int w = 0; w < 7; w++)
#pragma unroll
    for (int v = 0; v < 16; v++) acc[w][v] = 0;
  __syncthreads();
  v4i ca[4], cd[4];
  frags(ca, cd, 0);
  // PD divides JB (both powers of two, PD <= JB): no partial ring pass.
#pragma unroll
  for (int jb0 = 0; jb0 < JB; jb0 += PD) {
#pragma unroll
    for (int k = 0; k < PD; k++) {
      const int jb = jb0 + k;
      {
        __syncthreads();  // stage (jb + 1) % 3 written; stage (jb + 2) % 3 no longer read
        v4i na[4], nd[4];
        frags(na, nd, (jb + 1) % 3);  // past the last block: a stale stage, never used
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
          for (int b = 0; b < 4; b++)
            acc[a + b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ca[a], cd[b], acc[a + b], 0, 0, 0);
        // block jb + 2 (ring slot (jb + 2) % PD) into its stage, then that slot
        // loads block jb + 2 + PD
        Piece &sl = ring[(k + 2) % PD];
        stage(sl, (jb + 2) % 3);
        load(sl, min(jb + 2 + PD, JB - 1));
        // keep the loads here: sunk toward their use they would leave too few in flight
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int a = 0; a < 4; a++) { ca[a] = na[a]; cd[a] = nd[a]; }
      }
    }
  }
  const int row0 = ig * RB * 32 + rw * 32, col0 = cg * 64 + cw * 32;
  if (row0 >= I || col0 >= C) return;
  const ModConst &mc = m ? mc1 : mc0;
  const uint32_t *pwm = pw + m * 8;
  const int c = col0 + l32;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int i = row0 + 8 * (v >> 2) + 4 * h + (v & 3);
    outp[outp_word(m, p, t, i, c, T, I, C)] = limb_combine(acc, v, pwm, mc);
  }
}

// Persistent form of k_dbmul_mfma_ring: a grid of BPC blocks per CU, each
// looping over the tiles x TPX + y + k ny of its XCD x (y = blockIdx.x / 8,
// ny = gridDim.x / 8), so every XCD still sweeps one contiguous range of tiles
// (the tiles of a position share its L2).  The prefetch ring and the LDS
// stages run on across tile boundaries: the first j blocks of tile k + 1 are
// loaded and staged during the last ones of tile k, so only the first tile
// pays the pipeline's fill, and no block waits for a dispatch.  With 4-wave
// blocks (RB = 2) two blocks share a CU and drift apart, so one block's
// epilogue overlaps the other's multiplies.  Four LDS stages (so that the
// stage of j block jb is jb % 4 in every tile, JB a multiple of 4); the
// sequence is that of k_dbmul_mfma_ring: in step g the waves multiply block g,
// read the fragments of g + 1, stage g + 2 and load g + 2 + PD (PD + 2 <= JB,
// so a load runs at most one tile ahead).
template <int PD, int JB, int RB>
__global__ __launch_bounds__(RB * 128) __attribute__((amdgpu_waves_per_eu(2))) void k_dbmul_mfma_pers(const uint8_t *__restrict__ A, const uint8_t *__restrict__ D,
                                                              uint32_t *__restrict__ outp, int T, int C, int I, int n_ig,
                                                              int n_cg, int TPX, ModConst mc0, ModConst mc1,
                                                              const uint32_t *__restrict__ pw) {
  static_assert(JB % 4 == 0 && PD + 2 <= JB && JB % PD == 0, "tile pipeline shape");
  constexpr int NT = RB * 128, AP = RB * 32 * 4 * 2, DPT = 512 / NT;
  __shared__ uint4 sA[4][AP];
  __shared__ uint4 sD[4][512];
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int x = (int)(blockIdx.x & 7u), y = (int)(blockIdx.x >> 3), ny = (int)(gridDim.x >> 3);
  const int K = y < TPX ? (TPX - y + ny - 1) / ny : 0;  // tiles of this block
  if (K == 0) return;                                   // uniform per block
  const int ar = (tid >> 1) & (RB * 32 - 1), al = tid / (RB * 64), hh = tid & 1;
  const int dl = tid >> 7, dc = (tid >> 1) & 63;
  const size_t a_jb = (size_t)4 * I * 32, d_jb = (size_t)4 * C * 32;
  struct Tile {
    const uint8_t *a0, *d0;
    uint32_t ma, md;
    int t, ig, cg, p, m;
  };
  auto tile_of = [&](int k) {
    Tile tl;
    int bid = x * TPX + y + k * ny;
    tl.t = bid % T; bid /= T;
    tl.cg = bid % n_cg; bid /= n_cg;
    tl.ig = bid % n_ig; bid /= n_ig;
    tl.p = bid % N;
    tl.m = bid / N;
    const bool a_ok = tl.ig * RB * 32 + ar < I, d_ok = tl.cg * 64 + dc < C;
    tl.ma = a_ok ? ~0u : 0u;
    tl.md = d_ok ? ~0u : 0u;
    tl.a0 = A + (((size_t)tl.m * N + tl.p) * JB * 4 * I + (size_t)al * I + (a_ok ? tl.ig * RB * 32 + ar : 0)) * 32 +
            16 * hh;
    tl.d0 = D + ((((size_t)tl.t * 2 + tl.m) * N + tl.p) * JB * 4 * C + (size_t)dl * C + (d_ok ? tl.cg * 64 + dc : 0)) *
                    32 + 16 * hh;
    return tl;
  };
  struct Piece {
    uint4 a0, a1, d[DPT];
  };
  auto load = [&](Piece &pc, const Tile &tl, int jb) {
    const uint8_t *pa = tl.a0 + jb * a_jb;
    pc.a0 = *reinterpret_cast<const uint4 *>(pa);
    pc.a1 = *reinterpret_cast<const uint4 *>(pa + (size_t)2 * I * 32);
#pragma unroll
    for (int k = 0; k < DPT; k++)
      pc.d[k] = *reinterpret_cast<const uint4 *>(tl.d0 + jb * d_jb + (size_t)k * 2 * C * 32);
  };
  // masking as in k_dbmul_mfma_ring (it also keeps the ring's loads in place)
  auto masked = [](uint4 v, uint32_t mk) { return make_uint4(v.x & mk, v.y & mk, v.z & mk, v.w & mk); };
  const int st_slot = mfma_slot(tid);
  auto stage = [&](const Piece &pc, const Tile &tl, int st) {
    sA[st][st_slot] = masked(pc.a0, tl.ma);
    sA[st][st_slot + NT] = masked(pc.a1, tl.ma);
#pragma unroll
    for (int k = 0; k < DPT; k++) sD[st][st_slot + k * NT] = masked(pc.d[k], tl.md);
  };
  const int rw = wave >> 1, cw = wave & 1;
  const int h = lane >> 5, l32 = lane & 31;
  const int fa_slot = mfma_slot((rw * 32 + l32) * 2 + h), fd_slot = mfma_slot((cw * 32 + l32) * 2 + h);
  auto frags = [&](v4i (&fa)[4], v4i (&fd)[4], int st) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const uint4 u = sA[st][a * RB * 64 + fa_slot];
      const uint4 v = sD[st][a * 128 + fd_slot];
      fa[a] = v4i{(int)u.x, (int)u.y, (int)u.z, (int)u.w};
      fd[a] = v4i{(int)v.x, (int)v.y, (int)v.z, (int)v.w};
    }
  };
  Tile cur = tile_of(0);
  Tile nxt = K > 1 ? tile_of(1) : cur;
  const int last_jb_if_none = JB - 1;  // loads past the last tile repeat its last block (never used)
  {
    Piece p0, p1;
    load(p0, cur, 0);
    load(p1, cur, 1);
    stage(p0, cur, 0);
    stage(p1, cur, 1);
  }
  Piece ring[PD];
#pragma unroll
  for (int k = 0; k < PD; k++) {
    const int b = 2 + ((k - 2 + 2 * PD) % PD);  // the block of ring slot k among 2 .. PD + 1
    load(ring[k], cur, b);
  }
  __syncthreads();
  v4i ca[4], cd[4];
  frags(ca, cd, 0);
  const ModConst &mcm0 = mc0;
  for (int k = 0; k < K; k++) {
    const bool more = k + 1 < K;
    v16i acc[7];
#pragma unroll
    for (int w = 0; w < 7; w++)
#pragma unroll
      for (int v = 0; v < 16; v++) acc[w][v] = 0;
#pragma unroll
    for (int jb = 0; jb < JB; jb++) {
      __syncthreads();  // stage (jb + 1) % 4 written; stage (jb + 2) % 4 no longer read
      v4i na[4], nd[4];
      // block jb + 1 (the next tile's block 0 is read after the epilogue, so
      // that its fragments are not held through it)
      if (jb + 1 < JB) frags(na, nd, (jb + 1) % 4);
      // The slot's staging and refill go between the first SPLIT limbs' multiplies
      // and the rest, so that the wave reaches the next barrier while its last
      // multiplies still run.
#pragma unroll
      for (int a = 0; a < SP_PERS_SPLIT; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a + b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ca[a], cd[b], acc[a + b], 0, 0, 0);
      Piece &sl = ring[(jb + 2) % PD];
      // block jb + 2 of this tile, or block jb + 2 - JB of the next one
      if (jb + 2 < JB) stage(sl, cur, (jb + 2) % 4);
      else stage(sl, nxt, (jb + 2) % 4);
      // then the slot loads block jb + 2 + PD (of this tile or the next)
      if (jb + 2 + PD < JB) load(sl, cur, jb + 2 + PD);
      else if (more) load(sl, nxt, jb + 2 + PD - JB);
      else load(sl, cur, last_jb_if_none);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int a = SP_PERS_SPLIT; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a + b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ca[a], cd[b], acc[a + b], 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
      if (jb + 1 < JB)
#pragma unroll
        for (int a = 0; a < 4; a++) { ca[a] = na[a]; cd[a] = nd[a]; }
    }
    // epilogue of tile k
    const int row0 = cur.ig * RB * 32 + rw * 32, col0 = cur.cg * 64 + cw * 32;
    if (row0 < I && col0 < C) {
      const ModConst &mc = cur.m ? mc1 : mcm0;
      const uint32_t *pwm = pw + cur.m * 8;
      const int c = col0 + l32;
#pragma unroll
      for (int v = 0; v < 16; v++) {
        const int i = row0 + 8 * (v >> 2) + 4 * h + (v & 3);
        outp[outp_word(cur.m, cur.p, cur.t, i, c, T, I, C)] = limb_combine(acc, v, pwm, mc);
      }
    }
    cur = nxt;
    if (k + 2 < K) nxt = tile_of(k + 2);
    frags(ca, cd, 0);  // block 0 of tile k + 1 (stage 0, staged two steps ago; past the last tile: unused)
  }
}

// Streaming form for small batches (I <= 64 selector rows, 32 queries): the
// same products as k_dbmul_mfma, but a block holds only the waves that have
// rows (RT row tiles of 32 x 2 column tiles of 32 = one 64-column group), and
// each wave streams its operands straight from global memory through a ring
// of PD register slots, PD row blocks ahead.  k_dbmul_mfma keeps 8-wave
// blocks whose row waves past I return at once, so at I = 32 a CU streams
// the database limbs with 2 waves and one block of loads in flight each
// (about 2-3 TB/s); here 8 waves per CU keep PD blocks in flight each and
// reach 4.4-4.6 TB/s.  Ring depths 2, 4 and 8 measured the same (within 2%)
// and plain loads of the database limbs were 2-5% faster than non-temporal
// ones at (10, 7), so the bound is the 8 waves per CU that 112 accumulator
// registers per wave allow, not the latency each wave covers.
// JBC: row blocks known at compile time (whole loop unrolled) or 0.
template <int RT, int PD, int JBC>
__global__ __launch_bounds__(RT * 128) void k_dbmul_mfma_stream(const uint8_t *__restrict__ A,
                                                                const uint8_t *__restrict__ D,
                                                                uint32_t *__restrict__ outp, int T, int R, int C,
                                                                int I, ModConst mc0, ModConst mc1,
                                                                const uint32_t *__restrict__ pw) {
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  int bid = (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
  const int n_cg = (C + 63) / 64;
  const int t = bid % T; bid /= T;
  const int cg = bid % n_cg; bid /= n_cg;
  const int p = bid % N;
  const int m = bid / N;
  const int row0 = (wave >> 1) * 32;
  const int col0 = cg * 64 + (wave & 1) * 32;
  if (row0 >= I || col0 >= C) return;
  const int JB = JBC ? JBC : R / 32;
  const int h = lane >> 5, l32 = lane & 31;
  const uint8_t *a_base = A + ((((size_t)m * N + p) * JB) * 4 * I + row0 + l32) * 32 + 16 * h;
  const uint8_t *d_base = D + (((((size_t)t * 2 + m) * N + p) * JB) * 4 * C + col0 + l32) * 32 + 16 * h;
  const size_t a_limb = (size_t)I * 32, a_jb = 4 * a_limb;
  const size_t d_limb = (size_t)C * 32, d_jb = 4 * d_limb;
  v4i fa[PD][4], fd[PD][4];
  auto load = [&](v4i (&xa)[4], v4i (&xd)[4], int jb) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      xa[a] = *reinterpret_cast<const v4i *>(a_base + jb * a_jb + a * a_limb);
      xd[a] = *reinterpret_cast<const v4i *>(d_base + jb * d_jb + a * d_limb);
    }
  };
#pragma unroll
  for (int k = 0; k < PD; k++) load(fa[k], fd[k], min(k, JB - 1));
  v16i acc[7];
#pragma unroll
  for (int w = 0; w < 7; w++)
#pragma unroll
    for (int v = 0; v < 16; v++) acc[w][v] = 0;
  // PD divides JB (both powers of two, PD <= JB): no partial ring pass.
#pragma unroll
  for (int jb0 = 0; jb0 < JB; jb0 += PD) {
#pragma unroll
    for (int k = 0; k < PD; k++) {
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
          acc[a + b] = __builtin_amdgcn_mfma_i32_32x32x32_i8(fa[k][a], fd[k][b], acc[a + b], 0, 0, 0);
      load(fa[k], fd[k], min(jb0 + k + PD, JB - 1));
      // keep the refill here: sunk toward its use it would leave too few loads in flight
      __builtin_amdgcn_sched_barrier(0);
    }
  }
  const ModConst &mc = m ? mc1 : mc0;
  const uint32_t *pwm = pw + m * 8;
  const int c = col0 + l32;
#pragma unroll
  for (int v = 0; v < 16; v++) {
    const int i = row0 + 8 * (v >> 2) + 4 * h + (v & 3);
    outp[outp_word(m, p, t, i, c, T, I, C)] = limb_combine(acc, v, pwm, mc);
  }
}

// Products (outp_word layout, Montgomery form) -> per-query accumulators
// [b][t][c][row][m][P] (NTT domain, canonical: times 2^32 mod q), for readouts
// between the multiply and the fold.  Block: 64 (t, i, c) combinations x 64
// positions of residue m through an LDS tile.
__global__ __launch_bounds__(256) void k_outp_to_acc(const uint32_t *__restrict__ outp, int T, int I, int C, int nq,
                                                     uint32_t *__restrict__ acc, size_t acc_q, ModConst mc0,
                                                     ModConst mc1) {
  __shared__ uint32_t tile[64][65];
  __shared__ long long dst_of[64];  // accumulator word of combination x (-1: none)
  const int pb = blockIdx.x;        // 64 positions
  const int cb = blockIdx.y;        // 64 (t, i, c) combinations
  const int m = blockIdx.z;
  const ModConst &mc = m ? mc1 : mc0;
  const int tid = threadIdx.x;
  const int TIC = T * I * C;
  if (tid < 64) {  // one division per combination, not per element
    const int e = cb * 64 + tid;
    long long o = -1;
    if (e < TIC) {
      const int c = e % C, i = (e / C) % I, t = e / (C * I);
      const int b = i >> 1, row = i & 1;
      if (b < nq) o = (long long)((size_t)b * acc_q + (((size_t)t * C + c) * 2 + row) * 2 * N + (size_t)m * N);
    }
    dst_of[tid] = o;
  }
#pragma unroll
  for (int k = tid; k < 64 * 16; k += 256) {  // 16-byte loads: 4 positions of one combination
    const int q = k & 15, x = k >> 4;
    const int e = cb * 64 + x;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (e < TIC) {
      const int c = e % C, ti = e / C;  // ti = t * I + i
      v = *reinterpret_cast<const uint4 *>(outp + ((((size_t)m * T * I + ti) * (N / 4) + pb * 16 + q) * C + c) * 4);
    }
    tile[4 * q][x] = v.x;
    tile[4 * q + 1][x] = v.y;
    tile[4 * q + 2][x] = v.z;
    tile[4 * q + 3][x] = v.w;
  }
  __syncthreads();
  for (int k = tid; k < 64 * 64; k += 256) {
    const int pp = k & 63, x = k >> 6;
    const long long o = dst_of[x];
    if (o >= 0) acc[(size_t)o + pb * 64 + pp] = canon(shoup_mul(tile[pp][x], mc.two32, mc.two32_p, mc.q), mc.q);
  }
}

// from_ntt straight from the multiply's products (layout of outp_word, NTT
// domain, layout P, canonical) into the raw accumulators [b][t][c][row] (hi,
// lo planes; i = 2 b + row), in place of k_outp_to_acc + k_from_ntt.  One
// block of NCB consecutive columns of one (trial t, row i), wave 2 cc + m
// inverting residue m of column c0 + cc: lane l's register chunk k (layout P
// words 256 k + 4 l .. + 3) is one 16-byte piece of the products, so every
// wave loads its own polynomial with no staging.  Each transform's own tile
// then carries its canonical values to the CRT pass.  The columns of one
// (t, i) run on one XCD (logical block order), so the 128-byte lines that 8
// neighbouring columns share are fetched from HBM once.
template <int NCB>
__global__ __launch_bounds__(NCB * 128) void k_from_ntt_pm(const uint32_t *__restrict__ outp, int T, int I, int C,
                                                           int nq, uint32_t *__restrict__ acc, size_t acc_q,
                                                           const uint32_t *__restrict__ tw_base, ModConst mc0,
                                                           ModConst mc1) {
  __shared__ uint32_t tiles[2 * NCB][LDS_TILE];
  const int tid = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int cc = w >> 1, m = w & 1;
  int bid = (int)blockIdx.x;
  if ((gridDim.x & 7u) == 0) bid = (int)(blockIdx.x & 7u) * (int)(gridDim.x >> 3) + (int)(blockIdx.x >> 3);
  const int ncg = C / NCB;
  const int c = (bid % ncg) * NCB + cc;
  bid /= ncg;
  const int i = bid % (2 * nq), t = bid / (2 * nq);
  const ModConst &mc = m ? mc1 : mc0;
  const uint4 *src = reinterpret_cast<const uint4 *>(outp + outp_word(m, 0, t, i, c, T, I, C));
  uint32_t r[32];
  const int vl = vlane(lane);
#pragma unroll
  for (int k = 0; k < 8; k++) {  // position chunk 64 k + vl: words 4 (64 k + vl) .. + 3
    const uint4 v = src[(size_t)(64 * k + vl) * C];
    r[4 * k] = v.x; r[4 * k + 1] = v.y; r[4 * k + 2] = v.z; r[4 * k + 3] = v.w;
  }
  uint32_t *xm = tiles[w];
  ntt_inv_wave_lazy<false>(r, tw_base + (m * 2 + 1) * TW_WORDS, xm, mc.q, mc.mu1, lane);
#pragma unroll
  for (int ii = 0; ii < 32; ii++) xm[64 * ii + lane] = canon(shoup_mul(r[ii], mc.crt_c32, mc.crt_c32p, mc.q), mc.q);
  __syncthreads();
  const uint32_t *x0 = tiles[2 * cc], *x1 = tiles[2 * cc + 1];
  const int b = i >> 1, row = i & 1;
  uint32_t *dst = acc + (size_t)b * acc_q + (((size_t)t * C + c) * 2 + row) * 2 * N;
#pragma unroll
  for (int jj = 0; jj < 16; jj++) {  // wave m combines coefficients 64 (16 m + jj) + lane
    const int z = 64 * (16 * m + jj) + lane;
    uint64_t x = (uint64_t)x0[z] * Q1 + (uint64_t)x1[z] * Q0;
    x = (x >= QQ) ? x - QQ : x;
    dst[z] = (uint32_t)(x >> 32);
    dst[N + z] = (uint32_t)x;
  }
}

hipError_t launch_outp_to_acc(const uint32_t *outp, int T, int I, int C, int nq, uint32_t *acc, size_t acc_q,
                              const DevTables &t, hipStream_t s) {
  const int TIC = T * I * C;
  k_outp_to_acc<<<dim3(N / 64, (TIC + 63) / 64, 2), 256, 0, s>>>(outp, T, I, C, nq, acc, acc_q, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_from_ntt_pm(const uint32_t *outp, int T, int I, int C, int nq, uint32_t *acc, size_t acc_q,
                              int nc, const DevTables &t, hipStream_t s) {
  if (C % 8 || 2 * nq > I || (nc != 1 && nc != 2 && nc != 4)) return hipErrorInvalidValue;
  const unsigned blocks = (unsigned)(T * 2 * nq * (C / nc));
  if (nc == 4)
    k_from_ntt_pm<4><<<blocks, 512, 0, s>>>(outp, T, I, C, nq, acc, acc_q, t.tw, t.mc[0], t.mc[1]);
  else if (nc == 2)
    k_from_ntt_pm<2><<<blocks, 256, 0, s>>>(outp, T, I, C, nq, acc, acc_q, t.tw, t.mc[0], t.mc[1]);
  else
    k_from_ntt_pm<1><<<blocks, 128, 0, s>>>(outp, T, I, C, nq, acc, acc_q, t.tw, t.mc[0], t.mc[1]);
  return hipGetLastError();
}

hipError_t launch_db_to_limbs(const uint32_t *db, int T, int R, int C, uint8_t *D, hipStream_t s) {
  dim3 grid(N / 64, R / 32, T * 2 * (C / 4));
  k_db_to_limbs<<<grid, 256, 0, s>>>(db, R, C, D);
  return hipGetLastError();
}


hipError_t launch_dbmul_mfma(const uint32_t *v, size_t v_q, int nq, const uint8_t *D, int T, int R, int C, int I,
                             uint8_t *A, uint32_t *outp, uint32_t *acc, size_t acc_q, const uint32_t *pw,
                             const DevTables &t, int ring, hipStream_t s) {
  k_sel_to_limbs<<<dim3(N / 64, R / 32, 2 * (I / 4)), 256, 0, s>>>(v, v_q, nq, R, I, A);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int n_ig = (I + 127) / 128, n_cg = (C + 63) / 64;
  const unsigned blocks = (unsigned)(2 * N * n_ig * n_cg * T);  // 2N is a multiple of 8
  const int JB = R / 32;  // ring depth: 4 j blocks ahead, or all of them when there are fewer (a power of two)
  if (ring == 0)
    k_dbmul_mfma<<<blocks, 512, 0, s>>>(A, D, outp, T, R, C, I, n_ig, n_cg, t.mc[0], t.mc[1], pw);
  else if (ring < 0) {  // streaming form, I <= 64
    const unsigned sblocks = (unsigned)(2 * N * n_cg * T);
#define SP_STREAM(RTV, PDV, JBV)                                                                      \
  k_dbmul_mfma_stream<RTV, PDV, JBV><<<sblocks, RTV * 128, 0, s>>>(A, D, outp, T, R, C, I, t.mc[0], t.mc[1], pw)
#define SP_STREAM_JB(RTV)                  \
  switch (JB) {                            \
    case 1: SP_STREAM(RTV, 1, 1); break;   \
    case 2: SP_STREAM(RTV, 2, 2); break;   \
    case 4: SP_STREAM(RTV, 4, 4); break;   \
    case 8: SP_STREAM(RTV, 4, 8); break;   \
    case 16: SP_STREAM(RTV, 4, 16); break; \
    case 32: SP_STREAM(RTV, 4, 32); break; \
    default: SP_STREAM(RTV, 4, 0); break;  \
  }
    if (I <= 32) {
      SP_STREAM_JB(1)
    } else {
      SP_STREAM_JB(2)
    }
#undef SP_STREAM_JB
#undef SP_STREAM
  } else if (SP_PERS_RB && (JB == 4 || JB == 8 || JB == 16 || JB == 32)) {
    constexpr int RBP = SP_PERS_RB > 0 ? SP_PERS_RB : 2;
    const int p_ig = (I + RBP * 32 - 1) / (RBP * 32);
    const int tpx = 2 * N * p_ig * n_cg * T / 8;
    const unsigned pblocks = (unsigned)(std::max(t.n_cu, 8) / 8 * 8 * (RBP == 2 ? 2 : 1));
#define SP_PERS(PDV, JBV)                                                                                  \
  k_dbmul_mfma_pers<PDV, JBV, RBP><<<pblocks, RBP * 128, 0, s>>>(A, D, outp, T, C, I, p_ig, n_cg, tpx, t.mc[0], \
                                                                 t.mc[1], pw)
    switch (JB) {
      case 4: SP_PERS(2, 4); break;
      case 8: SP_PERS(SP_PERS_PD, 8); break;
      case 16: SP_PERS(SP_PERS_PD, 16); break;
      default: SP_PERS(SP_PERS_PD, 32); break;
    }
#undef SP_PERS
  } else {
    const int r_ig = (I + SP_RING_RB * 32 - 1) / (SP_RING_RB * 32);
    const unsigned rblocks = (unsigned)(2 * N * r_ig * n_cg * T);
#define SP_RING(PDV, JBV)                                                                                   \
  k_dbmul_mfma_ring<PDV, JBV, SP_RING_RB><<<rblocks, SP_RING_RB * 128, 0, s>>>(A, D, outp, T, R, C, I, r_ig, n_cg, \
                                                                           t.mc[0], t.mc[1], pw)
    switch (JB) {
      case 1: SP_RING(1, 1); break;
      case 2: SP_RING(2, 2); break;
      case 4: SP_RING(SP_RING_PD, 4); break;
      case 8: SP_RING(SP_RING_PD, 8); break;
      case 16: SP_RING(SP_RING_PD, 16); break;
      case 32: SP_RING(SP_RING_PD, 32); break;
      default: SP_RING(SP_RING_PD, 0); break;
    }
#undef SP_RING
  }
  e = hipGetLastError();
  if (e != hipSuccess || !acc) return e;
  return launch_outp_to_acc(outp, T, I, C, nq, acc, acc_q, t, s);
}

// ---- measurement: HBM streaming-read ceiling ----------------------------------
// Reads `n16` 16-byte words with non-temporal loads (the database multiply's
// load form), SR words in flight per thread, and xors them into one sink word
// per thread so that nothing is optimised away: the rate a pure streaming read
// reaches on this part, the practical ceiling of the HBM-bound multiply.
constexpr int STREAM_BLOCKS = 1024, STREAM_UNROLL = 8;
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_stream_read(const v4u *__restrict__ src, size_t n16,
                                                     uint32_t *__restrict__ sink) {
  uint32_t acc = 0;
  const size_t stride = (size_t)gridDim.x * 256 * STREAM_UNROLL;
  for (size_t base = (size_t)blockIdx.x * 256 * STREAM_UNROLL + threadIdx.x; base < n16; base += stride) {
    v4u w[STREAM_UNROLL];
#pragma unroll
    for (int u = 0; u < STREAM_UNROLL; u++) {
      const size_t k = base + (size_t)u * 256;
      w[u] = k < n16 ? __builtin_nontemporal_load(src + k) : v4u{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int u = 0; u < STREAM_UNROLL; u++) acc ^= w[u].x ^ w[u].y ^ w[u].z ^ w[u].w;
  }
  sink[(size_t)blockIdx.x * 256 + threadIdx.x] = acc;
}

size_t stream_sink_words() { return (size_t)STREAM_BLOCKS * 256; }

hipError_t launch_stream_read(const void *src, size_t bytes, uint32_t *sink, hipStream_t s) {
  k_stream_read<<<STREAM_BLOCKS, 256, 0, s>>>(reinterpret_cast<const v4u *>(src), bytes / 16, sink);
  return hipGetLastError();
}

}  // namespace sp
