// sp_api.hip — C ABI of libspiral_hip.so (declared in include/spiral_pir.h).
//
// Host side of the MI355X Spiral PIR answer path: context and device tables,
// reference-surface operators (NTT_R, INTT_R, from_ntt, to_ntt, automorph,
// inv_gadget_ntt_mult, db_mul_kernel, FOLD_CIPHERTEXTS,
// coefficient_expansion_r), the resident database and the per-query answer
// pipeline.  Every function returns a status code; nothing exits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spiral_pir.h"
#include "sp_kernels.hpp"

using namespace sp;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
static int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(SP_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                     \
  } while (0)

extern "C" const char *sp_last_error(void) { return g_err.c_str(); }
extern "C" const char *sp_version(void) {
  return "spiral_hip 0.2 gfx950: wave64 NTT (Shoup/Harvey, lazy digit stages), expansion/fold with digit-owning "
         "waves (8- and 16-wave blocks), db_mul u64-lazy VALU + int8-limb MFMA, fold fast+general";
}

// ---------------------------------------------------------------------------
// host arithmetic for table construction
// ---------------------------------------------------------------------------
typedef unsigned __int128 u128;
static uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = (uint64_t)((u128)r * b % m);
    b = (uint64_t)((u128)b * b % m);
    e >>= 1;
  }
  return r;
}
static uint32_t shoup_companion(uint32_t w, uint32_t q) { return (uint32_t)(((uint64_t)w << 32) / q); }
static uint32_t brv11_h(uint32_t x) {
  uint32_t r = 0;
  for (int i = 0; i < LOGN; i++) { r = (r << 1) | (x & 1); x >>= 1; }
  return r;
}
static const uint32_t kQ[2] = {Q0, Q1};
static const uint32_t kPsi[2] = {66687u, 158221u};  // gpu-pir/params.cuh:26,29

static ModConst make_modconst(int m) {
  ModConst c{};
  const uint32_t q = kQ[m], qo = kQ[1 - m];
  c.q = q;
  uint32_t inv = 1;  // q^-1 mod 2^32 by Newton iteration
  for (int i = 0; i < 6; i++) inv *= 2u - q * inv;
  c.qinv_neg = (uint32_t)(0u - inv);
  c.r2 = (uint32_t)(((u128)1 << 64) % q);
  const uint64_t ninv = powmod(N, q - 2, q);
  const uint64_t oinv = powmod(qo, q - 2, q);
  c.crt_c = (uint32_t)(ninv * oinv % q);
  c.crt_cp = shoup_companion(c.crt_c, q);
  c.ninv = (uint32_t)ninv;
  c.ninv_p = shoup_companion(c.ninv, q);
  c.two32 = (uint32_t)((1ull << 32) % q);
  c.two32_p = shoup_companion(c.two32, q);
  c.mu1 = (uint32_t)((1ull << 32) / q);
  c.crt_c32 = (uint32_t)((uint64_t)c.crt_c * c.two32 % q);
  c.crt_c32p = shoup_companion(c.crt_c32, q);
  return c;
}

// Twiddle table for (m, inverse): see the TW_* layout in sp_device.hpp.
static void build_tw(uint32_t *out, int m, int inv) {
  const uint32_t q = kQ[m];
  const uint64_t root = inv ? powmod(kPsi[m], q - 2, q) : kPsi[m];
  std::vector<uint32_t> T(N);
  for (uint32_t i = 0; i < (uint32_t)N; i++) T[i] = (uint32_t)powmod(root, brv11_h(i), q);
  std::fill(out, out + TW_WORDS, 0u);
  // forward tables: -w mod 2^32 in place of w (ct_bf* use y * (-w) + ...)
  const auto wv = [&](uint32_t w) { return inv ? w : 0u - w; };
  for (int idx = 1; idx < 32; idx++) {
    out[2 * idx] = wv(T[idx]);
    out[2 * idx + 1] = shoup_companion(T[idx], q);
  }
  for (int lane = 0; lane < 64; lane++) {
    const int L = lane & 31, b = lane >> 5;  // phase-B lane L + 32 b (sp_device.hpp)
    for (int s = 0; s < TW_B_SLOTS; s++) {
      int idx;
      if (s == 0) idx = 32 + L;
      else if (s < 3) idx = 64 + 2 * L + (s - 1);
      else if (s < 7) idx = 128 + 4 * L + (s - 3);
      else if (s < 15) idx = 256 + 8 * L + (s - 7);
      else if (s < 31) idx = 512 + 16 * L + (s - 15);
      else idx = 1024 + 32 * L + 16 * b + (s - 31);
      out[TW_A_WORDS + (2 * s) * 64 + lane] = wv(T[idx]);
      out[TW_A_WORDS + (2 * s + 1) * 64 + lane] = shoup_companion(T[idx], q);
    }
  }
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// Scratch of sp_fold_noalloc: one array per GPU for the life of the process
// (the converted fold keys and the split fold's partial sums).  Every call on
// a GPU shares that GPU's copy, so calls that overlap on one GPU race; nothing
// in it must start at zero.  Host code uses the address hipGetSymbolAddress
// gives for the context's device (sp_ctx::d_scratch), never the name.
constexpr size_t NOALLOC_SCRATCH_WORDS = (size_t)2 << 20;  // 8 MiB
static __device__ __align__(256) uint32_t scratch[NOALLOC_SCRATCH_WORDS];

struct sp_ctx {
  int device = 0;
  int variant[3] = {0, 0, 0};  // per stage (SP_STAGE_*): 0 = automatic choice
  uint32_t *d_scratch = nullptr;  // this device's copy of `scratch`
  hipStream_t stream = nullptr;
  uint32_t *d_tw = nullptr;
  uint32_t *d_neg1 = nullptr;
  uint32_t *d_pw = nullptr;  // [m][8]: 2^(7w) mod q_m, limb weights of the MFMA multiply
  DevTables tables{};
};

// The stream argument of device-pointer functions is a hipStream_t as is:
// NULL is HIP's null stream (torch's default stream), not the context's.
static hipStream_t pick(sp_ctx *ctx, void *stream) {
  (void)ctx;
  return reinterpret_cast<hipStream_t>(stream);
}

extern "C" int sp_ctx_create(int device, sp_ctx **out) {
  if (!out) return fail(SP_ERR_ARG, "sp_ctx_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(SP_ERR_ARG, "sp_ctx_create: device %d of %d", device, ndev);
  HIP_TRY(hipSetDevice(device));
  sp_ctx *c = new sp_ctx();
  c->device = device;
  hipError_t e = hipGetSymbolAddress(reinterpret_cast<void **>(&c->d_scratch), HIP_SYMBOL(scratch));
  if (e != hipSuccess || !c->d_scratch) {
    delete c;
    return fail(SP_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return fail(SP_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
  std::vector<uint32_t> tw((size_t)4 * TW_WORDS);
  for (int m = 0; m < 2; m++)
    for (int inv = 0; inv < 2; inv++) build_tw(tw.data() + (size_t)(m * 2 + inv) * TW_WORDS, m, inv);
  // v_neg1[r] = NTT(-x^(N - 2^r)) = x^(-2^r): value psi^(-2^r (2 brv(z) + 1)) at position z.
  std::vector<uint32_t> neg1((size_t)LOGN * 2 * 2 * N);
  for (int r = 0; r < LOGN; r++)
    for (int m = 0; m < 2; m++) {
      const uint32_t q = kQ[m];
      const uint64_t psi_inv = powmod(kPsi[m], q - 2, q);
      uint32_t *w = neg1.data() + ((size_t)r * 2 + m) * 2 * N;
      for (int p = 0; p < N; p++) {
        const uint32_t z = (uint32_t)z_of_p(p);
        const uint64_t e = ((uint64_t)(2 * brv11_h(z) + 1) << r) % (2 * N);
        const uint32_t val = (uint32_t)powmod(psi_inv, e, q);
        w[p] = val;
        w[N + p] = shoup_companion(val, q);
      }
    }
  uint32_t pw[16] = {0};
  for (int m = 0; m < 2; m++)
    for (int w = 0; w < 7; w++) pw[m * 8 + w] = (uint32_t)powmod(2, 7 * w, kQ[m]);
  e = hipMalloc(&c->d_tw, tw.size() * 4);
  if (e == hipSuccess) e = hipMalloc(&c->d_neg1, neg1.size() * 4);
  if (e == hipSuccess) e = hipMalloc(&c->d_pw, sizeof pw);
  if (e == hipSuccess) e = hipMemcpy(c->d_pw, pw, sizeof pw, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_tw, tw.data(), tw.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->d_neg1, neg1.data(), neg1.size() * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    int code = fail(SP_ERR_HIP, "sp_ctx_create tables: %s", hipGetErrorString(e));
    sp_ctx_destroy(c);
    return code;
  }
  c->tables.tw = c->d_tw;
  c->tables.neg1 = c->d_neg1;
  c->tables.mc[0] = make_modconst(0);
  c->tables.mc[1] = make_modconst(1);
  c->tables.n_cu = 256;
  {
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && n_cu > 0)
      c->tables.n_cu = n_cu;
  }
  *out = c;
  return SP_OK;
}

extern "C" int sp_ctx_destroy(sp_ctx *ctx) {
  if (!ctx) return SP_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  if (ctx->d_tw) (void)hipFree(ctx->d_tw);
  if (ctx->d_neg1) (void)hipFree(ctx->d_neg1);
  if (ctx->d_pw) (void)hipFree(ctx->d_pw);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SP_OK;
}

extern "C" int sp_ctx_set_variant(sp_ctx *ctx, int stage, int variant) {
  const bool ok = variant >= 0 && (variant <= 4 || (stage == SP_STAGE_EXPAND && variant >= 10 && variant <= 10 + LOGN));
  if (!ctx || stage < 0 || stage > 2 || !ok)
    return fail(SP_ERR_ARG, "sp_ctx_set_variant: stage %d variant %d", stage, variant);
  ctx->variant[stage] = variant;
  return SP_OK;
}

extern "C" int sp_ctx_synchronize(sp_ctx *ctx) {
  if (!ctx) return fail(SP_ERR_ARG, "ctx is NULL");
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return SP_OK;
}

// RAII device buffer for temporaries of boundary operators.
struct DevBuf {
  void *p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 4); }
  uint32_t *u32() const { return static_cast<uint32_t *>(p); }
};

// ---------------------------------------------------------------------------
// reference-surface operators
// ---------------------------------------------------------------------------
extern "C" int sp_ntt_fwd(sp_ctx *ctx, uint32_t *dev_polys, size_t n_polys, int residue, void *stream) {
  if (!ctx || (!dev_polys && n_polys) || residue < 0 || residue > 1) return fail(SP_ERR_ARG, "sp_ntt_fwd: bad argument");
  if (!n_polys) return SP_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_ntt_natural(false, dev_polys, n_polys, ctx->tables, residue, 0, pick(ctx, stream)));
  return SP_OK;
}

extern "C" int sp_ntt_inv(sp_ctx *ctx, uint32_t *dev_polys, size_t n_polys, int residue, int scale, void *stream) {
  if (!ctx || (!dev_polys && n_polys) || residue < 0 || residue > 1) return fail(SP_ERR_ARG, "sp_ntt_inv: bad argument");
  if (!n_polys) return SP_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_ntt_natural(true, dev_polys, n_polys, ctx->tables, residue, scale ? 1 : 0, pick(ctx, stream)));
  return SP_OK;
}

extern "C" int sp_from_ntt(sp_ctx *ctx, const uint32_t *dev_ntt, uint32_t *dev_raw, size_t n_polys, void *stream) {
  if (!ctx || ((!dev_ntt || !dev_raw) && n_polys)) return fail(SP_ERR_ARG, "sp_from_ntt: bad argument");
  if (!n_polys) return SP_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_from_ntt(false, dev_ntt, dev_raw, n_polys, 2 * N, 2 * N, ctx->tables, pick(ctx, stream)));
  return SP_OK;
}

extern "C" int sp_to_ntt(sp_ctx *ctx, const uint32_t *dev_raw, uint32_t *dev_ntt, size_t n_polys, void *stream) {
  if (!ctx || ((!dev_ntt || !dev_raw) && n_polys)) return fail(SP_ERR_ARG, "sp_to_ntt: bad argument");
  if (!n_polys) return SP_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_to_ntt(dev_raw, dev_ntt, n_polys, ctx->tables, pick(ctx, stream)));
  return SP_OK;
}

extern "C" int sp_automorph(sp_ctx *ctx, const uint64_t *dev_in, uint64_t *dev_out, size_t n_polys, uint32_t t,
                            void *stream) {
  if (!ctx || ((!dev_in || !dev_out) && n_polys)) return fail(SP_ERR_ARG, "sp_automorph: bad argument");
  if ((t & 1u) == 0) return fail(SP_ERR_ARG, "sp_automorph: t = %u must be odd", t);
  if (dev_in == dev_out) return fail(SP_ERR_ARG, "sp_automorph: in-place is not supported");
  if (!n_polys) return SP_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_automorph_u64(dev_in, dev_out, n_polys, t % (2 * N), pick(ctx, stream)));
  return SP_OK;
}

static int gadget_bits(int num_elems) { return (56 + num_elems - 1) / num_elems; }

extern "C" int sp_gadget_extprod(sp_ctx *ctx, const uint32_t *dev_in, int rdim, int mx, const uint32_t *dev_keys,
                                 uint32_t *dev_acc, size_t n_items, size_t in_stride, size_t key_stride,
                                 size_t acc_stride, void *stream) {
  if (!ctx || ((!dev_in || !dev_keys || !dev_acc) && n_items)) return fail(SP_ERR_ARG, "sp_gadget_extprod: NULL");
  if (rdim < 1 || rdim > 2 || mx < 2 * rdim || mx % rdim || mx > 56)
    return fail(SP_ERR_ARG, "sp_gadget_extprod: rdim %d, mx %d unsupported (rdim 1|2, 2 <= mx/rdim, mx <= 56)",
                rdim, mx);
  if (!n_items) return SP_OK;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_extprod(dev_in, rdim, mx, gadget_bits(mx / rdim), dev_keys, dev_acc, n_items, in_stride,
                         key_stride, acc_stride, ctx->tables, pick(ctx, stream)));
  return SP_OK;
}

static bool db_shape_ok(int nu1, int nu2) { return nu1 >= 2 && nu1 <= 14 && nu2 >= 1 && nu2 <= 12; }

extern "C" int sp_db_mul(sp_ctx *ctx, const uint32_t *dev_db, const uint32_t *dev_first_dim, uint32_t *dev_out,
                         int nu1, int nu2, void *stream) {
  if (!ctx || !dev_db || !dev_first_dim || !dev_out) return fail(SP_ERR_ARG, "sp_db_mul: NULL");
  if (!db_shape_ok(nu1, nu2)) return fail(SP_ERR_ARG, "sp_db_mul: shape (%d,%d) unsupported", nu1, nu2);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  const int R = 1 << nu1, C = 1 << nu2;
  DevBuf stage, db, sel, out;
  HIP_TRY(stage.alloc((size_t)2 * R * C * N * 4));
  HIP_TRY(db.alloc(db_trial_words(R, C) * 4));
  HIP_TRY(sel.alloc((size_t)R * 4 * N * 4));
  HIP_TRY(out.alloc((size_t)C * 4 * N * 4));
  HIP_TRY(launch_db_from_ref(dev_db, stage.u32(), R, C, ctx->tables, s));
  HIP_TRY(launch_db_pack(stage.u32(), db.u32(), 1, R, 0, R, C, s));
  HIP_TRY(launch_nat_to_p(dev_first_dim, sel.u32(), (size_t)R * 4, 1, 0, ctx->tables, s));
  HIP_TRY(launch_db_mul(db.u32(), sel.u32(), 4 * N, 0, out.u32(), 0, 1, 1, R, C, ctx->tables, s));
  HIP_TRY(launch_p_to_nat(out.u32(), dev_out, (size_t)C * 4, s));
  HIP_TRY(hipStreamSynchronize(s));
  return SP_OK;
}

static int check_fold_args(const sp_ctx *ctx, const uint32_t *dev_accs, int nu2, int t_gsw,
                           const uint32_t *dev_v_folding, const char *who) {
  if (!ctx || !dev_accs || !dev_v_folding) return fail(SP_ERR_ARG, "%s: NULL", who);
  if (nu2 < 1 || nu2 > 12 || t_gsw < 2 || t_gsw > 28) return fail(SP_ERR_ARG, "%s: nu2 %d, t_gsw %d", who, nu2, t_gsw);
  return SP_OK;
}

// Device words of the fold's temporaries: keys in layout P (and the given
// negated keys), then the split fold's partial sums ((2^nu2 / 2) pairs x
// [m][h][row][N]) when that form runs.
struct FoldScratch {
  size_t keys = 0, neg = 0, part = 0;
  size_t total() const { return keys + neg + part; }
};
static FoldScratch fold_scratch(const sp_ctx *ctx, int nu2, int t_gsw, bool fast) {
  FoldScratch f;
  f.keys = (size_t)2 * (2 * t_gsw) * 2 * nu2 * N;
  f.neg = fast ? 0 : f.keys;
  f.part = fast && t_gsw == 8 && ctx->variant[SP_STAGE_FOLD] != 1 ? (size_t)((1 << nu2) / 2) * 8 * N : 0;
  return f;
}

// The fold stages of sp_fold / sp_fold_noalloc on stream s, temporaries in
// kp (keys), kn (negated keys, general form) and part (partial sums).
static int fold_enqueue(sp_ctx *ctx, uint32_t *dev_accs, int nu2, int t_gsw, const uint32_t *dev_v_folding,
                        const uint32_t *dev_v_folding_neg, uint32_t *kp, uint32_t *kn, uint32_t *part, hipStream_t s) {
  const int mx = 2 * t_gsw;
  const size_t level_polys = (size_t)2 * mx * 2;
  const bool fast = dev_v_folding_neg == nullptr;
  HIP_TRY(launch_nat_to_p(dev_v_folding, kp, level_polys * nu2, 2 * mx, 1, ctx->tables, s));
  if (!fast) HIP_TRY(launch_nat_to_p(dev_v_folding_neg, kn, level_polys * nu2, 2 * mx, 1, ctx->tables, s));
  int num_per = 1 << nu2;
  for (int cur = 0; cur < nu2; cur++) {
    num_per >>= 1;
    const size_t lvl = (size_t)(nu2 - 1 - cur) * level_polys * N;
    if (fast && t_gsw == 8 && ctx->variant[SP_STAGE_FOLD] != 1)
      HIP_TRY(launch_fold8(dev_accs, 0, 1, 1, num_per, kp + lvl, 0, ctx->tables, ctx->variant[SP_STAGE_FOLD], part, 0,
                           s));
    else
      HIP_TRY(launch_fold_level(fast, dev_accs, 0, 1, 1, num_per, kp + lvl, fast ? nullptr : kn + lvl, 0, t_gsw,
                                gadget_bits(t_gsw), ctx->tables, s));
  }
  return SP_OK;
}

extern "C" int sp_fold(sp_ctx *ctx, uint32_t *dev_accs, int nu2, int t_gsw, const uint32_t *dev_v_folding,
                       const uint32_t *dev_v_folding_neg, void *stream) {
  int rc = check_fold_args(ctx, dev_accs, nu2, t_gsw, dev_v_folding, "sp_fold");
  if (rc) return rc;
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  const FoldScratch f = fold_scratch(ctx, nu2, t_gsw, dev_v_folding_neg == nullptr);
  DevBuf kp, kn, part;
  HIP_TRY(kp.alloc(f.keys * 4));
  if (f.neg) HIP_TRY(kn.alloc(f.neg * 4));
  if (f.part) HIP_TRY(part.alloc(f.part * 4));
  rc = fold_enqueue(ctx, dev_accs, nu2, t_gsw, dev_v_folding, dev_v_folding_neg, kp.u32(), kn.u32(), part.u32(), s);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(s));
  return SP_OK;
}

extern "C" int sp_fold_noalloc(sp_ctx *ctx, uint32_t *dev_accs, int nu2, int t_gsw, const uint32_t *dev_v_folding,
                               const uint32_t *dev_v_folding_neg, void *stream) {
  int rc = check_fold_args(ctx, dev_accs, nu2, t_gsw, dev_v_folding, "sp_fold_noalloc");
  if (rc) return rc;
  const FoldScratch f = fold_scratch(ctx, nu2, t_gsw, dev_v_folding_neg == nullptr);
  if (f.total() > NOALLOC_SCRATCH_WORDS)
    return fail(SP_ERR_ARG, "sp_fold_noalloc: nu2 %d, t_gsw %d needs %zu bytes of scratch, the maximum is %zu", nu2,
                t_gsw, f.total() * 4, NOALLOC_SCRATCH_WORDS * 4);
  HIP_TRY(hipSetDevice(ctx->device));
  uint32_t *kp = ctx->d_scratch, *kn = kp + f.keys, *part = kn + f.neg;
  return fold_enqueue(ctx, dev_accs, nu2, t_gsw, dev_v_folding, dev_v_folding_neg, kp, kn, part, pick(ctx, stream));
}

extern "C" int sp_regev_to_gsw(sp_ctx *ctx, const uint32_t *dev_v, size_t n_ct, int nu2, int t_gsw, int t_conv,
                               const uint32_t *dev_conv, uint32_t *dev_v_folding, void *stream) {
  if (!ctx || !dev_v || !dev_conv || !dev_v_folding) return fail(SP_ERR_ARG, "sp_regev_to_gsw: NULL");
  if (nu2 < 1 || nu2 > 12 || t_gsw < 2 || t_gsw > 28)
    return fail(SP_ERR_ARG, "sp_regev_to_gsw: nu2 %d, t_gsw %d", nu2, t_gsw);
  if (t_conv != SP_T_CONV) return fail(SP_ERR_ARG, "sp_regev_to_gsw: t_conv %d unsupported (%d)", t_conv, SP_T_CONV);
  const size_t need = (size_t)2 * nu2 * t_gsw;
  if (n_ct < need) return fail(SP_ERR_ARG, "sp_regev_to_gsw: %zu ciphertexts, %zu needed", n_ct, need);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  const size_t conv_polys = (size_t)2 * (2 * t_conv) * 2;
  const size_t fold_polys = (size_t)nu2 * 2 * (2 * t_gsw) * 2;
  DevBuf v, kc, out;
  HIP_TRY(v.alloc(need * 4 * N * 4));
  HIP_TRY(kc.alloc(conv_polys * N * 4));
  HIP_TRY(out.alloc(fold_polys * N * 4));
  HIP_TRY(launch_nat_to_p(dev_v, v.u32(), need * 4, 1, 0, ctx->tables, s));
  HIP_TRY(launch_nat_to_p(dev_conv, kc.u32(), conv_polys, 2 * (2 * t_conv), 1, ctx->tables, s));
  HIP_TRY(launch_regev_to_gsw(false, v.u32(), 0, 1, nu2, t_gsw, kc.u32(), 0, out.u32(), 0, nullptr, ctx->tables, s));
  HIP_TRY(launch_p_to_nat(out.u32(), dev_v_folding, fold_polys, s));
  HIP_TRY(hipStreamSynchronize(s));
  return SP_OK;
}

static int check_packing(int n2, int q2_bits, const char *who) {
  if (n2 != SP_PACK_N * SP_PACK_N) return fail(SP_ERR_ARG, "%s: packing needs n2 = %d (got %d)", who, SP_PACK_N * SP_PACK_N, n2);
  if (q2_bits < SP_PACK_Q1_BITS + 1 || q2_bits > 32) return fail(SP_ERR_ARG, "%s: q2_bits %d out of [11, 32]", who, q2_bits);
  return SP_OK;
}

extern "C" int sp_packed_response_bytes(int n2, int q2_bits, size_t *bytes) {
  if (!bytes) return fail(SP_ERR_ARG, "sp_packed_response_bytes: NULL");
  int rc = check_packing(n2, q2_bits, "sp_packed_response_bytes");
  if (rc) return rc;
  *bytes = packed_response_bytes(q2_bits);
  return SP_OK;
}

extern "C" int sp_pack_response(sp_ctx *ctx, const uint32_t *dev_raw, int n2, const uint32_t *dev_w, int t_conv,
                                int q2_bits, uint8_t *dev_out, void *stream) {
  if (!ctx || !dev_raw || !dev_w || !dev_out) return fail(SP_ERR_ARG, "sp_pack_response: NULL");
  int rc = check_packing(n2, q2_bits, "sp_pack_response");
  if (rc) return rc;
  if (t_conv != SP_T_CONV) return fail(SP_ERR_ARG, "sp_pack_response: t_conv %d unsupported (%d)", t_conv, SP_T_CONV);
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  const size_t w_polys = (size_t)SP_PACK_N * 2 * t_conv * (SP_PACK_N + 1);
  DevBuf kw;
  HIP_TRY(kw.alloc(w_polys * N * 4));
  HIP_TRY(launch_nat_to_p(dev_w, kw.u32(), w_polys, t_conv * (SP_PACK_N + 1), 1, ctx->tables, s));
  HIP_TRY(launch_pack_response(dev_raw, 0, 4 * N, 1, kw.u32(), 0, nullptr, dev_out, 0, q2_bits, ctx->tables, s));
  HIP_TRY(hipStreamSynchronize(s));
  return SP_OK;
}

static int ceil_log2(uint64_t x) {
  int v = 0;
  uint64_t p = 1;
  while (p < x) { p <<= 1; v++; }
  return v;
}

extern "C" int sp_expansion_depth(int nu1, int nu2, int t_gsw, int *g, int *stop_round) {
  if (nu1 < 1 || nu2 < 1 || t_gsw < 1) return fail(SP_ERR_ARG, "sp_expansion_depth: bad shape");
  const int gg = ceil_log2((uint64_t)t_gsw * nu2 + (1ull << nu1));
  const int st = ceil_log2((uint64_t)t_gsw * nu2);
  if (gg > LOGN) return fail(SP_ERR_ARG, "expansion depth %d exceeds log2(N)", gg);
  if (g) *g = gg;
  if (stop_round) *stop_round = st;
  return SP_OK;
}

// Runs `rounds` expansion rounds on internal-layout v (n_queries sets; query
// b's ciphertexts at v + b*q_stride, its keys at left/right + b*stride).
// Largest number of working steps of any expansion round (digit scratch rows).
static int expansion_max_work(int nu1, int nu2, int t_gsw) {
  int g = 0, stop = 0;
  if (sp_expansion_depth(nu1, nu2, t_gsw, &g, &stop)) return 0;
  int mx = 0;
  for (int r = 0; r < g; r++) mx = std::max(mx, expand_split_work(r, stop, t_gsw * nu2));
  return mx;
}

// Expansion kernel family for a round: 1 = one 4-wave block per step (any
// gadget), 2/3 = fused 8/16-wave step kernels, 4 = split digits + update
// kernels (needs the digit scratch).  Automatic: the split form when scratch
// is there, else the fused one.
static int expansion_family(const sp_ctx *ctx, int t_exp, const uint32_t *dig) {
  const int var = ctx->variant[SP_STAGE_EXPAND];
  if (t_exp != 8 || var == 1) return 1;
  if (var == 2 || var == 3) return var;
  return dig ? 4 : 2;
}

// Snapshot sets of the latency-form rounds: room for `queries` queries per set.
struct ExpandSnap {
  uint32_t *p = nullptr;
  size_t q_stride = 0, set_stride = 0;
  int queries = 0;
};

static int run_expansion(sp_ctx *ctx, uint32_t *v, size_t q_stride, int n_queries, int nu1, int nu2, int t_gsw,
                         int t_exp, const uint32_t *left_p, const uint32_t *right_p, size_t left_qstride,
                         size_t right_qstride, int rounds, uint32_t *dig, size_t dig_qstride, const ExpandSnap &snap,
                         const uint32_t *query, size_t query_qstride, hipStream_t s, bool live_only = false,
                         const int *odd_live = nullptr, bool even_only = false) {
  int g = 0, stop = 0;
  int rc = sp_expansion_depth(nu1, nu2, t_gsw, &g, &stop);
  if (rc) return rc;
  if (rounds < 0 || rounds > g) rounds = g;
  const size_t key_level = (size_t)2 * t_exp * 2 * N;
  const int family = expansion_family(ctx, t_exp, dig);
  if (family == 4 && !dig) return fail(SP_ERR_STATE, "split expansion without digit scratch");
  // Latency form for the leading rounds: first the half-position form (four
  // 4-wave blocks per step) while its grid gives each CU at most one block,
  // then the 8-wave form (two blocks per step) while its grid gives each CU at
  // most four (two resident at a time), then the split form.
  // At (9,6), one query: rounds 0-5 in halves 4 us faster each than split,
  // round 6 (448 half blocks) 0.7 us slower; round 6 in the 8-wave form (224
  // blocks) faster than split.  Variant 10 + R forces the half-position form
  // for the first R rounds (those with N >> r >= 4), then the split form.  For
  // the 8-wave form the blocks counted are those that work: a live-only run
  // forms no children of skipped elements, and one whose queries all carry
  // fold vectors (even_only) runs only the even steps, 2^r per round.  At
  // (9,6), one query with fold vectors: rounds 0-5 in halves, rounds 6-9 in
  // the 8-wave form (580.4-580.9 us per answer, against 583.2-583.6 us with
  // rounds 6-8 and 583.6-584.7 us with rounds 6-7 in it, the rest split).
  const int var = ctx->variant[SP_STAGE_EXPAND];
  int halves = 0, fused = 0;  // rounds in the half-position form, in either latency form
  if (family == 4 && snap.p && n_queries <= snap.queries) {
    const long cu = ctx->tables.n_cu;
    auto blocks = [&](int r) {
      const long steps = even_only   ? (1L << r)
                         : live_only ? expand_split_work(r, stop, t_gsw * nu2)
                                     : expand_fused_blocks(r, stop, t_gsw * nu2);
      return steps * n_queries;
    };
    auto blocks_all = [&](int r) { return (long)expand_fused_blocks(r, stop, t_gsw * nu2) * n_queries; };
    if (var == 0) {
      while (halves < rounds && (N >> halves) >= 4 && 4 * blocks_all(halves) <= cu) halves++;
      fused = halves;
      while (fused < rounds && 2 * blocks(fused) <= 4 * cu) fused++;
    } else if (var >= 10) {
      while (halves < std::min(rounds, var - 10) && (N >> halves) >= 4) halves++;
      fused = halves;
    }
  }
  // The queries (if given) go to v[0]; with the latency form one launch also readies round 0.
  if (fused > 0)
    HIP_TRY(launch_expand_prep(query, query_qstride, v, q_stride, n_queries, snap.p, snap.q_stride, ctx->tables, s));
  else if (query)
    HIP_TRY(launch_copy_queries(query, query_qstride, v, q_stride, n_queries, s));
  for (int r = 0; r < fused; r++) {
    const uint32_t *lr = left_p + (size_t)r * key_level, *rr = right_p + (size_t)std::min(r, stop) * key_level;
    if (r < halves)
      HIP_TRY(launch_expand_fused_h(v, q_stride, n_queries, r, rounds, stop, t_gsw * nu2, lr, rr, left_qstride,
                                    right_qstride, snap.p, snap.q_stride, snap.set_stride, ctx->tables, s,
                                    live_only, odd_live));
    else
      HIP_TRY(launch_expand_fused(v, q_stride, n_queries, r, rounds, stop, t_gsw * nu2, lr, rr, left_qstride,
                                  right_qstride, snap.p, snap.q_stride, snap.set_stride, ctx->tables, s, live_only,
                                  odd_live));
  }
  for (int r = fused; r < rounds; r++) {
    const int num_in = 1 << r;
    const uint32_t *neg1_r = ctx->tables.neg1 + (size_t)r * 2 * 2 * N;
    const uint32_t *lr = left_p + (size_t)r * key_level;
    const uint32_t *rr = right_p + (size_t)std::min(r, stop) * key_level;
    if (family == 4) {
      HIP_TRY(launch_expand_split(v, q_stride, n_queries, r, rounds, stop, t_gsw * nu2, lr, rr, left_qstride,
                                  right_qstride, dig, dig_qstride, ctx->tables, s, live_only, odd_live));
      continue;
    }
    HIP_TRY(launch_expand_neg1(v, q_stride, n_queries, num_in, neg1_r, ctx->tables, s));
    if (family != 1)
      HIP_TRY(launch_expand8(v, q_stride, n_queries, r, 2 * num_in, stop, t_gsw * nu2, lr, rr, left_qstride,
                             right_qstride, ctx->tables, family == 3 ? 3 : ctx->variant[SP_STAGE_EXPAND], s));
    else
      HIP_TRY(launch_expand_action(v, q_stride, n_queries, r, 2 * num_in, stop, t_gsw * nu2, lr, rr, left_qstride,
                                   right_qstride, t_exp, gadget_bits(t_exp), ctx->tables, s));
  }
  return SP_OK;
}

extern "C" int sp_expand(sp_ctx *ctx, uint32_t *dev_v, int nu1, int nu2, int t_gsw, int t_exp,
                         const uint32_t *dev_left, const uint32_t *dev_right, int rounds, void *stream) {
  if (!ctx || !dev_v || !dev_left || !dev_right) return fail(SP_ERR_ARG, "sp_expand: NULL");
  if (t_exp < 2 || t_exp > 56 || (t_exp & 1)) return fail(SP_ERR_ARG, "sp_expand: t_exp %d unsupported", t_exp);
  int g = 0, stop = 0;
  int rc = sp_expansion_depth(nu1, nu2, t_gsw, &g, &stop);
  if (rc) return rc;
  if (stop < 1) return fail(SP_ERR_ARG, "sp_expand: stop_round must be >= 1");
  HIP_TRY(hipSetDevice(ctx->device));
  hipStream_t s = pick(ctx, stream);
  const size_t n_ct = (size_t)1 << g;
  const size_t key_level_polys = (size_t)2 * t_exp * 2;
  DevBuf v, kl, kr;
  HIP_TRY(v.alloc(n_ct * 4 * N * 4));
  HIP_TRY(kl.alloc((size_t)g * key_level_polys * N * 4));
  HIP_TRY(kr.alloc((size_t)(stop + 1) * key_level_polys * N * 4));
  HIP_TRY(launch_nat_to_p(dev_v, v.u32(), n_ct * 4, 1, 0, ctx->tables, s));
  HIP_TRY(launch_nat_to_p(dev_left, kl.u32(), (size_t)g * key_level_polys, 2 * t_exp, 1, ctx->tables, s));
  HIP_TRY(launch_nat_to_p(dev_right, kr.u32(), (size_t)(stop + 1) * key_level_polys, 2 * t_exp, 1, ctx->tables, s));
  DevBuf dig, snapbuf;
  ExpandSnap snap;
  if (t_exp == 8 && (ctx->variant[SP_STAGE_EXPAND] == 0 || ctx->variant[SP_STAGE_EXPAND] == 4 ||
                     ctx->variant[SP_STAGE_EXPAND] >= 10)) {
    HIP_TRY(dig.alloc((size_t)expansion_max_work(nu1, nu2, t_gsw) * 8 * 512 * 4));
    snap.q_stride = snap.set_stride = n_ct * 2 * N;
    snap.queries = 1;
    HIP_TRY(snapbuf.alloc(2 * snap.set_stride * 4));
    snap.p = snapbuf.u32();
  }
  rc = run_expansion(ctx, v.u32(), 0, 1, nu1, nu2, t_gsw, t_exp, kl.u32(), kr.u32(), 0, 0, rounds, dig.u32(), 0, snap,
                     nullptr, 0, s);
  if (rc) return rc;
  HIP_TRY(launch_p_to_nat(v.u32(), dev_v, n_ct * 4, s));
  HIP_TRY(hipStreamSynchronize(s));
  return SP_OK;
}

// ---------------------------------------------------------------------------
// resident database
// ---------------------------------------------------------------------------
struct sp_db {
  sp_ctx *ctx = nullptr;
  int nu1 = 0, nu2 = 0, n2 = 0;
  uint32_t *d_db = nullptr;  // [n2][2][R][C/CG][W][P], 28-bit residues packed (db_group_cols)
  size_t bytes = 0;
  uint8_t *d_limbs = nullptr;  // int8-limb copy for the MFMA multiply, built on first use
  size_t limb_bytes = 0;
  bool limbs_valid = false;
};

static bool mfma_shape_ok(const sp_db *db) { return db->nu2 >= 5 && db->nu1 >= 5; }

// Builds (once per load) the int8-limb copy of the database on stream s.
static int db_limbs(sp_db *db, hipStream_t s) {
  if (db->limbs_valid) return SP_OK;
  if (!db->d_limbs) {
    hipError_t e = hipMalloc(&db->d_limbs, db->limb_bytes);
    if (e != hipSuccess)
      return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "database limbs: %s", hipGetErrorString(e));
  }
  HIP_TRY(launch_db_to_limbs(db->d_db, db->n2, 1 << db->nu1, 1 << db->nu2, db->d_limbs, s));
  db->limbs_valid = true;
  return SP_OK;
}

extern "C" int sp_db_create(sp_ctx *ctx, int nu1, int nu2, int n2, sp_db **out) {
  if (!ctx || !out) return fail(SP_ERR_ARG, "sp_db_create: NULL");
  *out = nullptr;
  if (!db_shape_ok(nu1, nu2) || n2 < 1 || n2 > 16) return fail(SP_ERR_ARG, "sp_db_create: shape (%d,%d,%d)", nu1, nu2, n2);
  HIP_TRY(hipSetDevice(ctx->device));
  sp_db *d = new sp_db();
  d->ctx = ctx;
  d->nu1 = nu1; d->nu2 = nu2; d->n2 = n2;
  d->bytes = (size_t)n2 * db_trial_words(1 << nu1, 1 << nu2) * 4;
  d->limb_bytes = (size_t)n2 * 2 * ((size_t)1 << (nu1 + nu2)) * N * 4;  // 4 limbs per residue
  hipError_t e = hipMalloc(&d->d_db, d->bytes);
  if (e != hipSuccess) {
    delete d;
    return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "sp_db_create: hipMalloc(%zu): %s", (size_t)0,
                hipGetErrorString(e));
  }
  *out = d;
  return SP_OK;
}

extern "C" int sp_db_destroy(sp_db *db) {
  if (!db) return SP_OK;
  (void)hipSetDevice(db->ctx->device);
  (void)hipStreamSynchronize(db->ctx->stream);
  if (db->d_db) (void)hipFree(db->d_db);
  if (db->d_limbs) (void)hipFree(db->d_limbs);
  delete db;
  return SP_OK;
}

extern "C" size_t sp_db_device_bytes(const sp_db *db) { return db ? db->bytes : 0; }

extern "C" int sp_db_build_limbs(sp_db *db, size_t *limb_bytes) {
  if (!db) return fail(SP_ERR_ARG, "sp_db_build_limbs: NULL");
  if (limb_bytes) *limb_bytes = 0;
  if (!mfma_shape_ok(db)) return SP_OK;
  HIP_TRY(hipSetDevice(db->ctx->device));
  int rc = db_limbs(db, db->ctx->stream);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(db->ctx->stream));
  if (limb_bytes) *limb_bytes = db->limb_bytes;
  return SP_OK;
}

extern "C" int sp_db_load_bytes(sp_db *db, const uint8_t *host_bytes) {
  if (!db || !host_bytes) return fail(SP_ERR_ARG, "sp_db_load_bytes: NULL");
  db->limbs_valid = false;
  sp_ctx *ctx = db->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  const int R = 1 << db->nu1, C = 1 << db->nu2;
  const size_t nrec = (size_t)R * C;
  const size_t rec_bytes = (size_t)db->n2 * N;
  // Chunks of whole rows (C records): bytes -> u32 NTT residues -> packed rows.
  size_t chunk = std::max<size_t>(1, ((size_t)64 << 20) / rec_bytes);
  chunk = std::max<size_t>(C, chunk - chunk % C);
  chunk = std::min(chunk, nrec);
  DevBuf stage, wide;
  HIP_TRY(stage.alloc(chunk * rec_bytes));
  HIP_TRY(wide.alloc(chunk * rec_bytes * 2 * 4));
  for (size_t r0 = 0; r0 < nrec; r0 += chunk) {
    const size_t cnt = std::min(chunk, nrec - r0);
    HIP_TRY(hipMemcpyAsync(stage.p, host_bytes + r0 * rec_bytes, cnt * rec_bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_db_encode(static_cast<const uint8_t *>(stage.p), r0, cnt, wide.u32(), db->n2, C, ctx->tables,
                             ctx->stream));
    HIP_TRY(launch_db_pack(wide.u32(), db->d_db, db->n2, (int)(cnt / C), (int)(r0 / C), R, C, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return SP_OK;
}

extern "C" int sp_db_load_ntt(sp_db *db, const uint32_t *host_ntt) {
  if (!db || !host_ntt) return fail(SP_ERR_ARG, "sp_db_load_ntt: NULL");
  db->limbs_valid = false;
  sp_ctx *ctx = db->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  const int R = 1 << db->nu1, C = 1 << db->nu2;
  const size_t trial_words = (size_t)2 * R * C * N;
  DevBuf stage, wide;
  HIP_TRY(stage.alloc(trial_words * 4));
  HIP_TRY(wide.alloc(trial_words * 4));
  for (int t = 0; t < db->n2; t++) {
    HIP_TRY(hipMemcpyAsync(stage.p, host_ntt + t * trial_words, trial_words * 4, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(launch_db_from_ref(stage.u32(), wide.u32(), R, C, ctx->tables, ctx->stream));
    HIP_TRY(launch_db_pack(wide.u32(), db->d_db + t * db_trial_words(R, C), 1, R, 0, R, C, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
  }
  return SP_OK;
}

// ---------------------------------------------------------------------------
// batched answer pipeline (a single query is a batch of capacity 1)
// ---------------------------------------------------------------------------
// What a captured stage chain depends on besides the buffers (which never move
// while the batch lives): the run size, conversion, packing, and the kernel
// variant of every stage.  A change of any of them captures a new graph.
struct GraphKey {
  int n = 0, packed = 0, pack_q2 = 0, dbmul = 0, expand = 0, fold = 0;
  bool operator==(const GraphKey &o) const {
    return n == o.n && packed == o.packed && pack_q2 == o.pack_q2 && dbmul == o.dbmul && expand == o.expand &&
           fold == o.fold;
  }
};

struct sp_batch {
  sp_db *db = nullptr;
  int cap = 0, t_gsw = 0, t_exp = 0, g = 0, stop = 0;
  int fold_mode = -1;             // -1 unset, 1 fast (G - C), 0 general (given C_neg)
  size_t v_q = 0, l_q = 0, r_q = 0, f_q = 0, acc_q = 0;  // per-query strides in words
  uint32_t *query_ct = nullptr;   // [cap][4N]  queries, layout P (expansion input)
  uint32_t *v = nullptr;          // [cap][2^g][4N]
  uint32_t *left = nullptr;       // [cap][g][m][t_exp][row][P], 2^32-scaled
  uint32_t *right = nullptr;      // [cap][stop+1][...]
  uint32_t *fkeys = nullptr;      // [cap][nu2][m][2 t_gsw][row][P], 2^32-scaled
  uint32_t *fkeys_neg = nullptr;  // general fold mode only
  uint32_t *acc = nullptr;        // [cap][n2][C][4N]: NTT (P) after db_mul, raw after from_ntt
  uint32_t *stage = nullptr;      // host-transfer staging
  size_t stage_words = 0;
  hipStream_t io = nullptr;       // uploads / readouts
  hipStream_t cap_stream = nullptr;
  hipEvent_t done = nullptr;      // recorded after every run (and asynchronous readout)
  hipEvent_t loaded = nullptr;    // recorded on io after every load; runs wait for it on the device
  std::vector<std::pair<GraphKey, hipGraphExec_t>> graphs;
  std::vector<hipGraph_t> graph_defs;
  int last_dbmul = 0;             // database-multiply variant of the last run (1 VALU, 2 MFMA)
  int pm_n = 0;                   // queries whose products wait in mfma_out (0: they are in acc)
  uint32_t *bulk = nullptr;       // bulk-load staging: [cap][4N] queries then [cap][f_q] v_folding
  uint32_t *dig = nullptr;        // split-expansion digit scratch [cap][dig_q]
  uint32_t *snap = nullptr;       // latency-form expansion snapshots, 2 sets x snap_q x SNAP_QUERIES
  size_t dig_q = 0;
  uint32_t *part = nullptr;       // split-fold partial sums [cap][n2][C/2][m][h][row][N]
  uint32_t *fold_alt = nullptr;   // latency-form fold: [FOLD_QUERIES][n2][max(C/4,1)] raw ciphertexts
  uint32_t *part_alt = nullptr;   //   and [FOLD_QUERIES][n2][max(C/4,1)][m][group 4][row][N] partial sums
  uint32_t *conv = nullptr;       // conversion keys of packed slots [cap][m][8][row][P], 2^32-scaled
  size_t conv_q = 0;
  std::vector<int> packed;        // per slot: 1 = packed query (v_folding converted on the device)
  int *d_packed = nullptr;        // device copy of `packed`
  int pack_q2 = 0;                // response packing: 0 off, else bits of the rescaled row 0
  int pack_q2_run = 0;            // pack_q2 of the last run (what the packed responses hold)
  uint32_t *pk = nullptr;         // packing keys per slot [cap][r][m][k][row][P], 2^32-scaled
  size_t pk_q = 0;
  std::vector<int> haspk;         // per slot: 1 = packing keys loaded
  int *d_haspk = nullptr;
  uint8_t *pout = nullptr;        // packed responses [cap][bytes(q2)]: sized once for q2 = 32, never moved
  uint32_t *rstage = nullptr;     // asynchronous raw readout: responses gathered [cap][n2][4N]
  uint8_t *mfma_a = nullptr;      // selector limbs workspace (MFMA multiply)
  uint32_t *mfma_out = nullptr;   // products workspace [m][t][I][N/4][C][4]
  std::vector<std::pair<int, int>> dbmul_choice;  // n -> variant picked by the autotuner
  std::vector<std::pair<int, float>> dbmul_ms[5];  // measurements per variant, for reporting
  float tune_ms[2] = {-1, -1};                     // last tuning: multi-query VALU forms (2, 4 queries per read)
};

// Queries a batch keeps expansion snapshots for (the latency form runs only
// for runs of at most this many queries; 2 x 16 MiB per query at (9,6)).
constexpr int SNAP_QUERIES = 2;
// Queries a batch keeps latency-form fold buffers for (runs of at most this
// many queries fold with k_fold_fused; 10 MiB per query at (9,6)).
constexpr int FOLD_QUERIES = 2;

static int mfma_rows(int n) { return ((2 * n + 31) / 32) * 32; }

struct sp_query {
  sp_batch *b = nullptr;
};

static void batch_free(sp_batch *b) {
  if (!b) return;
  for (auto &ge : b->graphs) (void)hipGraphExecDestroy(ge.second);
  for (auto &gd : b->graph_defs) (void)hipGraphDestroy(gd);
  uint32_t *bufs[] = {b->query_ct, b->v,        b->left, b->right, b->fkeys, b->fkeys_neg, b->acc,
                      b->stage,    b->mfma_out, b->bulk, b->dig,   b->part,  b->conv,
                      b->fold_alt, b->part_alt};
  for (uint32_t *p : bufs)
    if (p) (void)hipFree(p);
  if (b->mfma_a) (void)hipFree(b->mfma_a);
  if (b->d_packed) (void)hipFree(b->d_packed);
  if (b->d_haspk) (void)hipFree(b->d_haspk);
  if (b->snap) (void)hipFree(b->snap);
  if (b->pk) (void)hipFree(b->pk);
  if (b->pout) (void)hipFree(b->pout);
  if (b->rstage) (void)hipFree(b->rstage);
  if (b->done) (void)hipEventDestroy(b->done);
  if (b->loaded) (void)hipEventDestroy(b->loaded);
  if (b->io) (void)hipStreamDestroy(b->io);
  if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
  delete b;
}

extern "C" int sp_batch_create(sp_db *db, int capacity, int t_gsw, int t_exp, sp_batch **out) {
  if (!db || !out) return fail(SP_ERR_ARG, "sp_batch_create: NULL");
  *out = nullptr;
  if (capacity < 1 || capacity > 4096) return fail(SP_ERR_ARG, "sp_batch_create: capacity %d", capacity);
  if (t_gsw < 2 || t_gsw > 28 || t_exp < 2 || t_exp > 56 || (t_exp & 1))
    return fail(SP_ERR_ARG, "sp_batch_create: t_gsw %d / t_exp %d unsupported", t_gsw, t_exp);
  int g = 0, stop = 0;
  int rc = sp_expansion_depth(db->nu1, db->nu2, t_gsw, &g, &stop);
  if (rc) return rc;
  if (stop < 1) return fail(SP_ERR_ARG, "sp_batch_create: stop_round must be >= 1");
  if ((1 << g) < 2 * (1 << db->nu1)) return fail(SP_ERR_ARG, "sp_batch_create: expansion too shallow");
  sp_ctx *ctx = db->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  sp_batch *b = new sp_batch();
  b->db = db; b->cap = capacity; b->t_gsw = t_gsw; b->t_exp = t_exp; b->g = g; b->stop = stop;
  const int C = 1 << db->nu2;
  const size_t kl = (size_t)2 * t_exp * 2 * N;
  const size_t kf = (size_t)2 * (2 * t_gsw) * 2 * N;
  b->v_q = ((size_t)1 << g) * 4 * N;
  b->l_q = (size_t)g * kl;
  b->r_q = (size_t)(stop + 1) * kl;
  b->f_q = (size_t)db->nu2 * kf;
  b->acc_q = (size_t)db->n2 * C * 4 * N;
  b->conv_q = (size_t)2 * (2 * SP_T_CONV) * 2 * N;
  b->packed.assign((size_t)capacity, 0);
  b->haspk.assign((size_t)capacity, 0);
  b->pk_q = (size_t)SP_PACK_N * 2 * SP_T_CONV * (SP_PACK_N + 1) * N;
  b->stage_words = std::max<size_t>({b->l_q, b->r_q, b->f_q, b->acc_q, b->v_q});
  const size_t cap = (size_t)capacity;
  hipError_t e = hipSuccess;
  auto A = [&](uint32_t **p, size_t words) {
    if (e == hipSuccess) e = hipMalloc(p, words * 4);
  };
  A(&b->query_ct, cap * 4 * N);
  A(&b->v, cap * b->v_q);
  A(&b->left, cap * b->l_q);
  A(&b->right, cap * b->r_q);
  A(&b->fkeys, cap * b->f_q);
  A(&b->acc, cap * b->acc_q);
  A(&b->stage, b->stage_words);
  if (e == hipSuccess) e = hipMalloc(&b->d_packed, cap * sizeof(int));
  if (e == hipSuccess) e = hipMemset(b->d_packed, 0, cap * sizeof(int));
  if (e == hipSuccess) e = hipMalloc(&b->d_haspk, cap * sizeof(int));
  if (e == hipSuccess) e = hipMemset(b->d_haspk, 0, cap * sizeof(int));
  if (t_exp == 8) {  // digit scratch of the split expansion
    b->dig_q = (size_t)expansion_max_work(db->nu1, db->nu2, t_gsw) * 8 * 512;
    A(&b->dig, cap * b->dig_q);
    A(&b->snap, 2 * (size_t)std::min(capacity, SNAP_QUERIES) * ((size_t)1 << g) * 2 * N);
  }
  if (t_gsw == 8) {
    A(&b->part, cap * db->n2 * (size_t)(C / 2) * 8 * N);  // split-fold partial sums
    const size_t fq = (size_t)std::min(capacity, FOLD_QUERIES) * db->n2 * std::max(C / 4, 1);
    A(&b->fold_alt, fq * 4 * N);
    A(&b->part_alt, fq * 16 * N);
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->io, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&b->loaded, hipEventDisableTiming);
  if (e != hipSuccess) {
    batch_free(b);
    return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "sp_batch_create: %s", hipGetErrorString(e));
  }
  *out = b;
  return SP_OK;
}

extern "C" int sp_batch_destroy(sp_batch *b) {
  if (!b) return SP_OK;
  (void)hipSetDevice(b->db->ctx->device);
  (void)hipDeviceSynchronize();
  batch_free(b);
  return SP_OK;
}

// Host -> device (natural layout) -> internal P (2^32-scaled when mont), on the batch's io stream.
static int upload_p(sp_batch *b, const uint32_t *host, uint32_t *dst, size_t n_polys, int m_div, int mont) {
  sp_ctx *ctx = b->db->ctx;
  if (n_polys * N > b->stage_words) return fail(SP_ERR_ARG, "upload_p: staging too small");
  HIP_TRY(hipMemcpyAsync(b->stage, host, n_polys * N * 4, hipMemcpyHostToDevice, b->io));
  HIP_TRY(launch_nat_to_p(b->stage, dst, n_polys, m_div, mont, ctx->tables, b->io));
  HIP_TRY(hipStreamSynchronize(b->io));
  return SP_OK;
}

// Records the form of slots [first, first + n) (1 packed, 0 with fold vectors)
// on the host and, for the conversion kernel, on the device (io stream; the
// values travel as kernel arguments, so nothing on the host must outlive the call).
static int set_slot_form(sp_batch *b, int first, int n, int packed) {
  for (int i = 0; i < n; i++) b->packed[(size_t)first + i] = packed;
  HIP_TRY(launch_set_slot_flags(b->d_packed + first, n, b->packed.data() + first, b->io));
  return SP_OK;
}

// Ends a load: later runs wait for the io stream's work on the device; a
// blocking load also waits on the host, after which the host buffers may be reused.
static int finish_load(sp_batch *b, bool async) {
  HIP_TRY(hipEventRecord(b->loaded, b->io));
  if (!async) HIP_TRY(hipStreamSynchronize(b->io));
  return SP_OK;
}

static int load_slot_packing(sp_batch *b, int first, int n, sp_keys *const *keys);

static int batch_load(sp_batch *b, int slot, const uint32_t *host_query, const uint32_t *host_left,
                      const uint32_t *host_right, const uint32_t *host_v_folding, const uint32_t *host_v_folding_neg) {
  if (!b || !host_query || !host_left || !host_right || !host_v_folding) return fail(SP_ERR_ARG, "load: NULL");
  if (slot < 0 || slot >= b->cap) return fail(SP_ERR_ARG, "load: slot %d of %d", slot, b->cap);
  const int mode = host_v_folding_neg ? 0 : 1;
  if (b->cap > 1 && mode == 0) return fail(SP_ERR_ARG, "load: batches derive v_folding_neg = G - v_folding");
  HIP_TRY(hipSetDevice(b->db->ctx->device));
  HIP_TRY(hipEventSynchronize(b->done));  // the previous run must be done with the slot
  const size_t kl_polys = (size_t)2 * b->t_exp * 2;
  const size_t kf_polys = (size_t)2 * (2 * b->t_gsw) * 2;
  const size_t sl = (size_t)slot;
  int rc;
  if ((rc = upload_p(b, host_query, b->query_ct + sl * 4 * N, 4, 1, 0))) return rc;
  if ((rc = upload_p(b, host_left, b->left + sl * b->l_q, (size_t)b->g * kl_polys, 2 * b->t_exp, 1))) return rc;
  if ((rc = upload_p(b, host_right, b->right + sl * b->r_q, (size_t)(b->stop + 1) * kl_polys, 2 * b->t_exp, 1)))
    return rc;
  if ((rc = upload_p(b, host_v_folding, b->fkeys + sl * b->f_q, (size_t)b->db->nu2 * kf_polys, 4 * b->t_gsw, 1)))
    return rc;
  if (mode == 0) {
    if (!b->fkeys_neg) HIP_TRY(hipMalloc(&b->fkeys_neg, (size_t)b->cap * b->f_q * 4));
    if ((rc = upload_p(b, host_v_folding_neg, b->fkeys_neg + sl * b->f_q, (size_t)b->db->nu2 * kf_polys,
                       4 * b->t_gsw, 1)))
      return rc;
  }
  if ((rc = set_slot_form(b, slot, 1, 0))) return rc;
  if ((rc = load_slot_packing(b, slot, 1, nullptr))) return rc;
  if ((rc = finish_load(b, false))) return rc;
  b->fold_mode = mode;
  return SP_OK;
}

extern "C" int sp_batch_load(sp_batch *b, int slot, const uint32_t *host_query, const uint32_t *host_left,
                             const uint32_t *host_right, const uint32_t *host_v_folding) {
  return batch_load(b, slot, host_query, host_left, host_right, host_v_folding, nullptr);
}

// ---- device-resident client keys ------------------------------------------------
struct sp_keys {
  sp_db *db = nullptr;
  int t_gsw = 0, t_exp = 0;
  size_t l_words = 0, r_words = 0;
  uint32_t *left = nullptr, *right = nullptr;  // layout P, 2^32-scaled: as a batch slot's keys
  uint32_t *conv = nullptr;                    // conversion key [m][2 t_conv][row][P], 2^32-scaled
  uint32_t *pk = nullptr;                      // packing keys [r][m][k][row][P], 2^32-scaled
};

// Packing keys of slots [first, first + n): copied from keys[i] when it has
// them (device to device, io stream), flags recorded on host and device.
static int load_slot_packing(sp_batch *b, int first, int n, sp_keys *const *keys) {
  bool any = false;
  for (int i = 0; i < n; i++) any |= keys && keys[i]->pk;
  if (any && !b->pk) {
    hipError_t e = hipMalloc(&b->pk, (size_t)b->cap * b->pk_q * 4);
    if (e != hipSuccess) {
      b->pk = nullptr;
      return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "packing keys: %s", hipGetErrorString(e));
    }
  }
  for (int i = 0; i < n; i++) {
    const bool has = keys && keys[i]->pk;
    b->haspk[(size_t)first + i] = has;
    if (has)
      HIP_TRY(hipMemcpyAsync(b->pk + ((size_t)first + i) * b->pk_q, keys[i]->pk, b->pk_q * 4,
                             hipMemcpyDeviceToDevice, b->io));
  }
  HIP_TRY(launch_set_slot_flags(b->d_haspk + first, n, b->haspk.data() + first, b->io));
  return SP_OK;
}

extern "C" int sp_keys_create(sp_db *db, int t_gsw, int t_exp, const uint32_t *host_left, const uint32_t *host_right,
                              sp_keys **out) {
  if (!db || !host_left || !host_right || !out) return fail(SP_ERR_ARG, "sp_keys_create: NULL");
  *out = nullptr;
  if (t_exp < 2 || t_exp > 56 || (t_exp & 1) || t_gsw < 2 || t_gsw > 28)
    return fail(SP_ERR_ARG, "sp_keys_create: t_gsw %d / t_exp %d unsupported", t_gsw, t_exp);
  int g = 0, stop = 0;
  int rc = sp_expansion_depth(db->nu1, db->nu2, t_gsw, &g, &stop);
  if (rc) return rc;
  sp_ctx *ctx = db->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t kl = (size_t)2 * t_exp * 2 * N;
  sp_keys *k = new sp_keys();
  k->db = db; k->t_gsw = t_gsw; k->t_exp = t_exp;
  k->l_words = (size_t)g * kl;
  k->r_words = (size_t)(stop + 1) * kl;
  DevBuf stage;
  hipError_t e = hipMalloc(&k->left, k->l_words * 4);
  if (e == hipSuccess) e = hipMalloc(&k->right, k->r_words * 4);
  if (e == hipSuccess) e = stage.alloc(std::max(k->l_words, k->r_words) * 4);
  if (e == hipSuccess) e = hipMemcpyAsync(stage.p, host_left, k->l_words * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_nat_to_p(stage.u32(), k->left, k->l_words / N, 2 * t_exp, 1, ctx->tables, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(stage.p, host_right, k->r_words * 4, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = launch_nat_to_p(stage.u32(), k->right, k->r_words / N, 2 * t_exp, 1, ctx->tables, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    if (k->left) (void)hipFree(k->left);
    if (k->right) (void)hipFree(k->right);
    delete k;
    return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "sp_keys_create: %s", hipGetErrorString(e));
  }
  *out = k;
  return SP_OK;
}

extern "C" int sp_keys_destroy(sp_keys *k) {
  if (!k) return SP_OK;
  (void)hipSetDevice(k->db->ctx->device);
  (void)hipDeviceSynchronize();
  if (k->left) (void)hipFree(k->left);
  if (k->right) (void)hipFree(k->right);
  if (k->conv) (void)hipFree(k->conv);
  if (k->pk) (void)hipFree(k->pk);
  delete k;
  return SP_OK;
}

extern "C" int sp_keys_set_conversion(sp_keys *k, int t_conv, const uint32_t *host_conv) {
  if (!k || !host_conv) return fail(SP_ERR_ARG, "sp_keys_set_conversion: NULL");
  if (t_conv != SP_T_CONV) return fail(SP_ERR_ARG, "sp_keys_set_conversion: t_conv %d unsupported (%d)", t_conv, SP_T_CONV);
  sp_ctx *ctx = k->db->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t polys = (size_t)2 * (2 * t_conv) * 2;
  if (!k->conv) {
    hipError_t e = hipMalloc(&k->conv, polys * N * 4);
    if (e != hipSuccess) {
      k->conv = nullptr;
      return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "sp_keys_set_conversion: %s",
                  hipGetErrorString(e));
    }
  }
  DevBuf stage;
  HIP_TRY(stage.alloc(polys * N * 4));
  HIP_TRY(hipMemcpyAsync(stage.p, host_conv, polys * N * 4, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(launch_nat_to_p(stage.u32(), k->conv, polys, 2 * (2 * t_conv), 1, ctx->tables, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return SP_OK;
}

// Slots [first, first + n): query i (host, [4N], NTT natural) from
// host_queries[i], its resident keys keys[i], its v_folding (host,
// [nu2][2][2 t_gsw][2][N]) from host_v_folding[i].  The transfers go to a
// device staging area (asynchronous DMA when the host buffers are pinned),
// one layout conversion per array covers all n queries, keys are copied
// device to device.
static int batch_load_many(sp_batch *b, int first, int n, const uint32_t *const *host_queries, sp_keys *const *keys,
                           const uint32_t *const *host_v_folding, bool async) {
  if (!b || !host_queries || !keys || !host_v_folding) return fail(SP_ERR_ARG, "sp_batch_load_many: NULL");
  if (n < 1 || first < 0 || first + n > b->cap)
    return fail(SP_ERR_ARG, "sp_batch_load_many: slots [%d, %d) of %d", first, first + n, b->cap);
  for (int i = 0; i < n; i++) {
    const sp_keys *k = keys[i];
    if (!host_queries[i] || !host_v_folding[i]) return fail(SP_ERR_ARG, "sp_batch_load_many: query %d NULL", i);
    if (!k || k->db != b->db || k->t_exp != b->t_exp || k->t_gsw != b->t_gsw || k->l_words != b->l_q ||
        k->r_words != b->r_q)
      return fail(SP_ERR_ARG, "sp_batch_load_many: keys %d do not match this batch's database and gadget", i);
  }
  sp_ctx *ctx = b->db->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t qw = (size_t)4 * N;
  if (!b->bulk) {
    hipError_t e = hipMalloc(&b->bulk, (size_t)b->cap * (qw + b->f_q) * 4);
    if (e != hipSuccess)
      return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "bulk staging: %s", hipGetErrorString(e));
  }
  HIP_TRY(hipStreamWaitEvent(b->io, b->done, 0));  // the previous run must be done with the slots
  uint32_t *sq = b->bulk, *sf = b->bulk + (size_t)b->cap * qw;
  const size_t f0 = (size_t)first;
  for (int i = 0; i < n; i++) {
    HIP_TRY(hipMemcpyAsync(sq + (size_t)i * qw, host_queries[i], qw * 4, hipMemcpyHostToDevice, b->io));
    HIP_TRY(hipMemcpyAsync(sf + (size_t)i * b->f_q, host_v_folding[i], b->f_q * 4, hipMemcpyHostToDevice, b->io));
  }
  HIP_TRY(launch_nat_to_p(sq, b->query_ct + f0 * qw, (size_t)n * 4, 1, 0, ctx->tables, b->io));
  HIP_TRY(launch_nat_to_p(sf, b->fkeys + f0 * b->f_q, (size_t)n * b->f_q / N, 4 * b->t_gsw, 1, ctx->tables, b->io));
  for (int i = 0; i < n; i++) {
    HIP_TRY(hipMemcpyAsync(b->left + (f0 + i) * b->l_q, keys[i]->left, b->l_q * 4, hipMemcpyDeviceToDevice, b->io));
    HIP_TRY(hipMemcpyAsync(b->right + (f0 + i) * b->r_q, keys[i]->right, b->r_q * 4, hipMemcpyDeviceToDevice, b->io));
  }
  int rc = set_slot_form(b, first, n, 0);
  if (!rc) rc = load_slot_packing(b, first, n, keys);
  if (!rc) rc = finish_load(b, async);
  if (rc) return rc;
  b->fold_mode = 1;
  return SP_OK;
}

extern "C" int sp_batch_load_many(sp_batch *b, int first, int n, const uint32_t *const *host_queries,
                                  sp_keys *const *keys, const uint32_t *const *host_v_folding) {
  return batch_load_many(b, first, n, host_queries, keys, host_v_folding, false);
}

extern "C" int sp_batch_load_many_async(sp_batch *b, int first, int n, const uint32_t *const *host_queries,
                                        sp_keys *const *keys, const uint32_t *const *host_v_folding) {
  return batch_load_many(b, first, n, host_queries, keys, host_v_folding, true);
}

extern "C" int sp_keys_set_packing(sp_keys *k, int t_conv, const uint32_t *host_w) {
  if (!k || !host_w) return fail(SP_ERR_ARG, "sp_keys_set_packing: NULL");
  if (k->db->n2 != SP_PACK_N * SP_PACK_N) return fail(SP_ERR_ARG, "sp_keys_set_packing: packing needs n2 = 4");
  if (t_conv != SP_T_CONV) return fail(SP_ERR_ARG, "sp_keys_set_packing: t_conv %d unsupported (%d)", t_conv, SP_T_CONV);
  sp_ctx *ctx = k->db->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t polys = (size_t)SP_PACK_N * 2 * t_conv * (SP_PACK_N + 1);
  if (!k->pk) {
    hipError_t e = hipMalloc(&k->pk, polys * N * 4);
    if (e != hipSuccess) {
      k->pk = nullptr;
      return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "sp_keys_set_packing: %s", hipGetErrorString(e));
    }
  }
  DevBuf stage;
  HIP_TRY(stage.alloc(polys * N * 4));
  HIP_TRY(hipMemcpyAsync(stage.p, host_w, polys * N * 4, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(launch_nat_to_p(stage.u32(), k->pk, polys, t_conv * (SP_PACK_N + 1), 1, ctx->tables, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return SP_OK;
}

extern "C" int sp_batch_set_packing(sp_batch *b, int q2_bits) {
  if (!b) return fail(SP_ERR_ARG, "sp_batch_set_packing: NULL");
  if (q2_bits == 0) {
    b->pack_q2 = 0;
    return SP_OK;
  }
  int rc = check_packing(b->db->n2, q2_bits, "sp_batch_set_packing");
  if (rc) return rc;
  HIP_TRY(hipSetDevice(b->db->ctx->device));
  HIP_TRY(hipEventSynchronize(b->done));
  // Sized once for the largest response (q2_bits = 32) and never reallocated:
  // captured graphs keep pointing at it whatever q2_bits later runs use.
  if (!b->pout) {
    hipError_t e = hipMalloc(&b->pout, (size_t)b->cap * packed_response_bytes(32));
    if (e != hipSuccess) {
      b->pout = nullptr;
      return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "sp_batch_set_packing: %s", hipGetErrorString(e));
    }
  }
  if (!b->pk) {
    hipError_t e = hipMalloc(&b->pk, (size_t)b->cap * b->pk_q * 4);
    if (e != hipSuccess) {
      b->pk = nullptr;
      return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "sp_batch_set_packing: %s", hipGetErrorString(e));
    }
  }
  b->pack_q2 = q2_bits;
  return SP_OK;
}

// Checks for a readout of the packed responses of slots [0, n) of the last run.
static int check_read_packed(const sp_batch *b, int n, const char *who) {
  if (n < 1 || n > b->cap) return fail(SP_ERR_ARG, "%s: n = %d", who, n);
  if (!b->pack_q2_run) return fail(SP_ERR_STATE, "%s: the last run did not pack its responses", who);
  if (b->pack_q2_run != b->pack_q2)  // callers size the readout by the current setting
    return fail(SP_ERR_STATE, "%s: packing changed to q2_bits %d after a run at %d; run again", who, b->pack_q2,
                b->pack_q2_run);
  for (int i = 0; i < n; i++)
    if (!b->haspk[i]) return fail(SP_ERR_STATE, "%s: slot %d has no packing keys", who, i);
  return SP_OK;
}

extern "C" int sp_batch_read_packed(sp_batch *b, int n, uint8_t *host_out) {
  if (!b || !host_out) return fail(SP_ERR_ARG, "sp_batch_read_packed: NULL");
  int rc = check_read_packed(b, n, "sp_batch_read_packed");
  if (rc) return rc;
  HIP_TRY(hipSetDevice(b->db->ctx->device));
  HIP_TRY(hipEventSynchronize(b->done));
  // packed responses of a run are contiguous, packed_response_bytes(q2 of that run) apart
  HIP_TRY(hipMemcpyAsync(host_out, b->pout, (size_t)n * packed_response_bytes(b->pack_q2_run), hipMemcpyDeviceToHost,
                         b->io));
  HIP_TRY(hipStreamSynchronize(b->io));
  return SP_OK;
}

extern "C" int sp_batch_read_packed_async(sp_batch *b, int n, uint8_t *host_out, void *stream) {
  if (!b || !host_out) return fail(SP_ERR_ARG, "sp_batch_read_packed_async: NULL");
  int rc = check_read_packed(b, n, "sp_batch_read_packed_async");
  if (rc) return rc;
  HIP_TRY(hipSetDevice(b->db->ctx->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  HIP_TRY(hipStreamWaitEvent(s, b->done, 0));
  HIP_TRY(hipMemcpyAsync(host_out, b->pout, (size_t)n * packed_response_bytes(b->pack_q2_run), hipMemcpyDeviceToHost,
                         s));
  HIP_TRY(hipEventRecord(b->done, s));
  return SP_OK;
}

static int batch_load_packed(sp_batch *b, int first, int n, const uint32_t *const *host_queries,
                             sp_keys *const *keys, bool async) {
  if (!b || !host_queries || !keys) return fail(SP_ERR_ARG, "sp_batch_load_packed: NULL");
  if (n < 1 || first < 0 || first + n > b->cap)
    return fail(SP_ERR_ARG, "sp_batch_load_packed: slots [%d, %d) of %d", first, first + n, b->cap);
  for (int i = 0; i < n; i++) {
    const sp_keys *k = keys[i];
    if (!host_queries[i]) return fail(SP_ERR_ARG, "sp_batch_load_packed: query %d NULL", i);
    if (!k || k->db != b->db || k->t_exp != b->t_exp || k->t_gsw != b->t_gsw || k->l_words != b->l_q ||
        k->r_words != b->r_q)
      return fail(SP_ERR_ARG, "sp_batch_load_packed: keys %d do not match this batch's database and gadget", i);
    if (!k->conv) return fail(SP_ERR_ARG, "sp_batch_load_packed: keys %d have no conversion key", i);
  }
  sp_ctx *ctx = b->db->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  const size_t qw = (size_t)4 * N;
  hipError_t e = hipSuccess;
  if (!b->bulk) e = hipMalloc(&b->bulk, (size_t)b->cap * (qw + b->f_q) * 4);
  if (e == hipSuccess && !b->conv) e = hipMalloc(&b->conv, (size_t)b->cap * b->conv_q * 4);
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "packed-load buffers: %s", hipGetErrorString(e));
  HIP_TRY(hipStreamWaitEvent(b->io, b->done, 0));  // the previous run must be done with the slots
  const size_t f0 = (size_t)first;
  for (int i = 0; i < n; i++)
    HIP_TRY(hipMemcpyAsync(b->bulk + (size_t)i * qw, host_queries[i], qw * 4, hipMemcpyHostToDevice, b->io));
  HIP_TRY(launch_nat_to_p(b->bulk, b->query_ct + f0 * qw, (size_t)n * 4, 1, 0, ctx->tables, b->io));
  for (int i = 0; i < n; i++) {
    HIP_TRY(hipMemcpyAsync(b->left + (f0 + i) * b->l_q, keys[i]->left, b->l_q * 4, hipMemcpyDeviceToDevice, b->io));
    HIP_TRY(hipMemcpyAsync(b->right + (f0 + i) * b->r_q, keys[i]->right, b->r_q * 4, hipMemcpyDeviceToDevice, b->io));
    HIP_TRY(hipMemcpyAsync(b->conv + (f0 + i) * b->conv_q, keys[i]->conv, b->conv_q * 4, hipMemcpyDeviceToDevice,
                           b->io));
  }
  int rc = set_slot_form(b, first, n, 1);
  if (!rc) rc = load_slot_packing(b, first, n, keys);
  if (!rc) rc = finish_load(b, async);
  if (rc) return rc;
  b->fold_mode = 1;
  return SP_OK;
}

extern "C" int sp_batch_load_packed(sp_batch *b, int first, int n, const uint32_t *const *host_queries,
                                    sp_keys *const *keys) {
  return batch_load_packed(b, first, n, host_queries, keys, false);
}

extern "C" int sp_batch_load_packed_async(sp_batch *b, int first, int n, const uint32_t *const *host_queries,
                                          sp_keys *const *keys) {
  return batch_load_packed(b, first, n, host_queries, keys, true);
}

// Makes stream s wait on the device for the batch's uploads, unless they are
// known to be done (a wait on a completed event still costs the stream a
// barrier packet, a few microseconds per run).
static hipError_t wait_loaded(sp_batch *b, hipStream_t s) {
  const hipError_t q = hipEventQuery(b->loaded);
  if (q == hipSuccess) return hipSuccess;
  if (q == hipErrorNotReady) (void)hipGetLastError();  // a pending event is not an error for later launch checks
  return hipStreamWaitEvent(s, b->loaded, 0);
}

// Whether any of slots [0, n) holds a packed query (the run then converts).
static int slots_packed(const sp_batch *b, int n) {
  for (int i = 0; i < n; i++)
    if (b->packed[i]) return 1;
  return 0;
}

static int mfma_workspace(sp_batch *b) {
  if (b->mfma_a) return SP_OK;
  const sp_db *db = b->db;
  const size_t I = (size_t)mfma_rows(b->cap);
  const size_t a_bytes = (size_t)2 * N * ((size_t)1 << db->nu1) * 4 * I;
  const size_t o_words = (size_t)2 * N * db->n2 * I * ((size_t)1 << db->nu2);
  hipError_t e = hipMalloc(&b->mfma_a, a_bytes);
  if (e == hipSuccess) e = hipMalloc(&b->mfma_out, o_words * 4);
  if (e != hipSuccess)
    return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "MFMA workspace: %s", hipGetErrorString(e));
  return SP_OK;
}

// Whether the multi-query VALU multiply takes this shape (whole packed groups of 8 columns, 8+ rows).
static bool multi_shape_ok(const sp_db *db) { return db->nu2 >= 3 && db->nu1 >= 3; }

// Whether variant v of the database multiply can run n queries on this database.
static bool dbmul_applies(const sp_db *db, int v, int n) {
  if (v == 1) return true;
  if (v == 2) return mfma_shape_ok(db);
  return (v == 3 || v == 4) && n >= 2 && multi_shape_ok(db);
}

// Database-multiply variant for n queries: 1 VALU one query per database read,
// 2 int8-limb MFMA, 3 / 4 VALU 2 / 4 queries per database read.  Forced by the
// context where it applies, else the autotuner's cached choice, else the
// multi-query VALU form for n >= 2 and the one-query form for n = 1.
static int dbmul_variant(sp_batch *b, int n) {
  const int forced = b->db->ctx->variant[SP_STAGE_DB_MUL];
  if (forced && dbmul_applies(b->db, forced, n)) return forced;
  if (!forced)
    for (auto &c : b->dbmul_choice)
      if (c.first == n) return c.second;
  return dbmul_applies(b->db, 4, n) ? 4 : 1;
}

// Columns per block of the from_ntt that reads the MFMA products (1, 2 or 4).
#ifndef SP_PM_NC
#define SP_PM_NC 1
#endif

// Products left in mfma_out by the MFMA multiply -> the NTT accumulators (what the
// VALU variants write), for runs and readouts that need them there.
static hipError_t pm_to_acc(sp_batch *b, hipStream_t s) {
  if (b->pm_n == 0) return hipSuccess;
  const sp_db *db = b->db;
  const hipError_t e = launch_outp_to_acc(b->mfma_out, db->n2, mfma_rows(b->pm_n), 1 << db->nu2, b->pm_n, b->acc,
                                          b->acc_q, db->ctx->tables, s);
  b->pm_n = 0;
  return e;
}

// Enqueues the stages for slots [0, n) on stream s (no synchronization).
static int batch_enqueue(sp_batch *b, int n, int stages, hipStream_t s) {
  sp_db *db = b->db;
  sp_ctx *ctx = db->ctx;
  const int R = 1 << db->nu1, C = 1 << db->nu2;
  const int packed = slots_packed(b, n);
  if (stages & 1) {
    ExpandSnap snap;
    snap.p = b->snap;
    snap.queries = std::min(b->cap, SNAP_QUERIES);
    snap.q_stride = ((size_t)1 << b->g) * 2 * N;
    snap.set_stride = (size_t)snap.queries * snap.q_stride;
    int rc = run_expansion(ctx, b->v, b->v_q, n, db->nu1, db->nu2, b->t_gsw, b->t_exp, b->left, b->right, b->l_q,
                           b->r_q, -1, b->dig, b->dig_q, snap, b->query_ct, 4 * N, s, (stages & 2) != 0,
                           b->d_packed, (stages & 2) != 0 && !packed);
    if (rc) return rc;
    if (packed)
      HIP_TRY(launch_regev_to_gsw(true, b->v, b->v_q, n, db->nu2, b->t_gsw, b->conv, b->conv_q, b->fkeys, b->f_q,
                                  b->d_packed, ctx->tables, s));
  }
  if (stages & 2) {
    const int var = dbmul_variant(b, n);
    b->last_dbmul = var;
    if (var == 2) {
      int rc = db_limbs(db, s);
      if (rc) return rc;
      rc = mfma_workspace(b);
      if (rc) return rc;
      // The products stay in mfma_out: stage 4 takes them out of the NTT domain
      // straight from there (or a readout of the accumulators transposes them).
      HIP_TRY(launch_dbmul_mfma(b->v, b->v_q, n, db->d_limbs, db->n2, R, C, mfma_rows(n), b->mfma_a, b->mfma_out,
                                nullptr, b->acc_q, ctx->d_pw, ctx->tables, mfma_rows(n) > 64 ? 4 : -1, s));
      b->pm_n = n;
    } else if (var == 3 || var == 4) {
      HIP_TRY(launch_db_mul_multi(db->d_db, b->v, 2 * 4 * N, b->v_q, b->acc, b->acc_q, n, db->n2, R, C, var - 2,
                                  ctx->tables, s));
      b->pm_n = 0;
    } else {
      HIP_TRY(launch_db_mul(db->d_db, b->v, 2 * 4 * N, b->v_q, b->acc, b->acc_q, n, db->n2, R, C, ctx->tables, s));
      b->pm_n = 0;
    }
  }
  if (stages & 4) {
    if (b->pm_n >= n) {
      HIP_TRY(launch_from_ntt_pm(b->mfma_out, db->n2, mfma_rows(b->pm_n), C, n, b->acc, b->acc_q, SP_PM_NC,
                                 ctx->tables, s));
    } else {
      HIP_TRY(pm_to_acc(b, s));
      HIP_TRY(launch_from_ntt(true, b->acc, b->acc, (size_t)n * db->n2 * C * 2, 2 * N, 2 * N, ctx->tables, s));
    }
    b->pm_n = 0;
    const bool fast = b->fold_mode != 0;
    const int mx = 2 * b->t_gsw;
    const size_t level = (size_t)2 * mx * 2 * N;
    // Latency form for batches of at most FOLD_QUERIES queries (automatic choice only).
    const bool latency = fast && b->t_gsw == 8 && ctx->variant[SP_STAGE_FOLD] == 0 && n <= FOLD_QUERIES &&
                         b->fold_alt && b->part_alt;
    if (latency) {
      const size_t alt = (size_t)std::max(C / 4, 1);
      HIP_TRY(launch_fold_latency(b->acc, (size_t)C * 4 * N, b->fold_alt, alt * 4 * N, b->part, (size_t)(C / 2) * 8 * N,
                                  b->part_alt, alt * 16 * N, db->n2, n, db->nu2, b->fkeys, level, b->f_q, ctx->tables,
                                  s));
    }
    int num_per = C;
    for (int cur = 0; cur < db->nu2 && !latency; cur++) {
      num_per >>= 1;
      const size_t off = (size_t)(db->nu2 - 1 - cur) * level;
      if (fast && b->t_gsw == 8 && ctx->variant[SP_STAGE_FOLD] != 1)
        HIP_TRY(launch_fold8(b->acc, (size_t)C * 4 * N, db->n2, n, num_per, b->fkeys + off, b->f_q, ctx->tables,
                             ctx->variant[SP_STAGE_FOLD], b->part, (size_t)(C / 2) * 8 * N, s));
      else
        HIP_TRY(launch_fold_level(fast, b->acc, (size_t)C * 4 * N, db->n2, n, num_per, b->fkeys + off,
                                  fast ? nullptr : b->fkeys_neg + off, b->f_q, b->t_gsw, gadget_bits(b->t_gsw),
                                  ctx->tables, s));
    }
    if (b->pack_q2)
      HIP_TRY(launch_pack_response(b->acc, b->acc_q, (size_t)C * 4 * N, n, b->pk, b->pk_q, b->d_haspk, b->pout,
                                   packed_response_bytes(b->pack_q2), b->pack_q2, ctx->tables, s));
  }
  return SP_OK;
}

extern "C" int sp_batch_run(sp_batch *b, int n, int stages, void *stream) {
  if (!b) return fail(SP_ERR_ARG, "sp_batch_run: NULL");
  if (n < 1 || n > b->cap) return fail(SP_ERR_ARG, "sp_batch_run: n = %d, capacity %d", n, b->cap);
  if (b->fold_mode < 0) return fail(SP_ERR_STATE, "sp_batch_run: nothing loaded");
  HIP_TRY(hipSetDevice(b->db->ctx->device));
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  HIP_TRY(wait_loaded(b, s));  // the slots' uploads, on the device
  int rc = batch_enqueue(b, n, stages, s);
  if (rc) return rc;
  if (stages & 4) b->pack_q2_run = b->pack_q2;
  HIP_TRY(hipEventRecord(b->done, s));
  return SP_OK;
}

// Replays a captured graph of all stages for n queries (captured on first use).
extern "C" int sp_batch_graph_launch(sp_batch *b, int n, void *stream) {
  if (!b) return fail(SP_ERR_ARG, "sp_batch_graph_launch: NULL");
  if (n < 1 || n > b->cap) return fail(SP_ERR_ARG, "sp_batch_graph_launch: n = %d, capacity %d", n, b->cap);
  if (b->fold_mode < 0) return fail(SP_ERR_STATE, "sp_batch_graph_launch: nothing loaded");
  GraphKey gkey;
  gkey.n = n;
  gkey.packed = slots_packed(b, n);
  gkey.pack_q2 = b->pack_q2;
  gkey.dbmul = dbmul_variant(b, n);
  gkey.expand = b->db->ctx->variant[SP_STAGE_EXPAND];
  gkey.fold = b->db->ctx->variant[SP_STAGE_FOLD];
  HIP_TRY(hipSetDevice(b->db->ctx->device));
  hipGraphExec_t exec = nullptr;
  for (auto &ge : b->graphs)
    if (ge.first == gkey) exec = ge.second;
  if (!exec) {
    // One-time preparations stay out of the graph.
    if (dbmul_variant(b, n) == 2) {
      int rc = db_limbs(b->db, b->cap_stream);
      if (!rc) rc = mfma_workspace(b);
      if (rc) return rc;
      HIP_TRY(hipStreamSynchronize(b->cap_stream));
    }
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = batch_enqueue(b, n, 7, b->cap_stream);
    hipError_t e = hipStreamEndCapture(b->cap_stream, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(SP_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(SP_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    b->graphs.push_back({gkey, exec});
    b->graph_defs.push_back(graph);
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  HIP_TRY(wait_loaded(b, s));  // the slots' uploads, on the device
  HIP_TRY(hipGraphLaunch(exec, s));
  b->last_dbmul = gkey.dbmul;
  b->pm_n = 0;  // the replay ran every stage
  b->pack_q2_run = gkey.pack_q2;
  HIP_TRY(hipEventRecord(b->done, s));
  return SP_OK;
}

extern "C" int sp_batch_last_run(const sp_batch *b, int *dbmul_variant, int *graphs_cached) {
  if (!b) return fail(SP_ERR_ARG, "sp_batch_last_run: NULL");
  if (dbmul_variant) *dbmul_variant = b->last_dbmul;
  if (graphs_cached) *graphs_cached = (int)b->graphs.size();
  return SP_OK;
}

// ---- several batches (each on its own database) in one launch --------------------
struct sp_group {
  std::vector<sp_batch *> b;
  hipStream_t fork = nullptr;
  hipEvent_t ev_fork = nullptr;
  std::vector<hipEvent_t> ev_join;
  std::vector<std::pair<std::vector<int>, hipGraphExec_t>> graphs;  // ns -> executable graph
  std::vector<hipGraph_t> graph_defs;
  hipEvent_t done = nullptr;
};

static void group_free(sp_group *g) {
  for (auto &ge : g->graphs) (void)hipGraphExecDestroy(ge.second);
  for (auto &gd : g->graph_defs) (void)hipGraphDestroy(gd);
  for (auto e : g->ev_join)
    if (e) (void)hipEventDestroy(e);
  if (g->ev_fork) (void)hipEventDestroy(g->ev_fork);
  if (g->done) (void)hipEventDestroy(g->done);
  if (g->fork) (void)hipStreamDestroy(g->fork);
  delete g;
}

extern "C" int sp_group_create(sp_batch *const *batches, int count, sp_group **out) {
  if (!batches || !out || count < 1 || count > 64) return fail(SP_ERR_ARG, "sp_group_create: bad argument");
  *out = nullptr;
  const int dev = batches[0] ? batches[0]->db->ctx->device : -1;
  for (int i = 0; i < count; i++) {
    if (!batches[i] || batches[i]->db->ctx->device != dev) return fail(SP_ERR_ARG, "sp_group_create: batch %d", i);
    for (int j = 0; j < i; j++)
      if (batches[j] == batches[i]) return fail(SP_ERR_ARG, "sp_group_create: batch %d repeated", i);
  }
  HIP_TRY(hipSetDevice(dev));
  sp_group *g = new sp_group();
  g->b.assign(batches, batches + count);
  g->ev_join.assign(count, nullptr);
  hipError_t e = hipStreamCreateWithFlags(&g->fork, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&g->done, hipEventDisableTiming);
  for (int i = 0; i < count && e == hipSuccess; i++) e = hipEventCreateWithFlags(&g->ev_join[i], hipEventDisableTiming);
  if (e != hipSuccess) {
    group_free(g);
    return fail(SP_ERR_HIP, "sp_group_create: %s", hipGetErrorString(e));
  }
  *out = g;
  return SP_OK;
}

extern "C" int sp_group_destroy(sp_group *g) {
  if (!g) return SP_OK;
  (void)hipSetDevice(g->b[0]->db->ctx->device);
  (void)hipDeviceSynchronize();
  group_free(g);
  return SP_OK;
}

// All stages of slots [0, ns[i]) of every batch i with ns[i] > 0, as one graph
// whose branches (one per batch) run concurrently; captured on first use for
// this ns and replayed on `stream`.
extern "C" int sp_group_launch(sp_group *g, const int *ns, void *stream) {
  if (!g || !ns) return fail(SP_ERR_ARG, "sp_group_launch: NULL");
  const int count = (int)g->b.size();
  std::vector<int> key(ns, ns + count);
  int active = 0;
  for (int i = 0; i < count; i++) {
    if (key[i] < 0 || key[i] > g->b[i]->cap) return fail(SP_ERR_ARG, "sp_group_launch: ns[%d] = %d", i, key[i]);
    if (key[i] > 0 && g->b[i]->fold_mode < 0) return fail(SP_ERR_STATE, "sp_group_launch: batch %d not loaded", i);
    active += key[i] > 0;
  }
  for (int i = 0; i < count; i++) {  // the graph also depends on conversion, packing and kernel variants
    const sp_batch *b = g->b[i];
    const bool on = key[i] > 0;
    key.push_back(on ? slots_packed(b, key[i]) : 0);
    key.push_back(on ? b->pack_q2 : 0);
    key.push_back(on ? dbmul_variant(g->b[i], key[i]) : 0);
    key.push_back(on ? b->db->ctx->variant[SP_STAGE_EXPAND] : 0);
    key.push_back(on ? b->db->ctx->variant[SP_STAGE_FOLD] : 0);
  }
  if (!active) return SP_OK;
  HIP_TRY(hipSetDevice(g->b[0]->db->ctx->device));
  hipGraphExec_t exec = nullptr;
  for (auto &ge : g->graphs)
    if (ge.first == key) exec = ge.second;
  if (!exec) {
    for (int i = 0; i < count; i++) {  // one-time preparations stay out of the graph
      sp_batch *b = g->b[i];
      if (key[i] > 0 && dbmul_variant(b, key[i]) == 2) {
        int rc = db_limbs(b->db, b->cap_stream);
        if (!rc) rc = mfma_workspace(b);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(b->cap_stream));
      }
    }
    hipGraph_t graph = nullptr;
    HIP_TRY(hipStreamBeginCapture(g->fork, hipStreamCaptureModeThreadLocal));
    int rc = SP_OK;
    hipError_t e = hipEventRecord(g->ev_fork, g->fork);
    for (int i = 0; i < count && !rc && e == hipSuccess; i++) {
      if (key[i] == 0) continue;
      sp_batch *b = g->b[i];
      e = hipStreamWaitEvent(b->cap_stream, g->ev_fork, 0);
      if (e == hipSuccess) rc = batch_enqueue(b, key[i], 7, b->cap_stream);
      if (!rc && e == hipSuccess) e = hipEventRecord(g->ev_join[i], b->cap_stream);
      if (!rc && e == hipSuccess) e = hipStreamWaitEvent(g->fork, g->ev_join[i], 0);
    }
    hipError_t e2 = hipStreamEndCapture(g->fork, &graph);
    if (rc || e != hipSuccess || e2 != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (rc) return rc;
      return fail(SP_ERR_HIP, "sp_group_launch capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(graph);
      return fail(SP_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    g->graphs.push_back({key, exec});
    g->graph_defs.push_back(graph);
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  for (int i = 0; i < count; i++)
    if (key[i] > 0) HIP_TRY(wait_loaded(g->b[i], s));  // uploads of every active batch
  HIP_TRY(hipGraphLaunch(exec, s));
  HIP_TRY(hipEventRecord(g->done, s));
  for (int i = 0; i < count; i++)
    if (key[i] > 0) {
      sp_batch *b = g->b[i];
      b->last_dbmul = dbmul_variant(b, key[i]);
      b->pm_n = 0;  // the replay ran every stage
      b->pack_q2_run = b->pack_q2;
      HIP_TRY(hipEventRecord(b->done, s));
    }
  return SP_OK;
}

extern "C" int sp_group_synchronize(sp_group *g) {
  if (!g) return fail(SP_ERR_ARG, "sp_group_synchronize: NULL");
  HIP_TRY(hipEventSynchronize(g->done));
  return SP_OK;
}

// Times the database multiply of n queries with each variant (VALU u64 and,
// where the shape allows, int8-limb MFMA) on the batch's io stream, keeps
// the faster one for n.  Needs the expansion (stage 1) to have run.
extern "C" int sp_batch_tune(sp_batch *b, int n, float *ms_valu, float *ms_mfma, int *chosen) {
  if (!b) return fail(SP_ERR_ARG, "sp_batch_tune: NULL");
  if (n < 1 || n > b->cap) return fail(SP_ERR_ARG, "sp_batch_tune: n = %d", n);
  if (b->fold_mode < 0) return fail(SP_ERR_STATE, "sp_batch_tune: nothing loaded");
  sp_ctx *ctx = b->db->ctx;
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(hipEventSynchronize(b->done));
  const int saved = ctx->variant[SP_STAGE_DB_MUL];
  float best = 0;
  int pick = 1;
  float ms[5] = {-1, -1, -1, -1, -1};
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int rc = SP_OK;
  // MFMA first (the usual winner for larger runs), then the multi-query and
  // one-query VALU forms; a form whose first timed run is over twice the best
  // so far is not repeated.
  static const int order[4] = {2, 4, 3, 1};
  for (int oi = 0; oi < 4 && !rc; oi++) {
    const int var = order[oi];
    if (!dbmul_applies(b->db, var, n)) continue;
    ctx->variant[SP_STAGE_DB_MUL] = var;
    if (var == 2) rc = batch_enqueue(b, n, 2, b->io);  // warm-up: builds the limbs and workspaces
    float t = 0;
    int reps = 0;
    for (int k = 0; k < 3 && !rc; k++) {
      if (hipEventRecord(e0, b->io) != hipSuccess) rc = fail(SP_ERR_HIP, "event");
      if (!rc) rc = batch_enqueue(b, n, 2, b->io);
      if (!rc && hipEventRecord(e1, b->io) != hipSuccess) rc = fail(SP_ERR_HIP, "event");
      if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = fail(SP_ERR_HIP, "event sync");
      float dt = 0;
      if (!rc) (void)hipEventElapsedTime(&dt, e0, e1);
      t += dt;
      reps++;
      if (best > 0 && dt > 2 * best) break;
    }
    if (!rc) {
      ms[var] = t / reps;
      b->dbmul_ms[var].push_back({n, ms[var]});
      if (best == 0 || ms[var] < best) { best = ms[var]; pick = var; }
    }
  }
  ctx->variant[SP_STAGE_DB_MUL] = saved;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  bool found = false;
  for (auto &c : b->dbmul_choice)
    if (c.first == n) { c.second = pick; found = true; }
  if (!found) b->dbmul_choice.push_back({n, pick});
  if (ms_valu) *ms_valu = ms[1];
  if (ms_mfma) *ms_mfma = ms[2];
  if (chosen) *chosen = pick;
  b->tune_ms[0] = ms[3];
  b->tune_ms[1] = ms[4];
  return SP_OK;
}

extern "C" int sp_batch_set_dbmul(sp_batch *b, int n, int variant) {
  if (!b) return fail(SP_ERR_ARG, "sp_batch_set_dbmul: NULL");
  if (n < 1 || n > b->cap) return fail(SP_ERR_ARG, "sp_batch_set_dbmul: n = %d", n);
  if (variant < 1 || variant > 4 || !dbmul_applies(b->db, variant, n))
    return fail(SP_ERR_ARG, "sp_batch_set_dbmul: variant %d does not apply to %d queries of this database", variant, n);
  for (auto &c : b->dbmul_choice)
    if (c.first == n) { c.second = variant; return SP_OK; }
  b->dbmul_choice.push_back({n, variant});
  return SP_OK;
}

extern "C" int sp_batch_tune_multi_ms(const sp_batch *b, float *ms_q2, float *ms_q4) {
  if (!b) return fail(SP_ERR_ARG, "sp_batch_tune_multi_ms: NULL");
  if (ms_q2) *ms_q2 = b->tune_ms[0];
  if (ms_q4) *ms_q4 = b->tune_ms[1];
  return SP_OK;
}

// what: 0 = response of `slot` (n2 raw Ciphertexts), 1 = its expanded v (NTT,
// natural), 2 = its database-multiply accumulators (NTT, natural; between
// stages 2 and 4), 3 = its n2*C raw accumulators.
static int batch_read(sp_batch *b, int slot, int what, uint32_t *host_out) {
  if (!b || !host_out) return fail(SP_ERR_ARG, "read: NULL");
  if (slot < 0 || slot >= b->cap) return fail(SP_ERR_ARG, "read: slot %d of %d", slot, b->cap);
  sp_db *db = b->db;
  HIP_TRY(hipSetDevice(db->ctx->device));
  HIP_TRY(hipEventSynchronize(b->done));
  const int C = 1 << db->nu2;
  const uint32_t *acc = b->acc + (size_t)slot * b->acc_q;
  if (what == 0) {
    for (int t = 0; t < db->n2; t++)
      HIP_TRY(hipMemcpyAsync(host_out + (size_t)t * 4 * N, acc + (size_t)t * C * 4 * N, 4 * N * 4,
                             hipMemcpyDeviceToHost, b->io));
  } else if (what == 1) {
    const size_t n = ((size_t)1 << b->g) * 4;
    HIP_TRY(launch_p_to_nat(b->v + (size_t)slot * b->v_q, b->stage, n, b->io));
    HIP_TRY(hipMemcpyAsync(host_out, b->stage, n * N * 4, hipMemcpyDeviceToHost, b->io));
  } else if (what == 2) {
    const size_t n = (size_t)db->n2 * C * 4;
    HIP_TRY(pm_to_acc(b, b->io));
    HIP_TRY(launch_p_to_nat(acc, b->stage, n, b->io));
    HIP_TRY(hipMemcpyAsync(host_out, b->stage, n * N * 4, hipMemcpyDeviceToHost, b->io));
  } else if (what == 3) {
    HIP_TRY(hipMemcpyAsync(host_out, acc, b->acc_q * 4, hipMemcpyDeviceToHost, b->io));
  } else if (what == 4) {
    const size_t n = b->f_q / N;
    HIP_TRY(launch_p_to_nat_unscale(b->fkeys + (size_t)slot * b->f_q, b->stage, n, 4 * b->t_gsw, db->ctx->tables,
                                    b->io));
    HIP_TRY(hipMemcpyAsync(host_out, b->stage, b->f_q * 4, hipMemcpyDeviceToHost, b->io));
  } else {
    return fail(SP_ERR_ARG, "read: what = %d", what);
  }
  HIP_TRY(hipStreamSynchronize(b->io));
  return SP_OK;
}

extern "C" int sp_batch_read(sp_batch *b, int slot, int what, uint32_t *host_out) {
  return batch_read(b, slot, what, host_out);
}

// Response of slots [0, n) into host_out [n][n2][4N] in one call.
extern "C" int sp_batch_read_responses(sp_batch *b, int n, uint32_t *host_out) {
  if (!b || !host_out) return fail(SP_ERR_ARG, "sp_batch_read_responses: NULL");
  if (n < 1 || n > b->cap) return fail(SP_ERR_ARG, "sp_batch_read_responses: n = %d", n);
  sp_db *db = b->db;
  HIP_TRY(hipSetDevice(db->ctx->device));
  HIP_TRY(hipEventSynchronize(b->done));
  const int C = 1 << db->nu2;
  // Response t of slot q is the first ciphertext of (q, t), C ciphertexts apart.
  // They are gathered into the staging buffer by a kernel and copied out
  // contiguously: strided 2-D device-to-pageable copies have returned stale
  // bytes under kernel tracing, plain contiguous copies have not.
  const size_t rows = (size_t)n * db->n2;
  const size_t chunk = std::min<size_t>(b->stage_words / (4 * N), 65535);
  for (size_t r0 = 0; r0 < rows; r0 += chunk) {
    const size_t nr = std::min(chunk, rows - r0);
    HIP_TRY(launch_gather_rows(b->acc + r0 * C * 4 * N, (size_t)C * 4 * N, (int)nr, b->stage, b->io));
    HIP_TRY(hipMemcpyAsync(host_out + r0 * 4 * N, b->stage, nr * 4 * N * 4, hipMemcpyDeviceToHost, b->io));
    HIP_TRY(hipStreamSynchronize(b->io));
  }
  return SP_OK;
}

// Responses of slots [0, n) of the last run, copied to host_out on `stream`
// after that run (no host wait): gathered on the device into a staging buffer
// of the batch, then one contiguous copy.  sp_batch_poll / _synchronize cover
// the copy; host_out (pinned for a true DMA) must stay valid until then.
extern "C" int sp_batch_read_responses_async(sp_batch *b, int n, uint32_t *host_out, void *stream) {
  if (!b || !host_out) return fail(SP_ERR_ARG, "sp_batch_read_responses_async: NULL");
  if (n < 1 || n > b->cap) return fail(SP_ERR_ARG, "sp_batch_read_responses_async: n = %d", n);
  sp_db *db = b->db;
  HIP_TRY(hipSetDevice(db->ctx->device));
  const size_t rows = (size_t)n * db->n2;
  if (!b->rstage) {
    hipError_t e = hipMalloc(&b->rstage, (size_t)b->cap * db->n2 * 4 * N * 4);
    if (e != hipSuccess) {
      b->rstage = nullptr;
      return fail(e == hipErrorOutOfMemory ? SP_ERR_NOMEM : SP_ERR_HIP, "response staging: %s", hipGetErrorString(e));
    }
  }
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  HIP_TRY(hipStreamWaitEvent(s, b->done, 0));
  const int C = 1 << db->nu2;
  for (size_t r0 = 0; r0 < rows; r0 += 65535) {
    const size_t nr = std::min<size_t>(65535, rows - r0);
    HIP_TRY(launch_gather_rows(b->acc + r0 * C * 4 * N, (size_t)C * 4 * N, (int)nr, b->rstage + r0 * 4 * N, s));
  }
  HIP_TRY(hipMemcpyAsync(host_out, b->rstage, rows * 4 * N * 4, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipEventRecord(b->done, s));
  return SP_OK;
}

extern "C" int sp_batch_poll(sp_batch *b, int *done) {
  if (!b || !done) return fail(SP_ERR_ARG, "sp_batch_poll: NULL");
  HIP_TRY(hipSetDevice(b->db->ctx->device));
  const hipError_t e = hipEventQuery(b->done);
  if (e == hipSuccess) { *done = 1; return SP_OK; }
  if (e == hipErrorNotReady) { (void)hipGetLastError(); *done = 0; return SP_OK; }
  return fail(SP_ERR_HIP, "sp_batch_poll: %s", hipGetErrorString(e));
}

extern "C" int sp_batch_synchronize(sp_batch *b) {
  if (!b) return fail(SP_ERR_ARG, "sp_batch_synchronize: NULL");
  HIP_TRY(hipEventSynchronize(b->done));
  return SP_OK;
}

// ---- measurement ----------------------------------------------------------------
extern "C" size_t sp_stream_sink_words(void) { return stream_sink_words(); }

extern "C" int sp_stream_read(sp_ctx *ctx, const void *dev_buf, size_t bytes, uint32_t *dev_sink, void *stream) {
  if (!ctx || !dev_buf || !dev_sink) return fail(SP_ERR_ARG, "sp_stream_read: NULL");
  if (bytes == 0 || bytes % 16 || reinterpret_cast<uintptr_t>(dev_buf) % 16)
    return fail(SP_ERR_ARG, "sp_stream_read: %zu bytes (a multiple of 16, 16-byte aligned)", bytes);
  HIP_TRY(hipSetDevice(ctx->device));
  HIP_TRY(launch_stream_read(dev_buf, bytes, dev_sink, pick(ctx, stream)));
  return SP_OK;
}

// ---- single query = batch of capacity 1 -------------------------------------
extern "C" int sp_query_create(sp_db *db, int t_gsw, int t_exp, sp_query **out) {
  if (!out) return fail(SP_ERR_ARG, "sp_query_create: NULL");
  *out = nullptr;
  sp_batch *b = nullptr;
  int rc = sp_batch_create(db, 1, t_gsw, t_exp, &b);
  if (rc) return rc;
  *out = new sp_query{b};
  return SP_OK;
}

extern "C" int sp_query_destroy(sp_query *q) {
  if (!q) return SP_OK;
  sp_batch_destroy(q->b);
  delete q;
  return SP_OK;
}

extern "C" int sp_query_load(sp_query *q, const uint32_t *host_query, const uint32_t *host_left,
                             const uint32_t *host_right, const uint32_t *host_v_folding,
                             const uint32_t *host_v_folding_neg) {
  if (!q) return fail(SP_ERR_ARG, "sp_query_load: NULL");
  return batch_load(q->b, 0, host_query, host_left, host_right, host_v_folding, host_v_folding_neg);
}

extern "C" int sp_query_load_packed(sp_query *q, const uint32_t *host_query, sp_keys *keys) {
  if (!q || !host_query || !keys) return fail(SP_ERR_ARG, "sp_query_load_packed: NULL");
  return sp_batch_load_packed(q->b, 0, 1, &host_query, &keys);
}

extern "C" int sp_query_set_packing(sp_query *q, int q2_bits) {
  if (!q) return fail(SP_ERR_ARG, "sp_query_set_packing: NULL");
  return sp_batch_set_packing(q->b, q2_bits);
}

extern "C" int sp_query_read_packed(sp_query *q, uint8_t *host_out) {
  if (!q) return fail(SP_ERR_ARG, "sp_query_read_packed: NULL");
  return sp_batch_read_packed(q->b, 1, host_out);
}

extern "C" int sp_query_run(sp_query *q, int stages, void *stream) {
  if (!q) return fail(SP_ERR_ARG, "sp_query_run: NULL");
  return sp_batch_run(q->b, 1, stages, stream);
}

extern "C" int sp_query_read(sp_query *q, int what, uint32_t *host_out) {
  if (!q) return fail(SP_ERR_ARG, "sp_query_read: NULL");
  return batch_read(q->b, 0, what, host_out);
}

extern "C" int sp_answer(sp_db *db, int t_gsw, int t_exp, const uint32_t *host_query, const uint32_t *host_left,
                         const uint32_t *host_right, const uint32_t *host_v_folding, uint32_t *host_out) {
  sp_query *q = nullptr;
  int rc = sp_query_create(db, t_gsw, t_exp, &q);
  if (rc) return rc;
  rc = sp_query_load(q, host_query, host_left, host_right, host_v_folding, nullptr);
  if (!rc) rc = sp_query_run(q, 7, q->b->io);
  if (!rc) rc = sp_query_read(q, 0, host_out);
  sp_query_destroy(q);
  return rc;
}
