// sp_client.cpp — Spiral PIR client for the answer path: key generation,
// query generation and response decoding (host code, no GPU).
//
// The reference takes its queries and keys from spiral-rs dumps (the
// test_data/*.dat fixtures, gpu-pir/corrected_store_vars.rs); it has no
// client of its own.  This client produces inputs with the same structure,
// consistent with the server computation the reference defines:
//   * Regev ciphertext (a, a*s + e + mu) in the NTT domain, s ternary,
//     e discrete Gaussian (sigma = 6.4), a uniform mod Q.
//   * Expansion key for round r (t = N/2^r + 1), column k:
//       (a_k, a_k*s + e_k - tau_t(s) * 2^(bits*k))
//     so that v + W g^-1(tau(v).a) + (0, tau(v).as_e) encrypts mu + tau(mu)
//     (gpu-pir/expansion.cuh:141-165).
//   * Query plaintext mu: coefficient 2*i1 = floor(Q/p) * 2^-g mod Q, so that
//     expansion output v[2*i1] (first-dimension selector i1) encrypts
//     floor(Q/p) and every other selector encrypts 0.
//   * Folding key of level l: GSW encryption b*G + Z of bit l of i2, with
//     G[r][r + 2k] = 2^(bits*k) and Z a matrix of encryptions of zero; the
//     server derives v_folding_neg = G - v_folding (gpu-pir/folding.cuh:9-27).
//   * Decoding: round(p * (b - a*s) / Q) mod p per coefficient.
// Randomness: ChaCha20 keystream keyed by the 64-bit seed (synthetic-data
// generator; keys from it are only as secret as the seed).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/spiral_pir.h"

namespace {

constexpr int N = 2048;
constexpr int LOGN = 11;
constexpr uint32_t QM[2] = {268369921u, 249561089u};
constexpr uint32_t PSI[2] = {66687u, 158221u};
constexpr uint64_t QQ = (uint64_t)268369921u * 249561089u;
typedef unsigned __int128 u128;

thread_local std::string g_cerr;
int cfail(int code, const char *msg) {
  g_cerr = msg;
  return code;
}

uint64_t powmod(uint64_t b, uint64_t e, uint64_t m) {
  uint64_t r = 1 % m;
  b %= m;
  while (e) {
    if (e & 1) r = (uint64_t)((u128)r * b % m);
    b = (uint64_t)((u128)b * b % m);
    e >>= 1;
  }
  return r;
}
uint32_t brv11(uint32_t x) {
  uint32_t r = 0;
  for (int i = 0; i < LOGN; i++) { r = (r << 1) | (x & 1); x >>= 1; }
  return r;
}

// ---- ChaCha20 keystream -----------------------------------------------------
struct ChaCha {
  uint32_t key[8];
  uint64_t counter = 0;
  uint32_t block[16];
  int pos = 16;
  explicit ChaCha(uint64_t seed) {
    uint64_t z = seed;
    for (int i = 0; i < 8; i += 2) {  // splitmix64 to spread the seed into a key
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      x ^= x >> 31;
      key[i] = (uint32_t)x;
      key[i + 1] = (uint32_t)(x >> 32);
    }
  }
  static uint32_t rotl(uint32_t v, int c) { return (v << c) | (v >> (32 - c)); }
  static void qr(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d) {
    a += b; d ^= a; d = rotl(d, 16);
    c += d; b ^= c; b = rotl(b, 12);
    a += b; d ^= a; d = rotl(d, 8);
    c += d; b ^= c; b = rotl(b, 7);
  }
  void refill() {
    uint32_t s[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};
    for (int i = 0; i < 8; i++) s[4 + i] = key[i];
    s[12] = (uint32_t)counter;
    s[13] = (uint32_t)(counter >> 32);
    s[14] = 0;
    s[15] = 0;
    uint32_t x[16];
    memcpy(x, s, sizeof x);
    for (int i = 0; i < 10; i++) {
      qr(x[0], x[4], x[8], x[12]); qr(x[1], x[5], x[9], x[13]);
      qr(x[2], x[6], x[10], x[14]); qr(x[3], x[7], x[11], x[15]);
      qr(x[0], x[5], x[10], x[15]); qr(x[1], x[6], x[11], x[12]);
      qr(x[2], x[7], x[8], x[13]); qr(x[3], x[4], x[9], x[14]);
    }
    for (int i = 0; i < 16; i++) block[i] = x[i] + s[i];
    counter++;
    pos = 0;
  }
  uint32_t u32() {
    if (pos == 16) refill();
    return block[pos++];
  }
  uint64_t u64() { return (uint64_t)u32() | ((uint64_t)u32() << 32); }
  uint32_t below(uint32_t bound) {  // uniform in [0, bound), rejection sampling
    const uint32_t lim = (uint32_t)(0x100000000ull - 0x100000000ull % bound);
    for (;;) {
      const uint32_t v = u32();
      if (lim == 0 || v < lim) return v % bound;
    }
  }
  double unit() { return (double)(u64() >> 11) * (1.0 / 9007199254740992.0); }
};

// ---- CPU negacyclic NTT (natural Spiral order) ------------------------------
struct Ntt {
  std::vector<uint32_t> tw[2], twp[2], itw[2], itwp[2];
  uint32_t ninv[2], ninvp[2];
  static uint32_t comp(uint32_t w, uint32_t q) { return (uint32_t)(((uint64_t)w << 32) / q); }
  static uint32_t mulsh(uint32_t y, uint32_t w, uint32_t wp, uint32_t q) {
    uint32_t r = y * w - (uint32_t)(((uint64_t)y * wp) >> 32) * q;
    return r >= q ? r - q : r;
  }
  Ntt() {
    for (int m = 0; m < 2; m++) {
      const uint32_t q = QM[m];
      const uint64_t pinv = powmod(PSI[m], q - 2, q);
      tw[m].resize(N); twp[m].resize(N); itw[m].resize(N); itwp[m].resize(N);
      for (uint32_t i = 0; i < (uint32_t)N; i++) {
        tw[m][i] = (uint32_t)powmod(PSI[m], brv11(i), q);
        itw[m][i] = (uint32_t)powmod(pinv, brv11(i), q);
        twp[m][i] = comp(tw[m][i], q);
        itwp[m][i] = comp(itw[m][i], q);
      }
      ninv[m] = (uint32_t)powmod(N, q - 2, q);
      ninvp[m] = comp(ninv[m], q);
    }
  }
  void fwd(uint32_t *a, int m) const {
    const uint32_t q = QM[m];
    int t = N;
    for (int mm = 1; mm < N; mm <<= 1) {
      t >>= 1;
      for (int i = 0; i < mm; i++) {
        const uint32_t w = tw[m][mm + i], wp = twp[m][mm + i];
        for (int j = 2 * i * t; j < 2 * i * t + t; j++) {
          const uint32_t u = a[j], v = mulsh(a[j + t], w, wp, q);
          const uint32_t s = u + v;
          a[j] = s >= q ? s - q : s;
          a[j + t] = u >= v ? u - v : u + q - v;
        }
      }
    }
  }
  void inv(uint32_t *a, int m) const {
    const uint32_t q = QM[m];
    int t = 1;
    for (int mm = N; mm > 1; mm >>= 1) {
      const int h = mm >> 1;
      int j1 = 0;
      for (int i = 0; i < h; i++) {
        const uint32_t w = itw[m][h + i], wp = itwp[m][h + i];
        for (int j = j1; j < j1 + t; j++) {
          const uint32_t u = a[j], v = a[j + t];
          const uint32_t s = u + v;
          a[j] = s >= q ? s - q : s;
          a[j + t] = mulsh(u >= v ? u - v : u + q - v, w, wp, q);
        }
        j1 += 2 * t;
      }
      t <<= 1;
    }
    for (int j = 0; j < N; j++) a[j] = mulsh(a[j], ninv[m], ninvp[m], q);
  }
};

const Ntt &ntt() {
  static const Ntt k;
  return k;
}

uint64_t crt(uint32_t a, uint32_t b) {
  static const uint64_t c0 = powmod(QM[1], QM[0] - 2, QM[0]);
  static const uint64_t c1 = powmod(QM[0], QM[1] - 2, QM[1]);
  const uint64_t x0 = a * c0 % QM[0], x1 = b * c1 % QM[1];
  const u128 s = (u128)x0 * QM[1] + (u128)x1 * QM[0];
  return (uint64_t)(s % QQ);
}

// Signed small coefficients -> NTT residues (both moduli), out[m][k].
void small_to_ntt(const int32_t *c, uint32_t *out) {
  for (int m = 0; m < 2; m++) {
    uint32_t *o = out + m * N;
    for (int i = 0; i < N; i++) o[i] = c[i] >= 0 ? (uint32_t)c[i] : QM[m] - (uint32_t)(-c[i]);
    ntt().fwd(o, m);
  }
}
// Raw values mod Q -> NTT residues.
void raw_to_ntt(const uint64_t *c, uint32_t *out) {
  for (int m = 0; m < 2; m++) {
    uint32_t *o = out + m * N;
    for (int i = 0; i < N; i++) o[i] = (uint32_t)(c[i] % QM[m]);
    ntt().fwd(o, m);
  }
}

}  // namespace

struct sp_client {
  int nu1, nu2, n2, t_gsw, t_exp, g, stop;
  ChaCha rng;
  std::vector<int32_t> s;        // ternary secret, raw
  std::vector<uint32_t> s_ntt;   // [m][N]
  std::vector<double> cdt;       // discrete Gaussian CDF over [-B, B]
  int gauss_bound = 0;
  // Matrix secret of the packed response (n ternary polynomials), drawn on
  // first use from a stream of its own so the other keys do not depend on it.
  uint64_t seed;
  int n_mat = 0;                   // n with n * n = n2, 0 if n2 is not a square
  std::vector<int32_t> s_mat;      // [n][N]
  std::vector<uint32_t> s_mat_ntt; // [n][m][N]
  explicit sp_client(uint64_t sd) : rng(sd), seed(sd) {}

  void make_matrix_secret() {
    if (!s_mat.empty()) return;
    ChaCha r2(seed ^ 0x6d61747269787365ull);
    s_mat.resize((size_t)n_mat * N);
    s_mat_ntt.resize((size_t)n_mat * 2 * N);
    for (auto &x : s_mat) x = (int32_t)r2.below(3) - 1;
    for (int i = 0; i < n_mat; i++) small_to_ntt(s_mat.data() + (size_t)i * N, s_mat_ntt.data() + (size_t)i * 2 * N);
  }

  int32_t gauss() {
    const double u = rng.unit();
    size_t lo = 0, hi = cdt.size() - 1;
    while (lo < hi) {
      const size_t mid = (lo + hi) / 2;
      if (cdt[mid] > u) hi = mid; else lo = mid + 1;
    }
    return (int32_t)lo - gauss_bound;
  }

  // ct (NTT form, [row][m][N]) = (a, a*s + e + mu_ntt); mu_ntt may be null.
  void encrypt(uint32_t *ct, const uint32_t *mu_ntt) {
    std::vector<int32_t> e(N);
    for (int i = 0; i < N; i++) e[i] = gauss();
    std::vector<uint32_t> e_ntt(2 * N);
    small_to_ntt(e.data(), e_ntt.data());
    for (int m = 0; m < 2; m++) {
      const uint32_t q = QM[m];
      uint32_t *a = ct + m * N, *b = ct + (2 + m) * N;
      for (int k = 0; k < N; k++) {
        a[k] = rng.below(q);
        uint64_t v = (uint64_t)a[k] * s_ntt[m * N + k] % q;
        v = (v + e_ntt[m * N + k]) % q;
        if (mu_ntt) v = (v + mu_ntt[m * N + k]) % q;
        b[k] = (uint32_t)v;
      }
    }
  }
};

namespace {
int gadget_bits(int k) { return (56 + k - 1) / k; }
int ceil_log2(uint64_t x) {
  int v = 0;
  uint64_t p = 1;
  while (p < x) { p <<= 1; v++; }
  return v;
}
}  // namespace

extern "C" const char *sp_client_last_error(void) { return g_cerr.c_str(); }

extern "C" int sp_client_create(int nu1, int nu2, int n2, int t_gsw, int t_exp, uint64_t seed, sp_client **out) {
  if (!out) return cfail(SP_ERR_ARG, "sp_client_create: out is NULL");
  *out = nullptr;
  if (nu1 < 1 || nu2 < 1 || n2 < 1 || t_gsw < 2 || t_gsw > 28 || t_exp < 2 || t_exp > 56)
    return cfail(SP_ERR_ARG, "sp_client_create: bad parameters");
  const int g = ceil_log2((uint64_t)t_gsw * nu2 + (1ull << nu1));
  const int stop = ceil_log2((uint64_t)t_gsw * nu2);
  if (g > LOGN) return cfail(SP_ERR_ARG, "sp_client_create: expansion depth exceeds log2 N");
  sp_client *c = new sp_client(seed);
  c->nu1 = nu1; c->nu2 = nu2; c->n2 = n2; c->t_gsw = t_gsw; c->t_exp = t_exp; c->g = g; c->stop = stop;
  for (int k = 1; k * k <= n2; k++)
    if (k * k == n2) c->n_mat = k;
  const double sigma = 6.4;
  c->gauss_bound = (int)std::ceil(12 * sigma);
  double total = 0;
  std::vector<double> w(2 * c->gauss_bound + 1);
  for (int x = -c->gauss_bound; x <= c->gauss_bound; x++) {
    w[x + c->gauss_bound] = std::exp(-(double)x * x / (2 * sigma * sigma));
    total += w[x + c->gauss_bound];
  }
  c->cdt.resize(w.size());
  double acc = 0;
  for (size_t i = 0; i < w.size(); i++) { acc += w[i] / total; c->cdt[i] = acc; }
  c->cdt.back() = 1.0 + 1e-12;
  c->s.resize(N);
  for (int i = 0; i < N; i++) c->s[i] = (int32_t)c->rng.below(3) - 1;
  c->s_ntt.resize(2 * N);
  small_to_ntt(c->s.data(), c->s_ntt.data());
  *out = c;
  return SP_OK;
}

extern "C" int sp_client_destroy(sp_client *c) {
  delete c;
  return SP_OK;
}

// Sizes in u32 words of the buffers the client fills / reads.
extern "C" int sp_client_sizes(const sp_client *c, size_t *query_words, size_t *left_words, size_t *right_words,
                               size_t *fold_words, size_t *response_words) {
  if (!c) return cfail(SP_ERR_ARG, "sp_client_sizes: NULL");
  const size_t kl = (size_t)2 * c->t_exp * 2 * N;
  if (query_words) *query_words = (size_t)4 * N;
  if (left_words) *left_words = (size_t)c->g * kl;
  if (right_words) *right_words = (size_t)(c->stop + 1) * kl;
  if (fold_words) *fold_words = (size_t)c->nu2 * 2 * (2 * c->t_gsw) * 2 * N;
  if (response_words) *response_words = (size_t)c->n2 * 4 * N;
  return SP_OK;
}

// Expansion keys (public parameters): left[g], right[stop+1], each
// CiphertextHalf[2][t_exp] = u32[crt][col][row][N], standard form.
extern "C" int sp_client_keygen(sp_client *c, uint32_t *host_left, uint32_t *host_right) {
  if (!c || !host_left || !host_right) return cfail(SP_ERR_ARG, "sp_client_keygen: NULL");
  const int bits = gadget_bits(c->t_exp);
  std::vector<uint32_t> ct(4 * N), tau_s(2 * N);
  std::vector<int32_t> ts(N);
  for (int side = 0; side < 2; side++) {
    uint32_t *dst = side ? host_right : host_left;
    const int levels = side ? c->stop + 1 : c->g;
    for (int r = 0; r < levels; r++) {
      const uint32_t t = (uint32_t)(N >> r) + 1;
      for (int i = 0; i < N; i++) {  // tau_t(s) on signed coefficients
        const uint64_t it = (uint64_t)i * t;
        ts[it % N] = ((it / N) & 1) ? -c->s[i] : c->s[i];
      }
      small_to_ntt(ts.data(), tau_s.data());
      for (int k = 0; k < c->t_exp; k++) {
        c->encrypt(ct.data(), nullptr);
        for (int m = 0; m < 2; m++) {
          const uint32_t q = QM[m];
          const uint64_t gk = powmod(2, (uint64_t)bits * k, q);
          uint32_t *b = ct.data() + (2 + m) * N;
          for (int z = 0; z < N; z++) b[z] = (uint32_t)((b[z] + q - (uint64_t)tau_s[m * N + z] * gk % q) % q);
          uint32_t *o = dst + (((size_t)r * 2 + m) * c->t_exp + k) * 2 * N;
          memcpy(o, ct.data() + m * N, N * 4);
          memcpy(o + N, ct.data() + (2 + m) * N, N * 4);
        }
      }
    }
  }
  return SP_OK;
}

// Query for record `index` (= i1 * 2^nu2 + i2): the packed Regev query
// (NTT Ciphertext) and the GSW folding keys CiphertextHalf[nu2][2][2*t_gsw].
extern "C" int sp_client_query(sp_client *c, uint64_t index, uint32_t *host_query, uint32_t *host_v_folding) {
  if (!c || !host_query || !host_v_folding) return cfail(SP_ERR_ARG, "sp_client_query: NULL");
  if (index >> (c->nu1 + c->nu2)) return cfail(SP_ERR_ARG, "sp_client_query: index out of range");
  const uint64_t i1 = index >> c->nu2, i2 = index & ((1ull << c->nu2) - 1);
  // mu: coefficient 2*i1 = floor(Q/p) * 2^-g mod Q.
  const uint64_t delta = QQ / 256;
  const uint64_t inv2 = (QQ + 1) / 2;
  const uint64_t inv2g = powmod(inv2, c->g, QQ);
  std::vector<uint64_t> mu(N, 0);
  mu[2 * i1] = (uint64_t)((u128)delta * inv2g % QQ);
  std::vector<uint32_t> mu_ntt(2 * N);
  raw_to_ntt(mu.data(), mu_ntt.data());
  c->encrypt(host_query, mu_ntt.data());
  // GSW folding keys.
  const int mx = 2 * c->t_gsw;
  const int bits = gadget_bits(c->t_gsw);
  std::vector<uint32_t> ct(4 * N);
  for (int l = 0; l < c->nu2; l++) {
    const uint32_t bit = (uint32_t)((i2 >> l) & 1);
    for (int col = 0; col < mx; col++) {
      c->encrypt(ct.data(), nullptr);
      for (int m = 0; m < 2; m++) {
        const uint32_t q = QM[m];
        const int row = col & 1;  // G[row][row + 2k] = 2^(bits*k)
        if (bit) {
          const uint32_t gk = (uint32_t)powmod(2, (uint64_t)bits * (col >> 1), q);
          uint32_t *p = ct.data() + (row * 2 + m) * N;
          for (int z = 0; z < N; z++) p[z] = (uint32_t)(((uint64_t)p[z] + gk) % q);
        }
        uint32_t *o = host_v_folding + (((size_t)l * 2 + m) * mx + col) * 2 * N;
        memcpy(o, ct.data() + m * N, N * 4);
        memcpy(o + N, ct.data() + (2 + m) * N, N * 4);
      }
    }
  }
  return SP_OK;
}

// Conversion key (Spiral's RegevToGSW key; the reference declares it as
// PublicParameters.conversion, gpu-pir/structures.cuh:173-177, 2 x 2 t_conv):
// CiphertextHalf[2][2 t_conv] = u32 [m][col][row][N], standard form.  With
// bits = ceil(56 / t_conv), column 2k encrypts s^2 * 2^(bits k) and column
// 2k + 1 encrypts -s * 2^(bits k), so that for a ciphertext (a, b) of mu,
// key * g^-1(a, b) (rdim 2, digit k of poly j at row j + 2k) encrypts
// s^2 a - s b = -s mu - s e: the row-0 GSW column of mu.
extern "C" int sp_client_conversion_key(sp_client *c, int t_conv, uint32_t *host_conv) {
  if (!c || !host_conv) return cfail(SP_ERR_ARG, "sp_client_conversion_key: NULL");
  if (t_conv < 2 || t_conv > 28) return cfail(SP_ERR_ARG, "sp_client_conversion_key: t_conv out of range");
  const int bits = gadget_bits(t_conv);
  const int mx = 2 * t_conv;
  std::vector<uint32_t> ct(4 * N), mu(2 * N);
  for (int col = 0; col < mx; col++) {
    const int k = col >> 1;
    for (int m = 0; m < 2; m++) {
      const uint64_t q = QM[m];
      const uint64_t gk = powmod(2, (uint64_t)bits * k, q);
      for (int z = 0; z < N; z++) {
        const uint64_t s = c->s_ntt[m * N + z];
        const uint64_t v = (col & 1) ? (q - s) % q : s * s % q;
        mu[m * N + z] = (uint32_t)(v * gk % q);
      }
    }
    c->encrypt(ct.data(), mu.data());
    for (int m = 0; m < 2; m++) {
      uint32_t *o = host_conv + ((size_t)m * mx + col) * 2 * N;
      memcpy(o, ct.data() + m * N, N * 4);
      memcpy(o + N, ct.data() + (2 + m) * N, N * 4);
    }
  }
  return SP_OK;
}

// Packed query for record `index` (= i1 * 2^nu2 + i2): one Regev ciphertext
// (NTT Ciphertext) carrying both the first-dimension selector and the bits of
// i2 for the server-side Regev -> GSW conversion.  Plaintext mu:
//   coefficient 2 i1               = floor(Q/p) * 2^-g                 mod Q
//   coefficient 2 (l t_gsw + j) + 1 = bit l of i2 * 2^(bits j) * 2^-(stop+1) mod Q
// Expansion doubles a coefficient in every round it passes: the even slots
// pass all g rounds, the odd (GSW) slots rounds 0 .. stop_round.
extern "C" int sp_client_query_packed(sp_client *c, uint64_t index, uint32_t *host_query) {
  if (!c || !host_query) return cfail(SP_ERR_ARG, "sp_client_query_packed: NULL");
  if (index >> (c->nu1 + c->nu2)) return cfail(SP_ERR_ARG, "sp_client_query_packed: index out of range");
  if ((uint64_t)c->t_gsw * c->nu2 > ((uint64_t)1 << c->stop))
    return cfail(SP_ERR_ARG, "sp_client_query_packed: GSW bits do not fit the odd slots");
  const uint64_t i1 = index >> c->nu2, i2 = index & ((1ull << c->nu2) - 1);
  const uint64_t inv2 = (QQ + 1) / 2;
  std::vector<uint64_t> mu(N, 0);
  mu[2 * i1] = (uint64_t)((u128)(QQ / 256) * powmod(inv2, c->g, QQ) % QQ);
  const int bits = gadget_bits(c->t_gsw);
  const uint64_t inv_odd = powmod(inv2, c->stop + 1, QQ);
  for (int l = 0; l < c->nu2; l++) {
    if (!((i2 >> l) & 1)) continue;
    for (int j = 0; j < c->t_gsw; j++) {
      const uint64_t gj = powmod(2, (uint64_t)bits * j, QQ);
      mu[2 * (l * c->t_gsw + j) + 1] = (uint64_t)((u128)gj * inv_odd % QQ);
    }
  }
  std::vector<uint32_t> mu_ntt(2 * N);
  raw_to_ntt(mu.data(), mu_ntt.data());
  c->encrypt(host_query, mu_ntt.data());
  return SP_OK;
}

// Packing keys of the packed response (Spiral's response packing; the
// reference has none): for r < n, W_r is an (n+1) x t_conv matrix in NTT form,
// u32 [r][m][k][row][N], standard form.  Column k: row 0 = a_k uniform,
// row 1 + i = s~_i a_k + e_ik - [i == r] s 2^(bits k), with s~ the matrix
// secret and bits = ceil(56 / t_conv).  Then for a scalar ciphertext (a, b)
// of mu under s, W_r g^-1(a) + b e_{1+r} decrypts under S = [-s~ | I] to
// mu on row r.
extern "C" int sp_client_packing_keys(sp_client *c, int t_conv, uint32_t *host_w) {
  if (!c || !host_w) return cfail(SP_ERR_ARG, "sp_client_packing_keys: NULL");
  if (!c->n_mat) return cfail(SP_ERR_ARG, "sp_client_packing_keys: n2 is not a square");
  if (t_conv < 2 || t_conv > 28) return cfail(SP_ERR_ARG, "sp_client_packing_keys: t_conv out of range");
  c->make_matrix_secret();
  const int n = c->n_mat, bits = gadget_bits(t_conv);
  std::vector<int32_t> e(N);
  std::vector<uint32_t> e_ntt(2 * N);
  for (int r = 0; r < n; r++)
    for (int k = 0; k < t_conv; k++) {
      std::vector<uint32_t> a(2 * N);
      for (int m = 0; m < 2; m++)
        for (int z = 0; z < N; z++) a[m * N + z] = c->rng.below(QM[m]);
      for (int row = 0; row <= n; row++) {
        if (row > 0) {
          for (int z = 0; z < N; z++) e[z] = c->gauss();
          small_to_ntt(e.data(), e_ntt.data());
        }
        for (int m = 0; m < 2; m++) {
          const uint64_t q = QM[m];
          uint32_t *o = host_w + ((((size_t)r * 2 + m) * t_conv + k) * (n + 1) + row) * N;
          if (row == 0) {
            memcpy(o, a.data() + m * N, N * 4);
            continue;
          }
          const int i = row - 1;
          const uint64_t gk = powmod(2, (uint64_t)bits * k, q);
          for (int z = 0; z < N; z++) {
            uint64_t v = (uint64_t)c->s_mat_ntt[((size_t)i * 2 + m) * N + z] * a[m * N + z] % q;
            v = (v + e_ntt[m * N + z]) % q;
            if (i == r) v = (v + q - (uint64_t)c->s_ntt[m * N + z] * gk % q) % q;
            o[z] = (uint32_t)v;
          }
        }
      }
    }
  return SP_OK;
}

// Decodes a packed, modulus-switched response (layout: the n row-0 polys at
// q2_bits per coefficient, then poly (i, c) of rows 1..n at index (i-1) n + c
// at 10 bits = log2(4 p); coefficient k at bits [k b, (k+1) b), LSB first)
// into the record's n2 * 2048 bytes (trial t = i n + c):
//   u = round(2^10 (s~_i * C0_c mod 2^q2_bits) / 2^q2_bits),
//   record[i n + c] = round(((C_{1+i, c} - u) mod 2^10) / 4) mod 256.
extern "C" int sp_client_decode_packed(sp_client *c, const uint8_t *host_response, size_t bytes, int q2_bits,
                                       uint8_t *host_record) {
  if (!c || !host_response || !host_record) return cfail(SP_ERR_ARG, "sp_client_decode_packed: NULL");
  if (!c->n_mat) return cfail(SP_ERR_ARG, "sp_client_decode_packed: n2 is not a square");
  if (q2_bits < 11 || q2_bits > 32) return cfail(SP_ERR_ARG, "sp_client_decode_packed: q2_bits out of range");
  const int n = c->n_mat, q1_bits = 10;
  const size_t want = (size_t)n * N * q2_bits / 8 + (size_t)n * n * N * q1_bits / 8;
  if (bytes != want) return cfail(SP_ERR_ARG, "sp_client_decode_packed: response size");
  c->make_matrix_secret();
  auto unpack = [&](const uint8_t *src, int b, std::vector<uint64_t> &out) {
    out.assign(N, 0);
    for (int k = 0; k < N; k++) {
      uint64_t v = 0;
      for (int j = 0; j < b; j++) {
        const size_t pos = (size_t)k * b + j;
        v |= (uint64_t)((src[pos >> 3] >> (pos & 7)) & 1) << j;
      }
      out[k] = v;
    }
  };
  const uint8_t *rows = host_response + (size_t)n * N * q2_bits / 8;
  const uint64_t mask2 = (q2_bits == 64) ? ~0ull : ((1ull << q2_bits) - 1);
  std::vector<uint64_t> c0, ci;
  std::vector<uint32_t> c0_ntt(2 * N), prod(2 * N);
  for (int col = 0; col < n; col++) {
    unpack(host_response + (size_t)col * N * q2_bits / 8, q2_bits, c0);
    raw_to_ntt(c0.data(), c0_ntt.data());
    for (int i = 0; i < n; i++) {
      // s~_i * C0 as an exact integer polynomial (|coefficients| < N 2^32 < Q/2), via the NTT mod Q
      for (int m = 0; m < 2; m++)
        for (int z = 0; z < N; z++)
          prod[m * N + z] = (uint32_t)((uint64_t)c->s_mat_ntt[((size_t)i * 2 + m) * N + z] * c0_ntt[m * N + z] % QM[m]);
      ntt().inv(prod.data(), 0);
      ntt().inv(prod.data() + N, 1);
      unpack(rows + (size_t)(i * n + col) * N * q1_bits / 8, q1_bits, ci);
      for (int z = 0; z < N; z++) {
        const uint64_t x = crt(prod[z], prod[N + z]);
        const uint64_t t = (x > QQ / 2 ? (uint64_t)(x - QQ) : x) & mask2;  // mod 2^q2_bits (two's complement)
        const uint64_t u = ((t + (1ull << (q2_bits - q1_bits - 1))) >> (q2_bits - q1_bits)) & 1023u;
        const uint64_t v = (ci[z] + 1024 - u) & 1023u;
        host_record[(size_t)(i * n + col) * N + z] = (uint8_t)(((v + 2) >> 2) & 255u);
      }
    }
  }
  return SP_OK;
}

// Phase b - a*s mod Q (u64 [N], coefficient order) of one NTT Ciphertext
// (natural layout [row][m][N], standard form): the plaintext plus the noise.
extern "C" int sp_client_phase(const sp_client *c, const uint32_t *host_ct, uint64_t *host_phase) {
  if (!c || !host_ct || !host_phase) return cfail(SP_ERR_ARG, "sp_client_phase: NULL");
  std::vector<uint32_t> d(2 * N);
  for (int m = 0; m < 2; m++) {
    const uint32_t q = QM[m];
    for (int k = 0; k < N; k++) {
      const uint32_t a = host_ct[m * N + k], b = host_ct[(2 + m) * N + k];
      if (a >= q || b >= q) return cfail(SP_ERR_ARG, "sp_client_phase: residue out of range");
      const uint64_t as = (uint64_t)a * c->s_ntt[m * N + k] % q;
      d[m * N + k] = (uint32_t)((b + q - as) % q);
    }
    ntt().inv(d.data() + m * N, m);
  }
  for (int z = 0; z < N; z++) host_phase[z] = crt(d[z], d[N + z]);
  return SP_OK;
}

// The ternary secret s (int32 [N], coefficient order), for noise measurements.
extern "C" int sp_client_secret(const sp_client *c, int32_t *host_s) {
  if (!c || !host_s) return cfail(SP_ERR_ARG, "sp_client_secret: NULL");
  memcpy(host_s, c->s.data(), N * sizeof(int32_t));
  return SP_OK;
}

// Decode n2 raw response ciphertexts into the record's n2 * 2048 bytes.
extern "C" int sp_client_decode(sp_client *c, const uint32_t *host_response, uint8_t *host_record) {
  if (!c || !host_response || !host_record) return cfail(SP_ERR_ARG, "sp_client_decode: NULL");
  std::vector<uint32_t> as(2 * N);
  for (int t = 0; t < c->n2; t++) {
    const uint32_t *ct = host_response + (size_t)t * 4 * N;
    std::vector<uint64_t> a(N), b(N);
    for (int z = 0; z < N; z++) {
      a[z] = ((uint64_t)ct[z] << 32) | ct[N + z];
      b[z] = ((uint64_t)ct[2 * N + z] << 32) | ct[3 * N + z];
      if (a[z] > QQ || b[z] > QQ) return cfail(SP_ERR_ARG, "sp_client_decode: coefficient out of range");
    }
    raw_to_ntt(a.data(), as.data());
    for (int m = 0; m < 2; m++) {
      uint32_t *x = as.data() + m * N;
      for (int k = 0; k < N; k++) x[k] = (uint32_t)((uint64_t)x[k] * c->s_ntt[m * N + k] % QM[m]);
      ntt().inv(x, m);
    }
    for (int z = 0; z < N; z++) {
      const uint64_t as_z = crt(as[z], as[N + z]);
      const uint64_t bz = b[z] % QQ;
      const uint64_t v = bz >= as_z ? bz - as_z : bz + QQ - as_z;
      const u128 num = (u128)v * 256 + QQ / 2;
      host_record[(size_t)t * N + z] = (uint8_t)((uint64_t)(num / QQ) & 255);
    }
  }
  return SP_OK;
}
