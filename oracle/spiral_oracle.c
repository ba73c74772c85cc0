/*
 * spiral_oracle.c — textbook CPU restatement of the Spiral PIR answer path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.
 *
 * Every function is written for clarity, not speed: plain loops, u64/u128
 * arithmetic and a full reduction after every operation, so that no lazy
 * reduction range can hide a mistake.
 *
 * How each function is pinned (tests/test_oracle_*.py):
 *   or_ntt_fwd     KAT raw_poly.dat -> ntt_poly.dat (gpu-pir/cpu_version/ntt.py:379-403),
 *                  O(N^2) closed form or_ntt_brute, negacyclic convolution
 *                  against a schoolbook product
 *   or_ntt_inv     round trip (gpu-pir/ntt_test.cu:84-99), KAT ntt_poly.dat -> raw_poly.dat
 *   or_from_ntt    KAT ntt_poly.dat -> raw_poly.dat; CRT identity checks
 *   or_automorph   brute-force loop of gpu-pir/automorph_test.cu:14-22, composition
 *                  law tau_s(tau_t(x)) = tau_st(x), Q convention for zeros
 *   or_gadget_inv  digits recombine to x; bit width ceil(56/k) (gpu-pir/gadget.cuh:37-39)
 *   or_fold        KAT client_fold_input.dat + client_folding_keys(_neg).dat ->
 *                  client_fold_output.dat (gpu-pir/fold_test.cu:30-94)
 *   or_expand      KAT round 0: client_expand_input.dat + client_expand_right_keys.dat ->
 *                  client_expand_output_1round.dat (gpu-pir/cpu_version/expansion.py:262-294);
 *                  rounds >= 1 pinned by decryption of the expanded selectors
 *                  under the client's keys (no reference artifact covers them)
 *   or_db_mul      brute-force re-derivation per coefficient with Python ints;
 *                  linearity; decryption of the selected record
 *   or_answer      end to end: the client decodes the stored record
 *   or_pack_response  no reference artifact (packing and modulus switching are
 *                  spiral-rs code outside the reference): pinned by decryption
 *                  of the packed, rescaled response under the client's matrix
 *                  secret, an exact brute-force re-derivation of the rounding,
 *                  and a planted-mistake check
 *   or_regev_to_gsw  no reference artifact (the reference's conversion is a
 *                  stub): pinned by decryption of every converted column under
 *                  the client's secret (column 2j+1 -> bit * 2^(7j), column 2j
 *                  -> -s * bit * 2^(7j)) and by end-to-end decoding of packed
 *                  queries; its from_ntt / external-product parts are the
 *                  KAT-pinned functions above
 */
#include "spiral_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef unsigned __int128 u128;

/* gpu-pir/params.cuh:25-30 (C++ names Q0/Q1; this file says q1/q2 as m = 0/1). */
static const uint32_t Q_[2] = {268369921u, 249561089u};
static const uint32_t PSI_[2] = {66687u, 158221u};

static uint32_t g_tw[2][2][OR_N]; /* [m][inverse][i] = psi^(+-brv11(i)) */
static uint32_t g_ninv[2];
static uint32_t g_crt_c[2];       /* c0 = q2^-1 mod q1, c1 = q1^-1 mod q2 */
static int g_init = 0;
static int g_threads = 1; /* threads of or_expand / or_fold (or_set_threads) */

static uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = (uint64_t)((u128)r * b % q);
        b = (uint64_t)((u128)b * b % q);
        e >>= 1;
    }
    return r;
}

static uint32_t brv11(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < OR_LOGN; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
}

/* Twiddle tables psi^brv11(i): gpu-pir/host.cu:64-71 (before the Montgomery
 * scaling of host.cu:73-77 and the register restride of host.cu:91-111, which
 * are GPU-schedule details and not part of the mathematics). */
static void init(void) {
    if (g_init) return;
    for (int m = 0; m < 2; m++) {
        uint64_t q = Q_[m];
        uint64_t psi = PSI_[m];
        uint64_t psi_inv = powmod(psi, q - 2, q);
        for (uint32_t i = 0; i < OR_N; i++) {
            g_tw[m][0][i] = (uint32_t)powmod(psi, brv11(i), q);
            g_tw[m][1][i] = (uint32_t)powmod(psi_inv, brv11(i), q);
        }
        g_ninv[m] = (uint32_t)powmod(OR_N, q - 2, q);
    }
    g_crt_c[0] = (uint32_t)powmod(Q_[1], Q_[0] - 2, Q_[0]);
    g_crt_c[1] = (uint32_t)powmod(Q_[0], Q_[1] - 2, Q_[1]);
    g_init = 1;
}

uint64_t or_modulus(void) { return (uint64_t)Q_[0] * Q_[1]; }
uint32_t or_q(int m) { return Q_[m]; }
uint32_t or_root(int m) { return PSI_[m]; }
uint32_t or_twiddle(int m, int inverse, int i) { init(); return g_tw[m][inverse ? 1 : 0][i]; }

/* Forward negacyclic NTT, Cooley-Tukey with the psi^brv table; stage with
 * `mm` groups uses tw[mm + i] (gpu-pir/ntt_28bit.cuh:44-57 is the same
 * schedule for the first three stages).  Output is in the natural Spiral NTT
 * order: a[k] = A(psi^(2*brv11(k)+1)). */
void or_ntt_fwd(uint32_t *a, int m) {
    init();
    const uint64_t q = Q_[m];
    int t = OR_N;
    for (int mm = 1; mm < OR_N; mm <<= 1) {
        t >>= 1;
        for (int i = 0; i < mm; i++) {
            const uint64_t s = g_tw[m][0][mm + i];
            const int j1 = 2 * i * t;
            for (int j = j1; j < j1 + t; j++) {
                uint64_t u = a[j];
                uint64_t v = (uint64_t)a[j + t] * s % q;
                a[j] = (uint32_t)((u + v) % q);
                a[j + t] = (uint32_t)((u + q - v) % q);
            }
        }
    }
}

/* Inverse negacyclic NTT, Gentleman-Sande with the psi^-brv table
 * (gpu-pir/ntt_28bit.cuh:145-231), followed by the factor N^-1 that the
 * reference folds into its CRT constants (gpu-pir/params.cuh:33-35). */
void or_ntt_inv(uint32_t *a, int m) {
    init();
    const uint64_t q = Q_[m];
    int t = 1;
    for (int mm = OR_N; mm > 1; mm >>= 1) {
        const int h = mm >> 1;
        int j1 = 0;
        for (int i = 0; i < h; i++) {
            const uint64_t s = g_tw[m][1][h + i];
            for (int j = j1; j < j1 + t; j++) {
                uint64_t u = a[j], v = a[j + t];
                a[j] = (uint32_t)((u + v) % q);
                a[j + t] = (uint32_t)((u + q - v) % q * s % q);
            }
            j1 += 2 * t;
        }
        t <<= 1;
    }
    for (int j = 0; j < OR_N; j++) a[j] = (uint32_t)((uint64_t)a[j] * g_ninv[m] % q);
}

void or_ntt_brute(const uint32_t *a, uint32_t *out, int m) {
    const uint64_t q = Q_[m];
    for (uint32_t k = 0; k < OR_N; k++) {
        uint64_t pt = powmod(PSI_[m], 2 * brv11(k) + 1, q);
        uint64_t acc = 0, p = 1;
        for (int i = 0; i < OR_N; i++) {
            acc = (acc + (uint64_t)a[i] % q * p) % q;
            p = p * pt % q;
        }
        out[k] = (uint32_t)acc;
    }
}

static uint64_t raw_get(const uint32_t *raw, int z) {
    return ((uint64_t)raw[z] << 32) | raw[OR_N + z];
}
static void raw_set(uint32_t *raw, int z, uint64_t x) {
    raw[z] = (uint32_t)(x >> 32);
    raw[OR_N + z] = (uint32_t)x;
}

/* Inverse CRT: the unique x in [0, Q) with x = a mod q1, x = b mod q2
 * (gpu-pir/arith.cuh:51-59 without the Montgomery/N^-1 scaling). */
static uint64_t crt(uint32_t a, uint32_t b) {
    const uint64_t Q = or_modulus();
    uint64_t x0 = (uint64_t)a * g_crt_c[0] % Q_[0];
    uint64_t x1 = (uint64_t)b * g_crt_c[1] % Q_[1];
    u128 s = (u128)x0 * Q_[1] + (u128)x1 * Q_[0];
    return (uint64_t)(s % Q);
}

void or_from_ntt(const uint32_t *ntt, uint32_t *raw) {
    init();
    uint32_t r0[OR_N], r1[OR_N];
    memcpy(r0, ntt, sizeof r0);
    memcpy(r1, ntt + OR_N, sizeof r1);
    or_ntt_inv(r0, 0);
    or_ntt_inv(r1, 1);
    for (int z = 0; z < OR_N; z++) raw_set(raw, z, crt(r0[z], r1[z]));
}

/* gpu-pir/classes.cuh:140-173: reduce the raw coefficient per modulus, NTT. */
void or_to_ntt(const uint32_t *raw, uint32_t *ntt) {
    init();
    for (int m = 0; m < 2; m++) {
        uint32_t *r = ntt + m * OR_N;
        for (int z = 0; z < OR_N; z++) r[z] = (uint32_t)(raw_get(raw, z) % Q_[m]);
        or_ntt_fwd(r, m);
    }
}

/* gpu-pir/automorph_test.cu:14-22 / automorph.cuh:28-50. */
void or_automorph(const uint32_t *raw_in, uint32_t *raw_out, uint32_t t) {
    const uint64_t Q = or_modulus();
    uint64_t tmp[OR_N];
    for (uint64_t i = 0; i < OR_N; i++) {
        uint64_t x = raw_get(raw_in, (int)i);
        uint64_t idx = (i * t) % OR_N;
        uint64_t flip = ((i * t) / OR_N) & 1;
        tmp[idx] = flip ? Q - x : x;
    }
    for (int z = 0; z < OR_N; z++) raw_set(raw_out, z, tmp[z]);
}

/* gpu-pir/gadget.cuh:37-39: bits_per = ceil(MOD_BITS / num_elems), MOD_BITS = 56. */
/* Threads used by or_expand and or_fold (default 1); results do not depend on it. */
void or_set_threads(int n) { g_threads = n > 0 ? n : 1; }

int or_gadget_bits(int num_elems) { return (56 + num_elems - 1) / num_elems; }

/* gpu-pir/gadget.cuh:55-74 / cpu_version/gadget.py:116-134: row j + k*rdim of
 * the output holds bits [k*bits, (k+1)*bits) of input poly j. */
void or_gadget_inv(const uint32_t *raw_in, int rdim, int mx, uint32_t *digits) {
    const int num_elems = mx / rdim;
    const int bits = or_gadget_bits(num_elems);
    const uint64_t mask = (1ull << bits) - 1;
    for (int j = 0; j < rdim; j++) {
        const uint32_t *p = raw_in + (size_t)j * 2 * OR_N;
        for (int k = 0; k < num_elems; k++) {
            const int shift = k * bits;
            uint32_t *d = digits + (size_t)(j + k * rdim) * OR_N;
            for (int z = 0; z < OR_N; z++) {
                uint64_t x = raw_get(p, z);
                d[z] = shift >= 64 ? 0u : (uint32_t)((x >> shift) & mask);
            }
        }
    }
}

/* acc += M * NTT(g^-1(in)): gpu-pir/gadget.cuh:55-101 (digit -> NTT -> two
 * multiply-accumulates per residue), result kept canonical. */
void or_extprod_acc(const uint32_t *raw_in, int rdim, int mx,
                    const uint32_t *keys, uint32_t *acc) {
    init();
    uint32_t *digits = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)mx * OR_N);
    uint32_t d[OR_N];
    or_gadget_inv(raw_in, rdim, mx, digits);
    for (int col = 0; col < mx; col++) {
        for (int m = 0; m < 2; m++) {
            const uint64_t q = Q_[m];
            for (int z = 0; z < OR_N; z++) d[z] = (uint32_t)(digits[(size_t)col * OR_N + z] % q);
            or_ntt_fwd(d, m);
            for (int row = 0; row < 2; row++) {
                const uint32_t *k = keys + (((size_t)m * mx + col) * 2 + row) * OR_N;
                uint32_t *o = acc + ((size_t)row * 2 + m) * OR_N;
                for (int z = 0; z < OR_N; z++)
                    o[z] = (uint32_t)(((uint64_t)o[z] + (uint64_t)d[z] * k[z]) % q);
            }
        }
    }
    free(digits);
}

/* gpu-pir/db_mul_test.cu:240-282: u128 accumulation over rows, one %. */
void or_db_mul(const uint32_t *db, const uint32_t *first_dim, uint32_t *out,
               int nu1, int nu2) {
    const size_t rows = (size_t)1 << nu1, cols = (size_t)1 << nu2;
    for (size_t c = 0; c < cols; c++)
        for (int z = 0; z < OR_N; z++)
            for (int m = 0; m < 2; m++) {
                u128 acc_a = 0, acc_e = 0;
                for (size_t j = 0; j < rows; j++) {
                    uint32_t va = first_dim[((j * 2 + 0) * 2 + m) * OR_N + z];
                    uint32_t ve = first_dim[((j * 2 + 1) * 2 + m) * OR_N + z];
                    uint32_t d = db[((c * rows + j) * 2 + m) * OR_N + z];
                    acc_a += (u128)va * d;
                    acc_e += (u128)ve * d;
                }
                out[((c * 2 + 0) * 2 + m) * OR_N + z] = (uint32_t)(acc_a % Q_[m]);
                out[((c * 2 + 1) * 2 + m) * OR_N + z] = (uint32_t)(acc_e % Q_[m]);
            }
}

void or_db_mul_omp(const uint32_t *db, const uint32_t *first_dim, uint32_t *out,
                   int nu1, int nu2, int nthreads) {
    const long rows = 1L << nu1, cols = 1L << nu2;
    const long items = cols * OR_N;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : omp_get_max_threads())
#endif
    for (long it = 0; it < items; it++) {
        const long c = it / OR_N;
        const int z = (int)(it % OR_N);
        for (int m = 0; m < 2; m++) {
            u128 acc_a = 0, acc_e = 0;
            for (long j = 0; j < rows; j++) {
                uint32_t va = first_dim[((j * 2 + 0) * 2 + m) * OR_N + z];
                uint32_t ve = first_dim[((j * 2 + 1) * 2 + m) * OR_N + z];
                uint32_t d = db[((c * rows + j) * 2 + m) * OR_N + z];
                acc_a += (u128)va * d;
                acc_e += (u128)ve * d;
            }
            out[((c * 2 + 0) * 2 + m) * OR_N + z] = (uint32_t)(acc_a % Q_[m]);
            out[((c * 2 + 1) * 2 + m) * OR_N + z] = (uint32_t)(acc_e % Q_[m]);
        }
    }
    (void)nthreads;
}

/* gpu-pir/folding.cuh:9-27,45-53 (cpu_version/folding.py:19-44):
 * acc[i] <- from_ntt(C * NTT(g^-1(acc[num_per+i])) + C_neg * NTT(g^-1(acc[i]))),
 * C = v_folding[nu2-1-cur_dim]. */
void or_fold(uint32_t *accs, int nu2, int t_gsw,
             const uint32_t *v_folding, const uint32_t *v_folding_neg) {
    init();
    const int mx = 2 * t_gsw;
    const size_t ct_words = 4 * OR_N;
    const size_t level_words = (size_t)2 * mx * 2 * OR_N;
    int num_per = 1 << nu2;
    for (int cur_dim = 0; cur_dim < nu2; cur_dim++) {
        num_per /= 2;
        const uint32_t *c_pos = v_folding + (size_t)(nu2 - 1 - cur_dim) * level_words;
        const uint32_t *c_neg = v_folding_neg + (size_t)(nu2 - 1 - cur_dim) * level_words;
        /* pairs of a level are independent: pair i reads elements i and
         * num_per + i and writes element i only */
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic) num_threads(g_threads)
#endif
        for (int i = 0; i < num_per; i++) {
            uint32_t sum[4 * OR_N];
            memset(sum, 0, sizeof sum);
            or_extprod_acc(accs + (size_t)(num_per + i) * ct_words, 2, mx, c_pos, sum);
            or_extprod_acc(accs + (size_t)i * ct_words, 2, mx, c_neg, sum);
            or_from_ntt(sum, accs + (size_t)i * ct_words);
            or_from_ntt(sum + 2 * OR_N, accs + (size_t)i * ct_words + 2 * OR_N);
        }
    }
}

static int ceil_log2(uint64_t x) {
    int v = 0;
    uint64_t p = 1;
    while (p < x) { p <<= 1; v++; }
    return v;
}

/* gpu-pir/expansion.cuh:83-85. */
int or_expansion_g(int nu1, int nu2, int t_gsw) {
    return ceil_log2((uint64_t)t_gsw * nu2 + (1ull << nu1));
}
int or_expansion_stop_round(int nu2, int t_gsw) { return ceil_log2((uint64_t)t_gsw * nu2); }

void or_v_neg1(int r, uint32_t *out) {
    memset(out, 0, sizeof(uint32_t) * 2 * OR_N);
    for (int m = 0; m < 2; m++) {
        out[m * OR_N + (OR_N - (1 << r))] = Q_[m] - 1;
        or_ntt_fwd(out + m * OR_N, m);
    }
}

/* gpu-pir/expansion.cuh:87-166 (action_expand) and :171-219: for round r,
 * first v[2^r + i] = v_neg1[r] (.) v[i] for every i < 2^r, then every
 * element i < 2^(r+1) not skipped by the stop-round rule becomes
 *   v[i] + W * NTT(g^-1(tau_t(from_ntt(v[i]).a))) + (0, NTT(tau_t(from_ntt(v[i]).as_e)))
 * with t = N/2^r + 1 and W = left[r] if r != 0 and i is even, else right[r].
 * i is the element's global index (expansion.cuh:88). */
void or_expand(uint32_t *v, int nu1, int nu2, int t_gsw, int t_exp,
               const uint32_t *left, const uint32_t *right, int rounds) {
    init();
    const int g = or_expansion_g(nu1, nu2, t_gsw);
    const int stop = or_expansion_stop_round(nu2, t_gsw);
    const int max_bits = t_gsw * nu2;
    const size_t ct_words = 4 * OR_N;
    const size_t key_words = (size_t)2 * t_exp * 2 * OR_N;
    if (rounds < 0 || rounds > g) rounds = g;
    uint32_t neg1[2 * OR_N];
    for (int r = 0; r < rounds; r++) {
        const int num_in = 1 << r;
        const uint32_t t = (uint32_t)(OR_N / num_in + 1);
        or_v_neg1(r, neg1);
        /* the elements of a round are independent: each step reads and writes
         * only its own element (and, above, its own child) */
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(g_threads)
#endif
        for (int i = 0; i < num_in; i++) {
            const uint32_t *src = v + (size_t)i * ct_words;
            uint32_t *dst = v + (size_t)(num_in + i) * ct_words;
            for (int row = 0; row < 2; row++)
                for (int m = 0; m < 2; m++)
                    for (int z = 0; z < OR_N; z++) {
                        size_t o = ((size_t)row * 2 + m) * OR_N + z;
                        dst[o] = (uint32_t)((uint64_t)neg1[m * OR_N + z] * src[o] % Q_[m]);
                    }
        }
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic) num_threads(g_threads)
#endif
        for (int i = 0; i < 2 * num_in; i++) {
            uint32_t raw[4 * OR_N], au[4 * OR_N], acc[4 * OR_N], aue[2 * OR_N];
            if (stop > 0 && r > stop && (i % 2) == 1) continue;
            if (stop > 0 && r == stop && (i % 2) == 1 && (i / 2) >= max_bits) continue;
            const uint32_t *w = (r != 0 && (i % 2) == 0) ? left + (size_t)r * key_words
                                                         : right + (size_t)r * key_words;
            uint32_t *vi = v + (size_t)i * ct_words;
            or_from_ntt(vi, raw);
            or_from_ntt(vi + 2 * OR_N, raw + 2 * OR_N);
            or_automorph(raw, au, t);
            or_automorph(raw + 2 * OR_N, au + 2 * OR_N, t);
            memset(acc, 0, sizeof acc);
            or_extprod_acc(au, 1, t_exp, w, acc);
            or_to_ntt(au + 2 * OR_N, aue);
            for (int m = 0; m < 2; m++)
                for (int z = 0; z < OR_N; z++) {
                    const uint64_t q = Q_[m];
                    size_t oa = (size_t)m * OR_N + z, oe = (size_t)(2 + m) * OR_N + z;
                    vi[oa] = (uint32_t)(((uint64_t)vi[oa] + acc[oa]) % q);
                    vi[oe] = (uint32_t)(((uint64_t)vi[oe] + acc[oe] + aue[oa]) % q);
                }
        }
    }
}

/* Gadget matrix G (2 x 2t): G[r][r + 2k] = 2^(7k) (bits per digit for t
 * digits), constant polynomial, so its NTT is the same constant at every
 * evaluation point.  v_folding_neg = G - v_folding, as the reference fixtures
 * client_folding_keys_neg.dat = G - client_folding_keys.dat show. */
void or_folding_neg(const uint32_t *v_folding, uint32_t *v_folding_neg,
                    int nu2, int t_gsw) {
    const int mx = 2 * t_gsw;
    const int bits = or_gadget_bits(t_gsw);
    for (int l = 0; l < nu2; l++)
        for (int m = 0; m < 2; m++)
            for (int col = 0; col < mx; col++)
                for (int row = 0; row < 2; row++) {
                    size_t base = ((((size_t)l * 2 + m) * mx + col) * 2 + row) * OR_N;
                    uint64_t gval = 0;
                    if ((col % 2) == row) gval = powmod(2, (uint64_t)bits * (col / 2), Q_[m]);
                    for (int z = 0; z < OR_N; z++)
                        v_folding_neg[base + z] =
                            (uint32_t)((gval + Q_[m] - v_folding[base + z]) % Q_[m]);
                }
}

/* Regev -> GSW conversion.  The reference declares its pieces but leaves the
 * function a stub (gpu-pir/cpu_version/conversion.py:3): the conversion key
 * PublicParameters.conversion, 2 x 2*T_CONV (gpu-pir/structures.cuh:173-177,
 * cpu_version/structures.py:44-45), T_CONV = 4 (gpu-pir/params.cuh:65), and the
 * external-product shape "conversion: 2 by 1 -> 2t_conv by 1" with output
 * "2 by 1" (gpu-pir/gadget.cuh:41-53).  The steps are Spiral's RegevToGSW
 * (Menon-Wu 2022, section 3.3), as in spiral-rs's regev_to_gsw:
 *   GSW slot k = l*t_gsw + j (level l, gadget digit j) takes the expanded
 *   Regev ciphertext ct = v[2k + 1];
 *   column 2j+1 of level l = ct;
 *   column 2j   of level l = conv * NTT(g^-1(from_ntt(ct))), rdim 2 and
 *                            t_conv digits of ceil(56 / t_conv) bits.
 * v: NTT Ciphertext[>= 2 nu2 t_gsw]; conv: CiphertextHalf[2][2 t_conv];
 * v_folding: CiphertextHalf[nu2][2][2 t_gsw]. */
void or_regev_to_gsw(const uint32_t *v, int nu2, int t_gsw, int t_conv,
                     const uint32_t *conv, uint32_t *v_folding) {
    init();
    const int mx = 2 * t_gsw;
    const size_t ct_words = 4 * OR_N;
    uint32_t *raw = (uint32_t *)malloc(ct_words * sizeof(uint32_t));
    uint32_t *acc = (uint32_t *)malloc(ct_words * sizeof(uint32_t));
    for (int l = 0; l < nu2; l++)
        for (int j = 0; j < t_gsw; j++) {
            const uint32_t *ct = v + (size_t)(2 * (l * t_gsw + j) + 1) * ct_words;
            or_from_ntt(ct, raw);
            or_from_ntt(ct + 2 * OR_N, raw + 2 * OR_N);
            memset(acc, 0, ct_words * sizeof(uint32_t));
            or_extprod_acc(raw, 2, 2 * t_conv, conv, acc);
            for (int m = 0; m < 2; m++)
                for (int row = 0; row < 2; row++) {
                    uint32_t *lv = v_folding + (((size_t)l * 2 + m) * mx) * 2 * OR_N;
                    memcpy(lv + ((size_t)(2 * j) * 2 + row) * OR_N, acc + ((size_t)row * 2 + m) * OR_N,
                           OR_N * sizeof(uint32_t));
                    memcpy(lv + ((size_t)(2 * j + 1) * 2 + row) * OR_N, ct + ((size_t)row * 2 + m) * OR_N,
                           OR_N * sizeof(uint32_t));
                }
        }
    free(raw);
    free(acc);
}

/* x -> floor((x * 2^bits + floor(Q/2)) / Q) mod 2^bits: x * 2^bits / Q rounded
 * to the nearest integer (ties up), reduced mod 2^bits. */
uint32_t or_switch_modulus(uint64_t x, int bits) {
    const u128 Qm = (u128)Q_[0] * Q_[1];
    const u128 num = ((u128)x << bits) + Qm / 2;
    const uint64_t y = (uint64_t)(num / Qm);
    return (uint32_t)(y & ((1ull << bits) - 1));
}

size_t or_packed_response_bytes(int n, int q2_bits, int q1_bits) {
    return (size_t)n * OR_N * q2_bits / 8 + (size_t)n * n * OR_N * q1_bits / 8;
}

static void put_bits(uint8_t *dst, const uint32_t *vals, int bits) {
    memset(dst, 0, (size_t)OR_N * bits / 8);
    for (int k = 0; k < OR_N; k++)
        for (int b = 0; b < bits; b++)
            if ((vals[k] >> b) & 1u) {
                const size_t pos = (size_t)k * bits + b;
                dst[pos >> 3] |= (uint8_t)(1u << (pos & 7));
            }
}

/* Response packing + modulus switching (Spiral's response path; not in the
 * reference).  n^2 scalar Regev ciphertexts (trial t = r n + c, raw) become
 * one (n+1) x n matrix ciphertext whose column c is
 *     sum_r W_r * NTT(g^-1(a_{rn+c}))   (t_conv digits of ceil(56/t_conv) bits)
 * taken out of the NTT domain, plus b_{rn+c} on row 1 + r; W_r is the
 * (n+1) x t_conv packing key [2 crt][t_conv][n+1][N] (NTT form, standard).
 * Row 0 is rescaled to 2^q2_bits, rows 1..n to 2^q1_bits (or_switch_modulus)
 * and bit-packed: row-0 polys c = 0..n-1, then poly (i, c) of rows 1..n at
 * index (i-1) n + c; coefficient k at bits [k bits, (k+1) bits), LSB first. */
void or_pack_response(const uint32_t *raw_cts, int n, const uint32_t *w, int t_conv,
                      int q2_bits, int q1_bits, uint8_t *out) {
    init();
    const size_t ct_words = 4 * OR_N;
    uint32_t *digits = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)t_conv * OR_N);
    uint32_t *col = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n + 1) * 2 * OR_N); /* [row][m][N] */
    uint32_t raw[2 * OR_N], d[OR_N], vals[OR_N];
    uint8_t *row1 = out + (size_t)n * OR_N * q2_bits / 8;
    for (int c = 0; c < n; c++) {
        memset(col, 0, sizeof(uint32_t) * (size_t)(n + 1) * 2 * OR_N);
        for (int r = 0; r < n; r++) {
            const uint32_t *ct = raw_cts + (size_t)(r * n + c) * ct_words;
            const uint32_t *wr = w + (size_t)r * 2 * t_conv * (n + 1) * OR_N;
            or_gadget_inv(ct, 1, t_conv, digits);
            for (int k = 0; k < t_conv; k++)
                for (int m = 0; m < 2; m++) {
                    const uint64_t q = Q_[m];
                    for (int z = 0; z < OR_N; z++) d[z] = (uint32_t)(digits[(size_t)k * OR_N + z] % q);
                    or_ntt_fwd(d, m);
                    for (int row = 0; row <= n; row++) {
                        const uint32_t *kk = wr + (((size_t)m * t_conv + k) * (n + 1) + row) * OR_N;
                        uint32_t *o = col + ((size_t)row * 2 + m) * OR_N;
                        for (int z = 0; z < OR_N; z++)
                            o[z] = (uint32_t)(((uint64_t)o[z] + (uint64_t)d[z] * kk[z]) % q);
                    }
                }
        }
        for (int row = 0; row <= n; row++) {
            or_from_ntt(col + (size_t)row * 2 * OR_N, raw);
            for (int z = 0; z < OR_N; z++) {
                uint64_t x = raw_get(raw, z);
                if (row > 0) {
                    const uint32_t *b = raw_cts + (size_t)((row - 1) * n + c) * ct_words + 2 * OR_N;
                    const u128 sum = (u128)x + raw_get(b, z);
                    x = (uint64_t)(sum % ((u128)Q_[0] * Q_[1]));
                }
                vals[z] = or_switch_modulus(x, row ? q1_bits : q2_bits);
            }
            if (row == 0)
                put_bits(out + (size_t)c * OR_N * q2_bits / 8, vals, q2_bits);
            else
                put_bits(row1 + (size_t)((row - 1) * n + c) * OR_N * q1_bits / 8, vals, q1_bits);
        }
    }
    free(digits);
    free(col);
}

void or_answer(const uint32_t *db, int n2, int nu1, int nu2, int t_gsw, int t_exp,
               const uint32_t *query, const uint32_t *left, const uint32_t *right,
               const uint32_t *v_folding, const uint32_t *v_folding_neg,
               uint32_t *out) {
    init();
    const int g = or_expansion_g(nu1, nu2, t_gsw);
    const size_t ct_words = 4 * OR_N;
    const size_t rows = (size_t)1 << nu1, cols = (size_t)1 << nu2;
    uint32_t *v = (uint32_t *)calloc(((size_t)1 << g) * ct_words, sizeof(uint32_t));
    uint32_t *sel = (uint32_t *)malloc(rows * ct_words * sizeof(uint32_t));
    uint32_t *acc = (uint32_t *)malloc(cols * ct_words * sizeof(uint32_t));
    uint32_t *raw = (uint32_t *)malloc(cols * ct_words * sizeof(uint32_t));
    memcpy(v, query, ct_words * sizeof(uint32_t));
    or_expand(v, nu1, nu2, t_gsw, t_exp, left, right, -1);
    for (size_t j = 0; j < rows; j++)
        memcpy(sel + j * ct_words, v + 2 * j * ct_words, ct_words * sizeof(uint32_t));
    for (int t = 0; t < n2; t++) {
        /* same result as or_db_mul (tests check it), spread over the host cores */
        or_db_mul_omp(db + (size_t)t * cols * rows * 2 * OR_N, sel, acc, nu1, nu2, 0);
        for (size_t c = 0; c < cols; c++) {
            or_from_ntt(acc + c * ct_words, raw + c * ct_words);
            or_from_ntt(acc + c * ct_words + 2 * OR_N, raw + c * ct_words + 2 * OR_N);
        }
        or_fold(raw, nu2, t_gsw, v_folding, v_folding_neg);
        memcpy(out + (size_t)t * ct_words, raw, ct_words * sizeof(uint32_t));
    }
    free(v); free(sel); free(acc); free(raw);
}

void or_db_encode(const uint8_t *bytes, uint32_t *db, int n2, int nu1, int nu2) {
    init();
    const size_t rows = (size_t)1 << nu1, cols = (size_t)1 << nu2;
    const long nrec = (long)(rows * cols);
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long idx_l = 0; idx_l < nrec; idx_l++) {
        const size_t idx = (size_t)idx_l, i1 = idx / cols, i2 = idx % cols;
        for (int t = 0; t < n2; t++) {
            const uint8_t *p = bytes + (idx * n2 + t) * OR_N;
            uint32_t *o = db + (((size_t)t * cols + i2) * rows + i1) * 2 * OR_N;
            for (int m = 0; m < 2; m++) {
                for (int z = 0; z < OR_N; z++) o[m * OR_N + z] = p[z];
                or_ntt_fwd(o + m * OR_N, m);
            }
        }
    }
}
