/*
 * spiral_oracle.h — CPU oracle for the Spiral PIR server answer path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (gpu-pir_amd/) never links, loads or calls it.
 *
 * It restates, in plain textbook C, what the reference computes on this path
 * (citations are to EricArcherman/GPU-PIR, paths relative to the repo root):
 *
 *   negacyclic NTT/INTT            gpu-pir/ntt_28bit.cuh:31-231, host.cu:64-124
 *   from_ntt (INTT + inverse CRT)  gpu-pir/gadget.cuh:8-24, arith.cuh:51-59
 *   to_ntt                         gpu-pir/classes.cuh:140-173
 *   automorphism                   gpu-pir/automorph.cuh:28-50, automorph_test.cu:14-22
 *   gadget inverse + ext. product  gpu-pir/gadget.cuh:37-101
 *   coefficient expansion          gpu-pir/expansion.cuh:83-166,171-219
 *   first-dimension DB multiply    gpu-pir/db_mul.cuh:26-70, db_mul_test.cu:240-282
 *   GSW fold                       gpu-pir/folding.cuh:9-53
 *
 * Byte layouts are the reference's structs (gpu-pir/structures.cuh:150-210),
 * all little-endian u32, natural Spiral NTT order, standard (non-Montgomery)
 * form:
 *   NTT Poly       u32[2][2048]   crt[m][z] = residue mod q_m (m=0: 268369921,
 *                                 m=1: 249561089), canonical in [0, q_m)
 *   raw Poly       u32[2][2048]   crt[0][z] = hi 32 bits, crt[1][z] = lo 32 bits
 *                                 of a coefficient x in [0, Q]  (host.cu:168-173)
 *   Ciphertext     Poly[2]        polys[0] = a, polys[1] = a*s+e
 *   CiphertextHalf u32[2][2048]   one residue of both rows
 *   key matrices   CiphertextHalf[2 crt][cols]  -> u32 [crt][col][row][z]
 *
 * Pinning (what makes each function trustworthy) is listed in tests/ and in
 * the header comment of spiral_oracle.c.
 */
#ifndef SPIRAL_ORACLE_H
#define SPIRAL_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_N 2048
#define OR_LOGN 11

uint64_t or_modulus(void);                 /* Q = q1*q2 */
uint32_t or_q(int m);                      /* q_m */
uint32_t or_root(int m);                   /* primitive 4096-th root psi_m */
uint32_t or_twiddle(int m, int inverse, int i); /* psi^brv11(i) (or psi^-brv11(i)) mod q_m */

/* One residue, in place.  Input canonical; output canonical. */
void or_ntt_fwd(uint32_t *a, int m);
void or_ntt_inv(uint32_t *a, int m);       /* includes the factor N^-1 */
/* O(N^2) closed form: out[k] = sum_i a[i] * psi^((2*brv11(k)+1)*i) mod q_m. */
void or_ntt_brute(const uint32_t *a, uint32_t *out, int m);

/* NTT Poly -> raw Poly and back.  raw values may be up to 2^64-1 on input
 * to or_to_ntt (each residue is reduced first). */
void or_from_ntt(const uint32_t *ntt, uint32_t *raw);
void or_to_ntt(const uint32_t *raw, uint32_t *ntt);

/* raw Poly -> raw Poly, x -> x^t (t odd).  Negated entries are Q - x, so a
 * zero that is negated becomes Q (gpu-pir/automorph.cuh:40). */
void or_automorph(const uint32_t *raw_in, uint32_t *raw_out, uint32_t t);

/* Gadget inverse: rdim raw Polys -> mx digit polys (u32[mx][2048]); digit
 * row j + k*rdim = (x_j >> k*bits) & mask, bits = ceil(56 / (mx/rdim)). */
int or_gadget_bits(int num_elems);
void or_gadget_inv(const uint32_t *raw_in, int rdim, int mx, uint32_t *digits);

/* acc (NTT Ciphertext, canonical) += M * NTT(g^-1(raw_in)), M a 2 x mx key
 * matrix laid out [crt][col][row][z]. */
void or_extprod_acc(const uint32_t *raw_in, int rdim, int mx,
                    const uint32_t *keys, uint32_t *acc);

/* First-dimension DB multiply of one sub-database (u128 accumulation then %,
 * gpu-pir/db_mul_test.cu:240-282).
 *   db        [2^nu2 cols][2^nu1 rows] NTT Poly
 *   first_dim [2^nu1] NTT Ciphertext
 *   out       [2^nu2] NTT Ciphertext                                      */
void or_db_mul(const uint32_t *db, const uint32_t *first_dim, uint32_t *out,
               int nu1, int nu2);
/* Same result, OpenMP over (column, coefficient); nthreads<=0: all cores. */
void or_db_mul_omp(const uint32_t *db, const uint32_t *first_dim, uint32_t *out,
                   int nu1, int nu2, int nthreads);

/* In-place fold of 2^nu2 raw Ciphertexts into accs[0].  v_folding and
 * v_folding_neg are CiphertextHalf[nu2][2 crt][2*t_gsw]. */
/* Threads used by or_expand and or_fold (default 1; OpenMP over the
 * independent elements of a round / pairs of a level): results do not depend
 * on it. */
void or_set_threads(int n);
void or_fold(uint32_t *accs, int nu2, int t_gsw,
             const uint32_t *v_folding, const uint32_t *v_folding_neg);

/* Spiral expansion depth parameters (gpu-pir/expansion.cuh:83-85). */
int or_expansion_g(int nu1, int nu2, int t_gsw);
int or_expansion_stop_round(int nu2, int t_gsw);

/* v_neg1[r] = NTT(-x^(N - 2^r)) as an NTT Poly. */
void or_v_neg1(int r, uint32_t *out);

/* Coefficient expansion.  v: NTT Ciphertext[2^g], v[0] = query on entry,
 * the rest ignored.  left: CiphertextHalf[g][2][t_exp], right:
 * CiphertextHalf[stop_round+1][2][t_exp].  If rounds < g, only the first
 * `rounds` rounds are run (rounds < 0: all g). */
void or_expand(uint32_t *v, int nu1, int nu2, int t_gsw, int t_exp,
               const uint32_t *left, const uint32_t *right, int rounds);

/* Derive v_folding_neg = G - v_folding (G[r][r+2k] = 2^(7k) as a constant
 * polynomial).  Layout CiphertextHalf[nu2][2][2*t_gsw]. */
void or_folding_neg(const uint32_t *v_folding, uint32_t *v_folding_neg,
                    int nu2, int t_gsw);

/* Regev -> GSW conversion (Spiral's RegevToGSW; the reference leaves it a
 * stub, gpu-pir/cpu_version/conversion.py:3).  For GSW slot k = l*t_gsw + j:
 * column 2j+1 of level l = v[2k+1], column 2j = conv * NTT(g^-1(v[2k+1]))
 * with rdim 2 and t_conv digits.  v: NTT Ciphertext[>= 2 nu2 t_gsw];
 * conv: CiphertextHalf[2 crt][2 t_conv]; v_folding: CiphertextHalf[nu2][2][2 t_gsw]. */
void or_regev_to_gsw(const uint32_t *v, int nu2, int t_gsw, int t_conv,
                     const uint32_t *conv, uint32_t *v_folding);

/* Response packing + modulus switching (Spiral's response path, not in the
 * reference): n^2 raw Ciphertexts (trial t = r n + c) and packing keys
 * W [n][2 crt][t_conv][n+1][N] -> compact bytes (or_packed_response_bytes):
 * column c = from_ntt(sum_r W_r NTT(g^-1(a_{rn+c}))) + (0, b_{c}, b_{n+c}, ...),
 * row 0 rescaled to 2^q2_bits, rows 1..n to 2^q1_bits, bit-packed LSB first. */
uint32_t or_switch_modulus(uint64_t x, int bits);
size_t or_packed_response_bytes(int n, int q2_bits, int q1_bits);
void or_pack_response(const uint32_t *raw_cts, int n, const uint32_t *w, int t_conv,
                      int q2_bits, int q1_bits, uint8_t *out);

/* Whole answer for one query against an n^2-trial database:
 *   db   [n2][2^nu2][2^nu1] NTT Poly;  query: NTT Ciphertext;
 *   out  [n2] raw Ciphertext.                                              */
void or_answer(const uint32_t *db, int n2, int nu1, int nu2, int t_gsw, int t_exp,
               const uint32_t *query, const uint32_t *left, const uint32_t *right,
               const uint32_t *v_folding, const uint32_t *v_folding_neg,
               uint32_t *out);

/* Database ingest: plaintext bytes -> NTT form.
 *   bytes [2^(nu1+nu2) records][n2 trials][2048]; record idx = i1*2^nu2 + i2
 *   db    [n2][2^nu2 cols = i2][2^nu1 rows = i1] NTT Poly                  */
void or_db_encode(const uint8_t *bytes, uint32_t *db, int n2, int nu1, int nu2);

#ifdef __cplusplus
}
#endif
#endif
