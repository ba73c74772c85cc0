/*
 * spiral_pir.h — C ABI of libspiral_hip.so, the MI355X (gfx950) Spiral PIR
 * server answer path: coefficient expansion -> first-dimension database
 * multiply -> GSW fold, on the negacyclic 2048-point NTT over the CRT primes
 * q1 = 268369921 and q2 = 249561089 (Q = q1*q2), gadget base 2^7 x 8 digits.
 *
 * The reference (EricArcherman/GPU-PIR) has no FFI layer: its operator
 * surface is a set of CUDA kernels, the host helpers of gpu-pir/host.h and
 * the struct layouts of gpu-pir/structures.cuh.  Each entry point below names
 * the reference symbol it replaces.
 *
 * Conventions
 *  - Every function returns an int status: SP_OK (0) or a negative SP_ERR_*.
 *    Nothing calls exit(); sp_last_error() returns a message for the calling
 *    thread's most recent failure.
 *  - "dev" pointers are device memory (hipMalloc / torch) on the context's
 *    device; "host" pointers are host memory.  Device-pointer functions are
 *    asynchronous on `stream`, a hipStream_t taken as is (NULL is HIP's null
 *    stream, which is also torch's default stream).  Host-pointer functions
 *    work on the context's own stream and synchronize before returning.
 *  - Layouts (little-endian u32, structures.cuh:150-210):
 *      NTT Poly        u32[2][2048]  crt[m][k] = residue mod q_m at NTT position
 *                                    k (natural Spiral order), canonical [0,q_m)
 *      raw Poly        u32[2][2048]  [0][i] = high 32 bits, [1][i] = low 32 bits
 *                                    of coefficient i, a value in [0, Q]
 *      Ciphertext      Poly[2]       row 0 = a, row 1 = a*s + e
 *      CiphertextHalf  u32[2][2048]  [row][k], one residue
 *      key matrix      CiphertextHalf[2 crt][cols]  = u32[crt][col][row][k]
 *  - A context is single-threaded: the caller serializes calls on one ctx.
 */
#ifndef SPIRAL_PIR_H
#define SPIRAL_PIR_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SP_OK 0
#define SP_ERR_ARG (-1)     /* invalid argument (shape, null, range)      */
#define SP_ERR_HIP (-2)     /* HIP runtime error                          */
#define SP_ERR_NOMEM (-3)   /* device or host allocation failed           */
#define SP_ERR_STATE (-4)   /* object used in the wrong state             */

#define SP_POLY_LEN 2048
#define SP_T_CONV 4        /* conversion gadget digits (gpu-pir/params.cuh:65) */
#define SP_PACK_N 2        /* plaintext matrix dimension of response packing (n^2 = 4 sub-databases) */
#define SP_PACK_Q1_BITS 10 /* rows 1..n of a packed response are rescaled to 4 p = 2^10 */
#define SP_Q1 268369921u
#define SP_Q2 249561089u

typedef struct sp_ctx sp_ctx;
typedef struct sp_db sp_db;
typedef struct sp_client sp_client;

/* ---- context -------------------------------------------------------------
 * Replaces gen_twiddles_gpu (gpu-pir/host.cu:126-138, host.h:10): the
 * context owns the twiddle tables, automorphism permutations and constant
 * polynomials on its device. */
int sp_ctx_create(int device, sp_ctx **out);
int sp_ctx_destroy(sp_ctx *ctx);
int sp_ctx_synchronize(sp_ctx *ctx);
/* Kernel variant per stage, for tests and the autotuner; 0 = automatic.
 *   SP_STAGE_EXPAND, SP_STAGE_FOLD: 1 = one block of 4 waves per item (any
 *     gadget); for gadget 2^7 x 8 only: 2 = 8-wave blocks (two per CU),
 *     3 = 16-wave blocks (one per CU, half the passes), 4 = split in two
 *     launches (expansion: digits, then per-residue transforms and update;
 *     fold: partial products per residue and key-column half, then finish).
 *     Automatic: 4 for both, the expansion with its leading rounds in one
 *     launch each for runs of at most 2 queries.  SP_STAGE_EXPAND only:
 *     10 + R = the first R rounds in one launch each with 4-wave blocks per
 *     half of the NTT positions (rounds with 2048 >> r >= 4), then 4
 *   SP_STAGE_DB_MUL: 1 = VALU u64 multiply-accumulate (one query per database
 *     read), 2 = int8-limb MFMA, 3 / 4 = VALU with 2 / 4 queries per database
 *     read (runs of 2+ queries); a forced variant that does not apply to a run
 *     falls back to the automatic choice */
#define SP_STAGE_EXPAND 0
#define SP_STAGE_FOLD 1
#define SP_STAGE_DB_MUL 2
int sp_ctx_set_variant(sp_ctx *ctx, int stage, int variant);
const char *sp_last_error(void);
/* Library build string (kernel variant names); never NULL. */
const char *sp_version(void);

/* ---- transforms ----------------------------------------------------------
 * sp_ntt_fwd replaces NTT_R (gpu-pir/ntt_28bit.cuh:251): in place, n_polys
 * contiguous u32[2048] polynomials of one residue (0 -> q1, 1 -> q2),
 * coefficients canonical in, canonical NTT values out.
 * sp_ntt_inv replaces INTT_R (gpu-pir/ntt_28bit.cuh:292).  With scale = 0 it
 * keeps the reference's convention of leaving out N^-1 (the reference folds
 * it into inv_crt); scale = 1 gives the true inverse.  Outputs canonical. */
int sp_ntt_fwd(sp_ctx *ctx, uint32_t *dev_polys, size_t n_polys, int residue, void *stream);
int sp_ntt_inv(sp_ctx *ctx, uint32_t *dev_polys, size_t n_polys, int residue, int scale,
               void *stream);

/* NTT Poly -> raw Poly: from_ntt_regs (gpu-pir/gadget.cuh:8-24) / to_raw
 * (gpu-pir/classes.cuh:69-99).  n_polys NTT Polys in, n_polys raw Polys out. */
int sp_from_ntt(sp_ctx *ctx, const uint32_t *dev_ntt, uint32_t *dev_raw, size_t n_polys,
                void *stream);
/* raw Poly -> NTT Poly: to_ntt (gpu-pir/classes.cuh:140-173). */
int sp_to_ntt(sp_ctx *ctx, const uint32_t *dev_raw, uint32_t *dev_ntt, size_t n_polys,
              void *stream);

/* x -> x^t on u64 coefficients in [0, Q] (t odd): the `automorph` kernel
 * (gpu-pir/automorph.cuh:54-62), n_polys polynomials of u64[2048]; negated
 * entries are Q - x, so a negated zero becomes Q as in the reference. */
int sp_automorph(sp_ctx *ctx, const uint64_t *dev_in, uint64_t *dev_out, size_t n_polys,
                 uint32_t t, void *stream);

/* ---- gadget external product --------------------------------------------
 * inv_gadget_ntt_mult_rdim1/_rdim2 (gpu-pir/gadget.cuh:84-101):
 *   acc (NTT Ciphertext, canonical) += M * NTT(g^-1(in))
 * in: rdim (1 or 2) raw Polys; M: key matrix with mx columns, standard form,
 * layout u32[crt][mx][row][k]; g^-1 takes mx/rdim digits of
 * ceil(56/(mx/rdim)) bits.  n_items independent products, each operand
 * advanced by its stride (in u32 words; key stride 0 shares one matrix). */
int sp_gadget_extprod(sp_ctx *ctx, const uint32_t *dev_in, int rdim, int mx,
                      const uint32_t *dev_keys, uint32_t *dev_acc, size_t n_items,
                      size_t in_stride, size_t key_stride, size_t acc_stride, void *stream);

/* ---- first-dimension database multiply ----------------------------------
 * db_mul_kernel (gpu-pir/db_mul.cuh:26-70), reference layouts:
 *   db         Database: col[2^nu2].row[2^nu1] NTT Poly
 *   first_dim  FirstDim: selectors[2^nu1] NTT Ciphertext
 *   out        FoldingInputs: accumulators[2^nu2] NTT Ciphertext
 * out[c] = sum_j first_dim[j] * db[c][j] mod q, both ciphertext rows.
 * Converts the layouts on every call; the answer path uses a resident
 * sp_db instead. */
int sp_db_mul(sp_ctx *ctx, const uint32_t *dev_db, const uint32_t *dev_first_dim,
              uint32_t *dev_out, int nu1, int nu2, void *stream);

/* ---- fold ------------------------------------------------------------------
 * FOLD_CIPHERTEXTS / fold_all_ciphertexts (gpu-pir/folding.cuh:45-65):
 * in place on accs (2^nu2 raw Ciphertexts); the result is accs[0].
 * v_folding / v_folding_neg: CiphertextHalf[nu2][2 crt][2*t_gsw] (standard
 * form).  v_folding_neg may be NULL: it is then taken as G - v_folding, the
 * form a Spiral server derives, which allows the folding product to be
 * evaluated with half the NTTs. */
int sp_fold(sp_ctx *ctx, uint32_t *dev_accs, int nu2, int t_gsw, const uint32_t *dev_v_folding,
            const uint32_t *dev_v_folding_neg, void *stream);
/* sp_fold without any allocation: no reference counterpart; for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same arguments and result as
 * sp_fold; the converted keys and partial sums live in one static 8 MiB
 * device array per GPU, held until the process ends, so nu2 and t_gsw must
 * fit it (fast form, t_gsw = 8: nu2 <= 7; given v_folding_neg: up to 7 levels
 * of 16 key columns) or the call fails with SP_ERR_ARG.  Asynchronous on
 * `stream` (no host synchronization).  Calls that overlap on one GPU share
 * that array and race: serialize them (one stream, or events). */
int sp_fold_noalloc(sp_ctx *ctx, uint32_t *dev_accs, int nu2, int t_gsw, const uint32_t *dev_v_folding,
                    const uint32_t *dev_v_folding_neg, void *stream);

/* ---- coefficient expansion -----------------------------------------------
 * coefficient_expansion_r over all rounds (gpu-pir/expansion.cuh:222-260,
 * spiral-rs semantics described at expansion.cuh:16-28).
 * dev_v: NTT Ciphertext[2^g], v[0] = query on entry; g and stop_round as in
 * expansion.cuh:83-85.  left: CiphertextHalf[g][2][t_exp],
 * right: CiphertextHalf[stop_round+1][2][t_exp].  rounds < 0 runs all g. */
int sp_expansion_depth(int nu1, int nu2, int t_gsw, int *g, int *stop_round);
int sp_expand(sp_ctx *ctx, uint32_t *dev_v, int nu1, int nu2, int t_gsw, int t_exp,
              const uint32_t *dev_left, const uint32_t *dev_right, int rounds, void *stream);

/* ---- Regev -> GSW conversion ------------------------------------------------
 * The reference declares the conversion key (PublicParameters.conversion,
 * gpu-pir/structures.cuh:173-177; cpu_version/structures.py:44-45), its
 * external-product shape "conversion: 2 by 2t_conv" x "2 by 1 -> 2t_conv by
 * 1" (gpu-pir/gadget.cuh:41-53) and T_CONV = 4 (params.cuh:65), and leaves
 * regev_to_gsw a stub (gpu-pir/cpu_version/conversion.py:3).  This is
 * Spiral's RegevToGSW: GSW slot k = l*t_gsw + j takes the expanded Regev
 * ciphertext v[2k+1]; column 2j+1 of fold level l is that ciphertext, column
 * 2j is conv * NTT(g^-1(v[2k+1])) (rdim 2, t_conv digits of 14 bits).
 *   dev_v          NTT Ciphertext[n_ct] as sp_expand leaves it (n_ct >= 2 nu2 t_gsw)
 *   dev_conv       key matrix CiphertextHalf[2 crt][2 t_conv], standard form
 *   dev_v_folding  out: CiphertextHalf[nu2][2 crt][2 t_gsw], as sp_fold takes it
 * t_conv must be SP_T_CONV. */
int sp_regev_to_gsw(sp_ctx *ctx, const uint32_t *dev_v, size_t n_ct, int nu2, int t_gsw, int t_conv,
                    const uint32_t *dev_conv, uint32_t *dev_v_folding, void *stream);

/* ---- response packing + modulus switching -----------------------------------
 * Spiral's response path; the reference has none (it stops at the folded
 * ciphertexts).  The n^2 = 4 raw response ciphertexts of one query (trial
 * t = r n + c) become one (n+1) x n matrix ciphertext: column c is
 * from_ntt(sum_r W_r * NTT(g^-1(a_{rn+c}))) (t_conv digits of 14 bits) plus
 * b_{rn+c} on row 1 + r, W_r the client's packing keys.  Row 0 is then
 * rescaled to 2^q2_bits and rows 1..n to 2^10 (x -> floor((x 2^bits +
 * floor(Q/2)) / Q) mod 2^bits) and bit-packed: the n row-0 polynomials, then
 * poly (i, c) of rows 1..n at index (i-1) n + c; coefficient k of a poly at
 * bits [k bits, (k+1) bits), LSB first.  Size: sp_packed_response_bytes.
 *   dev_raw  n2 raw Ciphertexts (contiguous)
 *   dev_w    packing keys u32 [n][2 crt][t_conv][n+1][2048], standard form
 *   dev_out  the compact response (device bytes)
 * Needs n2 = 4, t_conv = SP_T_CONV, 11 <= q2_bits <= 32. */
int sp_packed_response_bytes(int n2, int q2_bits, size_t *bytes);
int sp_pack_response(sp_ctx *ctx, const uint32_t *dev_raw, int n2, const uint32_t *dev_w, int t_conv,
                     int q2_bits, uint8_t *dev_out, void *stream);

/* ---- resident database and answer path -------------------------------------
 * A database of 2^(nu1+nu2) records, each n2 polynomials of 2048 plaintext
 * bytes (n2 = n*n sub-databases; record index = i1 * 2^nu2 + i2).  It is
 * converted to NTT form on the GPU once and stays resident in HBM, its 28-bit
 * residues packed 8 to 7 words (2^nu2 >= 8; narrower databases pack 2^nu2
 * residues into ceil(28 * 2^nu2 / 32) words). */
int sp_db_create(sp_ctx *ctx, int nu1, int nu2, int n2, sp_db **out);
/* host_bytes: [records][n2][2048] u8, uploaded in chunks and NTT'd on device. */
int sp_db_load_bytes(sp_db *db, const uint8_t *host_bytes);
/* Reference layout (NTT Polys, [n2][2^nu2 cols][2^nu1 rows]), host memory. */
int sp_db_load_ntt(sp_db *db, const uint32_t *host_ntt);
int sp_db_destroy(sp_db *db);
/* HBM bytes of the resident (packed) NTT form. */
size_t sp_db_device_bytes(const sp_db *db);
/* Builds now the int8-limb copy of the database that the MFMA multiply reads
 * (4 bytes per residue: as much HBM as the u32 NTT form), which is otherwise
 * built on first use; *limb_bytes (may be NULL) gets its size, 0 for shapes
 * the MFMA multiply does not take (2^nu1 or 2^nu2 < 32).  Lets a server
 * count it before sizing its batches by free memory.  No reference
 * counterpart (the reference multiplies on the CUDA cores only). */
int sp_db_build_limbs(sp_db *db, size_t *limb_bytes);

/* Per-query device state: the query ciphertext, its expansion and folding
 * keys (resident in HBM) and the workspace of one answer.  Replaces the
 * QueryStorage / FoldingQueryStorage structs (gpu-pir/structures.cuh:179-190),
 * which the reference fills by hand in its test mains.
 *   sp_query_load   host pointers as for sp_answer; v_folding_neg may be NULL
 *                   (then G - v_folding, evaluated with half the NTTs)
 *   sp_query_run    asynchronous on `stream`; stage bits 1 = expansion,
 *                   2 = database multiply, 4 = from_ntt + fold.  Stages can
 *                   run one at a time; stage 1 restarts from the stored query
 *                   each time, so a run can be repeated without reloading.
 *                   A run of stage 1 alone forms every element of the
 *                   expanded v (expansion.cuh:87-166); a run of stages 1 and
 *                   2 together leaves out elements nothing after the
 *                   expansion reads: the odd i >= 2^(stop_round + 1) (v_neg1
 *                   products of elements skipped past stop_round; every such
 *                   i has i / 2 >= nu2 * t_gsw), which keep their previous
 *                   contents, and for a query loaded with fold vectors every
 *                   odd element (only the conversion of packed queries reads
 *                   them), whose contents are then unspecified.
 *   sp_query_read   synchronizes, then copies out: what = 0 the n2 response
 *                   ciphertexts (raw), 1 the expanded v (NTT, natural order),
 *                   2 the database-multiply accumulators (NTT, natural order,
 *                   between stages 2 and 4), 3 all n2*2^nu2 raw accumulators,
 *                   4 the fold vectors v_folding (CiphertextHalf[nu2][2][2 t_gsw],
 *                   standard form, natural order: given or converted). */
typedef struct sp_query sp_query;
int sp_query_create(sp_db *db, int t_gsw, int t_exp, sp_query **out);
int sp_query_load(sp_query *q, const uint32_t *host_query, const uint32_t *host_left,
                  const uint32_t *host_right, const uint32_t *host_v_folding,
                  const uint32_t *host_v_folding_neg);
int sp_query_run(sp_query *q, int stages, void *stream);
int sp_query_read(sp_query *q, int what, uint32_t *host_out);
int sp_query_destroy(sp_query *q);

/* Batches: `capacity` query slots against one resident database, run
 * together (one launch per stage for the whole batch).  The server derives
 * v_folding_neg = G - v_folding for every slot.
 *   sp_batch_load            upload one query's ciphertext and keys into a slot
 *                            (waits for the batch's previous run)
 *   sp_batch_run             enqueue stages for slots [0, n) on `stream`
 *   sp_batch_graph_launch    same as sp_batch_run(n, all stages) through a
 *                            hipGraph captured on first use for this n
 *   sp_batch_read            per-slot readouts as sp_query_read
 *   sp_batch_read_responses  responses of slots [0, n): [n][n2] raw Ciphertexts
 *   sp_batch_synchronize     wait for the batch's last run */
typedef struct sp_batch sp_batch;
int sp_batch_create(sp_db *db, int capacity, int t_gsw, int t_exp, sp_batch **out);
int sp_batch_load(sp_batch *b, int slot, const uint32_t *host_query, const uint32_t *host_left,
                  const uint32_t *host_right, const uint32_t *host_v_folding);
int sp_batch_run(sp_batch *b, int n, int stages, void *stream);
int sp_batch_graph_launch(sp_batch *b, int n, void *stream);
int sp_batch_read(sp_batch *b, int slot, int what, uint32_t *host_out);
int sp_batch_read_responses(sp_batch *b, int n, uint32_t *host_out);
int sp_batch_synchronize(sp_batch *b);
/* *done = 1 when the batch's last run has finished, 0 while it runs (no wait). */
int sp_batch_poll(sp_batch *b, int *done);
/* Autotuner: times the database multiply of n queries with every variant
 * that applies (SP_STAGE_DB_MUL: 1 VALU one query per database read; 2
 * int8-limb MFMA, needs 2^nu1 and 2^nu2 >= 32; 3 / 4 VALU reading the
 * database once for 2 / 4 queries, needs n >= 2 and 2^nu1, 2^nu2 >= 8), keeps
 * the fastest for later runs of n queries and reports the times of variants 1
 * and 2 (ms; -1 when a variant does not apply) and the choice.
 * sp_batch_tune_multi_ms gives the last tuning's times of variants 3 and 4.
 * Run it after the batch's expansion has run once; it overwrites the
 * accumulators. */
int sp_batch_tune(sp_batch *b, int n, float *ms_valu, float *ms_mfma, int *chosen);
int sp_batch_tune_multi_ms(const sp_batch *b, float *ms_q2, float *ms_q4);
/* Sets the database-multiply variant later runs of n queries use (as the
 * autotuner would), e.g. a choice tuned on another batch of the same
 * database; SP_ERR_ARG if the variant does not apply. */
int sp_batch_set_dbmul(sp_batch *b, int n, int variant);
int sp_batch_destroy(sp_batch *b);

/* Client expansion keys resident on the device, for one database shape and
 * gadget (the serving path keeps them for the client's session).  host_left /
 * host_right as for sp_expand.  No reference counterpart: the reference's
 * PublicParameters (gpu-pir/structures.cuh:173-177) live in its one test
 * query's storage. */
typedef struct sp_keys sp_keys;
int sp_keys_create(sp_db *db, int t_gsw, int t_exp, const uint32_t *host_left, const uint32_t *host_right,
                   sp_keys **out);
int sp_keys_destroy(sp_keys *k);
/* Adds the client's conversion key (CiphertextHalf[2 crt][2 t_conv], standard
 * form, host memory; t_conv = SP_T_CONV) for packed queries.  The reference
 * keeps it in PublicParameters.conversion (gpu-pir/structures.cuh:176). */
int sp_keys_set_conversion(sp_keys *k, int t_conv, const uint32_t *host_conv);
/* Adds the client's response packing keys (u32 [n][2][t_conv][n+1][2048],
 * standard form, host memory).  No reference counterpart. */
int sp_keys_set_packing(sp_keys *k, int t_conv, const uint32_t *host_w);
/* Response packing for later runs of a batch (q2_bits 0: off).  With packing
 * on, the fold stage ends with sp_pack_response for every slot whose keys
 * carry packing keys (loaded with sp_batch_load_many / _packed), and
 * sp_batch_read_packed copies the compact responses of slots [0, n) of the
 * last run ([n][sp_packed_response_bytes(q2_bits of that run)] bytes).  The
 * device buffer is sized for q2_bits = 32 on first use and never moves, so
 * graphs captured under one q2_bits stay valid when another is set.  No
 * reference counterpart. */
int sp_batch_set_packing(sp_batch *b, int q2_bits);
int sp_batch_read_packed(sp_batch *b, int n, uint8_t *host_out);
/* Loads slots [first, first + n) of a batch in one go: host_queries[i] the
 * query ciphertext of slot first + i ([4N], NTT natural order), keys[i] its
 * resident keys (same database and gadget as the batch; copied device to
 * device), host_v_folding[i] its fold vectors ([nu2][2][2 t_gsw][2][N]).
 * Waits for the batch's previous run; returns when the host buffers may be
 * reused (pinned buffers make the transfers asynchronous DMA). */
int sp_batch_load_many(sp_batch *b, int first, int n, const uint32_t *const *host_queries, sp_keys *const *keys,
                       const uint32_t *const *host_v_folding);
/* Packed queries: as sp_batch_load_many without fold vectors.  Each slot's
 * v_folding is derived on the device from its own expanded query by the
 * Regev -> GSW conversion (sp_regev_to_gsw), which then runs as part of the
 * expansion stage; keys[i] must carry a conversion key.  A run may mix
 * packed slots and slots loaded with fold vectors: only packed slots are
 * converted.  No reference counterpart (the reference's conversion is a stub). */
int sp_batch_load_packed(sp_batch *b, int first, int n, const uint32_t *const *host_queries,
                         sp_keys *const *keys);
/* Asynchronous forms of sp_batch_load_many / sp_batch_load_packed for
 * serving loops: the uploads are queued on the batch's io stream behind its
 * previous run (a device-side wait) and the call returns at once; the next
 * sp_batch_run / sp_batch_graph_launch / sp_group_launch waits for them on
 * the device.  The host buffers (pinned for true DMA) must stay valid until
 * that run has finished (sp_batch_poll).  No reference counterpart; the
 * reference plans this host/device overlap (gpu-pir/README.md:169-194,
 * gpu-pir/optplan.md:69-78) without building it. */
int sp_batch_load_many_async(sp_batch *b, int first, int n, const uint32_t *const *host_queries,
                             sp_keys *const *keys, const uint32_t *const *host_v_folding);
int sp_batch_load_packed_async(sp_batch *b, int first, int n, const uint32_t *const *host_queries,
                               sp_keys *const *keys);
/* Asynchronous readouts of the last run's responses of slots [0, n), queued
 * on `stream` after that run: raw ([n][n2] Ciphertexts, as
 * sp_batch_read_responses) or compact (as sp_batch_read_packed).  The copy
 * is complete when sp_batch_poll reports the batch done; host_out (pinned
 * for true DMA) must stay valid until then. */
int sp_batch_read_responses_async(sp_batch *b, int n, uint32_t *host_out, void *stream);
int sp_batch_read_packed_async(sp_batch *b, int n, uint8_t *host_out, void *stream);
/* Diagnostics: the database-multiply variant the batch's last run used
 * (1 VALU, 2 MFMA; 0 before any run) and the number of stage-chain graphs it
 * holds (one per run size, conversion / packing setting and kernel variants). */
int sp_batch_last_run(const sp_batch *b, int *dbmul_variant, int *graphs_cached);
/* A packed query into a single-query object (slot 0); response packing and
 * the compact response of a single query, as for batches. */
int sp_query_load_packed(sp_query *q, const uint32_t *host_query, sp_keys *keys);
int sp_query_set_packing(sp_query *q, int q2_bits);
int sp_query_read_packed(sp_query *q, uint8_t *host_out);

/* Several batches (each against its own database, e.g. of different shapes)
 * answered by one launch: sp_group_launch runs all stages of slots
 * [0, ns[i]) of every batch i with ns[i] > 0 as one hipGraph whose branches
 * (one per batch) run concurrently, captured on first use for this ns and
 * replayed on `stream`.  The batches' own readouts and loads synchronize with
 * it.  No reference counterpart (the reference answers one query at a time). */
typedef struct sp_group sp_group;
int sp_group_create(sp_batch *const *batches, int count, sp_group **out);
int sp_group_launch(sp_group *g, const int *ns, void *stream);
int sp_group_synchronize(sp_group *g);
int sp_group_destroy(sp_group *g);

/* One query against a resident database, stage by stage on the device:
 *   query        NTT Ciphertext (natural order)
 *   left/right   expansion keys as for sp_expand
 *   v_folding    CiphertextHalf[nu2][2][2*t_gsw]; v_folding_neg = G - v_folding
 *   out          n2 raw Ciphertexts
 * All host pointers. */
int sp_answer(sp_db *db, int t_gsw, int t_exp, const uint32_t *host_query,
              const uint32_t *host_left, const uint32_t *host_right,
              const uint32_t *host_v_folding, uint32_t *host_out);

/* ---- client ----------------------------------------------------------------
 * The reference has no client: its inputs are spiral-rs dumps
 * (the .dat files under gpu-pir/test_data, gpu-pir/corrected_store_vars.rs:1-55).  This host
 * client (no GPU) produces inputs of the same structure for the parameters
 * above: ternary secret, discrete Gaussian errors (sigma 6.4), expansion keys
 * for sp_expand / sp_query_load, the packed query, GSW folding keys, and the
 * decoding of a response.  Buffers are host u32 in the layouts above; sizes
 * in u32 words come from sp_client_sizes.  A client is single-threaded. */
int sp_client_create(int nu1, int nu2, int n2, int t_gsw, int t_exp, uint64_t seed,
                     sp_client **out);
int sp_client_destroy(sp_client *c);
int sp_client_sizes(const sp_client *c, size_t *query_words, size_t *left_words,
                    size_t *right_words, size_t *fold_words, size_t *response_words);
int sp_client_keygen(sp_client *c, uint32_t *host_left, uint32_t *host_right);
int sp_client_query(sp_client *c, uint64_t index, uint32_t *host_query,
                    uint32_t *host_v_folding);
int sp_client_decode(sp_client *c, const uint32_t *host_response, uint8_t *host_record);
/* Conversion key for packed queries: CiphertextHalf[2 crt][2 t_conv]; column
 * 2k encrypts s^2 * 2^(bits k), column 2k+1 encrypts -s * 2^(bits k),
 * bits = ceil(56 / t_conv). */
int sp_client_conversion_key(sp_client *c, int t_conv, uint32_t *host_conv);
/* Packed query: one NTT Ciphertext whose plaintext holds the first-dimension
 * selector at coefficient 2 i1 (scaled by 2^-g) and bit l of i2 times
 * 2^(bits j) at coefficient 2 (l t_gsw + j) + 1 (scaled by 2^-(stop_round+1)). */
int sp_client_query_packed(sp_client *c, uint64_t index, uint32_t *host_query);
/* Phase b - a*s mod Q (u64[2048], coefficient order) of one NTT Ciphertext,
 * and the ternary secret (int32[2048]): for noise measurements in tests. */
int sp_client_phase(const sp_client *c, const uint32_t *host_ct, uint64_t *host_phase);
/* Response packing keys u32 [n][2 crt][t_conv][n+1][2048] (n * n = n2): for
 * r < n, column k of W_r is (a_k, s~_i a_k + e_ik - [i == r] s 2^(bits k))
 * over rows 0, 1 + i, with s~ the client's matrix secret (n ternary polys). */
int sp_client_packing_keys(sp_client *c, int t_conv, uint32_t *host_w);
/* Decodes a compact response (sp_pack_response layout) into the record's
 * n2 * 2048 bytes. */
int sp_client_decode_packed(sp_client *c, const uint8_t *host_response, size_t bytes, int q2_bits,
                            uint8_t *host_record);
int sp_client_secret(const sp_client *c, int32_t *host_s);
const char *sp_client_last_error(void);

/* ---- measurement -----------------------------------------------------------
 * No counterpart in the reference.  The HBM streaming-read ceiling that the
 * database multiply is measured against: one launch reads `bytes` of device
 * memory (a multiple of 16, 16-byte aligned) with the multiply's
 * non-temporal 16-byte loads and writes sp_stream_sink_words() u32 words to
 * dev_sink.  Asynchronous on `stream`. */
size_t sp_stream_sink_words(void);
int sp_stream_read(sp_ctx *ctx, const void *dev_buf, size_t bytes, uint32_t *dev_sink, void *stream);

#ifdef __cplusplus
}
#endif
#endif
