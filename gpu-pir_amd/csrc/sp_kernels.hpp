// sp_kernels.hpp — host-side launchers for the kernels in sp_kernels.hip.
#pragma once
#include "sp_device.hpp"

namespace sp {

// Device-resident tables of a context.
struct DevTables {
  const uint32_t *tw;    // [m][dir][TW_WORDS]
  const uint32_t *neg1;  // [LOGN][m][2][N]  (w, w') in layout P
  ModConst mc[2];
  int n_cu;  // compute units of the device (launch-shape choices)
};

inline const uint32_t *tw_of(const DevTables &t, int m, int inv) { return t.tw + (size_t)(m * 2 + inv) * TW_WORDS; }

hipError_t launch_ntt_natural(bool inv, uint32_t *polys, size_t n, const DevTables &t, int m, int scale,
                              hipStream_t s);
hipError_t launch_from_ntt(bool layout_p, const uint32_t *ntt, uint32_t *raw, size_t n, size_t in_stride,
                           size_t out_stride, const DevTables &t, hipStream_t s);
hipError_t launch_to_ntt(const uint32_t *raw, uint32_t *ntt, size_t n, const DevTables &t, hipStream_t s);
hipError_t launch_automorph_u64(const uint64_t *in, uint64_t *out, size_t n, uint32_t tt, hipStream_t s);
hipError_t launch_extprod(const uint32_t *in, int rdim, int mx, int bits, const uint32_t *keys, uint32_t *acc,
                          size_t n, size_t in_stride, size_t key_stride, size_t acc_stride, const DevTables &t,
                          hipStream_t s);
hipError_t launch_nat_to_p(const uint32_t *in, uint32_t *out, size_t n, int m_div, int mont, const DevTables &t,
                           hipStream_t s);
hipError_t launch_p_to_nat(const uint32_t *in, uint32_t *out, size_t n, hipStream_t s);
// P -> natural with the 2^32 scaling removed (poly k is residue (k / m_div) & 1).
hipError_t launch_p_to_nat_unscale(const uint32_t *in, uint32_t *out, size_t n, int m_div, const DevTables &t,
                                   hipStream_t s);
// Resident database format: groups of CG = min(C, 8) columns, 28 bits per
// residue, packed into W = ceil(28 CG / 32) word planes: [t][m][R][C/CG][W][P].
__host__ __device__ inline int db_group_cols(int C) { return C < 8 ? C : 8; }
__host__ __device__ inline int db_group_words(int C) { return (28 * db_group_cols(C) + 31) / 32; }
inline size_t db_trial_words(int R, int C) {
  return (size_t)2 * R * (C / db_group_cols(C)) * db_group_words(C) * N;
}

// Reference Database layout [c][j][m][z] (one trial) -> u32 staging [m][j][c][P],
// every word reduced mod q_m (any 32-bit input is accepted).
hipError_t launch_db_from_ref(const uint32_t *in, uint32_t *out, int R, int C, const DevTables &t, hipStream_t s);
// Plaintext records rec0 .. rec0 + nrec - 1 (rec0 and nrec multiples of C) ->
// u32 staging [t][m][nrec / C][C][P].
hipError_t launch_db_encode(const uint8_t *bytes, size_t rec0, size_t nrec, uint32_t *db, int n2, int C,
                            const DevTables &t, hipStream_t s);
// u32 staging [t][m][rows][C][P] -> rows j0 .. j0 + rows - 1 of the resident format.
hipError_t launch_db_pack(const uint32_t *in, uint32_t *out, int n2, int rows, int j0, int R, int C, hipStream_t s);
// Batched over nq queries: selectors of query b at sel + b*sel_qstride, outputs at out + b*out_qstride.
hipError_t launch_db_mul(const uint32_t *db, const uint32_t *sel, size_t sel_stride, size_t sel_qstride,
                         uint32_t *out, size_t out_qstride, int nq, int T, int R, int C, const DevTables &t,
                         hipStream_t s);
// Several queries per database read (C >= 8, R >= 8): form 1 = 2 queries x 2
// positions per lane, 2 = 4 queries x 1 position per lane.
hipError_t launch_db_mul_multi(const uint32_t *db, const uint32_t *sel, size_t sel_stride, size_t sel_qstride,
                               uint32_t *out, size_t out_qstride, int nq, int T, int R, int C, int form,
                               const DevTables &t, hipStream_t s);
hipError_t launch_fold_level(bool fast, uint32_t *accs, size_t set_stride, int n_trials, int n_queries,
                             int num_per, const uint32_t *cpos, const uint32_t *cneg, size_t key_query_stride,
                             int t_gsw, int bits, const DevTables &t, hipStream_t s);
hipError_t launch_expand_neg1(uint32_t *v, size_t q_stride, int n_queries, int num_in, const uint32_t *neg1_r,
                              const DevTables &t, hipStream_t s);
hipError_t launch_expand_action(uint32_t *v, size_t q_stride, int n_queries, int r, int num_out, int stop,
                                int max_bits, const uint32_t *left, const uint32_t *right, size_t left_qstride,
                                size_t right_qstride, int t_exp, int bits, const DevTables &t, hipStream_t s);
hipError_t launch_copy_queries(const uint32_t *query, size_t query_qstride, uint32_t *v, size_t v_qstride, int nq,
                               hipStream_t s);
// dst[i] = vals[i] for i < n, on the device (vals is read on the host during the call).
constexpr int SLOT_FLAGS_PER_LAUNCH = 64;
hipError_t launch_set_slot_flags(int *dst, int n, const int *vals, hipStream_t s);
// rows of 4N words at src + r * src_stride -> contiguous dst (response gather for readout).
hipError_t launch_gather_rows(const uint32_t *src, size_t src_stride, int rows, uint32_t *dst, hipStream_t s);

// Response packing (n = 2, t_conv = 4) + modulus switching: per query, raw
// responses of the 4 trials at raw + trial * trial_stride, packing keys
// [r][m][k][row][P] (2^32-scaled); out: packed_response_bytes(q2_bits) bytes.
size_t packed_response_bytes(int q2_bits);
hipError_t launch_pack_response(const uint32_t *raw, size_t raw_q, size_t trial_stride, int n_queries,
                                const uint32_t *pk, size_t pk_q, const int *only, uint8_t *out, size_t out_q,
                                int q2_bits, const DevTables &t, hipStream_t s);

// Regev -> GSW conversion (t_conv = 4): GSW slots k < nu2 t_gsw of every query,
// v [query][>= 2 nu2 t_gsw][4N] layout P; conv [query][m][8][row][P] 2^32-scaled;
// fk [query][nu2][m][2 t_gsw][row][P], 2^32-scaled when mont, else standard.
// only: NULL, or per-query flags (device int[n_queries]); queries flagged 0 are skipped.
hipError_t launch_regev_to_gsw(bool mont, const uint32_t *v, size_t v_q, int n_queries, int nu2, int t_gsw,
                               const uint32_t *conv, size_t conv_q, uint32_t *fk, size_t fk_q, const int *only,
                               const DevTables &t, hipStream_t s);

// Gadget 2^7 x 8 (t_gsw = t_exp = 8) fold and expansion kernels with digit-owning waves.
// variant: 0 = by grid size, 2 = 8-wave blocks, 3 = 16-wave blocks, 4 = split
// (fold only: partial blocks per residue and column half, then a finish
// launch; part: scratch of part_set_stride >= num_per * 8 * N words per set).
hipError_t launch_fold8(uint32_t *accs, size_t set_stride, int n_trials, int n_queries, int num_per,
                        const uint32_t *cpos, size_t key_qstride, const DevTables &t, int variant, uint32_t *part,
                        size_t part_set_stride, hipStream_t s);
// Whole fold (fast form, gadget 2^7 x 8) in the latency form: k_fold_part<2> for
// level 0, one k_fold_fused per later level, one k_fold_finish; the response of
// set (query, trial) ends in accs[set][0].  alt_acc holds 2^nu2 / 4 ciphertexts
// per set and alt_part 2^nu2 / 4 pairs of partial sums of 16N words per set;
// part holds the first level's 2^nu2 / 2 pairs of 8N words.
hipError_t launch_fold_latency(uint32_t *accs, size_t set_stride, uint32_t *alt_acc, size_t alt_set_stride,
                               uint32_t *part, size_t part_set_stride, uint32_t *alt_part, size_t alt_part_set_stride,
                               int n_trials, int n_queries, int nu2, const uint32_t *fkeys, size_t level_words,
                               size_t key_qstride, const DevTables &t, hipStream_t s);
hipError_t launch_expand8(uint32_t *v, size_t q_stride, int n_queries, int r, int num_out, int stop, int max_bits,
                          const uint32_t *left, const uint32_t *right, size_t left_qstride, size_t right_qstride,
                          const DevTables &t, int variant, hipStream_t s);

// Expansion round in two launches (digits, then per-residue transforms and
// update); also forms v[2^r + i] = v_neg1[r] * v[i] (no separate neg1 launch).
//   dig: scratch, per query dig_qstride >= expand_split_work(r, ...) * 8 * 512 words
int expand_split_work(int r, int stop, int max_bits);
// Latency form of the expansion (one launch per round, for rounds with few
// steps; a prefix of the rounds).  snap: two snapshot sets, set k at
// snap + k * s_round, per query s_qstride words = 2^g * 2N (row 0 of every
// element, layout [m][P]).  launch_expand_prep readies round 0, first copying
// the queries into v[0] when `query` is given; expand_fused_blocks gives the
// blocks per query and residue.
hipError_t launch_expand_prep(const uint32_t *query, size_t query_qstride, uint32_t *v, size_t q_stride,
                              int n_queries, uint32_t *snap, size_t s_qstride, const DevTables &t, hipStream_t s);
int expand_fused_blocks(int r, int stop, int max_bits);
// end: the round after the last one this run performs (no epilogue for round end - 1).
hipError_t launch_expand_fused(uint32_t *v, size_t q_stride, int n_queries, int r, int end, int stop, int max_bits,
                               const uint32_t *left, const uint32_t *right, size_t left_qstride, size_t right_qstride,
                               uint32_t *snap, size_t s_qstride, size_t s_round, const DevTables &t, hipStream_t s,
                               bool live_only = false, const int *odd_live = nullptr);
// The latency form in half-position blocks (4-wave blocks per step, residue
// and half of the NTT positions); rounds with (N >> r) >= 4 only.
hipError_t launch_expand_fused_h(uint32_t *v, size_t q_stride, int n_queries, int r, int end, int stop, int max_bits,
                                 const uint32_t *left, const uint32_t *right, size_t left_qstride,
                                 size_t right_qstride, uint32_t *snap, size_t s_qstride, size_t s_round,
                                 const DevTables &t, hipStream_t s, bool live_only = false,
                                 const int *odd_live = nullptr);
// Round r of a run of `end` rounds in the split form (digits, then per-residue
// transforms and update; the update also forms round r + 1's children).
// live_only (all three forms): the odd elements past stop_round that no later
// step, selector or GSW slot reads (the v_neg1 children of skipped elements)
// are not formed; they stay as they were.  With live_only, odd_live (one int
// per query, may be NULL) marks the queries whose odd elements are read (the
// GSW slots of packed queries): the odd steps of the others are not run.
hipError_t launch_expand_split(uint32_t *v, size_t q_stride, int n_queries, int r, int end, int stop, int max_bits,
                               const uint32_t *left, const uint32_t *right, size_t left_qstride, size_t right_qstride,
                               uint32_t *dig, size_t dig_qstride, const DevTables &t, hipStream_t s,
                               bool live_only = false, const int *odd_live = nullptr);

// int8-limb MFMA database multiply (C >= 32, R multiple of 32).
//   D: database limbs [t][m][p][R/32][4][C][32] (k_db_to_limbs, one-time)
//   A: workspace for selector limbs, 2 * N * R * 4 * I bytes (I = 2*nq padded to 32)
//   outp: workspace, 2 * N * T * I * C words ([m][t][I][N/4][C][4]), each product times
//         2^-32 mod q_m (Montgomery form);  pw: [m][8] = 2^(7w) mod q_m
hipError_t launch_db_to_limbs(const uint32_t *db, int T, int R, int C, uint8_t *D, hipStream_t s);
// ring: 4 = staged through LDS with a 4-deep prefetch ring (I > 64); -1 =
// streaming form for I <= 64 (blocks of the row waves only, a register
// prefetch ring per wave); 0 = operands straight from global memory one block
// ahead (the first small-batch form, kept for comparison).
// acc = nullptr leaves the products in outp (for launch_from_ntt_pm);
// otherwise they are transposed into the accumulators [b][t][c][row][m][P].
hipError_t launch_dbmul_mfma(const uint32_t *v, size_t v_q, int nq, const uint8_t *D, int T, int R, int C, int I,
                             uint8_t *A, uint32_t *outp, uint32_t *acc, size_t acc_q, const uint32_t *pw,
                             const DevTables &t, int ring, hipStream_t s);
// Products of nq queries ([m][t][I][N/4][C][4]) -> NTT accumulators [b][t][c][row][m][P].
hipError_t launch_outp_to_acc(const uint32_t *outp, int T, int I, int C, int nq, uint32_t *acc, size_t acc_q,
                              const DevTables &t, hipStream_t s);
// Products of nq queries -> raw accumulators (the from_ntt of every
// accumulator, as launch_from_ntt(true, ...) after launch_outp_to_acc);
// nc = 1, 2 or 4 columns per block, C a multiple of 8, 2 nq <= I.
hipError_t launch_from_ntt_pm(const uint32_t *outp, int T, int I, int C, int nq, uint32_t *acc, size_t acc_q,
                              int nc, const DevTables &t, hipStream_t s);

// HBM streaming-read ceiling: reads bytes / 16 16-byte words of src (bytes a
// multiple of 16) into stream_sink_words() sink words.
size_t stream_sink_words();
hipError_t launch_stream_read(const void *src, size_t bytes, uint32_t *sink, hipStream_t s);

}  // namespace sp
