// CPU sanitizer driver: the oracle (oracle/spiral_oracle.c) and the host client
// (gpu-pir_amd/csrc/sp_client.cpp) built with AddressSanitizer and
// UndefinedBehaviorSanitizer, run through a whole answer at a small shape:
// keygen, query, database encoding, expansion, multiply, fold, decode; the
// packed query with the Regev -> GSW conversion, response packing and compact
// decode; the client's argument checks.  Test infrastructure: built and run
// by tests/test_sanitizers_cpu.py, never shipped.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

extern "C" {
#include "../../oracle/spiral_oracle.h"
}
#include "../../include/spiral_pir.h"

static int fails = 0;
#define CHECK(cond, what)                                   \
  do {                                                      \
    if (!(cond)) {                                          \
      std::fprintf(stderr, "sanitize: FAILED %s\n", what);  \
      fails++;                                              \
    }                                                       \
  } while (0)

int main() {
  const int nu1 = 4, nu2 = 2, n2 = 4, t = 8, N = 2048;
  const int rows = 1 << nu1, cols = 1 << nu2, records = rows * cols;
  sp_client *c = nullptr;
  CHECK(sp_client_create(nu1, nu2, n2, t, t, 77, &c) == SP_OK, "client create");
  size_t qw = 0, lw = 0, rw = 0, fw = 0, respw = 0;
  CHECK(sp_client_sizes(c, &qw, &lw, &rw, &fw, &respw) == SP_OK, "sizes");
  std::vector<uint32_t> left(lw), right(rw), query(qw), vf(fw), vneg(fw), resp(respw);
  CHECK(sp_client_keygen(c, left.data(), right.data()) == SP_OK, "keygen");

  // uniform records from a small LCG, NTT form by the oracle
  std::vector<uint8_t> recs((size_t)records * n2 * N);
  uint64_t st = 12345;
  for (auto &b : recs) { st = st * 6364136223846793005ull + 1442695040888963407ull; b = (uint8_t)(st >> 56); }
  std::vector<uint32_t> db((size_t)n2 * cols * rows * 2 * N);
  or_db_encode(recs.data(), db.data(), n2, nu1, nu2);

  const uint64_t index = 37;
  CHECK(sp_client_query(c, index, query.data(), vf.data()) == SP_OK, "query");
  or_folding_neg(vf.data(), vneg.data(), nu2, t);
  or_set_threads(2);
  or_answer(db.data(), n2, nu1, nu2, t, t, query.data(), left.data(), right.data(), vf.data(), vneg.data(),
            resp.data());
  std::vector<uint8_t> rec((size_t)n2 * N);
  CHECK(sp_client_decode(c, resp.data(), rec.data()) == SP_OK, "decode");
  CHECK(std::memcmp(rec.data(), recs.data() + index * n2 * N, rec.size()) == 0, "decoded record");

  // packed query: conversion on the "server", response packing, compact decode
  std::vector<uint32_t> conv((size_t)2 * 2 * SP_T_CONV * 2 * N), w((size_t)2 * 2 * SP_T_CONV * 3 * N);
  CHECK(sp_client_conversion_key(c, SP_T_CONV, conv.data()) == SP_OK, "conversion key");
  CHECK(sp_client_packing_keys(c, SP_T_CONV, w.data()) == SP_OK, "packing keys");
  const uint64_t index2 = 50;
  CHECK(sp_client_query_packed(c, index2, query.data()) == SP_OK, "packed query");
  const int g = or_expansion_g(nu1, nu2, t);
  std::vector<uint32_t> v(((size_t)1 << g) * 4 * N, 0);
  std::memcpy(v.data(), query.data(), qw * 4);
  or_expand(v.data(), nu1, nu2, t, t, left.data(), right.data(), -1);
  or_regev_to_gsw(v.data(), nu2, t, SP_T_CONV, conv.data(), vf.data());
  or_folding_neg(vf.data(), vneg.data(), nu2, t);
  or_answer(db.data(), n2, nu1, nu2, t, t, query.data(), left.data(), right.data(), vf.data(), vneg.data(),
            resp.data());
  const int q2 = 20;
  std::vector<uint8_t> packed(or_packed_response_bytes(2, q2, 10));
  or_pack_response(resp.data(), 2, w.data(), SP_T_CONV, q2, 10, packed.data());
  CHECK(sp_client_decode_packed(c, packed.data(), packed.size(), q2, rec.data()) == SP_OK, "decode packed");
  CHECK(std::memcmp(rec.data(), recs.data() + index2 * n2 * N, rec.size()) == 0, "decoded packed record");

  // transforms at the extremes
  std::vector<uint32_t> a(N), b(N);
  for (int m = 0; m < 2; m++) {
    for (int i = 0; i < N; i++) a[i] = or_q(m) - 1;
    b = a;
    or_ntt_fwd(b.data(), m);
    or_ntt_inv(b.data(), m);
    CHECK(b == a, "ntt round trip at q - 1");
  }
  std::vector<uint32_t> raw(2 * N), raw2(2 * N);
  const uint64_t Q = or_modulus();
  for (int i = 0; i < N; i++) { raw[i] = (uint32_t)(Q >> 32); raw[N + i] = (uint32_t)Q; }  // every coefficient = Q
  or_automorph(raw.data(), raw2.data(), 3);
  std::vector<uint32_t> digits((size_t)8 * N);
  or_gadget_inv(raw2.data(), 1, 8, digits.data());

  // the client's argument checks return errors
  CHECK(sp_client_create(0, 2, 4, 8, 8, 1, nullptr) != SP_OK, "create with NULL out");
  CHECK(sp_client_query(c, (uint64_t)records, query.data(), vf.data()) != SP_OK, "index out of range");
  CHECK(sp_client_decode_packed(c, packed.data(), packed.size() - 1, q2, rec.data()) != SP_OK, "short response");
  CHECK(std::strlen(sp_client_last_error()) > 0, "error message");
  CHECK(sp_client_destroy(c) == SP_OK, "destroy");
  if (fails) return 1;
  std::printf("sanitize: ok\n");
  return 0;
}
