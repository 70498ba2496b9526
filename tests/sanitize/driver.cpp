// Sanitizer driver: the CPU producers (codag_amd/csrc/host/producers.cpp,
// byte_producers.cpp) and the CPU oracle (oracle/*.c), built with AddressSanitizer and
// UndefinedBehaviorSanitizer, run over round trips, damaged and truncated
// streams. Exit status 0 = every round trip exact and no sanitizer report
// (a report aborts the run).
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "codag.h"
#include "oracle.h"

namespace codag {  // the library's error sink (capi.cpp), not linked here
void set_error(const std::string &) {}
void clear_error() {}
}  // namespace codag

static int failures = 0;
#define EXPECT(c)                                                        \
  do {                                                                   \
    if (!(c)) {                                                          \
      std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
      failures++;                                                        \
    }                                                                    \
  } while (0)

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t rnd() {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}

static void rlev1_round_trips() {
  const uint64_t chunk = 4096;
  for (uint32_t w : {1u, 2u, 4u, 8u}) {
    const uint64_t n = 3 * chunk / w + 17;
    std::vector<uint8_t> col(n * w);
    for (uint64_t i = 0; i < n; i++) {  // runs, steps and noise
      uint64_t v = (i / 37) % 3 == 0 ? rnd() % 300 : (i % 5 == 0 ? rnd() : i * 3);
      std::memcpy(&col[i * w], &v, w);
    }
    uint8_t *cont = nullptr;
    uint64_t cont_n = 0, *blk = nullptr, *cl = nullptr, nb = 0, nl = 0, nc = 0;
    EXPECT(codag_rlev1_encode(col.data(), col.size(), w, chunk, &cont, &cont_n, &blk, &nb, &cl,
                              &nl, &nc) == CODAG_OK);
    std::vector<uint8_t> out(nc * chunk);
    std::vector<int32_t> st(nc);
    EXPECT(oracle_rlev1_decode_chunks(cont + 8, blk, cl, nc, w, chunk, out.data(), st.data(), 2) == 0);
    EXPECT(std::memcmp(out.data(), col.data(), col.size()) == 0);
    // damaged: every byte of chunk 0 flipped in turn, then truncations
    const uint64_t len0 = cl[0];
    std::vector<uint8_t> s(cont + 8 + blk[0], cont + 8 + blk[0] + len0);
    std::vector<uint8_t> o(chunk);
    uint64_t produced = 0;
    for (uint64_t k = 0; k < len0; k += 7) {
      std::vector<uint8_t> d = s;
      d[k] ^= (uint8_t)(1 + rnd() % 255);
      (void)oracle_rlev1_decode_chunk(d.data(), d.size(), w, o.data(), chunk, &produced);
      EXPECT(produced <= chunk);
    }
    for (uint64_t cut = 0; cut < len0; cut += len0 / 9 + 1) {
      std::vector<uint8_t> d(s.begin(), s.begin() + cut);
      (void)oracle_rlev1_decode_chunk(d.data(), d.size(), w, o.data(), chunk, &produced);
      EXPECT(produced <= chunk);
    }
    codag_host_free(cont);
    codag_host_free(blk);
    codag_host_free(cl);
  }
}

static void rlev2_round_trips() {
  const uint64_t chunk = 8192;
  for (uint32_t w : {1u, 2u, 4u, 8u}) {
    uint8_t *raw = nullptr, *comp = nullptr;
    uint64_t comp_n = 0, *blk = nullptr, *cl = nullptr, nc = 0;
    const uint64_t n = 2 * chunk / w + 5;
    EXPECT(codag_synth_rlev2(n, w, chunk, 11 + w, &raw, &comp, &comp_n, &blk, &cl, &nc) == CODAG_OK);
    std::vector<uint8_t> out(nc * chunk);
    std::vector<int32_t> st(nc);
    EXPECT(oracle_rlev2_decode_chunks(comp, blk, cl, nc, w, chunk, out.data(), st.data(), 2) == 0);
    EXPECT(std::memcmp(out.data(), raw, n * w) == 0);
    // the reference-choice encoder on the same column; 64-bit values are
    // cut to 32 bits, since the reference decoder keeps only 32 bits of a
    // PATCHED_BASE patch (decoder_new.h:1105) while its encoder writes
    // patches up to 56 bits wide
    std::vector<uint8_t> col(raw, raw + n * w);
    if (w == 8)
      for (uint64_t i = 0; i < n; i++) std::memset(&col[i * 8 + 4], 0, 4);
    uint8_t *c2 = nullptr;
    uint64_t c2n = 0, *b2 = nullptr, *l2 = nullptr, n2 = 0;
    EXPECT(codag_rlev2_encode(col.data(), n * w, w, chunk, &c2, &c2n, &b2, &l2, &n2) == CODAG_OK);
    std::vector<uint8_t> out2(n2 * chunk);
    std::vector<int32_t> st2(n2);
    EXPECT(oracle_rlev2_decode_chunks(c2, b2, l2, n2, w, chunk, out2.data(), st2.data(), 2) == 0);
    EXPECT(std::memcmp(out2.data(), col.data(), n * w) == 0);
    // damaged chunk 0
    std::vector<uint8_t> s(comp + blk[0], comp + blk[0] + cl[0]);
    std::vector<uint8_t> o(chunk);
    uint64_t produced = 0;
    for (uint64_t k = 0; k < s.size(); k += 5) {
      std::vector<uint8_t> d = s;
      d[k] ^= (uint8_t)(1 + rnd() % 255);
      (void)oracle_rlev2_decode_chunk(d.data(), d.size(), w, o.data(), chunk / w, &produced);
      EXPECT(produced <= chunk / w);
    }
    codag_host_free(raw);
    codag_host_free(comp);
    codag_host_free(blk);
    codag_host_free(cl);
    codag_host_free(c2);
    codag_host_free(b2);
    codag_host_free(l2);
  }
}

static void deflate_round_trips() {
  const uint64_t chunk = 65536;
  uint8_t *raw = nullptr;
  const uint64_t n = 5 * chunk + 999;
  EXPECT(codag_synth_deflate_payload(n, chunk, 5, &raw) == CODAG_OK);
  uint8_t *comp = nullptr;
  uint64_t comp_n = 0, *blk = nullptr, *cl = nullptr, nc = 0;
  EXPECT(codag_deflate_compress_chunks(raw, n, chunk, 6, 2, &comp, &comp_n, &blk, &cl, &nc) ==
         CODAG_OK);
  std::vector<uint8_t> out(nc * chunk);
  std::vector<int32_t> st(nc);
  EXPECT(oracle_inflate_chunks(comp, blk + 1, nc, chunk, out.data(), st.data(), 2) == 0);
  EXPECT(std::memcmp(out.data(), raw, n) == 0);
  uint8_t *mixed = nullptr;
  uint64_t mixed_n = 0, counts[3] = {0, 0, 0};
  EXPECT(codag_deflate_compress_mixed(raw, n, 20, 5, 8192, 3, &mixed, &mixed_n, counts) == CODAG_OK);
  std::vector<uint8_t> o(n);
  uint64_t got = 0;
  EXPECT(oracle_zlib_inflate(mixed, mixed_n, o.data(), n, &got) == 0 && got == n);
  EXPECT(std::memcmp(o.data(), raw, n) == 0);
  codag_host_free(raw);
  codag_host_free(comp);
  codag_host_free(blk);
  codag_host_free(cl);
  codag_host_free(mixed);
}

static void byte_codec_round_trips() {
  for (uint32_t kind : {0u, 1u}) {
    const uint64_t chunk = kind ? 5000 : 7000, n = 4 * chunk + 321;
    uint8_t *raw = nullptr, *cont = nullptr;
    uint64_t cont_n = 0;
    EXPECT(codag_synth_byte_payload(n, kind, 17 + kind, &raw) == CODAG_OK);
    if (kind == 0) {
      EXPECT(codag_brle_encode(raw, n, chunk, 2, &cont, &cont_n) == CODAG_OK);
      uint32_t n_ptr = 0;
      std::memcpy(&n_ptr, cont, 4);
      const uint64_t nc = n_ptr - 1;
      std::vector<uint64_t> ptr(n_ptr);
      std::memcpy(ptr.data(), cont + 4, 8 * n_ptr);
      const uint8_t *data = cont + 12 + 8 * n_ptr;
      std::vector<uint8_t> out(nc * chunk);
      std::vector<uint64_t> len(nc);
      std::vector<int32_t> st(nc);
      EXPECT(oracle_brle_decode_chunks(data, ptr.data(), nc, chunk, out.data(), len.data(),
                                       st.data()) == 0);
      EXPECT(std::memcmp(out.data(), raw, n) == 0);
      std::vector<uint8_t> ref(chunk + chunk / 128 + 2);
      EXPECT(oracle_brle_encode_chunk(raw, chunk, ref.data()) == ptr[1] &&
             std::memcmp(ref.data(), data, ptr[1]) == 0);
      std::vector<uint8_t> s(data, data + ptr[1]), o(chunk);
      uint64_t produced = 0;
      for (uint64_t k = 0; k < s.size(); k += 3) {
        std::vector<uint8_t> d = s;
        d[k] ^= (uint8_t)(1 + rnd() % 255);
        (void)oracle_brle_decode_chunk(d.data(), d.size(), o.data(), chunk, &produced);
        EXPECT(produced <= chunk);
        (void)oracle_brle_decode_chunk(d.data(), k, o.data(), chunk, &produced);
        EXPECT(produced <= chunk);
      }
    } else {
      EXPECT(codag_lzss_encode(raw, n, chunk, 2, &cont, &cont_n) == CODAG_OK);
      uint32_t h[5];
      std::memcpy(h, cont + 8, 20);
      const uint64_t nc = h[1];
      std::vector<uint64_t> ends(nc);
      std::memcpy(ends.data(), cont + 28, 8 * nc);
      const uint8_t *data = cont + 28 + 8 * nc;
      std::vector<uint8_t> out(nc * chunk);
      std::vector<uint64_t> len(nc);
      std::vector<int32_t> st(nc);
      EXPECT(oracle_lzss_decode_chunks(data, ends.data(), nc, chunk, out.data(), len.data(),
                                       st.data()) == 0);
      EXPECT(std::memcmp(out.data(), raw, n) == 0);
      std::vector<uint8_t> ref(2 * chunk + chunk / 8 + 2);
      EXPECT(oracle_lzss_encode_chunk(raw, chunk, ref.data()) == ends[0] &&
             std::memcmp(ref.data(), data, ends[0]) == 0);
      std::vector<uint8_t> s(data, data + ends[0]), o(chunk);
      uint64_t produced = 0;
      for (uint64_t k = 0; k < s.size(); k += 3) {
        std::vector<uint8_t> d = s;
        d[k] ^= (uint8_t)(1 + rnd() % 255);
        (void)oracle_lzss_decode_chunk(d.data(), d.size(), o.data(), chunk, &produced);
        EXPECT(produced <= chunk);
        (void)oracle_lzss_decode_chunk(d.data(), k, o.data(), chunk, &produced);
        EXPECT(produced <= chunk);
      }
    }
    codag_host_free(raw);
    codag_host_free(cont);
  }
}

int main() {
  rlev1_round_trips();
  rlev2_round_trips();
  deflate_round_trips();
  byte_codec_round_trips();
  std::printf("sanitizer driver: %d failures\n", failures);
  return failures ? 1 : 0;
}
