/*
 * codag_compat.hpp -- the reference's templated host entry points, as
 * header-only C++ over libcodag's C ABI (codag.h). A caller written against
 * the reference (its CLI mains in src/rlev1/rlev1.cu, src/rlev2/rlev2.cu,
 * src/zlib/zlib.cu) includes this header instead of <rlev1/rlev1.h>,
 * <rlev2/rlev2.h> or <zlib/zlib.h>, links -lcodag, and keeps its calls:
 *
 *   rle_v1::compress_gpu_orig     include/rlev1/rlev1.h:103-192
 *   rle_v1::decompress_gpu_orig   include/rlev1/rlev1.h:419-592
 *   rlev2::compress_gpu_orig      include/rlev2/encoder_trans.h:482-564
 *   rlev2::decompress_gpu_orig    include/rlev2/decoder_trans.h:57-133
 *   deflate::decompress_gpu       include/zlib/zlib.h:1583-1700
 *   brle::compress_gpu / decompress_gpu        include/brle/brle.h:251-342
 *   lzss::compress_gpu / decompress_gpu_base   include/lzss/lzss.h:1060-1189
 *
 * Same template parameters, argument meaning, side files and stdout lines as
 * the reference. The launch-shape parameters (QUEUE_SIZE, NUM_CHUNK, NUM_WARP,
 * queue_size, NUM_CHUNKS, WRITE_COL_LEN, ...) are accepted and printed where
 * the reference prints them; the decode itself always runs the library's
 * tuned MI355X kernels. Output buffers are allocated the way each reference
 * caller releases them: rle_v1 and deflate outputs with malloc (rlev1.cu and
 * zlib.cu call free()), rlev2 outputs with new[] (rlev2.cu calls delete[]);
 * brle and lzss outputs with new[] as the reference allocates them (no
 * reference program calls these two).
 * Where the reference prints an error and continues (common.h:50-64), these
 * print "Fatal Error: ..." to stderr and leave the outputs empty.
 *
 * Each host thread gets its own library context, created on first use.
 */
#ifndef CODAG_COMPAT_HPP
#define CODAG_COMPAT_HPP

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>

#include "codag.h"

namespace codag_compat {

struct CtxHolder {
  codag_ctx *ctx = nullptr;
  ~CtxHolder() {
    if (ctx) codag_ctx_destroy(ctx);
  }
};

/* The calling thread's context (device 0, as the reference's single-GPU runs). */
inline codag_ctx *thread_ctx() {
  thread_local CtxHolder h;
  if (!h.ctx) {
    codag_status s = codag_ctx_create(0, &h.ctx);
    if (s != CODAG_OK) {
      std::fprintf(stderr, "Fatal Error: GPU context: %s (%s)\n", codag_status_string(s),
                   codag_last_error());
      h.ctx = nullptr;
    }
  }
  return h.ctx;
}

inline bool report(const char *what, codag_status s) {
  if (s == CODAG_OK) return true;
  std::fprintf(stderr, "Fatal Error: %s: %s (%s)\n", what, codag_status_string(s),
               codag_last_error());
  return false;
}

inline bool write_whole(const char *path, const void *p, uint64_t n) {
  FILE *f = std::fopen(path, "wb");
  if (!f) return false;
  const bool ok = n == 0 || std::fwrite(p, 1, n, f) == n;
  return std::fclose(f) == 0 && ok;
}

}  // namespace codag_compat

namespace rle_v1 {

/* rlev1.h:103-192: encodes the column into the container (*out: [u64 in
 * bytes][chunk streams], malloc'd) and writes ./input_data/rle_v1_col_len.bin
 * and ./input_data/rle_v1_blk_offset.bin. DATA_TYPE is the column element. */
template <typename READ_COL_TYPE, typename DATA_TYPE, typename OUT_COL_TYPE>
void compress_gpu_orig(const uint8_t *const in, uint8_t **out, const uint64_t in_n_bytes,
                       uint64_t *out_n_bytes, int COMP_COL_LEN, uint64_t CHUNK_SIZE) {
  (void)COMP_COL_LEN;
  *out = nullptr;
  *out_n_bytes = 0;
  uint8_t *cont = nullptr;
  uint64_t cont_n = 0, *blk = nullptr, *col = nullptr, nblk = 0, ncol = 0, nch = 0;
  if (!codag_compat::report("rle_v1::compress_gpu_orig",
                            codag_rlev1_encode(in, in_n_bytes, (uint32_t)sizeof(DATA_TYPE),
                                               CHUNK_SIZE, &cont, &cont_n, &blk, &nblk, &col,
                                               &ncol, &nch)))
    return;
  if (!codag_compat::write_whole("./input_data/rle_v1_col_len.bin", col, ncol * 8) ||
      !codag_compat::write_whole("./input_data/rle_v1_blk_offset.bin", blk, nblk * 8))
    std::fprintf(stderr, "Fatal Error: cannot write ./input_data side files\n");
  codag_host_free(blk);
  codag_host_free(col);
  *out = cont;
  *out_n_bytes = cont_n;
}

/* rlev1.h:419-592: reads ./input_data/rle_v1_{col_len,blk_offset}.bin,
 * decodes the container `in` on the GPU into *out (malloc'd, the original
 * length) and prints "NUM WARP: n" and "in_n\tmeta\tratio\tseconds". */
template <typename READ_COL_TYPE, typename DATA_TYPE, typename OUT_COL_TYPE, int QUEUE_SIZE = 8,
          int NUM_CHUNK = 1, int NUM_WARP = 1>
void decompress_gpu_orig(const uint8_t *const in, uint8_t **out, const uint64_t in_n_bytes,
                         uint64_t *out_n_bytes, int COMP_COL_LEN, uint64_t chunk_size) {
  (void)COMP_COL_LEN;
  *out = nullptr;
  *out_n_bytes = 0;
  codag_ctx *ctx = codag_compat::thread_ctx();
  if (!ctx) return;
  uint8_t *dec = nullptr;
  uint64_t dec_n = 0, meta = 0;
  double secs = 0;
  std::printf("NUM WARP: %i\n", NUM_WARP);
  if (!codag_compat::report(
          "rle_v1::decompress_gpu_orig",
          codag_rlev1_decode_files(ctx, in, in_n_bytes, "./input_data/rle_v1_col_len.bin",
                                   "./input_data/rle_v1_blk_offset.bin",
                                   (uint32_t)sizeof(DATA_TYPE), chunk_size, &dec, &dec_n, &meta,
                                   &secs)))
    return;
  const float ratio = (float)in_n_bytes / (float)(dec_n ? dec_n : 1);
  std::printf("%llu\t%llu\t%g\t%g\n", (unsigned long long)in_n_bytes, (unsigned long long)meta,
              ratio, secs);
  *out = dec;
  *out_n_bytes = dec_n;
}

}  // namespace rle_v1

namespace rlev2 {

typedef uint64_t col_len_t; /* include/rlev2/utils.h:83-84 */
typedef uint64_t blk_off_t;

/* encoder_trans.h:482-564: out (new[]), blk_off (new[], n_chunks+1 entries,
 * the last = in_n_bytes), col_len (new[], n_chunks*32 entries as the
 * reference allocates, the first n_chunks meaningful), meta_data_size as the
 * reference counts it. COMP_TYPE is the column element. */
template <int READ_UNIT, typename COMP_TYPE, typename INPUT_T>
void compress_gpu_orig(const INPUT_T *const in, const uint64_t in_n_bytes, uint8_t *&out,
                       uint64_t &out_n_bytes, uint64_t &meta_data_size, uint64_t &out_n_chunks,
                       blk_off_t *&blk_off, col_len_t *&col_len, uint64_t CHUNK_SIZE) {
  out = nullptr;
  blk_off = nullptr;
  col_len = nullptr;
  out_n_bytes = meta_data_size = out_n_chunks = 0;
  uint8_t *comp = nullptr;
  uint64_t comp_n = 0, *blk = nullptr, *col = nullptr, n = 0;
  if (!codag_compat::report("rlev2::compress_gpu_orig",
                            codag_rlev2_encode(reinterpret_cast<const uint8_t *>(in), in_n_bytes,
                                               (uint32_t)sizeof(COMP_TYPE), CHUNK_SIZE, &comp,
                                               &comp_n, &blk, &col, &n)))
    return;
  out = new uint8_t[comp_n ? comp_n : 1];
  if (comp_n) std::memcpy(out, comp, comp_n);
  blk_off = new blk_off_t[n + 1];
  std::memcpy(blk_off, blk, 8 * (n + 1));
  col_len = new col_len_t[n * 32 + 1]();
  if (n) std::memcpy(col_len, col, 8 * n);
  codag_host_free(comp);
  codag_host_free(blk);
  codag_host_free(col);
  out_n_bytes = comp_n;
  out_n_chunks = n;
  meta_data_size = sizeof(col_len_t) * n * 32 + sizeof(blk_off_t) * (n + 1);
}

/* decoder_trans.h:57-133: decodes n_chunks chunk streams (blk_off[n_chunks]
 * = original byte length) into `out` (new[] of INPUT_T; rlev2.cu releases it
 * with delete[]) and prints "<seconds> secs". */
template <int READ_UNIT, int queue_size, typename COMP_TYPE, int NUM_WARP = 2, int NUM_CHUNKS = 1,
          typename INPUT_T>
void decompress_gpu_orig(const uint8_t *in, const uint64_t in_n_bytes, const uint64_t n_chunks,
                         blk_off_t *blk_off, col_len_t *col_len, INPUT_T *&out,
                         uint64_t &out_n_bytes, uint64_t CHUNK_SIZE) {
  out = nullptr;
  out_n_bytes = 0;
  codag_ctx *ctx = codag_compat::thread_ctx();
  if (!ctx) return;
  const uint64_t orig = blk_off[n_chunks];
  std::unique_ptr<INPUT_T[]> buf(new INPUT_T[orig / sizeof(INPUT_T) + 1]);
  uint64_t got = 0;
  double secs = 0;
  if (!codag_compat::report(
          "rlev2::decompress_gpu_orig",
          codag_rlev2_decode_host(ctx, in, in_n_bytes, blk_off, col_len, n_chunks,
                                  (uint32_t)sizeof(COMP_TYPE), CHUNK_SIZE,
                                  reinterpret_cast<uint8_t *>(buf.get()), orig, &got, &secs,
                                  nullptr)))
    return;
  std::printf("%g secs\n", secs);
  out = buf.release();
  out_n_bytes = got;
}

}  // namespace rlev2

namespace deflate {

/* zlib.h:1583-1700: blk_offset_f is the whole _blk_offset.bin file
 * ([u64 original length][0][cumulative ends...], blk_n_bytes bytes); decodes
 * the chunked zlib container into *out (malloc'd, the original length) and
 * prints "\t<seconds>". */
template <typename READ_COL_TYPE, size_t WRITE_COL_LEN, uint16_t queue_depth,
          uint8_t NUM_SUBCHUNKS, bool All_Thread_Decoding = true>
void decompress_gpu(const uint8_t *const in, uint8_t **out, const uint64_t in_n_bytes,
                    uint64_t *out_n_bytes, uint64_t *blk_offset_f, const uint64_t blk_n_bytes,
                    uint64_t chunk_size) {
  *out = nullptr;
  *out_n_bytes = 0;
  codag_ctx *ctx = codag_compat::thread_ctx();
  if (!ctx) return;
  if (blk_n_bytes < 16) {
    std::fprintf(stderr, "Fatal Error: blk_offset file shorter than two entries\n");
    return;
  }
  const uint64_t orig = blk_offset_f[0];
  uint8_t *buf = static_cast<uint8_t *>(std::malloc(orig ? orig : 1));
  if (!buf) {
    std::fprintf(stderr, "Fatal Error: out of host memory\n");
    return;
  }
  uint64_t got = 0;
  double secs = 0;
  if (!codag_compat::report("deflate::decompress_gpu",
                            codag_inflate_chunks_host(ctx, in, in_n_bytes, blk_offset_f,
                                                      blk_n_bytes / 8, chunk_size, 0, buf, orig,
                                                      &got, &secs, nullptr))) {
    std::free(buf);
    return;
  }
  std::printf("\t%g\n", secs);
  *out = buf;
  *out_n_bytes = got;
}

}  // namespace deflate

namespace codag_compat {

/* Moves a library-allocated buffer into a new[] one. */
inline uint8_t *to_new(uint8_t *p, uint64_t n) {
  uint8_t *q = new uint8_t[n ? n : 1];
  if (n) std::memcpy(q, p, n);
  codag_host_free(p);
  return q;
}

}  // namespace codag_compat

namespace brle {

/* brle.h:251-303: *out = [u32 n+1][u64 ptr[n+1]][u64 in bytes][streams]. */
inline void compress_gpu(const uint8_t *in, uint8_t **out, const uint64_t in_n_bytes,
                         uint64_t *out_n_bytes) {
  *out = nullptr;
  *out_n_bytes = 0;
  uint8_t *c = nullptr;
  uint64_t n = 0;
  if (!codag_compat::report("brle::compress_gpu", codag_brle_encode(in, in_n_bytes, 0, 0, &c, &n)))
    return;
  *out = codag_compat::to_new(c, n);
  *out_n_bytes = n;
}

/* brle.h:305-342: decodes a compress_gpu container into *out (new[]). */
inline void decompress_gpu(const uint8_t *const in, uint8_t **out, const uint64_t in_n_bytes,
                           uint64_t *out_n_bytes) {
  *out = nullptr;
  *out_n_bytes = 0;
  codag_ctx *ctx = codag_compat::thread_ctx();
  if (!ctx || in_n_bytes < 4) return;
  uint32_t n_ptr = 0;
  std::memcpy(&n_ptr, in, 4);
  uint64_t orig = 0;
  if (in_n_bytes >= 12 + 8ull * n_ptr) std::memcpy(&orig, in + 4 + 8ull * n_ptr, 8);
  std::unique_ptr<uint8_t[]> buf(new uint8_t[orig ? orig : 1]);
  uint64_t got = 0;
  if (!codag_compat::report("brle::decompress_gpu",
                            codag_brle_decode_host(ctx, in, in_n_bytes, 0, buf.get(), orig, &got,
                                                   nullptr, nullptr)))
    return;
  *out = buf.release();
  *out_n_bytes = got;
}

}  // namespace brle

namespace lzss {

/* lzss.h:1060-1124: *out = [u64 in bytes][u32 chunk, n, 2048, 3, 34][u64
 * ends[n]][streams]. */
inline void compress_gpu(const uint8_t *const in, uint8_t **out, const uint64_t in_n_bytes,
                         uint64_t *out_n_bytes) {
  *out = nullptr;
  *out_n_bytes = 0;
  uint8_t *c = nullptr;
  uint64_t n = 0;
  if (!codag_compat::report("lzss::compress_gpu", codag_lzss_encode(in, in_n_bytes, 0, 0, &c, &n)))
    return;
  *out = codag_compat::to_new(c, n);
  *out_n_bytes = n;
}

/* lzss.h:1126-1189: decodes a compress_gpu container into *out (new[]). */
inline void decompress_gpu_base(const uint8_t *const in, uint8_t **out, const uint64_t in_n_bytes,
                                uint64_t *out_n_bytes) {
  *out = nullptr;
  *out_n_bytes = 0;
  codag_ctx *ctx = codag_compat::thread_ctx();
  if (!ctx || in_n_bytes < 8) return;
  uint64_t orig = 0;
  std::memcpy(&orig, in, 8);
  std::unique_ptr<uint8_t[]> buf(new uint8_t[orig ? orig : 1]);
  uint64_t got = 0;
  if (!codag_compat::report("lzss::decompress_gpu_base",
                            codag_lzss_decode_host(ctx, in, in_n_bytes, buf.get(), orig, &got,
                                                   nullptr, nullptr)))
    return;
  *out = buf.release();
  *out_n_bytes = got;
}

}  // namespace lzss

#endif /* CODAG_COMPAT_HPP */
