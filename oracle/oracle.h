/*
 * oracle.h -- CPU restatement of the CODAG decoders, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE. Nothing in the product (codag_amd/, include/codag.h,
 * the CLIs) includes, links or calls this library. Only tests/, the smoke()
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py load it.
 *
 * Each function cites the reference file:line it restates. Paths are relative
 * to the reference repository (illinois-impact/codag).
 *
 * Status values are shared with the product's per-chunk status codes
 * (include/codag.h: CODAG_CHUNK_*); tests/test_oracle.py checks they agree.
 */
#ifndef CODAG_ORACLE_H
#define CODAG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORACLE_OK = 0,
  ORACLE_TRUNCATED = 1, /* stream ended before the chunk/block was complete */
  ORACLE_OVERFLOW = 2,  /* decoded values exceed the chunk's output capacity */
  ORACLE_CORRUPT = 3,   /* invalid field (bad width index, patch out of block, bad zlib data) */
};

/* ---- RLE v1 (ORC v1 integer RLE) ------------------------------------------
 * Restates include/rlev1/decoder.h:671-791 (decoder_warp_orig_rdw_fsm).
 * Decodes values of `width` bytes (1, 2, 4, 8) until `chunk_bytes` output
 * bytes have been produced (the reference's loop condition, decoder.h:671).
 * `out` must hold chunk_bytes bytes. *produced gets the bytes written. */
int oracle_rlev1_decode_chunk(const uint8_t *in, uint64_t in_len, uint32_t width,
                              uint8_t *out, uint64_t chunk_bytes, uint64_t *produced);

/* Whole container: chunk i's stream is comp[blk_off[i] .. blk_off[i]+col_len[i])
 * (comp points just past the 8-byte original-length word, rlev1.h:521), its
 * output lands at out + i*chunk_bytes (decoder.h:934). Runs on `threads`
 * pthreads (0 = one). status[i] (nullable) receives each chunk's code.
 * Returns the number of chunks with a non-OK status. */
uint64_t oracle_rlev1_decode_chunks(const uint8_t *comp, const uint64_t *blk_off,
                                    const uint64_t *col_len, uint64_t n_chunks,
                                    uint32_t width, uint64_t chunk_bytes, uint8_t *out,
                                    int32_t *status, int threads);

/* ---- RLE v2 (ORC v2 integer RLE, reference dialect) -----------------------
 * Restates include/rlev2/decoder_new.h:909-1187 (decoder_warp_orig_rdw) with
 * the width tables of include/rlev2/utils.h:28-59. Consumes exactly col_len
 * input bytes' worth of blocks (decoder_new.h:1000); writes values of `width`
 * bytes at out[0..cap_values). Unwritten positions are left untouched (the
 * reference zero-fills the output first, decoder_trans.h:81; callers do the
 * same). *produced gets the number of values written. */
int oracle_rlev2_decode_chunk(const uint8_t *in, uint64_t col_len, uint32_t width,
                              uint8_t *out, uint64_t cap_values, uint64_t *produced);

/* Whole container (decoder_trans.h:57-133): chunk i starts at comp+blk_off[i],
 * is col_len[i] bytes long and decodes into out + i*chunk_bytes. `out` is
 * zero-filled by this call first, as the reference does. */
uint64_t oracle_rlev2_decode_chunks(const uint8_t *comp, const uint64_t *blk_off,
                                    const uint64_t *col_len, uint64_t n_chunks,
                                    uint32_t width, uint64_t chunk_bytes, uint8_t *out,
                                    int32_t *status, int threads);

/* Width tables (utils.h:28-59), exported so tests can pin them. */
uint32_t oracle_rlev2_decode_width(uint32_t code);  /* __BIT_WIDTH_DECODE_MAP */
uint32_t oracle_rlev2_closest_width(uint32_t bits); /* __CLOSEST_FIXED_BIT_MAP */

/* ---- Deflate / zlib (system zlib, RFC 1950/1951) --------------------------
 * The reference kernel (include/zlib/zlib.h:807-866) skips the 2-byte zlib
 * header and decodes the deflate blocks; the oracle for that arithmetic is
 * zlib's own inflate(). */
const char *oracle_zlib_version(void);
/* Full zlib stream (header + deflate + Adler-32 trailer checked by zlib). */
int oracle_zlib_inflate(const uint8_t *in, uint64_t in_len, uint8_t *out, uint64_t out_cap,
                        uint64_t *out_len);
/* gzip member (RFC 1952): header, deflate data, CRC-32 + ISIZE trailer,
 * all checked by zlib (windowBits 31). */
int oracle_gzip_inflate(const uint8_t *in, uint64_t in_len, uint8_t *out, uint64_t out_cap,
                        uint64_t *out_len);
/* Raw deflate data (no header/trailer). */
int oracle_raw_inflate(const uint8_t *in, uint64_t in_len, uint8_t *out, uint64_t out_cap,
                       uint64_t *out_len);
uint32_t oracle_crc32(const uint8_t *p, uint64_t n);
uint32_t oracle_adler32(const uint8_t *p, uint64_t n);

/* Chunked Deflate container (zlib.h:1583-1700, zlib.cu:25-132): stream i is
 * comp[off[i] .. off[i+1]) (zero padding after the trailer is ignored by
 * inflate), output at out + i*chunk_bytes. Multithreaded over chunks. */
uint64_t oracle_inflate_chunks(const uint8_t *comp, const uint64_t *off, uint64_t n_chunks,
                               uint64_t chunk_bytes, uint8_t *out, int32_t *status,
                               int threads);

/* Uneven batch: stream i is src[src_off[i] .. src_off[i]+src_len[i]) and
 * inflates into dst[dst_off[i] .. dst_off[i]+dst_cap[i]); out_len[i] gets the
 * bytes produced. Multithreaded over streams. */
uint64_t oracle_inflate_batch(const uint8_t *src, const uint64_t *src_off,
                              const uint64_t *src_len, uint64_t n, uint8_t *dst,
                              const uint64_t *dst_off, const uint64_t *dst_cap,
                              uint64_t *out_len, int32_t *status, int threads);

/* ---- byte RLE and LZSS (byte_oracle.c) ----------------------------------
 * BRLE restates include/brle/brle.h:26-100, LZSS include/lzss/lzss.h:234-305.
 * *produced gets the bytes written; the return value is the chunk status. */
int oracle_brle_decode_chunk(const uint8_t *in, uint64_t len, uint8_t *out, uint64_t cap,
                             uint64_t *produced);
int oracle_lzss_decode_chunk(const uint8_t *in, uint64_t len, uint8_t *out, uint64_t cap,
                             uint64_t *produced);
/* Whole containers: BRLE chunk i is data[ptr[i] .. ptr[i+1]); LZSS chunk i is
 * data[ends[i-1] .. ends[i]) (ends[-1] = 0). Output chunk i at out + i*chunk. */
uint64_t oracle_brle_decode_chunks(const uint8_t *data, const uint64_t *ptr, uint64_t n_chunks,
                                   uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                                   int32_t *status);
uint64_t oracle_lzss_decode_chunks(const uint8_t *data, const uint64_t *ends, uint64_t n_chunks,
                                   uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                                   int32_t *status);
/* One chunk's stream as the reference encoders write it (brle.h:129-230;
 * lzss.h:75-113, 310-419); out needs n + n/128 + 2 (BRLE) or
 * 2n + n/8 + 2 (LZSS) bytes. Returns the stream length. */
uint64_t oracle_brle_encode_chunk(const uint8_t *in, uint64_t n, uint8_t *out);
uint64_t oracle_lzss_encode_chunk(const uint8_t *in, uint64_t n, uint8_t *out);

#ifdef __cplusplus
}
#endif
#endif
