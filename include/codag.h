/*
 * codag.h -- C ABI of the MI355X-native CODAG decompressors (libcodag.so).
 *
 * Plain C: pointers, sizes and status codes only. Device pointers ("d_*") are
 * HIP device allocations (hipMalloc, or codag_dev_alloc below); a
 * codag_stream_t is a hipStream_t passed as void* (NULL = the context's own
 * stream). Every call returns a codag_status and prints nothing; decode calls
 * are stream-ordered and asynchronous unless the name ends in _host.
 *
 * Reference interfaces replaced (paths relative to illinois-impact/codag):
 *   rle_v1::decompress_gpu_orig   include/rlev1/rlev1.h:419-424
 *   rle_v1::compress_gpu_orig     include/rlev1/rlev1.h:103-192
 *   rlev2::decompress_gpu_orig    include/rlev2/decoder_trans.h:57-60
 *   rlev2::compress_gpu_orig      include/rlev2/encoder_trans.h:482-564
 *   deflate::decompress_gpu       include/zlib/zlib.h:1585-1587
 *   deflate_compressor.py         zlib_chunk_compress, deflate_compressor.py:44-126
 *   brle::kernel_decompress       include/brle/brle.h:26-110 (and compress_gpu,
 *                                 decompress_gpu 251-342)
 *   lzss::kernel_decompress       include/lzss/lzss.h:234-305, 1014-1020 (and
 *                                 compress_gpu, decompress_gpu_base 1060-1189)
 *   cuda_err_chk / PRINT_ERROR    include/common.h:50-64 (print-and-continue
 *                                 becomes a returned status)
 * The batch descriptor mirrors the shape of cuDF's gpu_inflate_input_s /
 * gpu_inflate_status_s (include/cudf_deflate.cuh:236-250).
 */
#ifndef CODAG_H
#define CODAG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef enum codag_status {
  CODAG_OK = 0,
  CODAG_ERR_INVALID = 1,     /* bad argument (null pointer, unsupported width, ...) */
  CODAG_ERR_HIP = 2,         /* a HIP runtime call failed */
  CODAG_ERR_NOMEM = 3,       /* host or device allocation failed */
  CODAG_ERR_IO = 4,          /* file could not be opened / read / written */
  CODAG_ERR_FORMAT = 5,      /* container or side file is inconsistent */
  CODAG_ERR_DECODE = 6,      /* at least one chunk/stream has a non-OK chunk status */
  CODAG_ERR_BUSY = 7,        /* server admission: budget or slots exhausted */
  CODAG_ERR_UNSUPPORTED = 8, /* feature not available (e.g. no GPU) */
  CODAG_ERR_TIMEOUT = 9,     /* a wait exceeded its deadline */
} codag_status;

/* Per-chunk / per-stream codes written by the decode kernels. */
enum {
  CODAG_CHUNK_OK = 0,
  CODAG_CHUNK_TRUNCATED = 1, /* stream ended before the chunk/block was complete */
  CODAG_CHUNK_OVERFLOW = 2,  /* decoded data exceeds the output capacity */
  CODAG_CHUNK_CORRUPT = 3,   /* invalid field, code, distance, header or checksum */
};

const char *codag_version(void);
const char *codag_status_string(codag_status s);
/* Message of the last failing call on this host thread ("" if none). */
const char *codag_last_error(void);

/* ----------------------------------------------------------------- context */
/* A context binds one device, one HIP stream (default priority), the tuned
 * launch parameters and grow-only scratch memory. One context per host
 * thread, or external locking; different contexts are independent. */
typedef struct codag_ctx codag_ctx;
typedef void *codag_stream_t;

codag_status codag_ctx_create(int device, codag_ctx **out);
codag_status codag_ctx_destroy(codag_ctx *ctx);
codag_stream_t codag_ctx_stream(codag_ctx *ctx);
/* Number of visible devices (0 when no GPU / no driver). Never fails. */
int codag_device_count(void);

/* ------------------------------------------------------------------ RLE v1 */
/* Replaces rle_v1::decompress_gpu_orig (include/rlev1/rlev1.h:419-592).
 * d_comp   : chunk streams (the container minus its 8-byte length word,
 *            rlev1.h:521 passes d_in + 8)
 * d_blk_off: n_chunks byte offsets of the streams inside d_comp
 * d_col_len: n_chunks stream lengths in bytes
 * Each chunk decodes exactly chunk_bytes/elem_bytes values (decoder.h:671)
 * into d_out + i*chunk_bytes. elem_bytes in {1,2,4,8}.
 * d_chunk_status (nullable) receives one CODAG_CHUNK_* per chunk.
 * comp_bytes bounds every stream: a chunk whose blk_off+col_len exceeds it is
 * reported CORRUPT without being read (the reference trusts the side files). */
codag_status codag_rlev1_decode(codag_ctx *ctx, const uint8_t *d_comp, uint64_t comp_bytes,
                                const uint64_t *d_blk_off, const uint64_t *d_col_len,
                                uint64_t n_chunks, uint32_t elem_bytes,
                                uint64_t chunk_bytes, uint8_t *d_out, int32_t *d_chunk_status,
                                codag_stream_t stream);

/* ------------------------------------------------------------------ RLE v2 */
/* Replaces rlev2::decompress_gpu_orig (include/rlev2/decoder_trans.h:57-133).
 * Chunk i's stream is d_comp[d_blk_off[i] .. +d_col_len[i]) and decodes into
 * d_out + i*chunk_bytes; positions a chunk does not produce are zero-filled
 * (the reference memsets the output first, decoder_trans.h:81). */
codag_status codag_rlev2_decode(codag_ctx *ctx, const uint8_t *d_comp, uint64_t comp_bytes,
                                const uint64_t *d_blk_off, const uint64_t *d_col_len,
                                uint64_t n_chunks, uint32_t elem_bytes,
                                uint64_t chunk_bytes, uint8_t *d_out, int32_t *d_chunk_status,
                                codag_stream_t stream);

/* ----------------------------------------------------------------- Deflate */
enum {
  CODAG_INFLATE_SKIP_ADLER = 1u << 0,   /* do not compare the Adler-32 (gzip: CRC-32) trailer
                                           (the trailer is still required) */
  CODAG_INFLATE_RAW = 1u << 1,          /* streams are raw deflate (no zlib header/trailer) */
  /* streams are gzip members (RFC 1952), checked as zlib's inflate() checks
   * them: ID 1f 8b, CM 8, no reserved FLG bits, FEXTRA/FNAME/FCOMMENT skipped,
   * FHCRC verified, trailer CRC-32 and ISIZE verified. The result's adler32
   * field then holds the CRC-32 of the bytes written. The header rules of
   * cuDF's parse_gzip_header (cudf_deflate.cuh:1180-1227) are a subset. */
  CODAG_INFLATE_GZIP = 1u << 2,
  /* per-stream / per-chunk status codes in cuDF's convention
   * (gpu_inflate_status_s, cudf_deflate.cuh:1243-1343): 0 ok, 1 output buffer
   * too small, 2 input ended early, -1 invalid data, instead of CODAG_CHUNK_*. */
  CODAG_INFLATE_CUDF_STATUS = 1u << 3,
};

/* Replaces deflate::decompress_gpu (include/zlib/zlib.h:1585-1700).
 * d_offsets: n_chunks+1 byte offsets into d_comp (the blk_offset file minus
 * its leading length word, zlib.h:1603); stream i is [off[i], off[i+1]) and
 * inflates into d_out + i*chunk_bytes (at most chunk_bytes).
 * d_out_len (nullable): bytes produced per chunk (on an error, the bytes
 * stored before it; the count does not depend on the launch shape). The
 * zlib header is checked as zlib does (the reference skips it unchecked,
 * zlib.h:819). */
codag_status codag_inflate_chunks(codag_ctx *ctx, const uint8_t *d_comp, uint64_t comp_bytes,
                                  const uint64_t *d_offsets, uint64_t n_chunks, uint64_t chunk_bytes, uint8_t *d_out,
                                  uint64_t *d_out_len, int32_t *d_chunk_status, uint32_t flags,
                                  codag_stream_t stream);

/* Uneven batch of independent zlib streams (no reference counterpart: the
 * reference has fixed-size chunks only; shape follows cudf_deflate.cuh:236-250).
 * One persistent launch pulls streams from a device work queue in descriptor
 * order (listing the largest streams first gives the best load balance).
 * The work queue's counter belongs to the HIP stream the batch is launched
 * on, so batches of one context on different streams may overlap; a context
 * serves up to 16 streams at once (a 17th waits for the least recently used
 * stream's last batch). */
typedef struct codag_stream_desc {
  const uint8_t *src; /* device */
  uint64_t src_bytes;
  uint8_t *dst; /* device */
  uint64_t dst_cap;
} codag_stream_desc;

typedef struct codag_stream_result {
  uint64_t bytes_written;  /* on an error: bytes stored before it (any launch shape) */
  int32_t status;    /* CODAG_CHUNK_* (cuDF codes with CODAG_INFLATE_CUDF_STATUS) */
  uint32_t adler32;  /* Adler-32 of the bytes written (CRC-32 with CODAG_INFLATE_GZIP) */
} codag_stream_result;

codag_status codag_inflate_batch(codag_ctx *ctx, const codag_stream_desc *d_descs, uint64_t n,
                                 codag_stream_result *d_results, uint32_t flags,
                                 codag_stream_t stream);

/* Allocation-free variant of codag_inflate_batch. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It takes no context and runs on the current HIP device: the caller passes
 * inputs and outputs, the work-queue counter lives in a static 256-byte
 * device array of the library (one per GPU, held until the process ends), and
 * the call makes no hipMalloc/hipFree and builds no graph of its own (a
 * caller's stream capture records its two launches). The counter is reset by
 * a kernel on `stream` (NULL = the device's null stream) before each launch.
 * All calls on one GPU in a process share that counter and it keeps the last
 * call's contents: calls that overlap on one GPU (different streams or host
 * threads) race and must be serialized by the caller. Limits: n <= 2^32-2 streams per call
 * (larger n is rejected with CODAG_ERR_INVALID); groups_per_cu 0..64 resident
 * waves per CU (0 = 28). */
codag_status codag_inflate_batch_noalloc(const codag_stream_desc *d_descs, uint64_t n,
                                         codag_stream_result *d_results, uint32_t flags,
                                         uint32_t groups_per_cu, codag_stream_t stream);

/* ------------------------------------------------------------- BRLE, LZSS */
/* The reference's two byte-oriented codecs, which no CLI of the reference
 * calls (include/brle/brle.h, include/lzss/lzss.h).
 *
 * Byte RLE: replaces brle::kernel_decompress (brle.h:26-110), one thread per
 * chunk there. A header byte h > 127 is followed by h-127 literal bytes,
 * h <= 127 by one byte repeated h+3 times. d_ptr: n_chunks+1 byte offsets
 * into d_comp (the container's ptr array, brle.h:281-296); chunk i is
 * [ptr[i], ptr[i+1]) and decodes into d_out + i*chunk_bytes, at most
 * chunk_bytes (the reference's chunk is 262144 bytes, brle.h:8).
 * d_out_len / d_chunk_status (nullable): bytes produced and CODAG_CHUNK_*
 * per chunk (TRUNCATED: the stream ends inside an item; OVERFLOW: output
 * past chunk_bytes; CORRUPT: the chunk lies outside comp_bytes or is
 * >= 2^31 bytes). chunk_bytes < 2^31. */
codag_status codag_brle_decode(codag_ctx *ctx, const uint8_t *d_comp, uint64_t comp_bytes,
                               const uint64_t *d_ptr, uint64_t n_chunks, uint64_t chunk_bytes,
                               uint8_t *d_out, uint64_t *d_out_len, int32_t *d_chunk_status,
                               codag_stream_t stream);

/* LZSS with a 2048-byte history: replaces lzss::kernel_decompress
 * (lzss.h:234-305, 1014-1020), one thread per chunk there. A flag byte
 * covers the next 8 items, least significant bit first: 1 = literal byte,
 * 0 = 2-byte little-endian v, length (v & 31)+3, offset v >> 5.
 * d_ends: n_chunks cumulative stream ends (the container's lens array,
 * lzss.h:1111, 1142); chunk i is [ends[i-1] (0 for i = 0), ends[i]) and
 * decodes into d_out + i*chunk_bytes (131072 in the reference, lzss.h:21).
 * Chunk codes as for BRLE, plus CORRUPT for an offset of 0 or one that
 * reaches before the chunk's first byte. */
codag_status codag_lzss_decode(codag_ctx *ctx, const uint8_t *d_comp, uint64_t comp_bytes,
                               const uint64_t *d_ends, uint64_t n_chunks, uint64_t chunk_bytes,
                               uint8_t *d_out, uint64_t *d_out_len, int32_t *d_chunk_status,
                               codag_stream_t stream);

/* ------------------------------------------------- host convenience (sync) */
/* Used by the drop-in CLIs. Inputs are host buffers laid out exactly as the
 * reference's files; outputs are host buffers of at least the original
 * length. kernel_seconds (nullable) gets the decode kernel time measured with
 * HIP events; total_seconds (nullable) includes the host<->device copies.
 *
 * RLE v1: container = [u64 original_len][streams] (rlev1.h:163,175);
 * blk_off_file = rle_v1_blk_offset.bin (num_chunks*32+1 u64, rlev1.h:469);
 * col_len_file = rle_v1_col_len.bin (>= num_chunks u64). */
codag_status codag_rlev1_decode_host(codag_ctx *ctx, const uint8_t *container,
                                     uint64_t container_bytes, const uint64_t *blk_off_file,
                                     uint64_t blk_off_entries, const uint64_t *col_len_file,
                                     uint64_t col_len_entries, uint32_t elem_bytes,
                                     uint64_t chunk_bytes, uint8_t *out, uint64_t out_cap,
                                     uint64_t *out_bytes, double *kernel_seconds,
                                     double *total_seconds);

/* RLE v1, side files read from disk as rle_v1::decompress_gpu_orig does
 * (rlev1.h:436-467, normally ./input_data/rle_v1_col_len.bin and
 * ./input_data/rle_v1_blk_offset.bin). *out receives the original length of
 * decoded bytes in a malloc'd buffer (free() or codag_host_free);
 * meta_bytes (nullable) the two files' sizes, as rlev1.h:501 counts them. */
codag_status codag_rlev1_decode_files(codag_ctx *ctx, const uint8_t *container,
                                      uint64_t container_bytes, const char *col_len_path,
                                      const char *blk_off_path, uint32_t elem_bytes,
                                      uint64_t chunk_bytes, uint8_t **out, uint64_t *out_bytes,
                                      uint64_t *meta_bytes, double *kernel_seconds);

/* RLE v2: comp = concatenated chunk streams; blk_off has n_chunks+1 entries
 * with blk_off[n_chunks] = original byte length (encoder_trans.h:545);
 * col_len has >= n_chunks entries. */
codag_status codag_rlev2_decode_host(codag_ctx *ctx, const uint8_t *comp, uint64_t comp_bytes,
                                     const uint64_t *blk_off, const uint64_t *col_len,
                                     uint64_t n_chunks, uint32_t elem_bytes, uint64_t chunk_bytes,
                                     uint8_t *out, uint64_t out_cap, uint64_t *out_bytes,
                                     double *kernel_seconds, double *total_seconds);

/* Deflate: blk_offset_file = [u64 original_len][0][cumulative offsets...]
 * (n_chunks+2 entries, deflate_compressor.py:59-60,118-124; zlib.h:1589). */
codag_status codag_inflate_chunks_host(codag_ctx *ctx, const uint8_t *comp, uint64_t comp_bytes,
                                       const uint64_t *blk_offset_file, uint64_t blk_entries,
                                       uint64_t chunk_bytes, uint32_t flags, uint8_t *out,
                                       uint64_t out_cap, uint64_t *out_bytes,
                                       double *kernel_seconds, double *total_seconds);

/* BRLE container as brle::compress_gpu writes it (brle.h:281-296):
 * [u32 n_chunks+1][u64 ptr[n_chunks+1]][u64 original_len][chunk streams],
 * replacing brle::decompress_gpu (brle.h:305-342). chunk_bytes 0 means the
 * reference's 262144. Each chunk must produce exactly its share of
 * original_len (checked; the reference checks nothing). */
codag_status codag_brle_decode_host(codag_ctx *ctx, const uint8_t *container,
                                    uint64_t container_bytes, uint64_t chunk_bytes, uint8_t *out,
                                    uint64_t out_cap, uint64_t *out_bytes, double *kernel_seconds,
                                    double *total_seconds);

/* LZSS container as lzss::compress_gpu writes it (lzss.h:1100-1119):
 * [u64 original_len][u32 chunk_size, n_chunks, history 2048, min match 3,
 * max match 34][u64 ends[n_chunks]][chunk streams], replacing
 * lzss::decompress_gpu_base (lzss.h:1126-1189). Other history/match
 * parameters are rejected (CODAG_ERR_FORMAT): the reference's decoder has
 * them compiled in. Each chunk must produce exactly its share of
 * original_len. */
codag_status codag_lzss_decode_host(codag_ctx *ctx, const uint8_t *container,
                                    uint64_t container_bytes, uint8_t *out, uint64_t out_cap,
                                    uint64_t *out_bytes, double *kernel_seconds,
                                    double *total_seconds);

/* --------------------------------------------------- producers (CPU, C++) */
/* Buffers returned by producers are allocated by the library with malloc;
 * release them with codag_host_free (or free). */
void codag_host_free(void *p);

/* RLE v1 container writer (role of rle_v1::compress_gpu_orig, rlev1.h:103-192):
 * pads the input to a multiple of 32 chunks with zeros (rlev1.h:111-127),
 * encodes each chunk, pads each stream to 4 bytes (encoder.h:780-783).
 * Outputs: container [u64 in_bytes][streams]; blk_off (n_chunks*32+1 u64, the
 * first n_chunks+1 meaningful, the rest repeat the total); col_len
 * (n_chunks*32*32 u64, the first n_chunks meaningful, the rest zero) -- the
 * sizes the reference writes to ./input_data/ (rlev1.h:178-186).
 * The chunk streams are the reference encoder's bytes (compression_warp_orig,
 * encoder.h:1352-1695): same runs, literal groups and header bytes; the
 * padding bytes between streams and the side-file entries past the first
 * n_chunks(+1), which the reference leaves uninitialised, are zero / repeat
 * the total here. */
codag_status codag_rlev1_encode(const uint8_t *in, uint64_t in_bytes, uint32_t elem_bytes,
                                uint64_t chunk_bytes, uint8_t **container,
                                uint64_t *container_bytes, uint64_t **blk_off,
                                uint64_t *blk_off_entries, uint64_t **col_len,
                                uint64_t *col_len_entries, uint64_t *n_chunks);

/* RLE v2 container writer (role of rlev2::compress_gpu_orig, encoder_trans.h:
 * 482-564): n_chunks rounded up to a multiple of 32, streams aligned to 128 B,
 * blk_off has n_chunks+1 entries with the last = in_bytes, col_len n_chunks.
 * The chunk streams are the reference encoder's bytes (block_encode_orig,
 * encoder_trans.h:184-357, over encoder.h:83-591), with one exception: a run
 * of 128 zeros, for which the reference writes a DIRECT block its decoder
 * cannot read, is written as a fixed DELTA block. As in the reference,
 * 64-bit columns whose PATCHED_BASE patches exceed 32 bits do not round-trip
 * (the decoder keeps 32 bits of a patch, decoder_new.h:1105). */
codag_status codag_rlev2_encode(const uint8_t *in, uint64_t in_bytes, uint32_t elem_bytes,
                                uint64_t chunk_bytes, uint8_t **comp, uint64_t *comp_bytes,
                                uint64_t **blk_off, uint64_t **col_len, uint64_t *n_chunks);

/* Deflate container writer (restates zlib_chunk_compress,
 * deflate_compressor.py:44-126, with a selectable level): zero-pads the input
 * to whole chunks, compresses each chunk as a zlib stream, pads each stream
 * to 8 B. blk_offset = [in_bytes][0][cumulative...] (n_chunks+2 entries),
 * col_len = padded sizes (n_chunks entries). */
codag_status codag_deflate_compress_chunks(const uint8_t *in, uint64_t in_bytes,
                                           uint64_t chunk_bytes, int level, int threads,
                                           uint8_t **comp, uint64_t *comp_bytes,
                                           uint64_t **blk_offset, uint64_t **col_len,
                                           uint64_t *n_chunks);

/* BRLE container writer (role of brle::compress_gpu, brle.h:251-303). The
 * chunk streams are the reference encoder's bytes (brle::encode,
 * brle.h:129-230, on its direct path: brle.h:23 leaves USE_LOOKAHEAD
 * undefined): literal groups of up to 128 bytes, runs of 3-130. chunk_bytes
 * 0 means 262144. ptr[0] is written as 0 (the reference scans an entry its
 * kernel never writes, brle.h:244-248, 267). */
codag_status codag_brle_encode(const uint8_t *in, uint64_t in_bytes, uint64_t chunk_bytes,
                               int threads, uint8_t **container, uint64_t *container_bytes);

/* LZSS container writer (role of lzss::compress_gpu, lzss.h:1060-1124). The
 * chunk streams are the reference encoder's bytes (compress_func and
 * find_match, lzss.h:75-113, 310-419): at each position the longest match
 * of 3-34 bytes that starts 1-2047 bytes back, lies wholly before the
 * position and does not start at the oldest history byte; of equal lengths
 * the oldest. Chunks are chunk_bytes (0 means 131072) with a shorter last
 * chunk (the reference requires whole chunks, lzss.h:1065-1068). The flag
 * byte the reference reserves after a full group at a chunk's end, which it
 * leaves uninitialised, is 0 here. threads 0 = all cores. */
codag_status codag_lzss_encode(const uint8_t *in, uint64_t in_bytes, uint64_t chunk_bytes,
                               int threads, uint8_t **container, uint64_t *container_bytes);

/* ------------------------------------------ synthetic workloads (CPU, C++) */
/* Seeded generators of the benchmark columns/streams. Each returns the raw
 * data and the encoded form ready for decode. */

/* RLE v1 column: alternating runs (length min(130, 3+Geometric(mean 5)),
 * 30% with a nonzero delta in [-128,127]) and literal groups of 1-128 values,
 * values uniform in [0, value_range); segments never cross chunks. */
codag_status codag_synth_rlev1(uint64_t n_values, uint32_t elem_bytes, uint64_t chunk_bytes,
                               uint64_t value_range, uint64_t seed, uint8_t **raw,
                               uint8_t **container, uint64_t *container_bytes, uint64_t **blk_off,
                               uint64_t **col_len, uint64_t *n_chunks);

/* RLE v2 column: a 25/25/25/25 mix (by value count) of SHORT_REPEAT (runs
 * 3-10), DIRECT (widths of the decode map up to the column width), PATCHED_BASE
 * (5% outliers, gaps <= 255) and DELTA (monotone, deltas in [0,31], first >= 1);
 * blocks never cross chunks; streams aligned to 128 B. blk_off has
 * n_chunks+1 entries, the last = raw bytes. */
codag_status codag_synth_rlev2(uint64_t n_values, uint32_t elem_bytes, uint64_t chunk_bytes,
                               uint64_t seed, uint8_t **raw, uint8_t **comp,
                               uint64_t *comp_bytes, uint64_t **blk_off, uint64_t **col_len,
                               uint64_t *n_chunks);

/* Deflate payload generator: 50% text (order-2 Markov over 64 symbols with
 * repeats of length 3-258 at distances up to 32 KiB), 50% little-endian int32
 * drawn from Zipf(1.1), alternating per `part_bytes` region. */
codag_status codag_synth_deflate_payload(uint64_t n_bytes, uint64_t part_bytes, uint64_t seed,
                                         uint8_t **raw);

/* One zlib stream whose deflate blocks follow a fixed/dynamic/stored mix
 * (percent, by block count), switching with deflateParams at block
 * boundaries of about block_bytes input bytes. Returns the stream and the
 * achieved block counts. */
codag_status codag_deflate_compress_mixed(const uint8_t *in, uint64_t in_bytes,
                                          uint32_t pct_fixed, uint32_t pct_stored,
                                          uint64_t block_bytes, uint64_t seed, uint8_t **stream,
                                          uint64_t *stream_bytes, uint64_t counts[3]);

/* Bytes for the BRLE / LZSS benchmarks. kind 0 (BRLE): runs of 3 +
 * Geometric(mean 30) equal bytes, at most 130, alternating with literal
 * stretches of 1-64 bytes of an order-1 Markov text. kind 1 (LZSS): text of
 * words from a 512-word vocabulary (2-9 letters, skewed choice), spaces,
 * numbers and newlines. */
codag_status codag_synth_byte_payload(uint64_t n_bytes, uint32_t kind, uint64_t seed,
                                      uint8_t **raw);

/* ---------------------------------------------------------------- autotune */
typedef enum codag_codec { CODAG_RLEV1 = 1, CODAG_RLEV2 = 2, CODAG_DEFLATE = 3 } codag_codec;

/* Launch shape of a codec's decode kernels. For RLE v1 / v2, queue_depth
 * selects the kernel family: 0 the per-chunk kernels (RLE v2 also 64 threads
 * = one wave per chunk, window 4096 or 8192), 2 the walker + expander kernel
 * (window 4096; 128, 192 (RLE v2 only) or 256 threads = 1, 2 or 3 expander
 * waves), 4 the RLE v1 pipeline (512 threads, window 2048). For Deflate's
 * batch launch (codag_inflate_batch), 64 threads = one wave per stream, 128
 * = a decoding and a writing wave per stream. codag_tune_set rejects shapes
 * that are not compiled in. */
typedef struct codag_tune_params {
  uint32_t threads_per_group; /* workgroup size (64..1024, multiple of 64) */
  uint32_t window_bytes;      /* compressed bytes staged in LDS per window */
  uint32_t groups_per_cu;     /* Deflate persistent launches: resident workgroups per CU */
  uint32_t queue_depth;       /* RLE: kernel family (above); Deflate: 1 */
} codag_tune_params;

codag_status codag_tune_get(codag_ctx *ctx, codag_codec codec, codag_tune_params *out);
codag_status codag_tune_set(codag_ctx *ctx, codag_codec codec, const codag_tune_params *p);
/* Time each candidate on the given device-resident sample (same arguments as
 * the matching decode call; for DEFLATE d_blk_off are the n+1 offsets and
 * d_col_len is ignored) and keep the fastest whose output equals the first
 * candidate's; the shape in use stays if it is within 2% of that one (a tie
 * in run-to-run noise does not flip the choice). The kept shape becomes the
 * context's shape. best_ms (nullable) gets its time. DEFLATE candidates differ in
 * the batch launch's workgroup (one or two waves per stream) and workgroups
 * per CU; the launch has min(CUs x groups_per_cu, n_chunks) workgroups, so
 * the sample needs more than CUs x 28 streams for all of them to differ. */
codag_status codag_tune_run(codag_ctx *ctx, codag_codec codec, const uint8_t *d_comp,
                            uint64_t comp_bytes, const uint64_t *d_blk_off, const uint64_t *d_col_len,
                            uint64_t n_chunks, uint32_t elem_bytes, uint64_t chunk_bytes,
                            uint8_t *d_out, int reps, codag_tune_params *best, float *best_ms);

/* ---------------------------------------------------------------- serving */
/* Continuous-batching decode server (no reference counterpart). max_inflight
 * worker slots, each with its own context and HIP stream, admit queued
 * requests against an HBM byte budget: a free slot takes the first request in
 * arrival order whose device footprint fits the budget left. A request that
 * does not fit can be passed by later, smaller requests at most 8 times;
 * after that nothing passes it until it is admitted. Host-I/O requests are
 * copied into stream-ordered buffers of the server's own device memory pool
 * (no device-wide synchronisation once the pool holds the sizes), each
 * admitted request's decode runs on its slot's stream so work overlaps on
 * the device, and every completion reports its latency. */
typedef struct codag_server codag_server;

typedef struct codag_server_config {
  int device;
  uint64_t hbm_budget_bytes; /* admission budget for device-resident request data */
  uint32_t max_inflight;     /* concurrent requests on the device (HIP streams) */
} codag_server_config;

enum {
  /* comp/blk_off/col_len/out are device pointers: the request's data already
   * lives in HBM and its output stays there (no host copies). */
  CODAG_REQ_DEVICE_IO = 1u << 0,
};

typedef struct codag_request {
  codag_codec codec;
  uint32_t flags;       /* CODAG_REQ_* */
  uint32_t elem_bytes;  /* RLE only */
  uint64_t chunk_bytes; /* bytes of output per chunk */
  uint64_t n_chunks;
  const uint8_t *comp;  /* chunk streams (RLE v1: without the 8-byte word) */
  uint64_t comp_bytes;
  const uint64_t *blk_off; /* n_chunks entries (DEFLATE: n_chunks+1 offsets) */
  const uint64_t *col_len; /* n_chunks entries (DEFLATE: unused) */
  uint8_t *out;            /* receives n_chunks*chunk_bytes bytes */
  uint64_t user_tag;
  /* host pointers unless flags has CODAG_REQ_DEVICE_IO; for device-resident
   * requests the metadata is trusted (bounded by comp_bytes in the kernels) */
} codag_request;

typedef struct codag_completion {
  uint64_t request_id;
  uint64_t user_tag;
  codag_status status;
  uint32_t failed_chunks;
  double queue_seconds;   /* submit -> admitted */
  double latency_seconds; /* submit -> result in host memory */
} codag_completion;

codag_status codag_server_create(const codag_server_config *cfg, codag_server **out);
/* Non-blocking. CODAG_ERR_BUSY never happens here: requests wait in the
 * queue until admitted; a request larger than the whole budget is rejected
 * with CODAG_ERR_INVALID. */
codag_status codag_server_submit(codag_server *s, const codag_request *req, uint64_t *request_id);
/* Wait up to timeout_ms (-1 = forever) for at least one completion; returns up
 * to max completions. */
codag_status codag_server_poll(codag_server *s, codag_completion *out, uint32_t max,
                               uint32_t *n_out, int timeout_ms);
/* Device bytes currently admitted, requests in flight, requests queued. */
codag_status codag_server_stats(codag_server *s, uint64_t *admitted_bytes, uint32_t *inflight,
                                uint32_t *queued);
codag_status codag_server_destroy(codag_server *s);

/* ------------------------------------------------- device plumbing helpers */
/* Thin wrappers so hosts without their own HIP bindings (the ctypes tests,
 * bench.py) can move data; no reference counterpart. */
codag_status codag_dev_alloc(codag_ctx *ctx, uint64_t bytes, void **d_ptr);
codag_status codag_dev_free(codag_ctx *ctx, void *d_ptr);
codag_status codag_copy_h2d(codag_ctx *ctx, void *d_dst, const void *h_src, uint64_t bytes);
codag_status codag_copy_d2h(codag_ctx *ctx, void *h_dst, const void *d_src, uint64_t bytes);
codag_status codag_memset_d(codag_ctx *ctx, void *d_dst, int value, uint64_t bytes);
/* Device-to-device copy by a plain 16-byte-vector kernel on the context
 * stream (asynchronous); both pointers 16-byte aligned. Used as the
 * reference point when HBM traffic counters are scaled. */
codag_status codag_copy_d2d(codag_ctx *ctx, void *d_dst, const void *d_src, uint64_t bytes);
codag_status codag_sync(codag_ctx *ctx);
/* Extra HIP streams (default priority, non-blocking) on the context's device,
 * for callers without their own HIP bindings that want decodes to overlap.
 * Destroying a stream waits for its work. */
codag_status codag_stream_create(codag_ctx *ctx, codag_stream_t *out);
codag_status codag_stream_destroy(codag_ctx *ctx, codag_stream_t stream);
codag_status codag_stream_sync(codag_ctx *ctx, codag_stream_t stream);
/* HIP events recorded on the context stream around a region. */
codag_status codag_timer_start(codag_ctx *ctx);
codag_status codag_timer_stop(codag_ctx *ctx, float *elapsed_ms);
/* 256 numbered events per context for timing individual launches on a
 * stream (NULL = the context stream); elapsed waits for event b. */
codag_status codag_event_record(codag_ctx *ctx, uint32_t slot, codag_stream_t stream);
codag_status codag_event_elapsed(codag_ctx *ctx, uint32_t a, uint32_t b, float *elapsed_ms);
/* Device name / CU count of the context's device. */
codag_status codag_device_info(codag_ctx *ctx, char *name, uint32_t name_cap, uint32_t *cu_count);

#ifdef __cplusplus
}
#endif
#endif /* CODAG_H */
