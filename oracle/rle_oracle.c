/*
 * rle_oracle.c -- CPU restatement of the reference RLE v1 / RLE v2 decoders.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker. The product never calls it.
 *
 * Shift semantics. The reference computes several shifts whose count can
 * reach or exceed the operand width (e.g. a 5-byte SHORT_REPEAT value into a
 * uint32, decoder_new.h:1011). On the reference's hardware a 32/64-bit shift
 * by >= the register width yields 0 (PTX shl/shr clamp the count), so bits
 * shifted past the value width are dropped. This file spells that out with
 * shl_n()/shr64() so the result is defined C.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static inline uint64_t width_mask(uint32_t width) {
  return width >= 8 ? ~(uint64_t)0 : (((uint64_t)1 << (8 * width)) - 1);
}

/* x << s with counts >= 64 giving 0. */
static inline uint64_t shl64(uint64_t x, uint32_t s) { return s >= 64 ? 0 : x << s; }
static inline uint64_t shr64(uint64_t x, uint32_t s) { return s >= 64 ? 0 : x >> s; }

static inline void store_value(uint8_t *out, uint64_t idx, uint32_t width, uint64_t v) {
  uint8_t *p = out + idx * width;
  switch (width) {
  case 1: p[0] = (uint8_t)v; break;
  case 2: { uint16_t t = (uint16_t)v; memcpy(p, &t, 2); } break;
  case 4: { uint32_t t = (uint32_t)v; memcpy(p, &t, 4); } break;
  default: memcpy(p, &v, 8); break;
  }
}

static inline uint64_t load_value(const uint8_t *out, uint64_t idx, uint32_t width) {
  const uint8_t *p = out + idx * width;
  switch (width) {
  case 1: return p[0];
  case 2: { uint16_t t; memcpy(&t, p, 2); return t; }
  case 4: { uint32_t t; memcpy(&t, p, 4); return t; }
  default: { uint64_t t; memcpy(&t, p, 8); return t; }
  }
}

/* ------------------------------------------------------------------------ */
/* RLE v1                                                                    */
/* ------------------------------------------------------------------------ */

/* Unsigned LEB128 as read by decoder.h:686-727 / 752-766: 7 payload bits per
 * byte, least significant group first, a byte < 0x80 terminates, no zigzag.
 * Payload bits at shift >= 64 are dropped; the value is later reduced to the
 * column width, which is what the reference's DATA_TYPE accumulator keeps. */
static int v1_varint(const uint8_t *in, uint64_t in_len, uint64_t *pos, uint64_t *val) {
  uint64_t v = 0;
  uint32_t shift = 0, nbytes = 0;
  for (;;) {
    if (*pos >= in_len) return ORACLE_TRUNCATED;
    uint8_t b = in[(*pos)++];
    v |= shl64((uint64_t)(b & 0x7f), shift);
    shift += 7;
    nbytes++;
    if (b < 0x80) break;
  }
  /* The reference keeps shifting past 64 bits on malformed input; this
   * project rejects varints longer than the 10 bytes a uint64 needs. */
  if (nbytes > 10) return ORACLE_CORRUPT;
  *val = v;
  return ORACLE_OK;
}

int oracle_rlev1_decode_chunk(const uint8_t *in, uint64_t in_len, uint32_t width,
                              uint8_t *out, uint64_t chunk_bytes, uint64_t *produced) {
  const uint64_t mask = width_mask(width);
  const uint64_t cap = chunk_bytes / width;
  uint64_t pos = 0, n = 0;
  int rc = ORACLE_OK;
  /* decoder.h:671 -- loop until CHUNK_SIZE output bytes exist. */
  while (n * width < chunk_bytes) {
    if (pos >= in_len) { rc = ORACLE_TRUNCATED; break; }
    int8_t h = (int8_t)in[pos++]; /* decoder.h:675-684 */
    if (h < 0) {
      /* literal group of -h varints (decoder.h:686-733) */
      uint32_t k = (uint32_t)(-(int32_t)h);
      for (uint32_t i = 0; i < k; i++) {
        uint64_t v;
        rc = v1_varint(in, in_len, &pos, &v);
        if (rc) goto done;
        if (n >= cap) { rc = ORACLE_OVERFLOW; goto done; }
        store_value(out, n++, width, v & mask);
      }
    } else {
      /* run of h+3 values: signed delta byte then varint base
       * (decoder.h:736-790, expansion at 643-652 in int64). */
      uint32_t len = (uint32_t)h + 3;
      if (pos >= in_len) { rc = ORACLE_TRUNCATED; break; }
      int64_t delta = (int8_t)in[pos++];
      uint64_t base;
      rc = v1_varint(in, in_len, &pos, &base);
      if (rc) break;
      for (uint32_t i = 0; i < len; i++) {
        if (n >= cap) { rc = ORACLE_OVERFLOW; goto done; }
        uint64_t v = base + (uint64_t)((int64_t)i * delta);
        store_value(out, n++, width, v & mask);
      }
    }
  }
done:
  if (produced) *produced = n * width;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* RLE v2                                                                    */
/* ------------------------------------------------------------------------ */

/* utils.h:28 __BIT_WIDTH_DECODE_MAP */
static const uint8_t k_decode_width[32] = {1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11,
                                           12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22,
                                           23, 24, 26, 28, 30, 32, 40, 48, 56, 64};

/* utils.h:42-48 __CLOSEST_FIXED_BIT_MAP (65 entries) */
static uint8_t closest_fixed(uint32_t bits) {
  if (bits <= 1) return 1;
  if (bits <= 24) return (uint8_t)bits;
  if (bits <= 26) return 26;
  if (bits <= 28) return 28;
  if (bits <= 30) return 30;
  if (bits <= 32) return 32;
  if (bits <= 40) return 40;
  if (bits <= 48) return 48;
  if (bits <= 56) return 56;
  return 64;
}

uint32_t oracle_rlev2_decode_width(uint32_t code) { return code < 32 ? k_decode_width[code] : 0; }
uint32_t oracle_rlev2_closest_width(uint32_t bits) { return bits <= 64 ? closest_fixed(bits) : 0; }

typedef struct {
  const uint8_t *in;
  uint64_t len;
  uint64_t used; /* decoder_new.h:963 used_bytes */
  int err;
} v2_reader;

/* decoder_new.h:966-977 read_byte */
static inline uint8_t v2_byte(v2_reader *r) {
  if (r->used >= r->len) {
    r->err = ORACLE_TRUNCATED;
    return 0;
  }
  return r->in[r->used++];
}

/* MSB-first bit unpacking used by every v2 sub-encoding (decoder_new.h:1022-1040):
 * a fresh reader starts at a byte boundary; leftover bits of the last byte
 * are discarded when the block ends. */
typedef struct {
  uint32_t bits_left;
  uint32_t cur;
} v2_bits;

static uint64_t v2_read_bits(v2_reader *r, v2_bits *b, uint32_t n) {
  uint64_t result = 0;
  while (n > b->bits_left) {
    result = shl64(result, b->bits_left) | (b->cur & ((1u << b->bits_left) - 1));
    n -= b->bits_left;
    b->cur = v2_byte(r);
    b->bits_left = 8;
  }
  if (n > 0) {
    result = shl64(result, n);
    b->bits_left -= n;
    result |= (b->cur >> b->bits_left) & ((1u << n) - 1);
  }
  return result;
}

/* decoder_new.h:979-991 read_uvarint: 7-bit groups, LSB first, into uint64. */
static uint64_t v2_uvarint(v2_reader *r) {
  uint64_t v = 0;
  uint32_t off = 0, n = 0;
  uint8_t b;
  do {
    b = v2_byte(r);
    v |= shl64((uint64_t)(b & 0x7f), off);
    off += 7;
    n++;
  } while (b >= 0x80 && !r->err);
  /* The reference keeps reading continuation bytes on malformed input; this
   * project rejects varints longer than the 10 bytes a 64-bit value needs. */
  if (n > 10 && !r->err) r->err = ORACLE_CORRUPT;
  return v;
}

/* decoder_new.h:993-997 read_svarint: (ret >> 1) ^ -(ret & 1) with ret an
 * int64_t, so the right shift is arithmetic. */
static uint64_t v2_svarint(v2_reader *r) {
  int64_t ret = (int64_t)v2_uvarint(r);
  int64_t half = ret >> 1; /* arithmetic on every target we build for */
  return (uint64_t)(half ^ -(ret & 1));
}

/* Big-endian multi-byte value (decoder_new.h:1009-1011, 1109-1112); bytes
 * shifted past 64 bits drop out. */
static uint64_t v2_be(v2_reader *r, uint32_t nbytes) {
  uint64_t v = 0;
  while (nbytes-- > 0) v |= shl64((uint64_t)v2_byte(r), nbytes * 8);
  return v;
}

int oracle_rlev2_decode_chunk(const uint8_t *in, uint64_t col_len, uint32_t width,
                              uint8_t *out, uint64_t cap_values, uint64_t *produced) {
  const uint64_t mask = width_mask(width);
  const uint32_t nbits = 8 * width;
  v2_reader r = {in, col_len, 0, 0};
  uint64_t op = 0; /* out_ptr */
  int rc = ORACLE_OK;

#define EMIT(val)                                                                  \
  do {                                                                             \
    if (op >= cap_values) { rc = ORACLE_OVERFLOW; goto done; }                     \
    store_value(out, op++, width, (uint64_t)(val) & mask);                         \
  } while (0)

  while (r.used < col_len) { /* decoder_new.h:1000 */
    uint8_t first = v2_byte(&r);
    switch (first & 0xc0) {
    case 0x00: { /* SHORT_REPEAT, decoder_new.h:1003-1015 */
      uint32_t nbytes = ((first >> 3) & 7) + 1;
      uint32_t count = (first & 7) + 3;
      uint64_t v = v2_be(&r, nbytes);
      if (r.err) { rc = r.err; goto done; }
      while (count-- > 0) EMIT(v);
    } break;
    case 0x40: { /* DIRECT, decoder_new.h:1016-1044 */
      uint32_t fbw = k_decode_width[(first >> 1) & 0x1f];
      /* `len` is a uint8_t in the reference (decoder_new.h:1019): 256 -> 0. */
      uint8_t len = (uint8_t)((((uint32_t)(first & 1) << 8) | v2_byte(&r)) + 1);
      v2_bits b = {0, 0};
      while (len-- > 0) {
        uint64_t v = v2_read_bits(&r, &b, fbw);
        if (r.err) { rc = r.err; goto done; }
        EMIT(v);
      }
      if (r.err) { rc = r.err; goto done; }
    } break;
    case 0xc0: { /* DELTA, decoder_new.h:1045-1093 */
      uint32_t enc = (first >> 1) & 0x1f;
      uint32_t fbw = k_decode_width[enc];
      uint8_t len = (uint8_t)((((uint32_t)(first & 1) << 8) | v2_byte(&r)) + 1);
      uint64_t base = v2_uvarint(&r) & mask;
      uint64_t base_delta = v2_svarint(&r) & mask;
      if (r.err) { rc = r.err; goto done; }
      EMIT(base);
      base = (base + base_delta) & mask;
      EMIT(base);
      len = (uint8_t)(len - 2);
      /* multiplier = (DATA_TYPE)base_delta > 0 ? 1 : -1 (decoder_new.h:1057):
       * with an unsigned DATA_TYPE only a zero base_delta subtracts. */
      int neg = (base_delta == 0);
      if (enc != 0) {
        v2_bits b = {0, 0};
        while (len-- > 0) {
          uint64_t d = v2_read_bits(&r, &b, fbw) & mask;
          if (r.err) { rc = r.err; goto done; }
          base = (neg ? base - d : base + d) & mask;
          EMIT(base);
        }
      } else {
        while (len-- > 0) {
          base = (base + base_delta) & mask;
          EMIT(base);
        }
      }
    } break;
    default: { /* 0x80 PATCHED_BASE, decoder_new.h:1094-1183 */
      uint32_t fbw = k_decode_width[(first >> 1) & 0x1f];
      uint8_t len = (uint8_t)((((uint32_t)(first & 1) << 8) | v2_byte(&r)) + 1);
      uint8_t third = v2_byte(&r);
      uint8_t fourth = v2_byte(&r);
      uint32_t bw = ((third >> 5) & 7) + 1;
      uint32_t pw = k_decode_width[third & 0x1f];
      uint32_t pgw = ((fourth >> 5) & 7) + 1;
      uint32_t pll = fourth & 0x1f;
      /* patch_mask is a uint32_t (decoder_new.h:1105); a 32-bit shift by
       * >= 32 gives 0 on the reference's hardware, so the mask saturates. */
      uint64_t patch_mask = pw >= 32 ? 0xffffffffull : ((1ull << pw) - 1);
      uint64_t block_start = op;
      uint64_t base = v2_be(&r, bw) & mask; /* no sign handling (1109-1112) */
      if (r.err) { rc = r.err; goto done; }
      uint32_t n = len;
      v2_bits b = {0, 0};
      while (n-- > 0) {
        uint64_t v = v2_read_bits(&r, &b, fbw);
        if (r.err) { rc = r.err; goto done; }
        EMIT(v + base);
      }
      /* Patch list (decoder_new.h:1134-1181): byte aligned, pll entries of
       * closest_bit(pw+pgw) bits; the gap field accumulates. */
      if (pw + pgw > 64) { rc = ORACLE_CORRUPT; goto done; } /* past the 65-entry table */
      uint32_t cfb = closest_fixed(pw + pgw);
      b.bits_left = 0;
      b.cur = 0;
      uint64_t gap = 0;
      while (pll-- > 0) {
        uint64_t e = v2_read_bits(&r, &b, cfb);
        if (r.err) { rc = r.err; goto done; }
        gap += shr64(e, pw);
        uint64_t at = block_start + gap;
        /* The reference would patch whatever sits there; a patch outside its
         * own block is never produced by the encoder and is rejected here. */
        if (at >= block_start + len) { rc = ORACLE_CORRUPT; goto done; }
        uint64_t cur = load_value(out, at, width);
        uint64_t add = fbw >= nbits ? 0 : ((e & patch_mask) << fbw);
        cur = (((cur - base) & mask) | (add & mask));
        cur = (cur + base) & mask;
        store_value(out, at, width, cur);
      }
    } break;
    }
    if (r.err) { rc = r.err; goto done; }
  }
#undef EMIT
done:
  if (produced) *produced = op;
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Multithreaded whole-container drivers (CPU baseline and test helpers)     */
/* ------------------------------------------------------------------------ */

typedef struct {
  int codec; /* 1 = v1, 2 = v2 */
  const uint8_t *comp;
  const uint64_t *blk_off, *col_len;
  uint64_t n_chunks;
  uint32_t width;
  uint64_t chunk_bytes;
  uint8_t *out;
  int32_t *status;
  uint64_t next; /* work counter */
  uint64_t failures;
  pthread_mutex_t mu;
} rle_job;

static void *rle_worker(void *arg) {
  rle_job *j = (rle_job *)arg;
  uint64_t fails = 0;
  for (;;) {
    uint64_t i = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
    if (i >= j->n_chunks) break;
    uint8_t *o = j->out + i * j->chunk_bytes;
    uint64_t produced = 0;
    int rc;
    if (j->codec == 1)
      rc = oracle_rlev1_decode_chunk(j->comp + j->blk_off[i], j->col_len[i], j->width, o,
                                     j->chunk_bytes, &produced);
    else
      rc = oracle_rlev2_decode_chunk(j->comp + j->blk_off[i], j->col_len[i], j->width, o,
                                     j->chunk_bytes / j->width, &produced);
    if (j->status) j->status[i] = rc;
    if (rc) fails++;
  }
  pthread_mutex_lock(&j->mu);
  j->failures += fails;
  pthread_mutex_unlock(&j->mu);
  return NULL;
}

static uint64_t run_rle(rle_job *j, int threads) {
  if (threads < 1) threads = 1;
  pthread_mutex_init(&j->mu, NULL);
  pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 1; t < threads; t++) pthread_create(&tid[t], NULL, rle_worker, j);
  rle_worker(j);
  for (int t = 1; t < threads; t++) pthread_join(tid[t], NULL);
  free(tid);
  pthread_mutex_destroy(&j->mu);
  return j->failures;
}

uint64_t oracle_rlev1_decode_chunks(const uint8_t *comp, const uint64_t *blk_off,
                                    const uint64_t *col_len, uint64_t n_chunks,
                                    uint32_t width, uint64_t chunk_bytes, uint8_t *out,
                                    int32_t *status, int threads) {
  rle_job j;
  memset(&j, 0, sizeof j);
  j.codec = 1;
  j.comp = comp; j.blk_off = blk_off; j.col_len = col_len; j.n_chunks = n_chunks;
  j.width = width; j.chunk_bytes = chunk_bytes; j.out = out; j.status = status;
  return run_rle(&j, threads);
}

uint64_t oracle_rlev2_decode_chunks(const uint8_t *comp, const uint64_t *blk_off,
                                    const uint64_t *col_len, uint64_t n_chunks,
                                    uint32_t width, uint64_t chunk_bytes, uint8_t *out,
                                    int32_t *status, int threads) {
  memset(out, 0, n_chunks * chunk_bytes); /* decoder_trans.h:81 */
  rle_job j;
  memset(&j, 0, sizeof j);
  j.codec = 2;
  j.comp = comp; j.blk_off = blk_off; j.col_len = col_len; j.n_chunks = n_chunks;
  j.width = width; j.chunk_bytes = chunk_bytes; j.out = out; j.status = status;
  return run_rle(&j, threads);
}
