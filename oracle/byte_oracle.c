/*
 * byte_oracle.c -- CPU restatement of the reference's byte-oriented
 * decoders, used ONLY as a checker (TEST INFRASTRUCTURE: tests/ and the
 * smoke check load it; the product never does).
 *
 *  BRLE  include/brle/brle.h:26-100 (brle::decode; USE_LOOKAHEAD is not
 *        defined -- brle.h:23 defines USE_LOOKAHEADd -- so the direct path
 *        at brle.h:76-90 is the live one): a header byte h > 127 is followed
 *        by h - 127 literal bytes, h <= 127 by one byte repeated h + 3 times.
 *        Container (brle.h:276-296, 305-318): [u32 n+1][u64 ptr[n+1]][u64 length]
 *        [chunk streams]; chunk i is data[ptr[i] .. ptr[i+1]) and decodes
 *        into out + i * 256 KiB.
 *  LZSS  include/lzss/lzss.h:234-305 (lzss::decompress_func): a flag byte
 *        covers the next 8 items, least significant bit first; bit 1 is a
 *        literal byte, bit 0 a 2-byte little-endian reference v with
 *        length (v & 31) + 3 and offset v >> 5 (OFFSET_SIZE = 11 bits for the
 *        2048-byte history, lzss.h:24-32), copied byte by byte from offset
 *        bytes back. Container (lzss.h:1072-1115): [u64 length][u32 chunk,
 *        n_chunks, history, min match, max match][u64 cumulative chunk
 *        ends x n][chunk streams]; chunk i is data[end[i-1] .. end[i]) and
 *        decodes into out + i * chunk.
 *
 * Encoders (for pinning the product's producers): brle_encode_chunk
 * restates brle::encode (brle.h:129-230) on its direct path, statement by
 * statement; lzss_encode_chunk restates compress_func and find_match
 * (lzss.h:75-113, 310-419) with the reference's 2048-byte history ring and
 * 512-byte lookahead ring and its exhaustive oldest-first match scan, so it
 * is slow (O(2047) per position) and meant for test-sized inputs.
 *
 * The reference checks nothing; these add the project's status rules: a
 * stream that ends inside an item is TRUNCATED, output past the chunk is
 * OVERFLOW, an LZSS reference to offset 0 or before the chunk start is
 * CORRUPT. The first error in stream order wins.
 */
#include "oracle.h"

#include <string.h>

int oracle_brle_decode_chunk(const uint8_t *in, uint64_t len, uint8_t *out, uint64_t cap,
                             uint64_t *produced) {
  uint64_t used = 0, o = 0;
  int st = ORACLE_OK;
  while (used < len) {
    const uint8_t h = in[used++];
    if (h > 127) { /* brle.h:79-82 */
      const uint64_t k = (uint64_t)h - 127;
      if (used + k > len) { st = ORACLE_TRUNCATED; break; }
      if (o + k > cap) { st = ORACLE_OVERFLOW; break; }
      memcpy(out + o, in + used, k);
      used += k;
      o += k;
    } else { /* brle.h:83-86 */
      const uint64_t k = (uint64_t)h + 3;
      if (used >= len) { st = ORACLE_TRUNCATED; break; }
      if (o + k > cap) { st = ORACLE_OVERFLOW; break; }
      memset(out + o, in[used++], k);
      o += k;
    }
  }
  if (produced) *produced = o;
  return st;
}

int oracle_lzss_decode_chunk(const uint8_t *in, uint64_t len, uint8_t *out, uint64_t cap,
                             uint64_t *produced) {
  uint64_t used = 0, o = 0;
  int st = ORACLE_OK;
  while (used < len && st == ORACLE_OK) {
    uint32_t flags = in[used++]; /* lzss.h:264-267 */
    for (int i = 0; i < 8 && used < len; i++, flags >>= 1) {
      if (flags & 1) { /* literal, lzss.h:297-303 */
        if (o + 1 > cap) { st = ORACLE_OVERFLOW; break; }
        out[o++] = in[used++];
        continue;
      }
      /* reference, lzss.h:272-295 */
      if (used + 2 > len) { st = ORACLE_TRUNCATED; break; }
      const uint32_t v = (uint32_t)in[used] | ((uint32_t)in[used + 1] << 8);
      used += 2;
      const uint32_t length = (v & 31u) + 3u, offset = v >> 5;
      if (offset == 0 || offset > o) { st = ORACLE_CORRUPT; break; }
      if (o + length > cap) { st = ORACLE_OVERFLOW; break; }
      for (uint32_t j = 0; j < length; j++, o++) out[o] = out[o - offset];
    }
  }
  if (produced) *produced = o;
  return st;
}

uint64_t oracle_brle_decode_chunks(const uint8_t *data, const uint64_t *ptr, uint64_t n_chunks,
                                   uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                                   int32_t *status) {
  uint64_t bad = 0;
  for (uint64_t i = 0; i < n_chunks; i++) {
    uint64_t got = 0;
    int st = oracle_brle_decode_chunk(data + ptr[i], ptr[i + 1] - ptr[i], out + i * chunk_bytes,
                                      chunk_bytes, &got);
    if (out_len) out_len[i] = got;
    if (status) status[i] = st;
    bad += st != ORACLE_OK;
  }
  return bad;
}

uint64_t oracle_lzss_decode_chunks(const uint8_t *data, const uint64_t *ends, uint64_t n_chunks,
                                   uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                                   int32_t *status) {
  uint64_t bad = 0;
  for (uint64_t i = 0; i < n_chunks; i++) {
    const uint64_t lo = i ? ends[i - 1] : 0; /* lzss.h:237-238 */
    uint64_t got = 0;
    int st = oracle_lzss_decode_chunk(data + lo, ends[i] - lo, out + i * chunk_bytes, chunk_bytes,
                                      &got);
    if (out_len) out_len[i] = got;
    if (status) status[i] = st;
    bad += st != ORACLE_OK;
  }
  return bad;
}

/* ---------------------------------------------------------------- encoders */

uint64_t oracle_brle_encode_chunk(const uint8_t *in, uint64_t n, uint8_t *out) {
  /* brle.h:145-150: the encoder's state */
  uint32_t pos = 0, num_literals = 0, tail_run = 0;
  int repeat = 0;
  uint64_t used_bytes = 0;
  uint8_t literals[128];
#define WRITE_VALUES_O                                         \
  if (num_literals != 0) {                                     \
    if (repeat) {                                              \
      out[pos++] = (uint8_t)(num_literals - 3);                \
      out[pos++] = literals[0];                                \
    } else {                                                   \
      out[pos++] = (uint8_t)(num_literals + 127);              \
      memcpy(out + pos, literals, num_literals);               \
      pos += num_literals;                                     \
    }                                                          \
    repeat = 0;                                                \
    tail_run = 0;                                              \
    num_literals = 0;                                          \
  }
  while (used_bytes < n) { /* brle.h:159-224, USE_LOOKAHEAD off */
    uint8_t c = in[used_bytes];
    used_bytes++;
    if (num_literals == 0) {
      literals[num_literals++] = c;
      tail_run = 1;
    } else if (repeat) {
      if (c == literals[0]) {
        num_literals++;
        if (num_literals == 130) { WRITE_VALUES_O; }
      } else {
        WRITE_VALUES_O;
        literals[num_literals++] = c;
        tail_run = 1;
      }
    } else {
      if (c == literals[num_literals - 1]) tail_run++;
      else tail_run = 1;
      if (tail_run == 3) {
        if (num_literals + 1 == 3) {
          num_literals += 1;
          repeat = 1;
        } else {
          num_literals -= 3 - 1;
          WRITE_VALUES_O;
          literals[0] = c;
          num_literals = 3;
          repeat = 1;
        }
      } else {
        literals[num_literals++] = c;
        if (num_literals == 128) { WRITE_VALUES_O; }
      }
    }
  }
  WRITE_VALUES_O;
#undef WRITE_VALUES_O
  return pos;
}

enum { HIST = 2048, LOOK = 512, MIN_MATCH = 3, MAX_MATCH = 34 };

/* find_match, lzss.h:75-113 */
static void find_match(const uint8_t *hist, uint32_t hist_head, uint32_t hist_count,
                       const uint8_t *la, uint32_t la_head, uint32_t la_count, uint32_t *offset,
                       uint32_t *length) {
  uint32_t hist_offset = 1, f_length = 0, max_len = 0, max_offset = 0;
  while (hist_offset < hist_count) {
    if (hist[(hist_head + hist_offset) % HIST] == la[la_head]) {
      f_length = 1;
      while ((hist_offset + f_length) < hist_count && f_length < la_count &&
             hist[(hist_head + hist_offset + f_length) % HIST] == la[(la_head + f_length) % LOOK]) {
        f_length++;
        if (f_length >= MAX_MATCH) break;
      }
      if (f_length > max_len) {
        max_len = f_length;
        max_offset = hist_offset;
      }
    }
    if (f_length >= MAX_MATCH) {
      f_length = MAX_MATCH;
      max_len = f_length;
      break;
    }
    hist_offset++;
  }
  if (max_len < MIN_MATCH) max_len = 0;
  *length = max_len;
  *offset = hist_count - max_offset;
}

uint64_t oracle_lzss_encode_chunk(const uint8_t *in, uint64_t n, uint8_t *out) {
  /* compress_func, lzss.h:310-419, for one chunk of n bytes */
  static __thread uint8_t hist[HIST];
  uint8_t la[LOOK];
  uint32_t hist_head = 0, hist_count = 0, la_head = 0, la_count = 0, blocks = 0;
  uint64_t consumed = 0, out_bytes = 1, cur_header = 0, used = 0;
  uint8_t header = 0;
  memset(hist, ' ', sizeof hist);
  out[0] = 0; /* written at the first full group or the end; left as is for n = 0 */
  while (used < n) {
    while (la_count < LOOK && consumed < n) la[(la_head + (la_count++)) % LOOK] = in[consumed++];
    uint32_t offset = 0, length = 0;
    find_match(hist, hist_head, hist_count, la, la_head, la_count, &offset, &length);
    if (length == 0) {
      uint8_t v = la[la_head];
      out[out_bytes++] = v;
      la_head = (la_head + 1) % LOOK;
      la_count--;
      hist[(hist_head + hist_count) % HIST] = v;
      hist_count += 1;
      if (hist_count > HIST) {
        hist_head = (hist_head + 1) % HIST;
        hist_count = HIST;
      }
      header = (uint8_t)(header | (1u << blocks));
      used++;
    } else {
      uint64_t v = ((uint64_t)offset << 5) | (length - MIN_MATCH);
      for (uint32_t i = 0; i < MIN_MATCH - 1; i++) {
        out[out_bytes++] = (uint8_t)(v & 0xff);
        v >>= 8;
      }
      uint32_t hist_start = hist_head + hist_count;
      for (uint32_t i = 0; i < length; i++) hist[(hist_start + i) % HIST] = la[(la_head + i) % LOOK];
      la_head = (la_head + length) % LOOK;
      la_count -= length;
      hist_count += length;
      if (hist_count > HIST) {
        hist_head = (hist_head + (hist_count - HIST)) % HIST;
        hist_count = HIST;
      }
      used += length;
    }
    if ((++blocks) == 8) {
      out[cur_header] = header;
      header = 0;
      blocks = 0;
      cur_header = out_bytes++;
      out[cur_header] = 0; /* the reference leaves this byte uninitialised */
    }
  }
  if (blocks != 0) out[cur_header] = header;
  return out_bytes;
}
