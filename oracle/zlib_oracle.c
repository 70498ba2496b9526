/*
 * zlib_oracle.c -- Deflate/zlib checker built on the system zlib.
 *
 * TEST INFRASTRUCTURE ONLY (tests/, __graft_entry__.smoke(), bench.py's
 * cpu_baseline leg). The reference's Deflate kernel (include/zlib/zlib.h:
 * 807-866) implements RFC 1951 inflate after skipping the 2-byte RFC 1950
 * header (zlib.h:819); zlib's inflate() is the arithmetic it must match.
 * The container follows deflate_compressor.py:44-126 and is read as in
 * include/zlib/zlib.h:1586-1603 (offset array after the length word).
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

const char *oracle_zlib_version(void) { return zlibVersion(); }

static int do_inflate(const uint8_t *in, uint64_t in_len, uint8_t *out, uint64_t out_cap,
                      uint64_t *out_len, int window_bits) {
  z_stream s;
  memset(&s, 0, sizeof s);
  if (inflateInit2(&s, window_bits) != Z_OK) return ORACLE_CORRUPT;
  s.next_in = (Bytef *)in;
  s.avail_in = (uInt)in_len;
  s.next_out = out;
  s.avail_out = (uInt)out_cap;
  int zr = inflate(&s, Z_FINISH);
  if (out_len) *out_len = s.total_out;
  inflateEnd(&s);
  if (zr == Z_STREAM_END) return ORACLE_OK;
  if (zr == Z_BUF_ERROR && s.avail_out == 0) return ORACLE_OVERFLOW;
  if (zr == Z_BUF_ERROR) return ORACLE_TRUNCATED;
  return ORACLE_CORRUPT;
}

int oracle_zlib_inflate(const uint8_t *in, uint64_t in_len, uint8_t *out, uint64_t out_cap,
                        uint64_t *out_len) {
  return do_inflate(in, in_len, out, out_cap, out_len, 15);
}

int oracle_gzip_inflate(const uint8_t *in, uint64_t in_len, uint8_t *out, uint64_t out_cap,
                        uint64_t *out_len) {
  return do_inflate(in, in_len, out, out_cap, out_len, 31);
}

int oracle_raw_inflate(const uint8_t *in, uint64_t in_len, uint8_t *out, uint64_t out_cap,
                       uint64_t *out_len) {
  return do_inflate(in, in_len, out, out_cap, out_len, -15);
}

uint32_t oracle_crc32(const uint8_t *p, uint64_t n) {
  uLong c = crc32(0L, Z_NULL, 0);
  while (n > 0) {
    uInt k = n > (1u << 30) ? (1u << 30) : (uInt)n;
    c = crc32(c, p, k);
    p += k;
    n -= k;
  }
  return (uint32_t)c;
}

uint32_t oracle_adler32(const uint8_t *p, uint64_t n) {
  uLong a = adler32(0L, Z_NULL, 0);
  while (n > 0) {
    uInt k = n > (1u << 30) ? (1u << 30) : (uInt)n;
    a = adler32(a, p, k);
    p += k;
    n -= k;
  }
  return (uint32_t)a;
}

typedef struct {
  const uint8_t *src;
  const uint64_t *src_off, *src_len;
  uint64_t n;
  uint8_t *dst;
  const uint64_t *dst_off, *dst_cap;
  uint64_t *out_len;
  int32_t *status;
  uint64_t next;
  uint64_t failures;
  pthread_mutex_t mu;
} inf_job;

static void *inf_worker(void *arg) {
  inf_job *j = (inf_job *)arg;
  uint64_t fails = 0;
  for (;;) {
    uint64_t i = __atomic_fetch_add(&j->next, 1, __ATOMIC_RELAXED);
    if (i >= j->n) break;
    uint64_t got = 0;
    int rc = oracle_zlib_inflate(j->src + j->src_off[i], j->src_len[i], j->dst + j->dst_off[i],
                                 j->dst_cap[i], &got);
    if (j->out_len) j->out_len[i] = got;
    if (j->status) j->status[i] = rc;
    if (rc) fails++;
  }
  pthread_mutex_lock(&j->mu);
  j->failures += fails;
  pthread_mutex_unlock(&j->mu);
  return NULL;
}

uint64_t oracle_inflate_batch(const uint8_t *src, const uint64_t *src_off,
                              const uint64_t *src_len, uint64_t n, uint8_t *dst,
                              const uint64_t *dst_off, const uint64_t *dst_cap,
                              uint64_t *out_len, int32_t *status, int threads) {
  inf_job j;
  memset(&j, 0, sizeof j);
  j.src = src; j.src_off = src_off; j.src_len = src_len; j.n = n;
  j.dst = dst; j.dst_off = dst_off; j.dst_cap = dst_cap; j.out_len = out_len;
  j.status = status;
  if (threads < 1) threads = 1;
  pthread_mutex_init(&j.mu, NULL);
  pthread_t *tid = (pthread_t *)calloc((size_t)threads, sizeof(pthread_t));
  for (int t = 1; t < threads; t++) pthread_create(&tid[t], NULL, inf_worker, &j);
  inf_worker(&j);
  for (int t = 1; t < threads; t++) pthread_join(tid[t], NULL);
  free(tid);
  pthread_mutex_destroy(&j.mu);
  return j.failures;
}

uint64_t oracle_inflate_chunks(const uint8_t *comp, const uint64_t *off, uint64_t n_chunks,
                               uint64_t chunk_bytes, uint8_t *out, int32_t *status,
                               int threads) {
  uint64_t *src_len = (uint64_t *)malloc(sizeof(uint64_t) * (n_chunks + 1));
  uint64_t *dst_off = (uint64_t *)malloc(sizeof(uint64_t) * (n_chunks + 1));
  uint64_t *dst_cap = (uint64_t *)malloc(sizeof(uint64_t) * (n_chunks + 1));
  for (uint64_t i = 0; i < n_chunks; i++) {
    src_len[i] = off[i + 1] - off[i]; /* zlib.h:894 col_len */
    dst_off[i] = i * chunk_bytes;      /* zlib.h:899 */
    dst_cap[i] = chunk_bytes;
  }
  uint64_t f = oracle_inflate_batch(comp, off, src_len, n_chunks, out, dst_off, dst_cap, NULL,
                                    status, threads);
  free(src_len);
  free(dst_off);
  free(dst_cap);
  return f;
}
