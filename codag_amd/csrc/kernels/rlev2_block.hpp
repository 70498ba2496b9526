// rlev2_block.hpp -- wave-cooperative expansion of one ORC RLE v2 block whose
// bytes sit in LDS. Semantics restate include/rlev2/decoder_new.h:1000-1185
// (decoder_warp_orig_rdw); every shift past the column width drops bits,
// as the reference's clamped 32/64-bit shifts do.
#pragma once

#include "device.hpp"

namespace codag {
namespace rle2 {

using namespace codag::dev;

// include/rlev2/utils.h:28 __BIT_WIDTH_DECODE_MAP
__constant__ uint8_t kDecodeWidth[32] = {1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11,
                                         12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22,
                                         23, 24, 26, 28, 30, 32, 40, 48, 56, 64};

__device__ __forceinline__ uint32_t decode_width(uint32_t code) {
  // same table as kDecodeWidth, computed so scalar code needs no memory load
  if (code < 24) return code + 1;
  if (code < 28) return 26 + 2 * (code - 24);
  return 40 + 8 * (code - 28);
}

// include/rlev2/utils.h:42-48 (__CLOSEST_FIXED_BIT_MAP)
__device__ __forceinline__ uint32_t closest_fixed(uint32_t bits) {
  if (bits <= 1) return 1;
  if (bits <= 24) return bits;
  if (bits <= 32) return (bits + 1) & ~1u;
  return (bits + 7) & ~7u;
}

// Unsigned LEB128 from LDS starting at *q (bits past 64 dropped); advances *q.
__device__ __forceinline__ uint64_t lds_uvarint(const uint8_t *s, uint32_t *q) {
  uint64_t v = 0;
  uint32_t sh = 0;
  uint8_t b;
  do {
    b = s[(*q)++];
    if (sh < 64) v |= (uint64_t)(b & 0x7f) << sh;
    sh += 7;
  } while (b >= 0x80);
  return v;
}

__device__ __forceinline__ uint64_t lds_be(const uint8_t *s, uint32_t nbytes) {
  uint64_t v = 0;
  for (uint32_t k = 0; k < nbytes; k++) v = (v << 8) | s[k];
  return v;
}

// Expands the DIRECT / DELTA / PATCHED_BASE block at LDS offset p into
// dst[o ...]. All 64 lanes of the wave must call it (cross-lane scans).
// Sets *bad for a patch outside its block.
__device__ __forceinline__ void expand_big(const uint8_t *sbuf, uint32_t p, uint32_t o,
                                           uint8_t *dst, uint32_t w, uint64_t mask,
                                           uint32_t nbits, uint32_t lane, uint32_t *bad) {
  const uint32_t first = sbuf[p];
  const uint32_t kind = first >> 6;
  const uint32_t code = (first >> 1) & 0x1f;
  const uint32_t fbw = decode_width(code);
  const uint32_t len8 = ((((first & 1) << 8) | sbuf[p + 1]) + 1) & 0xff;
  if (kind == 1) {  // DIRECT (decoder_new.h:1016-1044)
    const uint32_t b0 = (p + 2) * 8;
    for (uint32_t i = lane; i < len8; i += kWave)
      store_elem(dst + (uint64_t)(o + i) * w, w, lds_bits_be(sbuf, b0 + i * fbw, fbw) & mask);
  } else if (kind == 3) {  // DELTA (decoder_new.h:1045-1093)
    uint32_t q = p + 2;
    const uint64_t base = lds_uvarint(sbuf, &q) & mask;
    const int64_t zz = (int64_t)lds_uvarint(sbuf, &q);
    const uint64_t bd = (uint64_t)((zz >> 1) ^ -(zz & 1)) & mask;  // arithmetic >>
    const uint32_t nd8 = (len8 - 2) & 0xff;
    if (lane == 0) store_elem(dst + (uint64_t)o * w, w, base);
    if (lane == 1) store_elem(dst + (uint64_t)(o + 1) * w, w, (base + bd) & mask);
    if (code == 0) {  // fixed delta
      for (uint32_t i = 2 + lane; i < 2 + nd8; i += kWave)
        store_elem(dst + (uint64_t)(o + i) * w, w, (base + (uint64_t)i * bd) & mask);
    } else {
      const bool neg = (bd == 0);  // multiplier rule, decoder_new.h:1057
      uint64_t carry = base + bd;
      const uint32_t b0 = q * 8;
      for (uint32_t c = 0; c < nd8; c += kWave) {  // uniform loop: all lanes scan
        const uint32_t i = c + lane;
        const uint64_t d = i < nd8 ? (lds_bits_be(sbuf, b0 + i * fbw, fbw) & mask) : 0;
        const uint64_t s = wave_incl_scan64(d);
        const uint64_t v = neg ? carry - s : carry + s;
        if (i < nd8) store_elem(dst + (uint64_t)(o + 2 + i) * w, w, v & mask);
        const uint64_t tot = __shfl(s, kWave - 1, kWave);
        carry = neg ? carry - tot : carry + tot;
      }
    }
  } else {  // PATCHED_BASE (decoder_new.h:1094-1183)
    const uint32_t third = sbuf[p + 2], fourth = sbuf[p + 3];
    const uint32_t bw = (third >> 5) + 1;
    const uint32_t pw = decode_width(third & 0x1f);
    const uint32_t pgw = (fourth >> 5) + 1;
    const uint32_t pll = fourth & 0x1f;
    const uint64_t pmask = pw >= 32 ? 0xffffffffull : ((1ull << pw) - 1);  // uint32 mask
    const uint32_t cfb = closest_fixed(pw + pgw);
    const uint64_t base = lds_be(sbuf + p + 4, bw) & mask;
    const uint32_t vb0 = (p + 4 + bw) * 8;
    const uint32_t pb0 = (p + 4 + bw + (len8 * fbw + 7) / 8) * 8;
    // lane j < pll: entry j -> running gap (wave scan) and its OR-term
    const uint64_t ent = lane < pll ? lds_bits_be(sbuf, pb0 + lane * cfb, cfb) : 0;
    const uint32_t g = lane < pll ? (uint32_t)(pw >= 64 ? 0 : (ent >> pw)) : 0;
    const uint32_t gsum = wave_incl_scan(g);
    const uint64_t add = (fbw >= nbits) ? 0 : (((ent & pmask) << fbw) & mask);
    if (lane < pll && gsum >= len8) *bad = 1;
    // all lanes stay active around the cross-lane reads (a shuffle from an
    // inactive lane returns 0)
    for (uint32_t c = 0; c < len8; c += kWave) {
      const uint32_t i = c + lane;
      uint64_t v = i < len8 ? (lds_bits_be(sbuf, vb0 + i * fbw, fbw) & mask) : 0;
      for (uint32_t j = 0; j < pll; j++) {
        const uint32_t gj = __shfl(gsum, j, kWave);
        const uint64_t aj = __shfl(add, j, kWave);
        if (gj == i) v |= aj;
      }
      if (i < len8) store_elem(dst + (uint64_t)(o + i) * w, w, (v + base) & mask);
    }
  }
}

}  // namespace rle2
}  // namespace codag
