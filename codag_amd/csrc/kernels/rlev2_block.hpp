// rlev2_block.hpp -- wave-cooperative expansion of ORC RLE v2 blocks whose
// bytes sit in LDS. Semantics restate include/rlev2/decoder_new.h:1000-1185
// (decoder_warp_orig_rdw); every shift past the column width drops bits, as
// the reference's clamped 32/64-bit shifts do.
//
// The expansion is specialised on the element width W (1, 2, 4 or 8 bytes):
// columns of up to 4 bytes keep values in 32-bit registers (values are
// masked to the column width, so only the low 32 bits of any field matter),
// and every packed field is read with one or two LDS dword pairs instead of
// byte loops.
#pragma once

#include <type_traits>

#include "device.hpp"

namespace codag {
namespace rle2 {

using namespace codag::dev;

// include/rlev2/utils.h:28 __BIT_WIDTH_DECODE_MAP, in closed form so scalar
// code needs no memory load: codes 0..23 -> 1..24, 24..27 -> 26..32 step 2,
// 28..31 -> 40..64 step 8.
__device__ __forceinline__ uint32_t decode_width(uint32_t code) {
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

// MSB-first bit field of width f (1..32) starting at bit b of the LDS byte
// string s: the field lies inside the two dwords at b/32 (byte-swapped into a
// 64-bit big-endian window). s must extend 8 bytes past the field's dword.
__device__ __forceinline__ uint32_t bits32(const uint8_t *s, uint32_t b, uint32_t f) {
  const uint32_t *d = reinterpret_cast<const uint32_t *>(s) + (b >> 5);
  const uint64_t v = ((uint64_t)__builtin_bswap32(d[0]) << 32) | __builtin_bswap32(d[1]);
  return (uint32_t)((v << (b & 31)) >> (64 - f));
}

// Same for widths 33..64: three dwords.
__device__ __forceinline__ uint64_t bits64(const uint8_t *s, uint32_t b, uint32_t f) {
  const uint32_t *d = reinterpret_cast<const uint32_t *>(s) + (b >> 5);
  const uint32_t r = b & 31;
  const uint64_t ab = ((uint64_t)__builtin_bswap32(d[0]) << 32) | __builtin_bswap32(d[1]);
  const uint32_t c = __builtin_bswap32(d[2]);
  const uint64_t top = r ? ((ab << r) | (c >> (32 - r))) : ab;  // 64 bits from bit b on
  return f >= 64 ? top : (top >> (64 - f));
}

// Value type of a column of W-byte elements: 32-bit registers suffice up to
// W = 4 because every value is masked to the column width.
template <uint32_t W> struct Col {
  using T = typename std::conditional<(W <= 4), uint32_t, uint64_t>::type;
  static constexpr T kMask = W >= 8 ? ~0ull : (W == 4 ? 0xffffffffu : (T)((1ull << (8 * W)) - 1));
};

template <uint32_t W> __device__ __forceinline__ void store_w(uint8_t *dst, uint32_t idx, uint64_t v) {
  if (W == 1) dst[idx] = (uint8_t)v;
  else if (W == 2) reinterpret_cast<uint16_t *>(dst)[idx] = (uint16_t)v;
  else if (W == 4) reinterpret_cast<uint32_t *>(dst)[idx] = (uint32_t)v;
  else reinterpret_cast<uint64_t *>(dst)[idx] = v;
}

// Low bits of the MSB-first field (b, f) that a W-byte column keeps.
template <uint32_t W>
__device__ __forceinline__ typename Col<W>::T field(const uint8_t *s, uint32_t b, uint32_t f) {
  if (W <= 4) {  // only the field's last 32 bits survive the mask
    if (f > 32) return bits32(s, b + f - 32, 32);
    return bits32(s, b, f);
  }
  if (f <= 32) return bits32(s, b, f);
  return (typename Col<W>::T)bits64(s, b, f);
}

// Big-endian value of nb (1..8) bytes at LDS offset p, masked to W bytes.
template <uint32_t W>
__device__ __forceinline__ typename Col<W>::T be_bytes(const uint8_t *s, uint32_t p, uint32_t nb) {
  return field<W>(s, 8 * p, 8 * nb) & Col<W>::kMask;
}

// Expands the DIRECT / DELTA / PATCHED_BASE block at LDS offset p into
// element o onwards of dst. All 64 lanes of the wave must call it (cross-lane
// scans). Sets *bad for a patch outside its block.
template <uint32_t W>
__device__ __forceinline__ void expand_big(const uint8_t *sbuf, uint32_t p, uint32_t o, uint8_t *dst,
                                           uint32_t lane, uint32_t *bad) {
  using T = typename Col<W>::T;
  constexpr T mask = Col<W>::kMask;
  constexpr uint32_t nbits = 8 * W;
  const uint32_t first = sbuf[p];
  const uint32_t kind = first >> 6;
  const uint32_t code = (first >> 1) & 0x1f;
  const uint32_t fbw = decode_width(code);
  const uint32_t len8 = ((((first & 1) << 8) | sbuf[p + 1]) + 1) & 0xff;
  if (kind == 1) {  // DIRECT (decoder_new.h:1016-1044)
    const uint32_t b0 = (p + 2) * 8;
    for (uint32_t i = lane; i < len8; i += kWave)
      store_w<W>(dst, o + i, field<W>(sbuf, b0 + i * fbw, fbw) & mask);
  } else if (kind == 3) {  // DELTA (decoder_new.h:1045-1093)
    uint32_t q = p + 2;
    const T base = (T)lds_uvarint(sbuf, &q) & mask;
    const int64_t zz = (int64_t)lds_uvarint(sbuf, &q);
    const T bd = (T)((zz >> 1) ^ -(zz & 1)) & mask;  // arithmetic >>
    const uint32_t nd8 = (len8 - 2) & 0xff;
    if (lane == 0) store_w<W>(dst, o, base);
    if (lane == 1) store_w<W>(dst, o + 1, (T)(base + bd) & mask);
    if (code == 0) {  // fixed delta
      for (uint32_t i = 2 + lane; i < 2 + nd8; i += kWave)
        store_w<W>(dst, o + i, (T)(base + (T)i * bd) & mask);
    } else {
      const bool neg = (bd == 0);  // multiplier rule, decoder_new.h:1057
      T carry = base + bd;
      const uint32_t b0 = q * 8;
      for (uint32_t c = 0; c < nd8; c += kWave) {  // uniform loop: all lanes scan
        const uint32_t i = c + lane;
        const T d = i < nd8 ? (field<W>(sbuf, b0 + i * fbw, fbw) & mask) : (T)0;
        T s;
        if (W <= 4) s = (T)wave_incl_scan((uint32_t)d);
        else s = (T)wave_incl_scan64((uint64_t)d);
        const T v = neg ? (T)(carry - s) : (T)(carry + s);
        if (i < nd8) store_w<W>(dst, o + 2 + i, v & mask);
        const T tot = (T)__shfl(s, kWave - 1, kWave);
        carry = neg ? (T)(carry - tot) : (T)(carry + tot);
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
    const T base = be_bytes<W>(sbuf, p + 4, bw);
    const uint32_t vb0 = (p + 4 + bw) * 8;
    const uint32_t pb0 = (p + 4 + bw + (len8 * fbw + 7) / 8) * 8;
    // lane j < pll: entry j -> running gap (wave scan) and its OR-term
    const uint64_t ent = lane < pll ? bits64(sbuf, pb0 + lane * cfb, cfb) : 0;
    const uint32_t g = lane < pll ? (uint32_t)(pw >= 64 ? 0 : (ent >> pw)) : 0;
    const uint32_t gsum = wave_incl_scan(g);
    const T add = (fbw >= nbits) ? (T)0 : (T)(((ent & pmask) << fbw) & (uint64_t)mask);
    if (lane < pll && gsum >= len8) *bad = 1;
    // all lanes stay active around the cross-lane reads (a shuffle from an
    // inactive lane returns 0)
    for (uint32_t c = 0; c < len8; c += kWave) {
      const uint32_t i = c + lane;
      T v = i < len8 ? (field<W>(sbuf, vb0 + i * fbw, fbw) & mask) : (T)0;
      for (uint32_t j = 0; j < pll; j++) {
        const uint32_t gj = __shfl(gsum, j, kWave);
        const T aj = (T)__shfl(add, j, kWave);
        if (gj == i) v |= aj;
      }
      if (i < len8) store_w<W>(dst, o + i, (T)(v + base) & mask);
    }
  }
}

}  // namespace rle2
}  // namespace codag
