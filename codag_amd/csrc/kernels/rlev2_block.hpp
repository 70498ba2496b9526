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
  return code < 24 ? code + 1 : (code < 28 ? 2 * code - 22 : 8 * code - 184);  // selects, no branches
}

// include/rlev2/utils.h:42-48 (__CLOSEST_FIXED_BIT_MAP)
__device__ __forceinline__ uint32_t closest_fixed(uint32_t bits) {
  return bits <= 1 ? 1u : (bits <= 24 ? bits : (bits <= 32 ? (bits + 1) & ~1u : (bits + 7) & ~7u));
}

// MSB-first bit field of width f (1..32) starting at bit b of the LDS byte
// string s: the field lies inside the two dwords at b/32 (byte-swapped into a
// 64-bit big-endian window). s must extend 8 bytes past the field's dword.
__device__ __forceinline__ uint32_t bits32(const uint8_t *s, uint32_t b, uint32_t f) {
  const uint32_t *d = reinterpret_cast<const uint32_t *>(s) + (b >> 5);
  const uint64_t v = ((uint64_t)__builtin_bswap32(d[0]) << 32) | __builtin_bswap32(d[1]);
  return (uint32_t)((v << (b & 31)) >> (64 - f));
}

// Same for widths 33..64: three dwords, read together (no branch on the bit
// offset, so the three loads share one wait).
__device__ __forceinline__ uint64_t bits64(const uint8_t *s, uint32_t b, uint32_t f) {
  const uint32_t *d = reinterpret_cast<const uint32_t *>(s) + (b >> 5);
  const uint32_t r = b & 31;
  const uint64_t ab = ((uint64_t)__builtin_bswap32(d[0]) << 32) | __builtin_bswap32(d[1]);
  const uint32_t c = __builtin_bswap32(d[2]);
  const uint64_t top = (ab << r) | ((((uint64_t)c) << r) >> 32);  // 64 bits from bit b on
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
// scans). Sets *bad for a patch outside its block. The block was accepted by
// the header walk, so its two DELTA varints are at most 10 bytes each.
//
// LDS reads are issued unconditionally (lanes past a block's count read its
// element 0) and two 64-element groups at a time while more than 64 elements
// are left, so reads that do not depend on each other share one wait; the
// last group of 64 or fewer is read alone.
template <uint32_t W>
__device__ __forceinline__ void expand_big(const uint8_t *sbuf, uint32_t p, uint32_t o, uint8_t *dst,
                                           uint32_t lane, uint32_t *bad) {
  using T = typename Col<W>::T;
  constexpr T mask = Col<W>::kMask;
  constexpr uint32_t nbits = 8 * W;
  // Lane l reads dword p/4 + (l & 7) and byte p + (l & 31): the header,
  // PATCHED_BASE's base value and DELTA's varints come from these two reads.
  const uint32_t hw = reinterpret_cast<const uint32_t *>(sbuf)[(p >> 2) + (lane & 7)];
  const uint32_t hb = sbuf[p + (lane & 31)];
  const uint32_t w0 = (uint32_t)__builtin_amdgcn_readlane((int)hw, 0);
  const uint32_t w1 = (uint32_t)__builtin_amdgcn_readlane((int)hw, 1);
  const uint32_t sh = 8 * (p & 3);
  const uint32_t x0 = (uint32_t)((((uint64_t)w1 << 32) | w0) >> sh);  // bytes p .. p+3
  const uint32_t first = x0 & 0xff;
  const uint32_t kind = first >> 6;
  const uint32_t code = (first >> 1) & 0x1f;
  const uint32_t fbw = decode_width(code);
  const uint32_t len8 = ((x0 >> 8) + 1) & 0xff;  // 9-bit length + 1, kept in 8 bits
  if (kind == 1) {  // DIRECT (decoder_new.h:1016-1044)
    const uint32_t b0 = (p + 2) * 8;
    for (uint32_t c = 0; c < len8; c += 2 * kWave) {
      const uint32_t i0 = c + lane, i1 = c + kWave + lane;
      if (len8 - c > (uint32_t)kWave) {
        const T v0 = field<W>(sbuf, b0 + __umul24(i0, fbw), fbw) & mask;
        const T v1 = field<W>(sbuf, b0 + __umul24(i1 < len8 ? i1 : 0u, fbw), fbw) & mask;
        store_w<W>(dst, o + i0, v0);
        if (i1 < len8) store_w<W>(dst, o + i1, v1);
      } else {
        const T v0 = field<W>(sbuf, b0 + __umul24(i0 < len8 ? i0 : 0u, fbw), fbw) & mask;
        if (i0 < len8) store_w<W>(dst, o + i0, v0);
      }
    }
  } else if (kind == 3) {  // DELTA (decoder_new.h:1045-1093)
    // The varints of base and delta base start at byte p + 2 and end at the
    // first bytes below 0x80 (e1, e2); each lane shifts its byte's 7 value
    // bits into place and DPP OR-reductions assemble the varints
    // (decoder_new.h:978-989; a 10th byte keeps only its lowest bit, at bit
    // 63).
    const uint32_t b = lane < 32 ? hb : 0x80u;
    const uint64_t term = __ballot(lane >= 2 && b < 0x80);
    const uint32_t e1 = uni((uint32_t)__builtin_ctzll(term));
    const uint32_t e2 = uni((uint32_t)__builtin_ctzll(term & ~((2ull << e1) - 1)));
    const bool inb = lane >= 2 && lane <= e1;
    const bool inz = lane > e1 && lane <= e2;
    const uint32_t vsh = inb ? 7 * (lane - 2) : (inz ? 7 * (lane - e1 - 1) : 0u);  // <= 63
    const uint64_t g = (uint64_t)(b & 0x7f) << vsh;
    const uint64_t gb = inb ? g : 0ull, gz = inz ? g : 0ull;
    uint64_t base64 = lanes32_or((uint32_t)gb);
    if (W > 4) base64 |= (uint64_t)lanes32_or((uint32_t)(gb >> 32)) << 32;
    const int64_t zz = (int64_t)(((uint64_t)lanes32_or((uint32_t)(gz >> 32)) << 32) |
                                 lanes32_or((uint32_t)gz));
    const T base = (T)base64 & mask;
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
      const uint32_t b0 = (p + e2 + 1) * 8;  // first packed delta
      // 64 deltas of at most 26 bits sum below 2^32: a 32-bit scan is exact
      const bool narrow = W <= 4 || fbw <= 26;
      auto step = [&](uint32_t i, T d) {  // wave prefix sum of one group
        T sc;
        if (narrow) sc = (T)wave_scan_dpp((uint32_t)d);
        else sc = (T)wave_scan_dpp64((uint64_t)d);
        const T v = neg ? (T)(carry - sc) : (T)(carry + sc);
        if (i < nd8) store_w<W>(dst, o + 2 + i, v & mask);
        T tot;
        if (narrow) {
          tot = (T)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)sc, kWave - 1);
        } else {
          const uint64_t s64 = (uint64_t)sc;
          tot = (T)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(s64 >> 32), kWave - 1) << 32) |
                    (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)s64, kWave - 1));
        }
        carry = neg ? (T)(carry - tot) : (T)(carry + tot);
      };
      for (uint32_t c = 0; c < nd8; c += 2 * kWave) {  // uniform loop: all lanes scan
        const uint32_t i0 = c + lane, i1 = c + kWave + lane;
        if (nd8 - c > (uint32_t)kWave) {
          const T r0 = field<W>(sbuf, b0 + __umul24(i0, fbw), fbw) & mask;
          const T r1 = field<W>(sbuf, b0 + __umul24(i1 < nd8 ? i1 : 0u, fbw), fbw) & mask;
          step(i0, r0);
          step(i1, i1 < nd8 ? r1 : (T)0);
        } else {
          const T r0 = field<W>(sbuf, b0 + __umul24(i0 < nd8 ? i0 : 0u, fbw), fbw) & mask;
          step(i0, i0 < nd8 ? r0 : (T)0);
        }
      }
    }
  } else {  // PATCHED_BASE (decoder_new.h:1094-1183)
    const uint32_t third = (x0 >> 16) & 0xff, fourth = x0 >> 24;
    const uint32_t bw = (third >> 5) + 1;
    const uint32_t pw = decode_width(third & 0x1f);
    const uint32_t pgw = (fourth >> 5) + 1;
    const uint32_t pll = fourth & 0x1f;
    const uint64_t pmask = pw >= 32 ? 0xffffffffull : ((1ull << pw) - 1);  // uint32 mask
    const uint32_t cfb = closest_fixed(pw + pgw);
    // base: bw big-endian bytes at p + 4 (inside dwords p/4 + 1 .. p/4 + 3)
    const uint32_t w2 = (uint32_t)__builtin_amdgcn_readlane((int)hw, 2);
    const uint32_t w3 = (uint32_t)__builtin_amdgcn_readlane((int)hw, 3);
    const uint64_t y = (uint64_t)(uint32_t)((((uint64_t)w2 << 32) | w1) >> sh) |
                       ((uint64_t)(uint32_t)((((uint64_t)w3 << 32) | w2) >> sh) << 32);
    const T base = (T)(__builtin_bswap64(y) >> (64 - 8 * bw)) & mask;
    const uint32_t vb0 = (p + 4 + bw) * 8;
    const uint32_t pb0 = (p + 4 + bw + (len8 * fbw + 7) / 8) * 8;
    // lane j < pll: entry j -> running gap (wave scan) and its OR-term
    const uint64_t ent0 = bits64(sbuf, pb0 + __umul24(lane < pll ? lane : 0u, cfb), cfb);
    const uint64_t ent = lane < pll ? ent0 : 0;
    const uint32_t g = lane < pll ? (uint32_t)(pw >= 64 ? 0 : (ent >> pw)) : 0;
    const uint32_t gsum = wave_scan_dpp(g);
    const T add = (fbw >= nbits) ? (T)0 : (T)(((ent & pmask) << fbw) & (uint64_t)mask);
    if (lane < pll && gsum >= len8) *bad = 1;
    // the OR-term of patch entry j (held by lane j < pll)
    auto patch = [&](uint32_t j) -> T {
      T aj = (T)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)add, (int)j);
      if (W > 4)
        aj |= (T)((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)add >> 32), (int)j) << 32);
      return aj;
    };
    for (uint32_t c = 0; c < len8; c += 2 * kWave) {
      const uint32_t i0 = c + lane, i1 = c + kWave + lane;
      if (len8 - c > (uint32_t)kWave) {
        T v0 = field<W>(sbuf, vb0 + __umul24(i0, fbw), fbw) & mask;
        T v1 = field<W>(sbuf, vb0 + __umul24(i1 < len8 ? i1 : 0u, fbw), fbw) & mask;
        for (uint32_t j = 0; j < pll; j++) {  // patch entries read from their lanes
          const uint32_t gj = (uint32_t)__builtin_amdgcn_readlane((int)gsum, (int)j);
          const T aj = patch(j);
          if (gj == i0) v0 |= aj;
          if (gj == i1) v1 |= aj;
        }
        store_w<W>(dst, o + i0, (T)(v0 + base) & mask);
        if (i1 < len8) store_w<W>(dst, o + i1, (T)(v1 + base) & mask);
      } else {
        T v0 = field<W>(sbuf, vb0 + __umul24(i0 < len8 ? i0 : 0u, fbw), fbw) & mask;
        for (uint32_t j = 0; j < pll; j++) {
          const uint32_t gj = (uint32_t)__builtin_amdgcn_readlane((int)gsum, (int)j);
          const T aj = patch(j);
          if (gj == i0) v0 |= aj;
        }
        if (i0 < len8) store_w<W>(dst, o + i0, (T)(v0 + base) & mask);
      }
    }
  }
}

}  // namespace rle2
}  // namespace codag
