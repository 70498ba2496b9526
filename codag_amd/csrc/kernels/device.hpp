// device.hpp -- wave64 / LDS helpers shared by the gfx950 decode kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace codag {
namespace dev {

constexpr int kWave = 64;

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Marks a wave-uniform value as such (v_readfirstlane), so code that depends
// on it stays in scalar registers with scalar branches.
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t width_mask(uint32_t w) {
  return w >= 8 ? ~0ull : ((1ull << (8 * w)) - 1);
}

// Stores the low w bytes of v (w in {1,2,4,8}) at p.
__device__ __forceinline__ void store_elem(uint8_t *p, uint32_t w, uint64_t v) {
  switch (w) {
  case 1: *p = (uint8_t)v; break;
  case 2: *reinterpret_cast<uint16_t *>(p) = (uint16_t)v; break;
  case 4: *reinterpret_cast<uint32_t *>(p) = (uint32_t)v; break;
  default: *reinterpret_cast<uint64_t *>(p) = v; break;
  }
}

// Inclusive wave-wide prefix sum of a 32-bit value.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint32_t t = __shfl_up(v, d, kWave);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
}

// Number of set bits of m in the lanes below this one (v_mbcnt).
__device__ __forceinline__ uint32_t lanes_below(uint64_t m) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// Inclusive wave64 prefix sum with DPP row shifts and row broadcasts: six
// VALU operations, no LDS traffic.
__device__ __forceinline__ uint32_t wave_scan_dpp(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return v;
}

// 64-bit form of wave_scan_dpp: each step moves both halves by DPP and adds
// with carry.
__device__ __forceinline__ uint64_t wave_scan_dpp64(uint64_t v) {
#define CODAG_SCAN64_STEP(ctrl, rows)                                                   \
  {                                                                                     \
    const uint32_t tl = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v,      \
                                                              ctrl, rows, 0xf, false);  \
    const uint32_t th = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), \
                                                              ctrl, rows, 0xf, false);  \
    v += ((uint64_t)th << 32) | tl;                                                     \
  }
  CODAG_SCAN64_STEP(0x111, 0xf)  // row_shr:1
  CODAG_SCAN64_STEP(0x112, 0xf)  // row_shr:2
  CODAG_SCAN64_STEP(0x114, 0xf)  // row_shr:4
  CODAG_SCAN64_STEP(0x118, 0xf)  // row_shr:8
  CODAG_SCAN64_STEP(0x142, 0xa)  // row_bcast:15
  CODAG_SCAN64_STEP(0x143, 0xc)  // row_bcast:31
#undef CODAG_SCAN64_STEP
  return v;
}

// OR of v over lanes 0..31, returned to every lane (DPP row scan, then the
// row-0 total broadcast into row 1; read from lane 31).
__device__ __forceinline__ uint32_t lanes32_or(uint32_t v) {
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
  v |= (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
}

__device__ __forceinline__ uint64_t wave_incl_scan64(uint64_t v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t t = __shfl_up(v, d, kWave);
    if (lane >= (uint32_t)d) v += t;
  }
  return v;
}

// Copies stream bytes [pos, pos+n) into LDS. The copy starts at the 16-byte
// aligned address at or below src+pos; *lead receives how many bytes precede
// the first requested byte in `lds`. A 16-byte block is loaded only if it
// contains a requested byte, so no load leaves the pages the stream occupies.
// `lds` must hold n + 32 bytes.
__device__ __forceinline__ void stage_to_lds(uint8_t *lds, const uint8_t *src, uint64_t pos,
                                             uint32_t n, uint32_t tid, uint32_t nthreads,
                                             uint32_t *lead) {
  const uintptr_t a = reinterpret_cast<uintptr_t>(src + pos);
  const uintptr_t a0 = a & ~uintptr_t(15);
  const uint32_t ld = (uint32_t)(a - a0);
  const uint32_t nblk = (ld + n + 15) / 16;
  const uint4 *g = reinterpret_cast<const uint4 *>(a0);
  uint4 *l = reinterpret_cast<uint4 *>(lds);
  for (uint32_t k = tid; k < nblk; k += nthreads) l[k] = g[k];
  *lead = ld;
}

// MSB-first bit field of width w (1..64) starting at bit `bitoff` of the byte
// string in LDS at s. Reads three aligned dwords; the buffer must extend 12
// bytes past the last byte the field touches.
__device__ __forceinline__ uint64_t lds_bits_be(const uint8_t *s, uint32_t bitoff, uint32_t w) {
  const uint32_t byte = bitoff >> 3;
  const uint32_t *d = reinterpret_cast<const uint32_t *>(s) + (byte >> 2);
  const uint32_t a = __builtin_bswap32(d[0]);
  const uint32_t b = __builtin_bswap32(d[1]);
  const uint32_t c = __builtin_bswap32(d[2]);
  const uint32_t skip = (byte & 3) * 8 + (bitoff & 7);  // 0..31
  const uint64_t ab = ((uint64_t)a << 32) | b;
  const uint64_t bc = ((uint64_t)b << 32) | c;
  const uint32_t x = (uint32_t)(ab >> (32 - skip));
  const uint32_t y = (uint32_t)(bc >> (32 - skip));
  const uint64_t v = ((uint64_t)x << 32) | y;
  return w >= 64 ? v : (v >> (64 - w));
}

}  // namespace dev
}  // namespace codag
