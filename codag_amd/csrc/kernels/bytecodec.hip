// bytecodec.hip -- gfx950 decoders of the reference's two byte-oriented
// codecs:
//
//   BRLE  include/brle/brle.h:26-100 (brle::decode, direct path at 76-90;
//         kernel_decompress 102-110 runs one thread per chunk). A header
//         byte h > 127 is followed by h - 127 literal bytes, h <= 127 by one
//         byte repeated h + 3 times. Chunk i is data[ptr[i] .. ptr[i+1]) and
//         decodes into out + i * chunk (256 KiB in the reference, brle.h:8).
//   LZSS  include/lzss/lzss.h:234-305 (lzss::decompress_func, one thread per
//         chunk). A flag byte covers the next 8 items, least significant bit
//         first: bit 1 is a literal byte, bit 0 a 2-byte little-endian
//         reference v, length (v & 31) + 3, offset v >> 5 (11 offset bits for
//         the 2048-byte history, lzss.h:24-32), copied byte by byte from
//         offset bytes back. Chunk i is data[end[i-1] .. end[i]) and decodes
//         into out + i * chunk (128 KiB, lzss.h:21).
//
// Both formats are sequences of byte-aligned items whose sizes follow from
// one header byte (BRLE) or one flag byte per eight items (LZSS), so a batch
// is found by a scalar walk over header bytes that sit in two VGPRs
// (v_readlane, no memory round trip per step), each lane then takes one
// item, and the batch's output is assembled in an LDS ring and stored to HBM
// as 16-byte aligned vectors. The next batch's input window is loaded while
// the current batch is assembled. LZSS runs one wave64 per chunk; BRLE two,
// the walking wave one batch ahead of the writing wave.
//
// The reference checks nothing; these kernels report per chunk: TRUNCATED if
// the stream ends inside an item, OVERFLOW if the output would pass the
// chunk, CORRUPT for an LZSS offset of 0 or before the chunk start (or a
// chunk whose stream lies outside the input). The items before the first
// error are written; out_len gets the bytes produced.
#include <hip/hip_runtime.h>

#include "device.hpp"
#include "kernels.hpp"

namespace codag {
namespace bytecodec {

using namespace dev;

constexpr int32_t kOk = CODAG_CHUNK_OK;
constexpr int32_t kTrunc = CODAG_CHUNK_TRUNCATED;
constexpr int32_t kOverflow = CODAG_CHUNK_OVERFLOW;
constexpr int32_t kCorrupt = CODAG_CHUNK_CORRUPT;

// LDS input window: 64 lanes x 16 bytes from the 16-byte aligned address at
// or below the first unconsumed stream byte (`lead` bytes before it).
constexpr uint32_t kWin = 1024;
// Header bytes the walk reads lie in window bytes [0, kHead): two VGPRs.
constexpr uint32_t kHead = 512;
// Output ring. LZSS needs the last 2047 bytes as history plus one batch.
constexpr uint32_t kRing = 4096;
constexpr uint32_t kRingMask = kRing - 1;
// Batch output limits: BRLE has no history, LZSS keeps 2047 bytes of it.
// Both leave the < 16 bytes a flush keeps back intact.
constexpr uint32_t kBrleBatch = kRing - 32;
constexpr uint32_t kLzssBatch = kRing - 2048 - 16;
static_assert(kHead + 130 <= kWin && kHead + 17 <= kWin, "items must end inside the window");
static_assert(7 * 8 * 34 <= kLzssBatch, "seven LZSS groups always fit a batch");

__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t j) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)j);
}
// old with lane j's value replaced by the uniform v
__device__ __forceinline__ uint32_t wrlane(uint32_t old, uint32_t v, uint32_t j) {
  return lane_id() == j ? v : old;
}
// v of lane `src & 63` (ds_bpermute); called with every lane active.
__device__ __forceinline__ uint32_t bperm(uint32_t v, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)((src & 63u) << 2), (int)v);
}

// ------------------------------------------------------------------ input

// Stream bytes [at, at + kWin - lead) as one 16-byte load per lane. A block
// is loaded only if it holds a byte below the stream end, so no load leaves
// the 16-byte blocks (and pages) the stream occupies.
struct Window {
  uint4 v;
  uint32_t lead;
};

__device__ __forceinline__ Window window_load(const uint8_t *base, uint32_t len, uint32_t at,
                                              uint32_t lane) {
  typedef uint32_t v4u __attribute__((ext_vector_type(4)));
  typedef __attribute__((address_space(1))) const v4u gvec;
  const uintptr_t a = reinterpret_cast<uintptr_t>(base) + at;
  const uintptr_t a0 = a & ~uintptr_t(15);
  const uintptr_t end = reinterpret_cast<uintptr_t>(base) + len;
  Window w;
  w.lead = (uint32_t)(a - a0);
  w.v = make_uint4(0, 0, 0, 0);
  if (a0 + 16u * lane < end) {
    const v4u q = *reinterpret_cast<gvec *>(a0 + 16u * lane);
    w.v = make_uint4(q.x, q.y, q.z, q.w);
  }
  return w;
}

// Header byte at window index q < kHead from the VGPR copies of window bytes
// [0, 256) (h0) and [256, 512) (h1); q is wave-uniform.
__device__ __forceinline__ uint32_t head_byte(uint32_t h0, uint32_t h1, uint32_t q) {
  const uint32_t l = (q >> 2) & 63u;
  const uint32_t d0 = rdlane(h0, l), d1 = rdlane(h1, l);
  return ((q < 256u ? d0 : d1) >> ((q & 3u) * 8u)) & 0xffu;
}

// ------------------------------------------------------------------ output

// Byte x of the chunk's output sits at ring[(x + mis) & kRingMask], mis =
// dst & 15, so 16-byte aligned output addresses are 16-byte aligned in LDS.
struct Out {
  uint8_t *dst;
  uint8_t *ring;
  uint32_t cap;
  uint32_t pos;      // bytes produced
  uint32_t flushed;  // bytes [0, flushed) are in HBM
  uint32_t mis;
};

__device__ __forceinline__ uint32_t ridx(const Out &o, uint32_t x) { return (x + o.mis) & kRingMask; }

// Stores ring bytes [flushed, end) to HBM: single bytes up to a 16-byte
// aligned address, then 16-byte vectors; `last` also writes the tail,
// otherwise fewer than 16 bytes stay behind.
__device__ __forceinline__ void out_store(Out &o, uint32_t lane, uint32_t end, bool last) {
  uint32_t x = o.flushed;
  uint32_t head = (16u - ((o.mis + x) & 15u)) & 15u;
  if (head > end - x) head = end - x;
  if (head) {
    if (lane < head) o.dst[x + lane] = o.ring[ridx(o, x + lane)];
    x += head;
  }
  const uint32_t nvec = (end - x) / 16u;
  for (uint32_t k = lane; k < nvec; k += kWave) {
    const uint32_t at = x + 16u * k;
    *reinterpret_cast<uint4 *>(o.dst + at) = *reinterpret_cast<const uint4 *>(o.ring + ridx(o, at));
  }
  x += nvec * 16u;
  if (last && x < end) {
    if (lane < end - x) o.dst[x + lane] = o.ring[ridx(o, x + lane)];
    x = end;
  }
  o.flushed = x;
}

// floor(2^16 / d) + 1 for d = 1..63: (i * m[d]) >> 16 == i / d for i < 512
// (the overshoot i * (m[d] - 2^16/d) / 2^16 < 1/128 never reaches the next
// integer, since the fraction of i / d is at most 1 - 1/63).
struct Recip16 {
  uint32_t m[64];
  constexpr Recip16() : m() {
    for (uint32_t d = 1; d < 64; d++) m[d] = 0x10000u / d + 1;
  }
};
__constant__ Recip16 kRecip = Recip16();

__device__ __forceinline__ void chunk_bounds(const uint64_t lo, const uint64_t hi, uint64_t comp_bytes,
                                             uint64_t &len, int32_t &st) {
  st = kOk;
  len = 0;
  if (hi < lo || hi > comp_bytes || hi - lo >= (1ull << 31)) st = kCorrupt;
  else len = hi - lo;
}

// ------------------------------------------------------------------ BRLE

// Two waves per chunk, in lockstep (one workgroup barrier per batch): wave 0
// walks batch i (window staging, header chain, item checks, output offsets)
// while wave 1 writes batch i-1 into the ring and stores it to HBM. The walk
// is scalar-bound and the writing vector/LDS-bound, so the two overlap, and a
// 4096-chunk launch fills all 32 wave slots of a CU instead of 16. Batch i
// uses window buffer i % 3 (the writer still reads i-1's literal bytes while
// the walker stages i) and descriptor buffer i % 2.
struct __align__(16) BrleLds2 {
  uint8_t ring[kRing];
  uint8_t win[3][kWin];
  uint32_t item[2][kWave];  // header index | batch output offset << 16
  uint32_t nk[2];           // items in the batch
  uint32_t base[2];         // output position of the batch's first byte
  uint32_t last;            // the walker's last batch (kNoLast while walking)
};
constexpr uint32_t kNoLast = 0xffffffffu;

__global__ void __launch_bounds__(2 * kWave) brle_chunks2_kernel(const uint8_t *comp,
                                                                 uint64_t comp_bytes,
                                                                 const uint64_t *ptr, uint64_t n,
                                                                 uint64_t chunk_bytes, uint8_t *out,
                                                                 uint64_t *out_len, int32_t *status) {
  __shared__ BrleLds2 S;
  const uint32_t lane = lane_id();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x) / kWave;
  const uint64_t ci = blockIdx.x;
  if (ci >= n) return;
  uint64_t len64;
  int32_t err;
  chunk_bounds(ptr[ci], ptr[ci + 1], comp_bytes, len64, err);
  if (err != kOk) {  // both waves leave before any barrier
    if (threadIdx.x == 0) {
      if (out_len) out_len[ci] = 0;
      if (status) status[ci] = err;
    }
    return;
  }
  const uint8_t *src = comp + ptr[ci];
  const uint32_t len = (uint32_t)len64;
  uint8_t *dst = out + ci * chunk_bytes;
  const uint32_t cap = (uint32_t)chunk_bytes;
  if (threadIdx.x == 0) S.last = kNoLast;
  __syncthreads();
  // walker state
  uint32_t at = 0, pos = 0;
  Window w;
  if (wave == 0) w = window_load(src, len, 0, lane);
  // writer state
  Out o;
  o.dst = dst;
  o.ring = S.ring;
  o.cap = cap;
  o.pos = 0;
  o.flushed = 0;
  o.mis = (uint32_t)(reinterpret_cast<uintptr_t>(dst) & 15u);
  for (uint32_t i = 0;; i++) {
    if (wave == 0 && uni(S.last) == kNoLast) {
      uint8_t *win = S.win[i % 3];
      reinterpret_cast<uint4 *>(win)[lane] = w.v;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
      const uint32_t lead = uni(w.lead);
      const uint32_t h0 = reinterpret_cast<const uint32_t *>(win)[lane];
      const uint32_t h1 = reinterpret_cast<const uint32_t *>(win)[kWave + lane];
      const uint32_t qend = lead + (len - at);
      const uint32_t qlim = min(qend, kHead);
      uint32_t q = lead, nw = 0, iq = 0;
      while (nw < (uint32_t)kWave && q < qlim) {
        const uint32_t h = head_byte(h0, h1, q);
        iq = wrlane(iq, q, nw);
        nw++;
        q += h > 127u ? h - 126u : 2u;
      }
      const bool act = lane < nw;
      const uint32_t hq = act ? (uint32_t)win[iq] : 0u;
      const bool lit = hq > 127u;
      const uint32_t k = act ? (lit ? hq - 127u : hq + 3u) : 0u;
      const uint32_t inl = lit ? k + 1u : 2u;
      const uint32_t incl = wave_scan_dpp(k);
      const uint32_t bad = !act ? 0u : (iq + inl > qend ? 1u : (pos + incl > cap ? 2u : 0u));
      const uint64_t stop = __ballot(bad != 0u || (act && incl > kBrleBatch));
      uint32_t nk = nw;
      if (stop) {
        nk = (uint32_t)__builtin_ctzll(stop);
        const uint32_t why = rdlane(bad, nk);
        if (why) err = why == 1u ? kTrunc : kOverflow;
        at += rdlane(iq, nk) - lead;
      } else {
        at += q - lead;
      }
      const uint32_t T = nk ? rdlane(incl, nk - 1u) : 0u;
      S.item[i & 1][lane] = iq | ((incl - k) << 16);
      if (lane == 0) {
        S.nk[i & 1] = nk;
        S.base[i & 1] = pos;
        if (err != kOk || at >= len) S.last = i;
      }
      pos += T;
      if (err == kOk && at < len) w = window_load(src, len, at, lane);
    } else if (wave == 1 && i > 0) {
      // write batch i-1
      const uint32_t b = (i - 1) & 1;
      const uint8_t *win = S.win[(i - 1) % 3];
      const uint32_t nk = uni(S.nk[b]);
      const uint32_t it = S.item[b][lane];
      const uint32_t iq = it & 0xffffu, off = it >> 16;
      const uint32_t kk0 = lane < nk ? (uint32_t)win[iq] : 0u;
      const bool lit = kk0 > 127u;
      const uint32_t kk = lane < nk ? (lit ? kk0 - 127u : kk0 + 3u) : 0u;
      const uint32_t s0 = iq + 1u;
      const uint32_t c = kk ? (uint32_t)win[s0] : 0u;
      const uint32_t r0 = uni(S.base[b]) + off + o.mis;
      const uint32_t hb = min((4u - (r0 & 3u)) & 3u, kk);
#pragma unroll
      for (uint32_t t = 0; t < 3u; t++)
        if (t < hb) S.ring[(r0 + t) & kRingMask] = lit ? win[s0 + t] : (uint8_t)c;
      const uint32_t nd = (kk - hb) >> 2;
      const uint32_t c4 = c * 0x01010101u;
      uint32_t *ring32 = reinterpret_cast<uint32_t *>(S.ring);
      const uint32_t *win32 = reinterpret_cast<const uint32_t *>(win);
      for (uint32_t d = 0; d < nd; d++) {
        const uint32_t u = s0 + hb + 4u * d;
        uint32_t v = c4;
        if (lit) {
          const uint32_t a = u >> 2;
          v = __builtin_amdgcn_alignbit(win32[a + 1u], win32[a], (u & 3u) * 8u);
        }
        ring32[((r0 + hb + 4u * d) & kRingMask) >> 2] = v;
      }
      const uint32_t tb = hb + 4u * nd;
#pragma unroll
      for (uint32_t t = 0; t < 3u; t++)
        if (tb + t < kk) S.ring[(r0 + tb + t) & kRingMask] = lit ? win[s0 + tb + t] : (uint8_t)c;
      // the batch's last byte: the largest item end
      uint32_t e = lane < nk ? off + kk : 0u;
#pragma unroll
      for (uint32_t d = 32; d >= 1; d >>= 1) e = max(e, (uint32_t)__shfl_xor(e, (int)d, kWave));
      o.pos = uni(S.base[b]) + uni(e);
      out_store(o, lane, o.pos, false);
    }
    __syncthreads();
    const uint32_t last = uni(S.last);
    if (last != kNoLast && i >= last + 1u) break;
  }
  if (wave == 1) out_store(o, lane, o.pos, true);
  if (wave == 0 && lane == 0) {
    if (out_len) out_len[ci] = pos;
    if (status) status[ci] = err;
  }
}

// ------------------------------------------------------------------ LZSS

struct __align__(16) LzssLds {
  uint8_t ring[kRing];
  uint8_t win[kWin];
};

// Decodes one LZSS stream of `len` bytes into dst (at most cap bytes).
// A batch is up to eight groups (flag byte + 8 items), lane 8g + k holding
// item k of group g.
__device__ __forceinline__ int32_t lzss_stream(const uint8_t *src, uint32_t len, uint8_t *dst,
                                               uint32_t cap, LzssLds &S, uint32_t lane,
                                               uint32_t *produced) {
  Out o;
  o.dst = dst;
  o.ring = S.ring;
  o.cap = cap;
  o.pos = 0;
  o.flushed = 0;
  o.mis = (uint32_t)(reinterpret_cast<uintptr_t>(dst) & 15u);
  int32_t err = kOk;
  uint32_t at = 0;
  Window w = window_load(src, len, 0, lane);
  const uint32_t g = lane >> 3, k = lane & 7u;
  while (at < len) {
    __syncthreads();
    reinterpret_cast<uint4 *>(S.win)[lane] = w.v;
    __syncthreads();
    const uint32_t lead = uni(w.lead);
    const uint32_t h0 = reinterpret_cast<const uint32_t *>(S.win)[lane];
    const uint32_t h1 = reinterpret_cast<const uint32_t *>(S.win)[kWave + lane];
    const uint32_t qend = lead + (len - at);
    // walk over the flag bytes: group j's start -> lane j of gq
    uint32_t q = lead, ng = 0, gq = 0;
    while (ng < 8u && q < kHead && q < qend) {
      const uint32_t f = head_byte(h0, h1, q);
      gq = wrlane(gq, q, ng);
      ng++;
      q += 9u + (8u - (uint32_t)__builtin_popcount(f));
    }
    // items (lzss.h:268-303)
    const uint32_t gs = bperm(gq, g);
    const uint32_t f = (uint32_t)S.win[gs];
    const uint32_t ip = gs + 1u + 2u * k - (uint32_t)__builtin_popcount(f & ((1u << k) - 1u));
    const bool exists = g < ng && ip < qend;
    const bool lit = exists && ((f >> k) & 1u);
    const bool mat = exists && !lit;
    const uint32_t b0 = (uint32_t)S.win[ip], b1 = (uint32_t)S.win[ip + 1u];
    const uint32_t v = b0 | (b1 << 8);
    const uint32_t mlen = (v & 31u) + 3u, off = v >> 5;
    const uint32_t tl = lit ? 1u : (mat ? mlen : 0u);
    const uint32_t incl = wave_scan_dpp(tl);
    // an eighth group that would take the batch past kLzssBatch waits
    const uint32_t total = rdlane(incl, kWave - 1);
    uint32_t keep_g = ng;
    if (total > kLzssBatch) keep_g = 7u;
    const uint32_t x = o.pos + incl - tl;
    uint32_t e = 0;
    if (g < keep_g && exists) {
      if (mat && ip + 2u > qend) e = (uint32_t)kTrunc;
      else if (mat && (off == 0u || off > x)) e = (uint32_t)kCorrupt;
      else if (x + tl > cap) e = (uint32_t)kOverflow;
    }
    const uint64_t em = __ballot(e != 0u);
    uint32_t nk = keep_g * 8u;  // lanes kept (items past the stream end have tl 0)
    if (em) {
      const uint32_t first = (uint32_t)__builtin_ctzll(em);
      err = (int32_t)rdlane(e, first);
      nk = first;
    }
    const bool keep = lane < nk;
    const uint32_t pend = o.pos + (nk ? rdlane(incl, nk - 1u) : 0u);
    // consumed input: the kept groups (the last may end at the stream end)
    const uint32_t qn = keep_g < ng ? rdlane(gq, keep_g) : min(q, qend);
    at += qn - lead;
    if (err == kOk && at < len) w = window_load(src, len, at, lane);
    if (keep && lit) S.ring[ridx(o, x)] = (uint8_t)b0;
    // matches whose source ends before the batch: all lanes at once, 4
    // bytes per step (bytes past a match's end repeat its last byte's copy)
    const bool km = keep && mat;
    const bool ind = km && (x - off) + mlen <= o.pos;
    if (__ballot(ind)) {
      const uint32_t sp = x - off, last = mlen - 1u;
      for (uint32_t i = 0; i < 34u; i += 4u) {
        if (!__ballot(ind && i < mlen)) break;
        if (ind && i < mlen) {
          const uint32_t k1 = min(i + 1u, last), k2 = min(i + 2u, last), k3 = min(i + 3u, last);
          const uint8_t c0 = S.ring[ridx(o, sp + i)], c1 = S.ring[ridx(o, sp + k1)];
          const uint8_t c2 = S.ring[ridx(o, sp + k2)], c3 = S.ring[ridx(o, sp + k3)];
          S.ring[ridx(o, x + i)] = c0;
          S.ring[ridx(o, x + k1)] = c1;
          S.ring[ridx(o, x + k2)] = c2;
          S.ring[ridx(o, x + k3)] = c3;
        }
      }
    }
    // the others read bytes of this batch: in stream order, one per step
    uint64_t dm = __ballot(km && !ind);
    while (dm) {
      const uint32_t j = (uint32_t)__builtin_ctzll(dm);
      dm &= dm - 1;
      const uint32_t xj = rdlane(x, j), oj = rdlane(off, j), lj = rdlane(mlen, j);
      if (lane < lj) {
        // byte i copies source byte i, or i mod oj when the match overlaps
        // itself (lzss.h:285-288 copies forward one byte at a time)
        const uint32_t rm = kRecip.m[oj & 63u];  // used only when oj < lj <= 34
        const uint32_t m = oj >= lj ? lane : lane - ((lane * rm) >> 16) * oj;
        S.ring[ridx(o, xj + lane)] = S.ring[ridx(o, xj - oj + m)];
      }
    }
    o.pos = pend;
    out_store(o, lane, o.pos, false);
    if (err != kOk) break;
  }
  out_store(o, lane, o.pos, true);
  *produced = o.pos;
  return err;
}

__global__ void __launch_bounds__(kWave) lzss_chunks_kernel(const uint8_t *comp, uint64_t comp_bytes,
                                                            const uint64_t *ends, uint64_t n,
                                                            uint64_t chunk_bytes, uint8_t *out,
                                                            uint64_t *out_len, int32_t *status) {
  __shared__ LzssLds S;
  const uint32_t lane = lane_id();
  const uint64_t i = blockIdx.x;
  if (i >= n) return;
  const uint64_t lo = i ? ends[i - 1] : 0;  // lzss.h:237-238
  uint64_t len;
  int32_t st;
  chunk_bounds(lo, ends[i], comp_bytes, len, st);
  uint32_t produced = 0;
  if (st == kOk)
    st = lzss_stream(comp + lo, (uint32_t)len, out + i * chunk_bytes, (uint32_t)chunk_bytes, S, lane,
                     &produced);
  if (lane == 0) {
    if (out_len) out_len[i] = produced;
    if (status) status[i] = st;
  }
}

}  // namespace bytecodec

hipError_t launch_brle_chunks(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *ptr,
                              uint64_t n, uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                              int32_t *status, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(bytecodec::brle_chunks2_kernel, dim3((uint32_t)n), dim3(2 * kWaveSize), 0,
                     stream, comp, comp_bytes, ptr, n, chunk_bytes, out, out_len, status);
  return hipGetLastError();
}

hipError_t launch_lzss_chunks(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *ends,
                              uint64_t n, uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                              int32_t *status, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(bytecodec::lzss_chunks_kernel, dim3((uint32_t)n), dim3(kWaveSize), 0, stream,
                     comp, comp_bytes, ends, n, chunk_bytes, out, out_len, status);
  return hipGetLastError();
}

}  // namespace codag
