// rle.hip -- gfx950 decode kernels for ORC-style RLE v1 and RLE v2 chunks.
//
// One workgroup (kNT threads = 4 wave64s) decodes one chunk. The chunk's
// compressed stream is processed in windows staged into LDS:
//   1. stage   : all threads copy up to kStage bytes of the stream into LDS
//                with 16-byte loads;
//   2. index   : (RLE v1 only) every byte is classified as varint terminator
//                (< 0x80) or not, and a rank table (terminators before each
//                byte) and a select table (position of the r-th terminator)
//                are built with a workgroup prefix sum;
//   3. walk    : one lane follows the header chain through the window and
//                records one descriptor (stream position, output index) per
//                segment/block. With the rank/select tables an RLE v1 literal
//                group of k varints is skipped with two LDS reads;
//   4. expand  : the four waves take descriptors round-robin and expand
//                them lane-parallel: one lane per value, coalesced stores.
// The next window starts at the first segment that did not fit. A segment
// that cannot fit in an empty window is malformed (CODAG_CHUNK_CORRUPT).
//
// Semantics restated from the reference decoders:
//   RLE v1: include/rlev1/decoder.h:671-791 (decoder_warp_orig_rdw_fsm)
//   RLE v2: include/rlev2/decoder_new.h:909-1187 (decoder_warp_orig_rdw)
#include "device.hpp"
#include "kernels.hpp"

namespace codag {
namespace rle {

using namespace codag::dev;

// Launch shapes (threads per workgroup, LDS window bytes) are template
// parameters so the autotuner can choose among the instantiated ones.
constexpr uint32_t kPosBits = 13;             // descriptor: pos | out_idx << 13
constexpr uint32_t kMaxWaves = 16;

enum : int32_t { kOk = 0, kTrunc = 1, kOverflow = 2, kCorrupt = 3 };

struct Control {
  uint64_t next_abs;  // stream offset of the next segment/block
  uint32_t out_idx;   // values produced so far
  uint32_t ndesc;
  uint32_t nterm;
  uint32_t lead;
  uint32_t avail;     // lead + staged bytes
  int32_t status;
  uint32_t done;
  uint32_t bad;       // set by expanders (overlong varint / bad patch)
  uint32_t wave_tot[kMaxWaves];
};

// ------------------------------------------------------------------ RLE v1

template <uint32_t kStage> struct V1Lds {
  uint8_t sbuf[kStage + 32];
  uint16_t rank[kStage + 1];
  uint16_t sel[kStage];
  uint32_t desc[kStage / 2];  // a segment is at least 2 bytes
  Control c;
};

// Exclusive prefix sum over the workgroup (kNT values, one per thread).
template <uint32_t kNT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *wave_tot,
                                                    uint32_t *total) {
  constexpr uint32_t kWaves = kNT / kWave;
  const uint32_t lane = lane_id();
  const uint32_t wv = threadIdx.x / kWave;
  uint32_t incl = wave_incl_scan(v);
  if (lane == kWave - 1) wave_tot[wv] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
#pragma unroll
  for (uint32_t w = 0; w < kWaves; w++) {
    uint32_t t = wave_tot[w];
    if (w < wv) before += t;
    all += t;
  }
  *total = all;
  return before + incl - v;
}

template <uint32_t kNT, uint32_t kStage>
__global__ void __launch_bounds__(kNT)
rlev1_decode_kernel(const uint8_t *__restrict__ comp, uint64_t comp_bytes,
                    const uint64_t *__restrict__ blk_off, const uint64_t *__restrict__ col_len,
                    uint64_t n_chunks, uint32_t w, uint64_t chunk_bytes, uint8_t *__restrict__ out,
                    int32_t *__restrict__ status) {
  constexpr uint32_t kWaves = kNT / kWave;
  constexpr uint32_t kPer = kStage / kNT;  // index-build positions per thread
  constexpr uint32_t kMaxDesc = kStage / 2;
  static_assert(kPer % 4 == 0 && kPer <= 32 && kStage <= (1u << kPosBits), "launch shape");
  __shared__ V1Lds<kStage> L;
  const uint64_t chunk = blockIdx.x;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wv = tid / kWave;
  const uint64_t boff = blk_off[chunk];
  const uint64_t len = col_len[chunk];
  if (boff > comp_bytes || len > comp_bytes - boff) {  // stream outside the buffer
    if (tid == 0 && status) status[chunk] = kCorrupt;
    return;
  }
  const uint8_t *src = comp + boff;
  const uint32_t cap = (uint32_t)(chunk_bytes / w);
  const uint64_t mask = width_mask(w);
  uint8_t *dst = out + chunk * chunk_bytes;

  if (tid == 0) {
    L.c.next_abs = 0;
    L.c.out_idx = 0;
    L.c.status = kOk;
    L.c.done = 0;
    L.c.bad = 0;
  }
  for (;;) {
    __syncthreads();
    if (L.c.done || L.c.status != kOk) break;
    // ---- 1. stage
    const uint64_t pos = L.c.next_abs;
    const uint32_t n = (uint32_t)min<uint64_t>(len - pos, kStage - 16);
    uint32_t lead = 0;
    if (n) stage_to_lds(L.sbuf, src, pos, n, tid, kNT, &lead);
    else lead = 0;
    const uint32_t P = lead + n;
    __syncthreads();
    // ---- 2. index: terminator rank/select over positions [lead, P)
    {
      const uint32_t q0 = tid * kPer;
      const uint32_t *d = reinterpret_cast<const uint32_t *>(L.sbuf + q0);
      uint32_t m = 0;
#pragma unroll
      for (uint32_t k = 0; k < kPer / 4; k++) {
        uint32_t x = d[k];
        // bit j of t = (byte j of x) < 0x80
        uint32_t t = ~x & 0x80808080u;
        uint32_t nib = ((t >> 7) & 1) | ((t >> 14) & 2) | ((t >> 21) & 4) | ((t >> 28) & 8);
        m |= nib << (4 * k);
      }
      // restrict to [lead, P)
      uint32_t lo = lead > q0 ? min(lead - q0, kPer) : 0;
      uint32_t hi = P > q0 ? min(P - q0, kPer) : 0;
      uint32_t valid = (hi > lo) ? ((uint32_t)((1ull << (hi - lo)) - 1) << lo) : 0;
      m &= valid;
      uint32_t total;
      uint32_t excl = block_excl_scan<kNT>(__popc(m), L.c.wave_tot, &total);
#pragma unroll
      for (uint32_t j = 0; j < kPer; j++)
        L.rank[q0 + j] = (uint16_t)(excl + __popc(m & (uint32_t)((1ull << j) - 1)));
      uint32_t r = excl;
      while (m) {
        uint32_t j = __ffs(m) - 1;
        L.sel[r++] = (uint16_t)(q0 + j);
        m &= m - 1;
      }
      if (tid == 0) {
        L.c.nterm = total;
        L.c.lead = lead;
        L.c.avail = P;
      }
    }
    __syncthreads();
    // ---- 3. walk (one lane)
    if (tid == 0) {
      const uint32_t nterm = L.c.nterm;
      const bool at_end = (pos + n == len);
      uint32_t p = lead, o = L.c.out_idx, nd = 0;
      int32_t st = kOk;
      uint32_t done = 0;
      for (;;) {
        if (o >= cap) { done = 1; break; }
        if (p >= P) { if (at_end) st = kTrunc; break; }
        const int32_t h = (int8_t)L.sbuf[p];
        uint32_t cnt, t;
        if (h < 0) {  // literal group: -h varints (decoder.h:686-733)
          cnt = (uint32_t)(-h);
          t = (uint32_t)L.rank[p + 1] + cnt - 1;
        } else {      // run: delta byte + varint base (decoder.h:736-790)
          cnt = (uint32_t)h + 3;
          if (p + 2 > P) { if (at_end) st = kTrunc; break; }
          t = L.rank[p + 2];
        }
        if (t >= nterm) { if (at_end) st = kTrunc; break; }
        if (o + cnt > cap) { st = kOverflow; break; }
        L.desc[nd++] = p | (o << kPosBits);
        o += cnt;
        p = (uint32_t)L.sel[t] + 1;
        if (nd == kMaxDesc) break;
      }
      if (nd == 0 && !done && st == kOk) st = at_end ? kTrunc : kCorrupt;
      L.c.ndesc = nd;
      L.c.out_idx = o;
      L.c.next_abs = pos + (p - lead);
      L.c.status = st;
      L.c.done = done;
    }
    __syncthreads();
    // ---- 4. expand
    const uint32_t nd = L.c.ndesc;
    uint32_t bad = 0;
    for (uint32_t di = wv; di < nd; di += kWaves) {
      const uint32_t e = L.desc[di];
      const uint32_t p = e & ((1u << kPosBits) - 1);
      const uint32_t o = e >> kPosBits;
      const int32_t h = (int8_t)L.sbuf[p];
      if (h < 0) {
        const uint32_t k = (uint32_t)(-h);
        const uint32_t r0 = L.rank[p + 1];
        for (uint32_t i = lane; i < k; i += kWave) {
          const uint32_t end = L.sel[r0 + i];
          const uint32_t start = i ? (uint32_t)L.sel[r0 + i - 1] + 1 : p + 1;
          if (end - start >= 10) bad = 1;
          uint64_t v = 0;
          uint32_t sh = 0;
          for (uint32_t b = start; b <= end && sh < 64; b++, sh += 7)
            v |= (uint64_t)(L.sbuf[b] & 0x7f) << sh;
          store_elem(dst + (uint64_t)(o + i) * w, w, v & mask);
        }
      } else {
        const uint32_t cnt = (uint32_t)h + 3;
        const int64_t delta = (int8_t)L.sbuf[p + 1];
        const uint32_t start = p + 2;
        const uint32_t end = L.sel[L.rank[p + 2]];
        if (end - start >= 10) bad = 1;
        uint64_t v = 0;
        uint32_t sh = 0;
        for (uint32_t b = start; b <= end && sh < 64; b++, sh += 7)
          v |= (uint64_t)(L.sbuf[b] & 0x7f) << sh;
        for (uint32_t i = lane; i < cnt; i += kWave)
          store_elem(dst + (uint64_t)(o + i) * w, w, (v + (uint64_t)((int64_t)i * delta)) & mask);
      }
    }
    if (bad) atomicOr(&L.c.bad, 1u);
    __syncthreads();
    if (tid == 0 && L.c.bad && L.c.status == kOk) L.c.status = kCorrupt;
  }
  if (tid == 0 && status) status[chunk] = L.c.status;
}

// ------------------------------------------------------------------ RLE v2

__constant__ uint8_t kDecodeWidth[32] = {1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11,
                                         12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22,
                                         23, 24, 26, 28, 30, 32, 40, 48, 56, 64};

// include/rlev2/utils.h:42-48 (__CLOSEST_FIXED_BIT_MAP)
__device__ __forceinline__ uint32_t closest_fixed(uint32_t bits) {
  if (bits <= 1) return 1;
  if (bits <= 24) return bits;
  if (bits <= 32) return (bits + 1) & ~1u;
  return (bits + 7) & ~7u;
}

template <uint32_t kStage> struct V2Lds {
  uint8_t sbuf[kStage + 32];
  uint32_t desc[kStage / 2];  // a block is at least 2 bytes
  Control c;
};

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

template <uint32_t kNT, uint32_t kStage>
__global__ void __launch_bounds__(kNT)
rlev2_decode_kernel(const uint8_t *__restrict__ comp, uint64_t comp_bytes,
                    const uint64_t *__restrict__ blk_off, const uint64_t *__restrict__ col_len,
                    uint64_t n_chunks, uint32_t w, uint64_t chunk_bytes, uint8_t *__restrict__ out,
                    int32_t *__restrict__ status) {
  constexpr uint32_t kWaves = kNT / kWave;
  constexpr uint32_t kMaxDesc = kStage / 2;
  static_assert(kStage <= (1u << kPosBits), "launch shape");
  __shared__ V2Lds<kStage> L;
  const uint64_t chunk = blockIdx.x;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wv = tid / kWave;
  const uint64_t boff = blk_off[chunk];
  const uint64_t len = col_len[chunk];
  if (boff > comp_bytes || len > comp_bytes - boff) {
    if (tid == 0 && status) status[chunk] = kCorrupt;
    return;
  }
  const uint8_t *src = comp + boff;
  const uint32_t cap = (uint32_t)(chunk_bytes / w);
  const uint64_t mask = width_mask(w);
  const uint32_t nbits = 8 * w;
  uint8_t *dst = out + chunk * chunk_bytes;

  if (tid == 0) {
    L.c.next_abs = 0;
    L.c.out_idx = 0;
    L.c.status = kOk;
    L.c.done = (len == 0);
    L.c.bad = 0;
  }
  for (;;) {
    __syncthreads();
    if (L.c.done || L.c.status != kOk) break;
    const uint64_t pos = L.c.next_abs;
    const uint32_t n = (uint32_t)min<uint64_t>(len - pos, kStage - 16);
    uint32_t lead = 0;
    if (n) stage_to_lds(L.sbuf, src, pos, n, tid, kNT, &lead);
    const uint32_t P = lead + n;
    __syncthreads();
    // ---- walk (one lane): block sizes from the headers
    if (tid == 0) {
      const bool at_end = (pos + n == len);
      uint32_t p = lead, o = L.c.out_idx, nd = 0;
      int32_t st = kOk;
      uint32_t done = 0;
      for (;;) {
        if (pos + (p - lead) >= len) { done = 1; break; }  // decoder_new.h:1000
        if (p >= P) break;
        const uint32_t first = L.sbuf[p];
        uint32_t size = 0, cnt = 0;
        bool fits = true;
        switch (first >> 6) {
        case 0:  // SHORT_REPEAT
          size = 1 + ((first >> 3) & 7) + 1;
          cnt = (first & 7) + 3;
          break;
        case 1: {  // DIRECT
          if (p + 2 > P) { fits = false; break; }
          const uint32_t fbw = kDecodeWidth[(first >> 1) & 0x1f];
          const uint32_t len8 = ((((first & 1) << 8) | L.sbuf[p + 1]) + 1) & 0xff;
          cnt = len8;
          size = 2 + (len8 * fbw + 7) / 8;
        } break;
        case 3: {  // DELTA
          if (p + 2 > P) { fits = false; break; }
          const uint32_t enc = (first >> 1) & 0x1f;
          const uint32_t fbw = kDecodeWidth[enc];
          const uint32_t len8 = ((((first & 1) << 8) | L.sbuf[p + 1]) + 1) & 0xff;
          const uint32_t nd8 = (len8 - 2) & 0xff;
          uint32_t q = p + 2;
          for (int v = 0; v < 2 && fits; v++) {  // skip uvarint base, svarint delta
            const uint32_t q0 = q;
            while (q < P && L.sbuf[q] >= 0x80) q++;
            if (q >= P) fits = false;
            q++;
            if (fits && q - q0 > 10) st = kCorrupt;  // longer than any 64-bit varint
          }
          cnt = 2 + nd8;
          size = (q - p) + (enc ? (nd8 * fbw + 7) / 8 : 0);
        } break;
        default: {  // PATCHED_BASE
          if (p + 4 > P) { fits = false; break; }
          const uint32_t fbw = kDecodeWidth[(first >> 1) & 0x1f];
          const uint32_t len8 = ((((first & 1) << 8) | L.sbuf[p + 1]) + 1) & 0xff;
          const uint32_t third = L.sbuf[p + 2], fourth = L.sbuf[p + 3];
          const uint32_t bw = (third >> 5) + 1;
          const uint32_t pw = kDecodeWidth[third & 0x1f];
          const uint32_t pgw = (fourth >> 5) + 1;
          const uint32_t pll = fourth & 0x1f;
          if (pw + pgw > 64) { st = kCorrupt; break; }
          cnt = len8;
          size = 4 + bw + (len8 * fbw + 7) / 8 + (pll * closest_fixed(pw + pgw) + 7) / 8;
        } break;
        }
        if (st != kOk) break;
        if (!fits || p + size > P) { if (at_end) st = kTrunc; break; }
        if (o + cnt > cap) { st = kOverflow; break; }
        L.desc[nd++] = p | (o << kPosBits);
        o += cnt;
        p += size;
        if (nd == kMaxDesc) break;
      }
      if (nd == 0 && !done && st == kOk) st = at_end ? kTrunc : kCorrupt;
      L.c.ndesc = nd;
      L.c.out_idx = o;
      L.c.next_abs = pos + (p - lead);
      L.c.status = st;
      L.c.done = done;
    }
    __syncthreads();
    // ---- expand
    const uint32_t nd = L.c.ndesc;
    uint32_t bad = 0;
    for (uint32_t di = wv; di < nd; di += kWaves) {
      const uint32_t e = L.desc[di];
      const uint32_t p = e & ((1u << kPosBits) - 1);
      const uint32_t o = e >> kPosBits;
      const uint32_t first = L.sbuf[p];
      const uint32_t kind = first >> 6;
      if (kind == 0) {  // SHORT_REPEAT: big-endian W bytes (decoder_new.h:1003-1015)
        const uint32_t nb = ((first >> 3) & 7) + 1;
        const uint32_t cnt = (first & 7) + 3;
        const uint64_t v = lds_be(L.sbuf + p + 1, nb) & mask;
        if (lane < cnt) store_elem(dst + (uint64_t)(o + lane) * w, w, v);
        continue;
      }
      const uint32_t code = (first >> 1) & 0x1f;
      const uint32_t fbw = kDecodeWidth[code];
      const uint32_t len8 = ((((first & 1) << 8) | L.sbuf[p + 1]) + 1) & 0xff;
      if (kind == 1) {  // DIRECT (decoder_new.h:1016-1044)
        const uint32_t b0 = (p + 2) * 8;
        for (uint32_t i = lane; i < len8; i += kWave)
          store_elem(dst + (uint64_t)(o + i) * w, w, lds_bits_be(L.sbuf, b0 + i * fbw, fbw) & mask);
      } else if (kind == 3) {  // DELTA (decoder_new.h:1045-1093)
        uint32_t q = p + 2;
        const uint64_t base = lds_uvarint(L.sbuf, &q) & mask;
        const int64_t zz = (int64_t)lds_uvarint(L.sbuf, &q);
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
          for (uint32_t c = 0; c < nd8; c += kWave) {
            const uint32_t i = c + lane;
            const uint64_t d = i < nd8 ? (lds_bits_be(L.sbuf, b0 + i * fbw, fbw) & mask) : 0;
            const uint64_t s = wave_incl_scan64(d);
            const uint64_t v = neg ? carry - s : carry + s;
            if (i < nd8) store_elem(dst + (uint64_t)(o + 2 + i) * w, w, v & mask);
            const uint64_t tot = __shfl(s, kWave - 1, kWave);
            carry = neg ? carry - tot : carry + tot;
          }
        }
      } else {  // PATCHED_BASE (decoder_new.h:1094-1183)
        const uint32_t third = L.sbuf[p + 2], fourth = L.sbuf[p + 3];
        const uint32_t bw = (third >> 5) + 1;
        const uint32_t pw = kDecodeWidth[third & 0x1f];
        const uint32_t pgw = (fourth >> 5) + 1;
        const uint32_t pll = fourth & 0x1f;
        const uint64_t pmask = pw >= 32 ? 0xffffffffull : ((1ull << pw) - 1);  // uint32 mask
        const uint32_t cfb = closest_fixed(pw + pgw);
        const uint64_t base = lds_be(L.sbuf + p + 4, bw) & mask;
        const uint32_t vb0 = (p + 4 + bw) * 8;
        const uint32_t pb0 = (p + 4 + bw + (len8 * fbw + 7) / 8) * 8;
        // lane j < pll: entry j -> running gap (wave scan) and its OR-term
        uint64_t ent = lane < pll ? lds_bits_be(L.sbuf, pb0 + lane * cfb, cfb) : 0;
        uint32_t g = lane < pll ? (uint32_t)(pw >= 64 ? 0 : (ent >> pw)) : 0;
        uint32_t gsum = wave_incl_scan(g);
        uint64_t add = (fbw >= nbits) ? 0 : (((ent & pmask) << fbw) & mask);
        if (lane < pll && gsum >= len8) bad = 1;
        // all lanes stay active around the cross-lane reads (a lane shuffle
        // from an inactive lane returns 0)
        for (uint32_t c = 0; c < len8; c += kWave) {
          const uint32_t i = c + lane;
          uint64_t v = i < len8 ? (lds_bits_be(L.sbuf, vb0 + i * fbw, fbw) & mask) : 0;
          for (uint32_t j = 0; j < pll; j++) {
            const uint32_t gj = __shfl(gsum, j, kWave);
            const uint64_t aj = __shfl(add, j, kWave);
            if (gj == i) v |= aj;
          }
          if (i < len8) store_elem(dst + (uint64_t)(o + i) * w, w, (v + base) & mask);
        }
      }
    }
    if (bad) atomicOr(&L.c.bad, 1u);
    __syncthreads();
    if (tid == 0 && L.c.bad && L.c.status == kOk) L.c.status = kCorrupt;
  }
  __syncthreads();
  // positions the stream did not produce are zero (decoder_trans.h:81)
  const uint32_t o_end = L.c.out_idx;
  if (L.c.status == kOk)
    for (uint32_t i = o_end + tid; i < cap; i += kNT) store_elem(dst + (uint64_t)i * w, w, 0);
  if (tid == 0 && status) status[chunk] = L.c.status;
}

}  // namespace rle

// ------------------------------------------------------------------ launchers

#define CODAG_RLE_SHAPES(X) X(256, 4096) X(128, 4096) X(256, 8192) X(512, 8192) X(512, 4096)

static bool rle_shape_ok(uint32_t nt, uint32_t stage) {
#define X(a, b) if (nt == a && stage == b) return true;
  CODAG_RLE_SHAPES(X)
#undef X
  return false;
}

hipError_t launch_rlev1_decode(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                               const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                               uint64_t chunk_bytes, uint8_t *out, int32_t *status,
                               uint32_t nt, uint32_t stage, hipStream_t stream) {
  if (n_chunks == 0) return hipSuccess;
  if (!rle_shape_ok(nt, stage)) return hipErrorInvalidValue;
#define X(a, b)                                                                                 \
  if (nt == a && stage == b)                                                                    \
    hipLaunchKernelGGL((rle::rlev1_decode_kernel<a, b>), dim3((uint32_t)n_chunks), dim3(a), 0,   \
                       stream, comp, comp_bytes, blk_off, col_len, n_chunks, w, chunk_bytes, out, \
                       status);
  CODAG_RLE_SHAPES(X)
#undef X
  return hipGetLastError();
}

hipError_t launch_rlev2_decode(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                               const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                               uint64_t chunk_bytes, uint8_t *out, int32_t *status,
                               uint32_t nt, uint32_t stage, hipStream_t stream) {
  if (n_chunks == 0) return hipSuccess;
  if (!rle_shape_ok(nt, stage)) return hipErrorInvalidValue;
#define X(a, b)                                                                                 \
  if (nt == a && stage == b)                                                                    \
    hipLaunchKernelGGL((rle::rlev2_decode_kernel<a, b>), dim3((uint32_t)n_chunks), dim3(a), 0,   \
                       stream, comp, comp_bytes, blk_off, col_len, n_chunks, w, chunk_bytes, out, \
                       status);
  CODAG_RLE_SHAPES(X)
#undef X
  return hipGetLastError();
}

bool rle_shape_supported(uint32_t nt, uint32_t stage) { return rle_shape_ok(nt, stage); }

}  // namespace codag
