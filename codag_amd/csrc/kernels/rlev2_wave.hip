// rlev2_wave.hip -- RLE v2 decode kernels built around a scalar header walk.
//
// The chunk's stream is staged into an LDS window (kWin bytes). The header
// walk is a serial chain (a block's size gives the next header), run by one
// wave in span steps: every lane sizes the block a header at p + lane would
// start, and the scalar unit follows the chain through those sizes with one
// v_readlane per block; a wave prefix sum places the chained blocks' values.
// Headers the span step does not take (errors, window edges, long DELTA
// varints) go through a one-block scalar path that reads them from a VGPR
// mirror of the 256 bytes around the walk position. A walk collects up to 64
// block descriptors in LDS; expansion then writes them:
//   * SHORT_REPEAT blocks (3..10 values) expand one lane per block;
//   * DIRECT / DELTA / PATCHED_BASE blocks expand across a wave (expand_big:
//     one lane per value, wave prefix sums for DELTA, a wave scan of the
//     patch gaps for PATCHED_BASE).
// A block that does not fit in the window restarts the window at that block.
//
// Two kernels share the walk and the expansion:
//   * rlev2_wave_kernel: one wave per chunk walks and expands in turn, with no
//     workgroup barriers; several chunks share each SIMD to hide latency.
//   * rlev2_pc_kernel: one walker wave and kExp expander waves per chunk. The
//     walker publishes batches of block descriptors through a four-deep LDS
//     queue and walks ahead while the expanders write earlier batches, so a
//     chunk's time approaches max(walk, expand) instead of their sum.
//
// Semantics: include/rlev2/decoder_new.h:909-1187 (decoder_warp_orig_rdw).
#include "kernels.hpp"
#include "rlev2_block.hpp"

namespace codag {
namespace rle2 {

enum : int32_t { kOk = 0, kTrunc = 1, kOverflow = 2, kCorrupt = 3 };

// Eight stream bytes starting at LDS offset p, little-endian, taken from the
// VGPR window `vw` (lane l = LDS dword vbase/4 + l). Needs p+12 <= vbase+256.
__device__ __forceinline__ uint64_t win8(uint32_t vw, uint32_t vbase, uint32_t p) {
  const uint32_t k = (p - vbase) >> 2;
  const uint32_t sh = (p & 3) * 8;
  const uint64_t lo = (uint64_t)(uint32_t)__builtin_amdgcn_readlane(vw, k) |
                      ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(vw, k + 1) << 32);
  if (!sh) return lo;
  const uint64_t hi = (uint32_t)__builtin_amdgcn_readlane(vw, k + 2);
  return (lo >> sh) | (hi << (64 - sh));
}

// Byte length of the LEB128 varint at LDS offset q: the first byte < 0x80
// ends it. Uses the VGPR window for up to 24 bytes, LDS beyond that.
__device__ __forceinline__ uint32_t varint_len(const uint8_t *sbuf, uint32_t vw, uint32_t vbase,
                                               uint32_t q, uint32_t limit) {
  for (uint32_t k = 0; k < 24; k += 8) {
    if (q + k + 12 > vbase + 256) break;
    const uint64_t t = ~win8(vw, vbase, q + k) & 0x8080808080808080ull;
    if (t) return k + (uint32_t)(__builtin_ctzll(t) >> 3) + 1;
  }
  uint32_t n = 0;  // long (malformed) varint: byte loop, bounded by the window
  while (q + n < limit && sbuf[q + n] >= 0x80) n++;
  return n + 1;
}

// Walk state of one chunk. Stream offsets are 32-bit: a chunk's stream is far
// below 4 GiB (longer streams are rejected up front).
struct Walk {
  const uint8_t *src;  // chunk stream
  uint32_t len;        // stream bytes
  uint32_t cap;        // values in the chunk
  uint32_t wpos;       // stream offset held at sbuf[lead]
  uint32_t lead, P;    // window: bytes [lead, P) of sbuf hold the stream
  uint32_t at_end;     // the window holds the stream's end
  // The one-block path stops at p >= pdone (stream consumed: P when the
  // window holds the stream's end) and restages at p > pstop (header bytes
  // might lie past the window: P - 32 when it does not).
  uint32_t pdone, pstop;
  uint32_t p, o;       // walk position (LDS offset) and values produced
  uint32_t stop;       // outcome of the last walk_batch (kStop*)
  int32_t err;         // error status of the last walk_batch
  uint32_t vbase, vw;  // VGPR mirror of LDS bytes [vbase, vbase + 256)
};

// Stages stream bytes [pos, ...) into the window with the 64 lanes of the
// calling wave and restarts the walk at pos. The caller orders it against
// other waves' reads of the window.
template <uint32_t kWin>
__device__ __forceinline__ void walk_stage(Walk &s, uint8_t *sbuf, uint32_t pos, uint32_t lane) {
  const uint32_t n = (s.len - pos) < (kWin - 16) ? (s.len - pos) : (kWin - 16);
  s.lead = 0;
  if (n) stage_to_lds(sbuf, s.src, pos, n, lane, kWave, &s.lead);
  s.P = s.lead + n;
  s.at_end = (pos + n == s.len);
  s.wpos = pos;
  s.pdone = s.at_end ? s.P : 0xffffffffu;
  s.pstop = s.at_end ? 0xffffffffu : s.P - 32;
  s.p = s.lead;
}

// Reloads the VGPR mirror at the walk position (after the window changed).
__device__ __forceinline__ void walk_mirror(Walk &s, const uint8_t *sbuf, uint32_t lane) {
  s.vbase = s.p & ~3u;
  s.vw = reinterpret_cast<const uint32_t *>(sbuf)[s.vbase / 4 + lane];
}

enum : uint32_t { kStopNone = 0, kStopDone = 1, kStopRestage = 2 };

constexpr uint32_t kSpan = kWave - 1;  // header positions one span step examines

// Byte size and value count of the block whose header would start at LDS
// offset q = p + lane (decoder_new.h:1003-1183 read the same header fields),
// and whether the span step may take it: the block ends inside the window
// (q + size <= P) and is none of the cases left to the one-block path (a
// PATCHED_BASE width sum past 64; a DELTA whose varints are not both found in
// the 64 bytes after its header, or one of them longer than 10 bytes). Bytes
// past P never make a position acceptable: any field they feed puts the
// block's end past P. Called by all 64 lanes (the varint search is a ballot
// over the span's bytes).
__device__ __forceinline__ void span_block(const uint8_t *sbuf, uint32_t p, uint32_t lane,
                                           uint32_t P, uint32_t &size, uint32_t &cnt, bool &ok) {
  const uint8_t *h = sbuf + p + lane;
  const uint32_t first = h[0], second = h[1], third = h[2], fourth = h[3];
  const uint32_t far = h[64];
  // varint terminators (bytes below 0x80) among bytes p .. p+127; lane l
  // looks at bytes q+2 .. q+65
  const uint64_t t0 = __ballot(first < 0x80), t1 = __ballot(far < 0x80);
  const uint32_t vs = lane + 2;  // 2 .. 65
  const uint64_t m = vs < 64 ? (t0 >> vs) | (t1 << (64 - vs)) : (t1 >> (vs - 64));
  // select forms of decode_width / closest_fixed: per-lane selects, no
  // divergent branches
  auto width = [](uint32_t k) { return k < 24 ? k + 1 : (k < 28 ? 2 * k - 22 : 8 * k - 184); };
  auto closest = [](uint32_t b) {  // b >= 2 here
    return b <= 24 ? b : (b <= 32 ? (b + 1) & ~1u : (b + 7) & ~7u);
  };
  const uint32_t kind = first >> 6;
  const uint32_t code = (first >> 1) & 0x1f;
  const uint32_t fbw = width(code);
  // the 9-bit length field plus one, kept in 8 bits (decoder_new.h:1018)
  const uint32_t len8 = (second + 1) & 0xff;
  const uint32_t body = (len8 * fbw + 7) >> 3;
  // PATCHED_BASE: base width, patch width, gap width, patch list length
  const uint32_t pw = width(third & 0x1f), pgw = (fourth >> 5) + 1;
  const uint32_t pb_size =
      5 + (third >> 5) + body + (((fourth & 0x1f) * closest(pw + pgw) + 7) >> 3);
  // DELTA: the varints of base and delta base end at the first two
  // terminators from byte q+2 (e1, e2: their indexes from q+2)
  const uint32_t e1 = m ? (uint32_t)__builtin_ctzll(m) : 64u;
  const uint64_t m2 = m & (m - 1);
  const uint32_t e2 = m2 ? (uint32_t)__builtin_ctzll(m2) : 64u;
  const uint32_t nd8 = (len8 - 2) & 0xff;
  const uint32_t delta_size = 3 + e2 + (code ? (nd8 * fbw + 7) >> 3 : 0u);
  uint32_t sz = 2 + ((first >> 3) & 7), c = (first & 7) + 3;  // SHORT_REPEAT
  bool good = true;
  sz = kind == 1 ? 2 + body : sz;
  sz = kind == 2 ? pb_size : sz;
  sz = kind == 3 ? delta_size : sz;
  c = kind == 0 ? c : (kind == 3 ? 2 + nd8 : len8);
  good = kind == 2 ? pw + pgw <= 64 : good;
  // a varint over 10 bytes is the one-block path's error to report
  good = kind == 3 ? e2 < 64 && e1 < 10 && e2 - e1 <= 10 : good;
  size = sz;
  cnt = c;
  ok = good && p + lane + sz <= P;
}

// Descriptor j of a batch: LDS offset and first value index in two arrays,
// or packed into one word (offset | index << 13: offsets below 2^13 in a
// 4 KiB window, value indexes below 2^19).
template <bool kPacked>
__device__ __forceinline__ void put_desc(uint32_t *ldpos, uint32_t *ldout, uint32_t j, uint32_t pos,
                                         uint32_t out) {
  if (kPacked) {
    ldpos[j] = pos | (out << 13);
  } else {
    ldpos[j] = pos;
    ldout[j] = out;
  }
}

// Collects up to 64 blocks into the LDS descriptor arrays: block j's LDS
// offset goes to dpos[j] and its first value index to dout[j]. Sets s.stop
// to kStopDone when the stream is consumed or kStopRestage when the next
// header or block may lie past the window, and s.err on an error. (Outcomes
// travel as integers in the walk state: flags passed by reference would be
// kept in scratch memory.)
//
// Span step: each lane l < 63 sizes the block a header at p + l would start
// (span_block), so the walk from p is a chain of lane indices -- next(l) = l
// + size, lane 63 standing for "past the span" and a position the step does
// not take pointing at itself. The scalar walker follows that chain with one
// v_readlane per block; a wave prefix sum of the chained blocks' value counts
// gives their output indices, and the chained lanes write their descriptors.
// The first block that would overflow the chunk or the batch ends the step;
// positions the step does not take go through the one-block path below,
// which owns every error and window decision.
template <bool kPacked = false>
__device__ __forceinline__ uint32_t walk_batch(Walk &s, const uint8_t *sbuf, uint32_t lane,
                                               uint32_t *ldpos, uint32_t *ldout) {
  const uint32_t *sdw = reinterpret_cast<const uint32_t *>(sbuf);
  uint32_t nd = 0;
  // The walk state is wave-uniform; saying so keeps the loop in scalar
  // registers with scalar branches (the compiler cannot prove it through the
  // window bookkeeping).
  uint32_t p = uni(s.p), o = uni(s.o), vbase = uni(s.vbase), vw = s.vw;
  uint32_t stop = kStopNone;
  int32_t err = kOk;
  s.P = uni(s.P);
  s.cap = uni(s.cap);
  s.pdone = uni(s.pdone);
  s.pstop = uni(s.pstop);
  while (nd < (uint32_t)kWave) {
    {  // ---- span step over the headers at p .. p + 62
      uint32_t sz, c;
      bool ok;
      span_block(sbuf, p, lane, s.P, sz, c, ok);
      const uint32_t nx = lane + sz;  // next header, relative to p
      const uint32_t tab = (lane < kSpan && ok) ? (nx < kSpan ? nx : kSpan) : lane;
      // tab rises strictly along the chain until a fixed point: four links
      // per test
      uint32_t r = 0, rf = kSpan;
      uint64_t mark = 0;
      for (;;) {
        const uint32_t r1 = (uint32_t)__builtin_amdgcn_readlane(tab, r);
        const uint32_t r2 = (uint32_t)__builtin_amdgcn_readlane(tab, r1);
        const uint32_t r3 = (uint32_t)__builtin_amdgcn_readlane(tab, r2);
        const uint32_t r4 = (uint32_t)__builtin_amdgcn_readlane(tab, r3);
        mark |= (1ull << r) | (1ull << r1) | (1ull << r2) | (1ull << r3);
        if (r4 == r3) {
          rf = r4;
          break;
        }
        r = r4;
      }
      mark &= ~(1ull << rf);  // the fixed point is not taken here
      const bool blk = (mark >> lane) & 1;
      const uint32_t cl = blk ? c : 0u;
      const uint32_t incl = wave_scan_dpp(cl);
      const uint32_t rank = lanes_below(mark);
      // the first block past the chunk's values or the batch's 64 slots ends
      // the step (the one-block path reports the overflow)
      const uint64_t cut = __ballot(blk && (o + incl > s.cap || nd + rank >= (uint32_t)kWave));
      if (cut) mark &= (1ull << __builtin_ctzll(cut)) - 1;
      if (mark) {
        if ((mark >> lane) & 1) {
          put_desc<kPacked>(ldpos, ldout, nd + rank, p + lane, o + incl - cl);
        }
        const uint32_t last = 63 - (uint32_t)__builtin_clzll(mark);
        o += (uint32_t)__builtin_amdgcn_readlane(incl, last);
        p += (uint32_t)__builtin_amdgcn_readlane(nx, last);
        nd += (uint32_t)__builtin_popcountll(mark);
      }
      if (nd >= (uint32_t)kWave) break;
      if (rf == kSpan && !cut) continue;  // the chain left the span
    }
    // ---- one-block path: the block at p, which the span step did not take
    if (p >= s.pdone) {  // decoder_new.h:1000 -- stream consumed
      stop = kStopDone;
      break;
    }
    if (p > s.pstop) {  // header bytes might lie past the window
      stop = kStopRestage;
      break;
    }
    if (p - vbase > 224) {
      vbase = p & ~3u;
      vw = sdw[vbase / 4 + lane];
    }
    // four header bytes from two adjacent dwords of the VGPR mirror
    const uint32_t k = (p - vbase) >> 2;
    const uint64_t pair = (uint64_t)(uint32_t)__builtin_amdgcn_readlane(vw, k) |
                          ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(vw, k + 1) << 32);
    const uint32_t x = (uint32_t)(pair >> ((p & 3) * 8));
    const uint32_t first = x & 0xff;
    uint32_t size, cnt;
    if (first < 64) {  // SHORT_REPEAT
      size = 2 + ((first >> 3) & 7);
      cnt = (first & 7) + 3;
    } else {
      const uint32_t kind = first >> 6;
      const uint32_t code = (first >> 1) & 0x1f;
      const uint32_t fbw = decode_width(code);
      const uint32_t len8 = ((((first & 1) << 8) | (uint32_t)((x >> 8) & 0xff)) + 1) & 0xff;
      if (kind == 1) {  // DIRECT
        size = 2 + (len8 * fbw + 7) / 8;
        cnt = len8;
      } else if (kind == 3) {  // DELTA: two varints then packed deltas
        const uint32_t nd8 = (len8 - 2) & 0xff;
        uint32_t q = p + 2;
        const uint32_t l1 = varint_len(sbuf, vw, vbase, q, s.P);
        const uint32_t l2 = q + l1 < s.P ? varint_len(sbuf, vw, vbase, q + l1, s.P) : 1;
        if (q + l1 + l2 <= s.P && (l1 > 10 || l2 > 10)) {  // longer than any 64-bit varint
          err = kCorrupt;
          break;
        }
        q += l1 + l2;
        size = (q - p) + (code ? (nd8 * fbw + 7) / 8 : 0);
        cnt = 2 + nd8;
      } else {  // PATCHED_BASE
        const uint32_t third = (x >> 16) & 0xff, fourth = (x >> 24) & 0xff;
        const uint32_t pw = decode_width(third & 0x1f);
        const uint32_t pgw = (fourth >> 5) + 1;
        if (pw + pgw > 64) {  // beyond the 65-entry closest-width table
          err = kCorrupt;
          break;
        }
        size = 4 + (third >> 5) + 1 + (len8 * fbw + 7) / 8 +
               ((fourth & 0x1f) * closest_fixed(pw + pgw) + 7) / 8;
        cnt = len8;
      }
    }
    if (p + size > s.P) {
      if (s.at_end) err = kTrunc;
      else stop = kStopRestage;
      break;
    }
    if (o + cnt > s.cap) {
      err = kOverflow;
      break;
    }
    if (lane == 0) {
      put_desc<kPacked>(ldpos, ldout, nd, p, o);
    }
    nd++;
    o += cnt;
    p += size;
  }
  s.stop = stop;
  s.err = err;
  s.p = p;
  s.o = o;
  s.vbase = vbase;
  s.vw = vw;
  return nd;
}

// Writes the SHORT_REPEAT blocks of a batch (lane j holds block j's
// dpos/dout, j < nd) that `sel` selects, for a column of W-byte elements, and
// returns the mask of the batch's other blocks (every lane < nd looks at its
// block). All 64 lanes of the wave must call it.
template <uint32_t W>
__device__ __forceinline__ uint64_t expand_sr(const uint8_t *sbuf, uint32_t nd, uint64_t sel,
                                              uint32_t dpos, uint32_t dout, uint8_t *dst,
                                              uint32_t lane) {
  const bool in = lane < nd;
  const bool mine = in && ((sel >> lane) & 1);
  // a SHORT_REPEAT block (header + at most 8 value bytes) lies inside the
  // three dwords from its header's dword: one read for header and value
  const uint32_t *sdw = reinterpret_cast<const uint32_t *>(sbuf) + ((in ? dpos : 0u) >> 2);
  const uint32_t d0 = sdw[0], d1 = sdw[1], d2 = sdw[2];
  const uint32_t k = (dpos & 3) + 1;  // first value byte, from the dword start
  const uint32_t first = in ? (d0 >> (8 * (dpos & 3))) & 0xff : 0u;
  if (mine && first < 64) {  // SHORT_REPEAT: big-endian value, 3..10 copies (decoder_new.h:1003-1015)
    const uint32_t nb = ((first >> 3) & 7) + 1;
    const uint32_t c = (first & 7) + 3;
    const uint32_t lo = k == 4 ? d1 : __builtin_amdgcn_alignbyte(d1, d0, k);
    const uint32_t hi = k == 4 ? d2 : __builtin_amdgcn_alignbyte(d2, d1, k);
    const auto v = (typename Col<W>::T)(__builtin_bswap64(((uint64_t)hi << 32) | lo) >> (64 - 8 * nb)) &
                   Col<W>::kMask;
    for (uint32_t q = 0; q < c; q++) store_w<W>(dst, dout + q, v);
  }
  return __ballot(in && first >= 64);
}

// Expands the blocks of a batch that `sel` selects.
template <uint32_t W>
__device__ __forceinline__ void expand_batch(const uint8_t *sbuf, uint32_t nd, uint64_t sel,
                                             uint32_t dpos, uint32_t dout, uint8_t *dst,
                                             uint32_t lane, uint32_t *bad) {
  uint64_t big = expand_sr<W>(sbuf, nd, sel, dpos, dout, dst, lane) & sel;
  while (big) {
    const uint32_t j = (uint32_t)__builtin_ctzll(big);
    big &= big - 1;
    const uint32_t bp = (uint32_t)__builtin_amdgcn_readlane(dpos, j);
    const uint32_t bo = (uint32_t)__builtin_amdgcn_readlane(dout, j);
    expand_big<W>(sbuf, bp, bo, dst, lane, bad);
  }
}

// Validates the chunk's metadata; false (status written) when the stream
// lies outside comp or exceeds the 32-bit offsets.
__device__ __forceinline__ bool chunk_ok(uint64_t boff, uint64_t len, uint64_t comp_bytes) {
  return boff <= comp_bytes && len <= comp_bytes - boff && len < 0xffffffffull;
}

template <uint32_t kWin> struct WaveLds {
  uint8_t sbuf[kWin + 288];  // window + slack for the 256-byte VGPR mirror
  uint32_t dpos[kWave];      // the batch's block descriptors
  uint32_t dout[kWave];
};

template <uint32_t kWin, uint32_t W>
__global__ void __launch_bounds__(kWave)
rlev2_wave_kernel(const uint8_t *__restrict__ comp, uint64_t comp_bytes,
                  const uint64_t *__restrict__ blk_off, const uint64_t *__restrict__ col_len,
                  uint64_t n_chunks, uint64_t chunk_bytes, uint8_t *__restrict__ out,
                  int32_t *__restrict__ status) {
  __shared__ WaveLds<kWin> L;
  const uint64_t chunk = blockIdx.x;
  const uint32_t lane = lane_id();
  const uint64_t boff = blk_off[chunk];
  const uint64_t len = col_len[chunk];
  if (!chunk_ok(boff, len, comp_bytes)) {
    if (lane == 0 && status) status[chunk] = kCorrupt;
    return;
  }
  Walk s;
  s.src = comp + boff;
  s.len = (uint32_t)len;
  s.cap = (uint32_t)(chunk_bytes / W);
  s.o = 0;
  uint8_t *dst = out + chunk * chunk_bytes;

  walk_stage<kWin>(s, L.sbuf, 0, lane);
  __syncthreads();
  walk_mirror(s, L.sbuf, lane);
  int32_t st = kOk;
  uint32_t bad = 0;
  bool done = (len == 0);
  while (!done && st == kOk) {
    const uint32_t p0 = s.p;
    const uint32_t nd = walk_batch(s, L.sbuf, lane, L.dpos, L.dout);
    asm volatile("" ::: "memory");  // descriptors written by other lanes of this wave
    const uint32_t dpos = L.dpos[lane], dout = L.dout[lane];
    done = s.stop == kStopDone;
    const bool restage = s.stop == kStopRestage;
    st = s.err;
    expand_batch<W>(L.sbuf, nd, ~0ull, dpos, dout, dst, lane, &bad);
    // A block never exceeds ~2.3 KB, so a fresh window always makes progress;
    // no progress would mean a malformed stream.
    if (restage && nd == 0 && p0 == s.lead) st = kCorrupt;
    if (restage && st == kOk) {
      __syncthreads();  // this wave's LDS reads are done before the overwrite
      walk_stage<kWin>(s, L.sbuf, s.wpos + (s.p - s.lead), lane);
      __syncthreads();
      walk_mirror(s, L.sbuf, lane);
    }
  }
  if (__ballot(bad) && st == kOk) st = kCorrupt;
  // positions the stream did not produce are zero (decoder_trans.h:81)
  if (st == kOk)
    for (uint32_t i = s.o + lane; i < s.cap; i += kWave) store_w<W>(dst, i, 0);
  if (lane == 0 && status) status[chunk] = st;
}

// ---------------------------------------------------------------- walker + expanders

constexpr uint32_t kBatchLast = 1u << 30;  // no batch follows this one
constexpr uint32_t kBatchWin = 1u << 28;   // the batch's blocks lie in window 1
constexpr uint32_t kQ = 4;                 // descriptor buffers in flight
// Waits sleep between polls: the scalar instructions of a spin are taken from
// the other waves of the CU. The walker sleeps 2048 cycles per poll (a full
// queue frees a buffer once per batch), an expander 1024 (a batch takes tens
// of thousands of cycles). A wait that outlasts its poll limit (about 1.3e8
// cycles, far past any batch) ends the chunk with kCorrupt.
constexpr uint32_t kWalkerSleep = 32, kWalkerSpinLimit = 1u << 16;
constexpr uint32_t kExpSleep = 16, kSpinLimit = 1u << 17;

template <uint32_t kWin> struct PcLds {
  uint8_t sbuf[2][kWin + 288];  // two stream windows
  uint32_t desc[kQ][kWave];     // packed descriptors (put_desc)
  uint32_t info[kQ];            // nd | kBatchLast | kBatchWin
  uint32_t pub;                 // batches published by the walker
  uint32_t done[3];             // batches finished, per expander wave
  uint32_t bad;                 // an expander found a patch outside its block
  uint32_t stuck;               // a wait ran past its poll limit
};

__device__ __forceinline__ uint32_t lds_acquire(const uint32_t *a) {
  return uni(__hip_atomic_load(a, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP));
}
__device__ __forceinline__ void lds_release(uint32_t *a, uint32_t v, uint32_t lane) {
  if (lane == 0) __hip_atomic_store(a, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Queue protocol (no workgroup barriers after the first): the walker writes
// batch k into descriptor buffer k % kQ and publishes it by raising `pub` to
// k + 1 (release: the descriptors are in LDS first); expander e waits for pub
// > k, writes the batch and raises done[e] to k + 1 once its reads of the
// buffer and the window are over. The walker reuses buffer k % kQ only when
// every done[] has passed k - kQ, and stages a window only when every batch
// that used it is done, so it can run up to kQ batches ahead and the
// per-batch differences between walking and expanding average out. Every
// wait is bounded (the poll limits); a wait that runs out sets `stuck`, which ends
// every wave's loop.
template <uint32_t kWin, uint32_t kExp, uint32_t W, bool kNoWait>
__global__ void __launch_bounds__(kWave *(1 + kExp))
rlev2_pc_kernel(const uint8_t *__restrict__ comp, uint64_t comp_bytes,
                const uint64_t *__restrict__ blk_off, const uint64_t *__restrict__ col_len,
                uint64_t n_chunks, uint64_t chunk_bytes, uint8_t *__restrict__ out,
                int32_t *__restrict__ status) {
  static_assert(kWin <= 4096 && kExp <= 3, "packed descriptors and the done[] array");
  // poll limits: compile-time constants (a runtime limit costs registers and
  // 12% on the 64-bit column); kNoWait, for a test, makes every wait not met
  // at once run out
  constexpr uint32_t walker_limit = kNoWait ? 0u : kWalkerSpinLimit;
  constexpr uint32_t exp_limit = kNoWait ? 0u : kSpinLimit;
  __shared__ PcLds<kWin> L;
  const uint64_t chunk = blockIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t boff = blk_off[chunk];
  const uint64_t len = col_len[chunk];
  if (!chunk_ok(boff, len, comp_bytes)) {  // uniform over the workgroup
    if (threadIdx.x == 0 && status) status[chunk] = kCorrupt;
    return;
  }
  const uint32_t cap = (uint32_t)(chunk_bytes / W);
  uint8_t *dst = out + chunk * chunk_bytes;
  if (threadIdx.x == 0) {
    L.pub = 0;
    for (uint32_t e = 0; e < 3; e++) L.done[e] = 0;
    L.bad = 0;
    L.stuck = 0;
  }
  __syncthreads();

  if (wave == 0) {  // ---- walker: the critical path; its instructions issue first
    __builtin_amdgcn_s_setprio(3);
    // waits until every expander has finished batch x - 1; false if stuck
    auto wait_done = [&](uint32_t x) -> bool {
      for (uint32_t spins = 0;; spins++) {
        uint32_t m = lds_acquire(&L.done[0]);
        for (uint32_t e = 1; e < kExp; e++) {
          const uint32_t d = lds_acquire(&L.done[e]);
          m = d < m ? d : m;
        }
        if (m >= x) return true;
        if (spins >= walker_limit || lds_acquire(&L.stuck)) {
          lds_release(&L.stuck, 1, lane);
          return false;
        }
        __builtin_amdgcn_s_sleep(kWalkerSleep);
      }
    };
    Walk s;
    s.src = comp + boff;
    s.len = (uint32_t)len;
    s.cap = cap;
    s.o = 0;
    uint32_t wi = 0;                       // window the walk is in
    uint32_t used[2] = {0u, 0u};           // 1 + last batch that used each window (0: none)
    walk_stage<kWin>(s, L.sbuf[0], 0, lane);
    walk_mirror(s, L.sbuf[0], lane);  // the walker's own LDS writes precede its reads
    int32_t st = kOk;
    bool done = (len == 0);
    for (uint32_t k = 0;; k++) {
      const uint32_t b = k % kQ;
      if (k >= kQ && !wait_done(k - kQ + 1)) {
        st = kCorrupt;
        break;
      }
      uint32_t nd = 0;
      bool restage = false;
      if (!done && st == kOk) {
        const uint32_t p0 = s.p;
        nd = walk_batch<true>(s, L.sbuf[wi], lane, L.desc[b], nullptr);
        done = s.stop == kStopDone;
        restage = s.stop == kStopRestage;
        st = s.err;
        if (restage && nd == 0 && p0 == s.lead) st = kCorrupt;
      }
      const bool last = done || st != kOk;
      if (last) restage = false;
      if (wi) used[1] = k + 1;
      else used[0] = k + 1;
      if (lane == 0) L.info[b] = nd | (last ? kBatchLast : 0u) | (wi ? kBatchWin : 0u);
      lds_release(&L.pub, k + 1, lane);
      if (last) {
        if (!wait_done(k + 1)) st = kCorrupt;
        break;
      }
      if (restage) {
        const uint32_t pos = s.wpos + (s.p - s.lead);
        wi ^= 1;
        if (!wait_done(wi ? used[1] : used[0])) {
          st = kCorrupt;
          break;
        }
        walk_stage<kWin>(s, L.sbuf[wi], pos, lane);
        walk_mirror(s, L.sbuf[wi], lane);
      }
    }
    if (lds_acquire(&L.bad) && st == kOk) st = kCorrupt;
    if (st == kOk)
      for (uint32_t i = s.o + lane; i < cap; i += kWave) store_w<W>(dst, i, 0);
    if (lane == 0 && status) status[chunk] = st;
  } else {  // ---- expanders: wave e takes blocks j with j % kExp == e - 1
    uint64_t sel = 0;
    for (uint32_t j = wave - 1; j < (uint32_t)kWave; j += kExp) sel |= 1ull << j;
    uint32_t bad = 0;
    for (uint32_t k = 0;; k++) {
      bool ready = false;
      for (uint32_t spins = 0;; spins++) {
        if (lds_acquire(&L.pub) > k) {
          ready = true;
          break;
        }
        if (spins >= exp_limit || lds_acquire(&L.stuck)) break;
        __builtin_amdgcn_s_sleep(kExpSleep);
      }
      if (!ready) {
        lds_release(&L.stuck, 1, lane);
        return;
      }
      const uint32_t b = k % kQ;
      const uint32_t info = L.info[b];
      const uint32_t nd = info & 0xffff;
      const uint32_t d = L.desc[b][lane];
      expand_batch<W>(L.sbuf[(info & kBatchWin) ? 1 : 0], nd, sel, d & 0x1fff, d >> 13, dst, lane,
                      &bad);
      const bool any_bad = __ballot(bad) != 0;
      if ((info & kBatchLast) && any_bad && lane == 0) L.bad = 1;
      lds_release(&L.done[wave - 1], k + 1, lane);
      if (info & kBatchLast) return;
    }
  }
}

}  // namespace rle2

namespace {

template <uint32_t kWin, uint32_t W>
void wave_launch(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                 const uint64_t *col_len, uint64_t n_chunks, uint64_t chunk_bytes, uint8_t *out,
                 int32_t *status, hipStream_t stream) {
  hipLaunchKernelGGL((rle2::rlev2_wave_kernel<kWin, W>), dim3((uint32_t)n_chunks), dim3(kWaveSize),
                     0, stream, comp, comp_bytes, blk_off, col_len, n_chunks, chunk_bytes, out,
                     status);
}

template <uint32_t kExp, uint32_t W>
void pc_launch(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
               const uint64_t *col_len, uint64_t n_chunks, uint64_t chunk_bytes, uint8_t *out,
               int32_t *status, uint32_t poll_shift, hipStream_t stream) {
  if (poll_shift)
    hipLaunchKernelGGL((rle2::rlev2_pc_kernel<4096, kExp, W, true>), dim3((uint32_t)n_chunks),
                       dim3((1 + kExp) * kWaveSize), 0, stream, comp, comp_bytes, blk_off, col_len,
                       n_chunks, chunk_bytes, out, status);
  else
    hipLaunchKernelGGL((rle2::rlev2_pc_kernel<4096, kExp, W, false>), dim3((uint32_t)n_chunks),
                       dim3((1 + kExp) * kWaveSize), 0, stream, comp, comp_bytes, blk_off, col_len,
                       n_chunks, chunk_bytes, out, status);
}

// Instantiates F<..., W> for the element width w (1, 2, 4 or 8).
#define CODAG_BY_WIDTH(F, ...)                     \
  switch (w) {                                     \
  case 1: F<__VA_ARGS__, 1>(ARGS); break;          \
  case 2: F<__VA_ARGS__, 2>(ARGS); break;          \
  case 4: F<__VA_ARGS__, 4>(ARGS); break;          \
  case 8: F<__VA_ARGS__, 8>(ARGS); break;          \
  default: return hipErrorInvalidValue;            \
  }

}  // namespace

hipError_t launch_rlev2_wave(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                             const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                             uint64_t chunk_bytes, uint8_t *out, int32_t *status, uint32_t win,
                             hipStream_t stream) {
  if (n_chunks == 0) return hipSuccess;
#define ARGS comp, comp_bytes, blk_off, col_len, n_chunks, chunk_bytes, out, status, stream
  if (win == 4096) {
    CODAG_BY_WIDTH(wave_launch, 4096)
  } else if (win == 8192) {
    CODAG_BY_WIDTH(wave_launch, 8192)
  } else {
    return hipErrorInvalidValue;
  }
#undef ARGS
  return hipGetLastError();
}

hipError_t launch_rlev2_pc(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                           const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                           uint64_t chunk_bytes, uint8_t *out, int32_t *status, uint32_t nexp,
                           uint32_t poll_shift, hipStream_t stream) {
  if (n_chunks == 0) return hipSuccess;
#define ARGS comp, comp_bytes, blk_off, col_len, n_chunks, chunk_bytes, out, status, poll_shift, stream
  if (nexp == 1) {
    CODAG_BY_WIDTH(pc_launch, 1)
  } else if (nexp == 2) {
    CODAG_BY_WIDTH(pc_launch, 2)
  } else if (nexp == 3) {
    CODAG_BY_WIDTH(pc_launch, 3)
  } else {
    return hipErrorInvalidValue;
  }
#undef ARGS
  return hipGetLastError();
}

#undef CODAG_BY_WIDTH

}  // namespace codag
