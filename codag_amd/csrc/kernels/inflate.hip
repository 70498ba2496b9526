// inflate.hip -- gfx950 zlib/Deflate (RFC 1950/1951) decoder, one wave64 per stream.
//
// Mapping onto CDNA:
//  * Bit reader: wave-uniform (SGPR) 64-bit bit buffer. The compressed input
//    is held in two VGPR windows -- lane i holds dword i of a 256-byte window
//    -- and refilled into the bit buffer with v_readlane, so a refill costs no
//    memory round trip; the next window is prefetched with one coalesced load.
//  * Huffman tables: per-wave LDS lookup tables built per block by all 64
//    lanes (literal/length table of 10 bits in the one-wave kernels and 11
//    in the decoding wave of a two-wave group, 8-bit distance table,
//    canonical count/symbol arrays for the rare longer codes). Table entries carry the
//    RFC 1951 length/distance base and extra-bit count directly; a literal's
//    entry has bit 15 set so the symbol loop tests it with one scalar compare.
//  * Two phases per batch of up to 64 symbols:
//      1. decode only. The window decoder (vector unit) has lane j decode the
//         token that would start at bit P + j of a 64-bit window, and a
//         scalar readlane chain through the token lengths keeps the real
//         token starts; the scalar symbol loop decodes the symbols that stop
//         the chain (codes longer than the LUT, end-of-block, invalid codes)
//         and the batches near the stream end. Token j of the batch is parked
//         in lane j of a VGPR;
//      2. the writer (this wave, or the second wave of a two-wave group)
//         materialises the batch in ring-sized pieces: a wave prefix sum
//         gives each token its output offset, literals go to an LDS ring in
//         one store, matches whose source has left the ring are gathered from
//         HBM (eight per memory round trip), short matches whose source
//         precedes the piece are copied together (one lane each), and the
//         remaining matches are copied ring -> ring in token order, 64 bytes
//         per step.
//    Decoding never depends on output bytes, so it runs without waiting on
//    memory beyond its own table reads and stream loads.
//  * The ring is written to HBM with 16-byte stores after each piece that
//    leaves at least kFlush unwritten bytes; Adler-32 partial sums (sum b,
//    sum pos*b) are accumulated on the way, so the trailer is verified without
//    a second pass over the output.
//  * Occupancy: a one-wave kernel uses a 1 KiB ring and 4736 B of LDS (32
//    waves per CU); a two-wave group gives its writer a 2 KiB ring.
//
// Reference counterpart: include/zlib/zlib.h:807-866 (full_warp_shared_lut),
// 531-707 (dynamic header / table build), 712-805 (symbol loop), 188-347
// (warp copy). Differences: stored blocks are decoded (the reference's branch
// is empty, zlib.h:840-842), BTYPE 11 is rejected (zlib.h:843 treats it as
// dynamic), the zlib header and Adler-32 trailer are checked as zlib does.
#include <cstddef>
#include <mutex>

#include "device.hpp"
#include "kernels.hpp"

namespace codag {
namespace inflate {

using namespace codag::dev;

// literal/length LUT bits: 10 for the one-wave kernels (32 waves fit a CU),
// 11 for the decoding wave of a two-wave group (fewer codes leave the
// window decoder for the symbol loop)
constexpr uint32_t kLitBitsOne = 10;
constexpr uint32_t kLitBitsPair = 11;
constexpr uint32_t kDistBits = 8;
constexpr uint32_t kCodeBits = 7;
constexpr uint32_t kDistMask = (1u << kDistBits) - 1;

enum : int32_t { kOk = 0, kTrunc = 1, kOverflow = 2, kCorrupt = 3 };
constexpr int32_t kEob = 16;  // symbol loop stop: end of block (not a status)
enum : uint32_t { kKindLit = 0, kKindLen = 1, kKindEob = 2, kKindBad = 3 };
enum : uint32_t { kTabLit = 0, kTabDist = 1, kTabCodes = 2 };

// 16-bit LUT entries (bits 3:0 always hold the code length; 0 = not in LUT):
//   literal      : bit 15 set, byte in bits 11:4
//   length code c: extra-bit count in bits 6:4 (0..5), base-3 in bits 14:7;
//                  bits 6:4 = 7 marks end-of-block, 6 an invalid symbol
//   distance d   : extra-bit count in bits 7:4, m in bits 9:8 with
//                  base = (m << extra) + 1; bit 13 = invalid (d >= 30)
//   code-length  : symbol in bits 12:8
constexpr uint32_t kLitFlag = 0x8000u;
constexpr uint32_t kDistBad = 0x2000u;
constexpr uint32_t kXEob = 7, kXBad = 6;

// RFC 1951 section 3.2.5 length and distance codes (the tables of
// include/zlib/zlib.h:35-51) in closed form: length code c (0..28) has
// (c - 4) / 4 extra bits from c = 8 on (none for c = 28, length 258) and base
// 3 + c below 8, else 3 + ((4 + c % 4) << extra); distance code d (0..29) has
// (d - 2) / 2 extra bits from d = 4 on and base 1 + d below 4, else
// 1 + ((2 + d % 2) << extra).
__device__ __forceinline__ uint32_t len_extra(uint32_t c) { return (c < 8 || c == 28) ? 0 : (c - 4) >> 2; }
__device__ __forceinline__ uint32_t len_base(uint32_t c, uint32_t extra) {
  return c < 8 ? c + 3 : (c == 28 ? 258 : ((4 + (c & 3)) << extra) + 3);
}
__device__ __forceinline__ uint32_t dist_extra(uint32_t d) { return d < 4 ? 0 : (d - 2) >> 1; }
__device__ __forceinline__ uint32_t dist_base(uint32_t d, uint32_t extra) {
  return d < 4 ? d + 1 : ((2 + (d & 1)) << extra) + 1;
}
// Order of the code-length code lengths (RFC 1951 3.2.7).
__constant__ uint8_t kClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

template <int NSYM> struct Huff {
  uint16_t count[16];
  uint16_t first[16];
  uint16_t index[16];
  uint16_t sym[NSYM];
};

// Output ring and piece limits. The writer materialises a batch (at most 64
// tokens) in pieces: a piece ends at the first token at which its bytes
// reach kBatch = R / 4, so it produces T < kBatch + 258 bytes, and the ring
// is flushed to HBM once kFlush = R / 8 bytes are pending, so fewer than
// kFlush bytes are unwritten at a piece start. Then (a) a piece never
// overwrites a ring byte that is not yet in HBM and (b) a match whose source
// lies before the ring's oldest intact byte reads only bytes that are
// already in HBM, when R >= T + 258 + kFlush. One-wave kernels use a 1 KiB
// ring, which leaves a wave under 5 KiB of LDS (32 waves per CU); the
// writing wave of a two-wave group has a 2 KiB ring (fewer far gathers).
template <uint32_t R> struct RingGeo {
  static constexpr uint32_t kRing = R, kMask = R - 1, kBatch = R / 4, kFlush = R / 8;
  static_assert((R & (R - 1)) == 0 && R >= kBatch + 258 + 258 + kFlush, "ring too small");
};
constexpr uint32_t kRingOne = 1024;
constexpr uint32_t kRingPair = 2048;
// A symbol reads at most 15 + 5 + 15 + 13 bits: a full batch (64 symbols)
// that starts this far from the stream end cannot read past it, and the
// 96-bit field the window decoder loads from the dword of any of its symbols
// lies inside the stream's dwords.
constexpr uint32_t kCarefulBits = 64 * 48 + 96;
constexpr uint32_t kGather = 8;  // far-match pieces in flight per memory round trip
// A wave's serial decode is latency-bound, so throughput grows with the waves
// resident per SIMD: a one-wave kernel's LDS (4736 B, ten 512-byte allocation
// blocks) allows 32 per CU, and its register budget is held to eight per SIMD
// (64 VGPRs), so a 1 GiB container of 128 KiB chunks is decoded in one round.
// Two-wave groups hold seven per SIMD.
constexpr int kWavesOne = 8;
constexpr int kWavesPair = 7;

template <uint32_t LB> struct LdsT {
  static constexpr uint32_t kLitBits = LB, kLitMask = (1u << LB) - 1;
  alignas(16) uint8_t ring[kRingOne];
  uint16_t lit[1u << LB];
  uint16_t dist[1u << kDistBits];
  Huff<288> hl;
  Huff<32> hd;
  alignas(4) uint8_t lens[320];
};
using Lds = LdsT<kLitBitsOne>;
using LdsPair = LdsT<kLitBitsPair>;
static_assert(sizeof(Lds::lens) >= 4 * (kWave + 16), "lens doubles as the token slots");
// The first 4 KiB (ring, lit, dist and the start of hl) hold the CRC-32 slice
// tables while no block is being decoded (gzip header and trailer)
static_assert(sizeof(Lds) >= 4096, "CRC tables need 4 KiB");

__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t j) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)j);
}

// Orders this wave's LDS accesses around a point where lanes exchange data
// through LDS. The decoder's tables, slots and ring are written and read by
// one wave, so fences (no workgroup barrier) are enough, and a second wave
// of the workgroup (the writing wave of the two-wave batch kernel) is never
// waited for.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// ------------------------------------------------------------- bit reader

struct Bits {
  const uint32_t *w;  // 4-aligned address at or below the stream start
  uint32_t limit;     // bit index (from w) one past the stream's last bit
  uint32_t nw;        // dwords that hold stream bytes
  uint32_t next;      // next dword to append to buf
  uint64_t buf;
  uint32_t cnt;
  uint32_t wbase;     // first dword of the current 64-dword block (a multiple of 64)
  // Blocks of 64 dwords, lane i holding dword i: the block at dword 64*m lives
  // in w0 for even m and in w1 for odd m. The current block and the next one
  // are resident; a register is reloaded only when its block is used up, so
  // the windows never move between registers.
  uint32_t w0, w1;
};

__device__ __forceinline__ uint32_t bits_block(const Bits &b, uint32_t first) {
  const uint32_t i = first + lane_id();
  return i < b.nw ? b.w[i] : 0u;
}

__device__ __forceinline__ void bits_windows(Bits &b) {
  const uint32_t cur = bits_block(b, b.wbase), nxt = bits_block(b, b.wbase + kWave);
  const bool odd = (b.wbase / kWave) & 1;
  b.w0 = odd ? nxt : cur;
  b.w1 = odd ? cur : nxt;
}

__device__ __forceinline__ void bits_refill(Bits &b) {
  if (b.cnt <= 32) {
    uint32_t k = b.next - b.wbase;
    if (k >= (uint32_t)kWave) {
      // the used-up block's register takes the block after the next one
      if ((b.wbase / kWave) & 1) b.w1 = bits_block(b, b.wbase + 2 * kWave);
      else b.w0 = bits_block(b, b.wbase + 2 * kWave);
      b.wbase += kWave;
      k -= kWave;
    }
    const uint32_t d = ((b.wbase / kWave) & 1) ? rdlane(b.w1, k) : rdlane(b.w0, k);
    b.buf |= (uint64_t)d << b.cnt;
    b.cnt += 32;
    b.next++;
  }
}

__device__ __forceinline__ void bits_drop(Bits &b, uint32_t n) {
  b.buf >>= n;
  b.cnt -= n;
}

// Moves the reader forward to bit P (P >= its position).
__device__ __forceinline__ void bits_set(Bits &b, uint32_t P) {
  const uint32_t nx = P >> 5;
  if (nx - b.wbase >= 2u * kWave) {  // past both resident blocks
    b.wbase = nx & ~(uint32_t)(kWave - 1);
    bits_windows(b);
  }
  b.next = nx;
  b.buf = 0;
  b.cnt = 0;
  bits_refill(b);
  bits_drop(b, P & 31);
}

__device__ __forceinline__ uint32_t bits_pos(const Bits &b) { return b.next * 32 - b.cnt; }

__device__ __forceinline__ bool bits_past_end(const Bits &b) { return bits_pos(b) > b.limit; }

// Positions the reader at stream byte `at` (byte granular).
__device__ __forceinline__ void bits_seek(Bits &b, uint64_t lead_plus_at) {
  const uint64_t bit = lead_plus_at * 8;
  b.next = (uint32_t)(bit >> 5);
  b.wbase = b.next & ~(uint32_t)(kWave - 1);
  bits_windows(b);
  b.buf = 0;
  b.cnt = 0;
  bits_refill(b);
  bits_drop(b, (uint32_t)(bit & 31));
  bits_refill(b);
}

// ------------------------------------------------------------- output side

// The most recent R output bytes live in an LDS ring. Byte x of the
// stream sits at ring[(x + mis) & (R - 1)], where mis = dst & 15, so a
// 16-byte aligned output address is also 16-byte aligned in the ring.
// Positions are 32-bit (outputs below 2^31 bytes) so the wave-uniform state
// stays in scalar registers.
template <uint32_t R> struct Out : RingGeo<R> {
  uint8_t *dst;      // global output
  uint8_t *ring;     // LDS ring
  uint32_t cap;
  uint32_t pos;      // bytes produced
  uint32_t flushed;  // bytes [0, flushed) have been stored to dst
  uint32_t mis;      // dst & 15
  // Per-lane Adler-32 partial sums: s1 = sum b, s2 = sum pos*b (mod 65521),
  // both reduced after every store. One store adds at most ~8e8 to s2 (at
  // most 15 single bytes and two 16-byte groups per lane), so 32 bits hold.
  uint32_t s1, s2;
};

template <class O> __device__ __forceinline__ uint32_t ridx(const O &o, uint32_t x) {
  return (x + o.mis) & O::kMask;
}

template <class O> __device__ __forceinline__ void adler_add(O &o, uint32_t at, uint32_t b) {
  o.s1 += b;
  o.s2 += (at % 65521u) * b;
}

// Stores ring bytes [flushed, end) to HBM; `last` also writes a partial
// 16-byte tail.
template <class O> __device__ __forceinline__ void out_store(O &o, uint32_t lane, uint32_t end, bool last) {
  uint32_t x = o.flushed;
  // head bytes up to a 16-byte aligned output address
  uint32_t head = (16 - ((o.mis + x) & 15)) & 15;
  if (head > end - x) head = end - x;
  if (head) {
    if (lane < head) {
      const uint32_t b = o.ring[ridx(o, x + lane)];
      o.dst[x + lane] = (uint8_t)b;
      adler_add(o, x + lane, b);
    }
    x += head;
  }
  const uint32_t nvec = (end - x) / 16;
  for (uint32_t k = lane; k < nvec; k += kWave) {
    const uint32_t at = x + 16 * k;
    const uint4 q = *reinterpret_cast<const uint4 *>(o.ring + ridx(o, at));
    *reinterpret_cast<uint4 *>(o.dst + at) = q;
    // sum b and sum i*b over the 16 bytes (v_dot4_u32_u8 against the byte
    // weights), then s2 += at*sum + sum(i*b)
    const uint32_t sum = __builtin_amdgcn_udot4(
        q.x, 0x01010101u,
        __builtin_amdgcn_udot4(q.y, 0x01010101u,
                               __builtin_amdgcn_udot4(q.z, 0x01010101u,
                                                      __builtin_amdgcn_udot4(q.w, 0x01010101u, 0u, false),
                                                      false),
                               false),
        false);
    const uint32_t wsum = __builtin_amdgcn_udot4(
        q.x, 0x03020100u,
        __builtin_amdgcn_udot4(q.y, 0x07060504u,
                               __builtin_amdgcn_udot4(q.z, 0x0b0a0908u,
                                                      __builtin_amdgcn_udot4(q.w, 0x0f0e0d0cu, 0u, false),
                                                      false),
                               false),
        false);
    o.s1 += sum;
    o.s2 += (at % 65521u) * sum + wsum;
  }
  x += nvec * 16;
  if (last && x < end) {
    const uint32_t tail = end - x;
    if (lane < tail) {
      const uint32_t b = o.ring[ridx(o, x + lane)];
      o.dst[x + lane] = (uint8_t)b;
      adler_add(o, x + lane, b);
    }
    x = end;
  }
  o.flushed = x;
  o.s1 %= 65521u;
  o.s2 %= 65521u;
}

// Copies far matches (lanes in fm) from HBM into the ring. The first 64 bytes
// of kGather matches are loaded before any is written, so a group costs one
// memory round trip; the rest of the (rare) longer far matches follows.
template <class O>
__device__ __forceinline__ void gather_far(O &o, uint32_t lane, uint64_t fm, uint32_t x,
                                           uint32_t dist, uint32_t mlen) {
  uint64_t g = fm;
  while (g) {
    uint32_t v[kGather];
    uint64_t h = g;
#pragma unroll
    for (uint32_t k = 0; k < kGather; k++) {
      v[k] = 0;
      if (h) {
        const uint32_t j = (uint32_t)__builtin_ctzll(h);
        h &= h - 1;
        const uint32_t sj = rdlane(x, j) - rdlane(dist, j), lj = rdlane(mlen, j);
        if (lane < lj) v[k] = o.dst[sj + lane];
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < kGather; k++) {
      if (g) {
        const uint32_t j = (uint32_t)__builtin_ctzll(g);
        g &= g - 1;
        const uint32_t xj = rdlane(x, j), lj = rdlane(mlen, j);
        if (lane < lj) o.ring[ridx(o, xj + lane)] = (uint8_t)v[k];
      }
    }
  }
  uint64_t lm = fm & __ballot(mlen > (uint32_t)kWave);
  while (lm) {
    const uint32_t j = (uint32_t)__builtin_ctzll(lm);
    lm &= lm - 1;
    const uint32_t xj = rdlane(x, j), dj = rdlane(dist, j), lj = rdlane(mlen, j);
#pragma clang loop vectorize(disable) interleave(disable) unroll(disable)
    for (uint32_t i = kWave + lane; i < lj; i += kWave) o.ring[ridx(o, xj + i)] = o.dst[xj - dj + i];
  }
}

// floor(2^16 / d) + 1 for d = 1..63: (i * m[d]) >> 16 == i / d for every
// i < 512. The product overshoots i / d by i * (m[d] - 2^16 / d) / 2^16 <
// 512 / 2^16 = 1/128, while the fractional part of i / d is at most
// 1 - 1/d <= 1 - 1/63, so the overshoot never carries into the next integer.
struct Recip16 {
  uint32_t m[64];
  constexpr Recip16() : m() {
    for (uint32_t d = 1; d < 64; d++) m[d] = 0x10000u / d + 1;
  }
};
__constant__ Recip16 kRecip16 = Recip16();

// Materialises the tokens of lanes [a, b) of a batch as one piece: rel is
// the token's inclusive byte count from the piece start, tl its byte count.
// Errors are reported for the first offending token, in stream order:
// output past cap (overflow) or a distance beyond the bytes produced so far.
// The tokens before it are still written, so the bytes written before an
// error do not depend on where batches and pieces end.
template <class O>
__device__ __forceinline__ int32_t out_piece(O &o, uint32_t lane, uint32_t tok, bool isl, uint32_t mlen,
                                             uint32_t dist, uint32_t tl, uint32_t rel, uint32_t a,
                                             uint32_t b) {
  const uint32_t x = o.pos + rel - tl;  // output position of this token
  const bool inp = lane >= a && lane < b;
  uint32_t err = 0;
  if (inp && x + tl > o.cap) err = kOverflow;
  else if (inp && !isl && dist > x) err = kCorrupt;  // zlib: "invalid distance too far back"
  const uint64_t em = __ballot(err != 0);
  int32_t st = kOk;
  if (em) {
    const uint32_t k = (uint32_t)__builtin_ctzll(em);
    st = (int32_t)rdlane(err, k);
    b = k;
    if (b == a) return st;
  }
  const bool act = lane >= a && lane < b;
  const bool lit = act && isl;
  const bool mat = act && !isl;
  const uint32_t pend = o.pos + rdlane(rel, b - 1);

  if (lit) o.ring[ridx(o, x)] = (uint8_t)(tok >> 4);
  // a source byte y is intact in the ring iff y >= pend - R
  const bool far = mat && (x - dist) + O::kRing < pend;
  const uint64_t fm = __ballot(far);
  if (fm) gather_far(o, lane, fm, x, dist, mlen);
  // A near match whose source ends before the piece's first byte reads only
  // bytes that were in the ring before the piece, and writes bytes that no
  // earlier token of the piece reads: such matches of up to 16 bytes are
  // copied together, each lane its own, 4 bytes per step (bytes past the
  // match's end repeat its last byte's copy). The other near matches follow
  // in token order.
  const bool ind = mat && !far && mlen <= 16u && (x - dist) + mlen <= o.pos;
  if (__ballot(ind)) {
    const uint32_t src = x - dist, last = mlen - 1;
#pragma unroll
    for (uint32_t i = 0; i < 16u; i += 4) {
      if (!__ballot(ind && i < mlen)) break;
      if (ind && i < mlen) {
        const uint32_t k1 = min(i + 1, last), k2 = min(i + 2, last), k3 = min(i + 3, last);
        const uint8_t b0 = o.ring[ridx(o, src + i)], b1 = o.ring[ridx(o, src + k1)];
        const uint8_t b2 = o.ring[ridx(o, src + k2)], b3 = o.ring[ridx(o, src + k3)];
        o.ring[ridx(o, x + i)] = b0;
        o.ring[ridx(o, x + k1)] = b1;
        o.ring[ridx(o, x + k2)] = b2;
        o.ring[ridx(o, x + k3)] = b3;
      }
    }
  }
  uint64_t nm = __ballot(mat && !far && !ind);
  while (nm) {  // in token order: a match may read bytes of earlier tokens
    const uint32_t j = (uint32_t)__builtin_ctzll(nm);
    nm &= nm - 1;
    const uint32_t xj = rdlane(x, j), dj = rdlane(dist, j), lj = rdlane(mlen, j);
    const uint32_t sj = xj - dj;
    if (dj >= (uint32_t)kWave || lj <= dj) {
      // byte i reads source byte sj + i: either the source ends before the
      // match (lj <= dj) or, for i >= dj, it repeats byte i - dj, written one
      // 64-byte step earlier. Most matches take the single first step.
      if (lane < lj) o.ring[ridx(o, xj + lane)] = o.ring[ridx(o, sj + lane)];
#pragma clang loop vectorize(disable) interleave(disable) unroll(disable)
      for (uint32_t i = kWave + lane; i < lj; i += kWave)
        o.ring[ridx(o, xj + i)] = o.ring[ridx(o, sj + i)];
    } else {
      // periodic copy: byte i = source byte i mod dj, the quotient by an
      // integer reciprocal multiply (exact for i < 512, dj < 64)
      const uint32_t rm = kRecip16.m[dj];
#pragma clang loop vectorize(disable) interleave(disable) unroll(disable)
      for (uint32_t i = lane; i < lj; i += kWave) {
        const uint32_t m = i - ((i * rm) >> 16) * dj;
        o.ring[ridx(o, xj + i)] = o.ring[ridx(o, sj + m)];
      }
    }
  }
  o.pos = pend;
  if (pend - o.flushed >= O::kFlush) out_store(o, lane, pend, false);
  return st;
}

// Materialises a batch of nt tokens (lane j holds token j):
//   literal: its LUT entry (bit 15 set, byte in bits 11:4);
//   match: length in bits 8:0 (bit 15 clear), distance in bits 31:16.
// A batch of up to 64 tokens can produce up to 64 x 258 bytes, more than the
// ring holds, so it is written in pieces: a piece ends at the first token at
// which its bytes reach O::kBatch, so a piece produces T < kBatch + 258
// bytes. The decoder only limits batches to 64 tokens.
template <class O>
__device__ __forceinline__ int32_t out_tokens(O &o, uint32_t lane, uint32_t tok, uint32_t nt) {
  if (nt == 0) return kOk;
  const bool isl = (tok & kLitFlag) != 0;
  const uint32_t mlen = tok & 0x1ff;
  const uint32_t dist = tok >> 16;
  const uint32_t tl = lane < nt ? (isl ? 1u : mlen) : 0u;
  const uint32_t incl = wave_scan_dpp(tl);
  uint32_t a = 0, base = 0;  // the piece's first lane; bytes of the tokens before it
  for (;;) {
    const uint64_t over = __ballot(lane >= a && lane < nt && incl - base >= O::kBatch);
    const uint32_t b = over ? (uint32_t)__builtin_ctzll(over) + 1 : nt;
    const int32_t st = out_piece(o, lane, tok, isl, mlen, dist, tl, incl - base, a, b);
    if (st != kOk || b == nt) return st;
    base = rdlane(incl, b - 1);
    a = b;
  }
}

// Copies ln stored-block bytes from src + at through the ring (RFC 1951
// 3.2.4), 256 bytes per step.
template <class O>
__device__ __forceinline__ int32_t store_copy(O &o, const uint8_t *src, uint64_t at, uint32_t ln,
                                              uint32_t lane) {
  if (o.pos + ln > o.cap) return kOverflow;
  for (uint32_t done = 0; done < ln;) {
    const uint32_t step = (ln - done) < 256u ? (ln - done) : 256u;
#pragma clang loop vectorize(disable) interleave(disable) unroll(disable)
    for (uint32_t i = lane; i < step; i += kWave) o.ring[ridx(o, o.pos + i)] = src[at + done + i];
    o.pos += step;
    done += step;
    if (o.pos - o.flushed >= O::kFlush) out_store(o, lane, o.pos, false);
  }
  return kOk;
}

// The last partial flush, then the Adler-32 of the bytes written (a wave
// reduction of the per-lane partial sums, each already < 65521).
template <class O> __device__ __forceinline__ uint32_t out_finish(O &o, uint32_t lane) {
  out_store(o, lane, o.pos, true);
  uint32_t s1r = o.s1, s2r = o.s2;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    s1r += __shfl_xor(s1r, d, kWave);
    s2r += __shfl_xor(s2r, d, kWave);
  }
  uint64_t s1 = s1r, s2 = s2r;
  s1 %= 65521u;
  s2 %= 65521u;
  const uint64_t n = o.pos % 65521u;
  const uint32_t A = (uint32_t)((1 + s1) % 65521u);
  const uint32_t B = (uint32_t)((n * ((1 + s1) % 65521u) + 65521u * 65521u - s2) % 65521u);
  return uni((B << 16) | A);
}

// ------------------------------------------------------------- table build

__device__ __forceinline__ uint32_t make_entry(uint32_t tab, uint32_t sym, uint32_t len) {
  if (tab == kTabCodes) return len | (sym << 8);
  if (tab == kTabDist) {
    if (sym >= 30) return len | kDistBad;
    const uint32_t m = sym < 4 ? sym : 2 + (sym & 1);  // dist_base == (m << extra) + 1
    return len | (dist_extra(sym) << 4) | (m << 8);
  }
  if (sym < 256) return len | (sym << 4) | kLitFlag;
  if (sym == 256) return len | (kXEob << 4);
  if (sym <= 285) {
    const uint32_t c = sym - 257, x = len_extra(c);
    return len | (x << 4) | ((len_base(c, x) - 3) << 7);
  }
  return len | (kXBad << 4);
}

// Builds the canonical tables for n code lengths in lens[] and the LUT of
// 2^bits entries. Applies zlib's acceptance rules (inftrees.c): an
// over-subscribed set is rejected; an incomplete set is rejected unless it
// has a single code of length 1 (literal/distance tables) or no codes at all.
// Out of line: it runs once per block, and keeping its unrolled code out of
// the symbol loop leaves that loop's uniform state in scalar registers.
template <int NSYM>
__device__ __noinline__ int32_t build_table(const uint8_t *lens, uint32_t n, uint32_t tab, Huff<NSYM> &h,
                               uint16_t *lut, uint32_t bits) {
  const uint32_t lane = lane_id();
  if (lane < 16) h.count[lane] = 0;
  for (uint32_t i = lane; i < (1u << bits); i += kWave) lut[i] = 0;
  wave_sync();
  for (uint32_t s = lane; s < n; s += kWave) {
    const uint32_t L = lens[s];
    if (L) atomicAdd(reinterpret_cast<uint32_t *>(&h.count[0]) + (L >> 1), 1u << (16 * (L & 1)));
  }
  wave_sync();
  // counts are packed two per dword above; all lanes walk them uniformly
  int32_t left = 1;
  uint32_t maxlen = 0, code = 0, idx = 0;
  uint32_t first_l = 0, index_l = 0;
  const uint32_t myL = lane < 16 ? lane : 0;
  for (uint32_t L = 1; L < 16; L++) {
    const uint32_t c = uni(h.count[L]);
    left = 2 * left - (int32_t)c;
    if (c) maxlen = L;
    if (L == myL) {
      first_l = code;
      index_l = idx;
    }
    code = (code + c) << 1;
    idx += c;
  }
  if (left < 0) return kCorrupt;
  if (left > 0) {
    if (tab == kTabCodes) return kCorrupt;
    if (maxlen > 1) return kCorrupt;
  }
  wave_sync();
  if (lane >= 1 && lane < 16) {
    h.first[lane] = (uint16_t)first_l;
    h.index[lane] = (uint16_t)index_l;
  }
  wave_sync();
  uint32_t run[16];
#pragma unroll
  for (int L = 0; L < 16; L++) run[L] = 0;
  for (uint32_t base = 0; base < n; base += kWave) {
    const uint32_t s = base + lane;
    const uint32_t L = s < n ? lens[s] : 0;
    uint32_t r = 0;
#pragma unroll
    for (int q = 1; q < 16; q++) {
      const uint64_t m = __ballot(L == (uint32_t)q);
      if (L == (uint32_t)q) r = run[q] + __popcll(m & ((1ull << lane) - 1));
      run[q] += __popcll(m);
    }
    if (L) {
      h.sym[h.index[L] + r] = (uint16_t)s;
      if (L <= bits) {
        const uint32_t c = h.first[L] + r;
        const uint32_t rev = __builtin_bitreverse32(c) >> (32 - L);
        const uint32_t e = make_entry(tab, s, L);
        for (uint32_t k = 0; k < (1u << (bits - L)); k++) lut[rev | (k << L)] = e;
      }
    }
  }
  wave_sync();
  return kOk;
}

// Canonical decode of codes longer than the LUT (or invalid ones), RFC 1951
// 3.2.2: the code of length L is the first L stream bits read MSB-first, and
// it is valid when it lies in [first[L], first[L] + count[L]). Lane L (1..15)
// tests length L, so the tables are read once instead of once per length; the
// shortest length that hits is the code (canonical codes are prefix-free).
// Called by the whole wave. Returns the symbol or -1; *len gets the length.
template <int NSYM>
__device__ __forceinline__ int32_t slow_decode(const Huff<NSYM> &h, uint64_t buf, uint32_t *len) {
  const uint32_t lane = lane_id();
  const uint32_t L = lane & 15u;
  const uint32_t c = h.count[L], f = h.first[L], ix = h.index[L];
  const uint32_t code = __builtin_bitreverse32((uint32_t)buf) >> ((32u - L) & 31u);  // L = 0: unused
  const uint64_t m = __ballot(lane >= 1u && lane < 16u && code - f < c);
  if (m == 0) return -1;
  const uint32_t hit = (uint32_t)__builtin_ctzll(m);
  *len = hit;
  return (int32_t)uni(h.sym[rdlane(ix + code - f, hit)]);
}

// ------------------------------------------------------------- stream decode

// Status in the caller's convention: CODAG_CHUNK_* by default, cuDF's
// gpu_inflate_status_s codes (cudf_deflate.cuh:1243-1343: 0 ok, 1 output
// buffer too small, 2 input ended early, -1 invalid data) with
// CODAG_INFLATE_CUDF_STATUS.
__device__ __forceinline__ int32_t status_code(int32_t st, uint32_t flags) {
  if (!(flags & CODAG_INFLATE_CUDF_STATUS) || st == kOk) return st;
  return st == kOverflow ? 1 : (st == kTrunc ? 2 : -1);
}

template <class LD>
__device__ __forceinline__ int32_t read_dynamic(Bits &b, LD &S, uint32_t *nlen_out) {
  const uint32_t lane = lane_id();
  bits_refill(b);
  const uint32_t nlen = (uint32_t)(b.buf & 31) + 257;
  const uint32_t ndist = (uint32_t)((b.buf >> 5) & 31) + 1;
  const uint32_t ncode = (uint32_t)((b.buf >> 10) & 15) + 4;
  bits_drop(b, 14);
  if (nlen > 286 || ndist > 30) return kCorrupt;  // zlib: "too many length or distance symbols"
  // code-length code lengths, 3 bits each, in permuted order (<= 57 bits)
  bits_refill(b);
  const uint32_t n1 = ncode < 10 ? ncode : 10;
  const uint64_t c1 = b.buf;
  bits_drop(b, 3 * n1);
  bits_refill(b);
  const uint32_t n2 = ncode - n1;
  const uint64_t c2 = b.buf;
  bits_drop(b, 3 * n2);
  if (lane < 19) {
    uint32_t v = 0;
    if (lane < n1) v = (uint32_t)((c1 >> (3 * lane)) & 7);
    else if (lane < ncode) v = (uint32_t)((c2 >> (3 * (lane - n1))) & 7);
    S.lens[kClOrder[lane]] = (uint8_t)v;
  }
  wave_sync();
  int32_t st = (int32_t)uni((uint32_t)build_table(S.lens, 19, kTabCodes, S.hl, S.lit, kCodeBits));
  if (st != kOk) return st;
  // literal/length + distance code lengths with repeat codes 16/17/18
  const uint32_t total = nlen + ndist;
  uint32_t i = 0;
  while (i < total) {
    bits_refill(b);
    const uint32_t e = uni(S.lit[b.buf & ((1u << kCodeBits) - 1)]);
    const uint32_t L = e & 15;
    if (L == 0) return bits_past_end(b) ? kTrunc : kCorrupt;
    bits_drop(b, L);
    const uint32_t sym = e >> 8;
    if (sym < 16) {
      if (lane == 0) S.lens[i] = (uint8_t)sym;
      i++;
      continue;
    }
    uint32_t rep, val = 0;
    if (sym == 16) {
      if (i == 0) return kCorrupt;  // zlib: "invalid bit length repeat"
      wave_sync();
      val = uni(S.lens[i - 1]);
      rep = 3 + (uint32_t)(b.buf & 3);
      bits_drop(b, 2);
    } else if (sym == 17) {
      rep = 3 + (uint32_t)(b.buf & 7);
      bits_drop(b, 3);
    } else {
      rep = 11 + (uint32_t)(b.buf & 127);
      bits_drop(b, 7);
    }
    if (i + rep > total) return kCorrupt;
    if (lane < rep) S.lens[i + lane] = (uint8_t)val;  // rep <= 138: two passes at most
    if (lane + kWave < rep) S.lens[i + kWave + lane] = (uint8_t)val;
    if (lane + 2 * kWave < rep) S.lens[i + 2 * kWave + lane] = (uint8_t)val;
    i += rep;
  }
  if (bits_past_end(b)) return kTrunc;
  wave_sync();
  if (uni(S.lens[256]) == 0) return kCorrupt;  // zlib: "missing end-of-block"
  st = (int32_t)uni((uint32_t)build_table(S.lens, nlen, kTabLit, S.hl, S.lit, LD::kLitBits));
  if (st != kOk) return st;
  st = (int32_t)uni((uint32_t)build_table(S.lens + nlen, ndist, kTabDist, S.hd, S.dist, kDistBits));
  *nlen_out = nlen;
  return st;
}

template <class LD> __device__ __forceinline__ void load_fixed(LD &S) {
  const uint32_t lane = lane_id();
  for (uint32_t s = lane; s < 288; s += kWave)
    S.lens[s] = (uint8_t)(s < 144 ? 8 : (s < 256 ? 9 : (s < 280 ? 7 : 8)));
  wave_sync();
  build_table(S.lens, 288, kTabLit, S.hl, S.lit, LD::kLitBits);
  for (uint32_t s = lane; s < 32; s += kWave) S.lens[s] = 5;
  wave_sync();
  build_table(S.lens, 32, kTabDist, S.hd, S.dist, kDistBits);
}

// The three stream dwords that hold the 64 bits from stream bit `at` (the
// window decoder's per-lane input): one load, its first dword clamped to
// nw - 3. The window decoder runs only for full batches, which start at
// least kCarefulBits before the stream end, so at every position that starts
// a symbol of the batch the three dwords are inside the stream and the clamp
// changes only lanes no symbol of the batch starts at. The bits are taken
// apart where they are used, so the load stays in flight until then.
struct Field {
  uint32_t d0, d1, d2;
};

__device__ __forceinline__ Field load_field(const Bits &b, uint32_t at) {
  typedef __attribute__((address_space(1))) const uint32_t gdword;
  const gdword *w = (const gdword *)b.w + min(at >> 5, b.nw - 3);
  Field f;
  f.d0 = w[0];
  f.d1 = w[1];
  f.d2 = w[2];
  return f;
}

// v_alignbit_b32 shifts by its third operand's low five bits.
__device__ __forceinline__ uint64_t field_bits(const Field &f, uint32_t at) {
  const uint32_t xl = __builtin_amdgcn_alignbit(f.d1, f.d0, at);
  const uint32_t xh = __builtin_amdgcn_alignbit(f.d2, f.d1, at);
  return ((uint64_t)xh << 32) | xl;
}

// Window decode on the vector unit. Lane j decodes the token (a literal, or a
// length code with its distance and extra bits) that would start at bit P + j
// of the stream, from its own three stream dwords and the same LUT entries as
// the symbol loop; a scalar chain through the token lengths then keeps the
// real token starts. A token whose code is not in the LUT, an end-of-block
// and an invalid code stop the chain (token length 128); the symbol loop
// decodes that symbol. A batch ends at its 64th token (the writer cuts it
// into pieces that fit its ring). Token k of the batch goes to slot[k] in
// LDS (the caller moves the slots into tok once per batch); slots 64..79
// take the writes of the lanes that start no token. The next window's loads
// are issued as soon as the chain has placed it, ahead of the slot work.
// Returns true when the batch is full, false at a symbol for the loop.
template <class LD>
__device__ __forceinline__ bool decode_window(Bits &b, const LD &S, uint32_t *slot, uint32_t lane,
                                              uint32_t &nt) {
  const uint32_t P0 = bits_pos(b);
  uint32_t P = P0;
  bool full;
  Field f = load_field(b, P + lane);
  for (;;) {
    const uint64_t x = field_bits(f, P + lane);
    const uint32_t xl = (uint32_t)x;
    const uint32_t e = S.lit[xl & LD::kLitMask];
    const uint32_t L = e & 15, lx = (e >> 4) & 7;
    // a length code's extra bits, its distance code and that code's extra
    // bits lie within the 32 bits from bit L + lx (< 32) on
    const uint32_t y = __builtin_amdgcn_alignbit((uint32_t)(x >> 32), xl, L + lx);
    const uint32_t d = S.dist[y & kDistMask];
    const uint32_t dl = d & 15, dx = (d >> 4) & 15;
    const uint32_t mlen = ((e >> 7) & 255) + 3 + __builtin_amdgcn_ubfe(xl, L, lx);
    const uint32_t dist = (((d >> 8) & 3) << dx) + 1 + __builtin_amdgcn_ubfe(y, dl, dx);
    const uint64_t litm = __ballot(e >= kLitFlag);
    const bool lit = __builtin_amdgcn_inverse_ballot_w64(litm);
    // a length code stops the chain if it or its distance code is not in the
    // LUT, if it is end-of-block or invalid (extra field 6, 7), or if the
    // distance code is 30 or 31
    const uint32_t bad = (((L - 1u) | (dl - 1u)) >> 31) | ((lx + 2u) >> 3) | ((d & kDistBad) >> 13);
    uint32_t tlm = bad ? 128u : L + lx + dl + dx;
    asm("" : "+v"(tlm));  // every lane computes it: no branch around the literal lanes
    const uint32_t tl = lit ? L : tlm;
    const uint32_t tk = lit ? e : (mlen | (dist << 16));
    // chain: token starts from window bit 0 while they lie in the window; a
    // stop lane (length 128) ends it at j >= 128
    // M |= 1 << j, then tl of lane j, while j < 64. Written out eight links
    // at a time with a fall-through test per link: a taken branch per link
    // (the loop's back edge) costs a lone wave about as much as the link
    // itself. j is kept as j - 64, so the add's carry is the test, and the
    // bitset and the lane select use only its low six bits. The scalar
    // bitset also gives the wait state a v_readlane needs after the vector
    // write of tl.
    uint64_t M = 0;
    uint32_t j = 0u - (uint32_t)kWave;  // j - 64 (mod 2^32): the add carries once j >= 64
    {
      uint32_t t;
      asm("1:\n"
          ".rept 8\n"
          "s_bitset1_b64 %[M], %[j]\n"
          "v_readlane_b32 %[t], %[tl], %[j]\n"
          "s_add_u32 %[j], %[j], %[t]\n"
          "s_cbranch_scc1 2f\n"
          ".endr\n"
          "s_branch 1b\n"
          "2:"
          : [j] "+s"(j), [M] "+s"(M), [t] "=&s"(t)
          : [tl] "v"(tl)
          : "scc");
    }
    j += (uint32_t)kWave;
    const bool stop = j >= 128u;
    uint32_t adv = j & 127u;  // bits used by the tokens before the stop / window end
    // the next window, unless this one stops or ends the batch
    f = load_field(b, P + adv + lane);
    // The batch's 64th token ends it (nt < 64 here). The token starts are
    // kept as wave masks in scalar registers; inverse ballots hand them to
    // the vector unit as conditions. The two exits and the common path each
    // store their own slots, so no flag has to live across a join.
    const uint64_t mkm = M & ~__ballot(tl >= 128u);  // a stop lane starts no token
    const uint32_t rank = lanes_below(mkm);
    const uint32_t cnt = (uint32_t)__builtin_popcountll(mkm);
    if (nt + cnt >= (uint32_t)kWave) {  // the batch ends at token lastc
      const uint64_t cm = mkm & __ballot(rank >= (uint32_t)kWave - 1 - nt);
      const uint32_t lastc = (uint32_t)__builtin_ctzll(cm);
      const uint64_t keepm = mkm & ((2ull << lastc) - 1);
      slot[__builtin_amdgcn_inverse_ballot_w64(keepm) ? nt + rank : kWave + (lane & 15)] = tk;
      nt += (uint32_t)__builtin_popcountll(keepm);
      P += lastc + rdlane(tl, lastc);
      full = true;
      break;
    }
    slot[__builtin_amdgcn_inverse_ballot_w64(mkm) ? nt + rank : kWave + (lane & 15)] = tk;
    nt += cnt;
    P += adv;
    if (stop) {
      full = false;
      break;
    }
  }
  if (P != P0) bits_set(b, P);
  return full;
}

// Symbol loop: decodes up to `limit` (<= 64) symbols of the current block
// into tok (lane j <- token j, nt tokens); every token passes through its
// LDS slot, written by lane 0 here and by decode_window. Returns kOk when
// the batch is full (`limit` tokens), kEob at end-of-block, or the status
// of an invalid symbol. The tokens decoded before
// an invalid symbol are valid and still have to be written. All stops leave
// through one exit so the loop's uniform state stays in scalar registers.
// A full batch (limit 64) is decoded by decode_window, and the loop takes
// only the symbols that stop it.
template <class LD>
__device__ __forceinline__ int32_t decode_tokens(Bits &b, const LD &S, uint32_t *slot, uint32_t lane,
                                                 uint32_t &tok, uint32_t &nt, uint32_t limit) {
  int32_t st = kOk;
  for (;;) {
    if (limit == (uint32_t)kWave && decode_window(b, S, slot, lane, nt)) break;
    bits_refill(b);  // >= 33 bits
    uint32_t e = uni(S.lit[(uint32_t)b.buf & LD::kLitMask]);
    if (e & kLitFlag) {  // literal: the token is the table entry
      bits_drop(b, e & 15);
      if (lane == 0) slot[nt] = e;
      if (++nt == limit) break;
      // >= 18 bits are left: enough for the next code without a refill
      e = uni(S.lit[(uint32_t)b.buf & LD::kLitMask]);
      if (e & kLitFlag) {
        bits_drop(b, e & 15);
        if (lane == 0) slot[nt] = e;
        if (++nt == limit) break;
        continue;
      }
    }
    // a length / end-of-block / invalid code, or one longer than the LUT;
    // the buffer holds >= 18 bits
    uint32_t L = e & 15;
    if (L == 0) {  // code longer than the LUT, or no such code
      const int32_t sym = slow_decode(S.hl, b.buf, &L);
      if (sym < 0) {
        st = bits_past_end(b) ? kTrunc : kCorrupt;
        break;
      }
      e = make_entry(kTabLit, (uint32_t)sym, L);
    }
    bits_drop(b, L);
    if (e & kLitFlag) {
      if (lane == 0) slot[nt] = e;
      if (++nt == limit) break;
      continue;
    }
    const uint32_t lx = (e >> 4) & 7;
    if (lx >= kXBad) {
      st = lx == kXEob ? kEob : kCorrupt;
      break;
    }
    // >= 33 bits after the refill: length extra (<= 5), distance code (<= 15)
    // and distance extra (<= 13)
    bits_refill(b);
    const uint32_t len = ((e >> 7) & 255) + 3 + ((uint32_t)b.buf & ((1u << lx) - 1));
    bits_drop(b, lx);
    uint32_t d = uni(S.dist[(uint32_t)b.buf & kDistMask]);
    uint32_t dl = d & 15;
    if (dl == 0) {
      const int32_t sym = slow_decode(S.hd, b.buf, &dl);
      if (sym < 0) {
        st = bits_past_end(b) ? kTrunc : kCorrupt;
        break;
      }
      d = make_entry(kTabDist, (uint32_t)sym, dl);
    }
    if (d & kDistBad) {  // distance codes 30 and 31
      st = kCorrupt;
      break;
    }
    bits_drop(b, dl);
    const uint32_t dx = (d >> 4) & 15;
    const uint32_t dist = (((d >> 8) & 3) << dx) + 1 + ((uint32_t)b.buf & ((1u << dx) - 1));
    bits_drop(b, dx);
    if (lane == 0) slot[nt] = len | (dist << 16);
    if (++nt == limit) break;
  }
  if (nt) {
    wave_sync();
    tok = slot[lane];
  }
  return st;
}

// Decodes symbols until end-of-block. Returns a status. The sink takes each
// batch (SinkOne writes it on this wave, SinkQueue hands it to the writing
// wave) and returns the first output error seen so far.
template <class Sink, class LD>
__device__ __forceinline__ int32_t decode_block(Bits &b, Sink &out, const LD &S, uint32_t *slot,
                                                uint32_t lane) {
  for (;;) {
    // A symbol reads at most 48 bits, so a batch of floor(left / 48) symbols
    // cannot pass the stream end; near the end batches shrink to one symbol,
    // and reading past the end is caught at the symbol that does it.
    // (left * 1365) >> 16 is at most left / 48 for left below 2^16.
    const uint32_t pos = bits_pos(b);
    const uint32_t left = pos < b.limit ? b.limit - pos : 0u;
    uint32_t limit = (uint32_t)kWave;
    if (left < kCarefulBits) {
      const uint32_t safe = (left * 1365u) >> 16;
      limit = safe == 0 ? 1u : (safe < (uint32_t)kWave ? safe : (uint32_t)kWave - 1);
    }
    uint32_t tok = 0, nt = 0;
    const int32_t sa = decode_tokens(b, S, slot, lane, tok, nt, limit);
    if (bits_past_end(b)) return kTrunc;
    const int32_t sb = out.tokens(lane, tok, nt);
    if (sb != kOk) return sb;
    if (sa == kEob) return kOk;
    if (sa != kOk) return sa;
  }
}

struct Result {
  int32_t status;
  uint64_t bytes;
  uint32_t adler;
};

// Output of a stream decoded by one wave: the wave writes its own batches
// through an LDS ring of R bytes.
template <uint32_t R> struct SinkOne {
  Out<R> o;
  __device__ __forceinline__ void init(uint8_t *dst, uint8_t *ring, uint64_t cap) {
    o.dst = dst;
    o.ring = ring;
    o.cap = (uint32_t)cap;
    o.pos = 0;
    o.flushed = 0;
    o.mis = (uint32_t)(reinterpret_cast<uintptr_t>(dst) & 15);
    o.s1 = 0;
    o.s2 = 0;
  }
  __device__ __forceinline__ int32_t tokens(uint32_t lane, uint32_t tok, uint32_t nt) {
    return out_tokens(o, lane, tok, nt);
  }
  // stored block data: ln bytes from src + at (RFC 1951 3.2.4)
  __device__ __forceinline__ int32_t stored(const uint8_t *src, uint64_t at, uint32_t ln,
                                            uint32_t lane) {
    return store_copy(o, src, at, ln, lane);
  }
  // the final flush; status st of the decode; fills bytes and Adler-32
  __device__ __forceinline__ int32_t finish(uint32_t lane, int32_t st, Result &res) {
    res.adler = out_finish(o, lane);
    res.bytes = o.pos;
    return st;
  }
};

// ------------------------------------------------------------- CRC-32 (gzip)

// CRC-32 of RFC 1952 (reflected polynomial 0xEDB88320). Polynomials mod P are
// held bit-reflected: bit 31 is x^0. mulmod multiplies two of them (shift
// and add over GF(2)); kX2n[k] is x^(2^k) mod P, so x^(8n) mod P is a product
// of kX2n entries over the bits of n (the combination rule zlib's
// crc32_combine uses: crc(A B) = crc(A) * x^(8|B|) + crc(B)).
constexpr uint32_t kCrcPoly = 0xEDB88320u;

__host__ __device__ constexpr uint32_t crc_mulmod(uint32_t a, uint32_t b) {
  uint32_t m = 1u << 31, p = 0;
  for (;;) {
    if (a & m) {
      p ^= b;
      if ((a & (m - 1)) == 0) break;
    }
    m >>= 1;
    b = (b & 1) ? (b >> 1) ^ kCrcPoly : b >> 1;
  }
  return p;
}

struct X2nTable {
  uint32_t t[32];
  constexpr X2nTable() : t() {
    uint32_t p = 1u << 30;  // x^1
    t[0] = p;
    for (int k = 1; k < 32; k++) t[k] = p = crc_mulmod(p, p);
  }
};
__constant__ X2nTable kX2n = X2nTable();

// x^(8n) mod P.
__device__ __forceinline__ uint32_t crc_x8n(uint32_t n) {
  uint32_t p = 1u << 31;  // x^0
  uint32_t k = 3;
  while (n) {
    if (n & 1) p = crc_mulmod(kX2n.t[k & 31], p);
    n >>= 1;
    k++;
  }
  return p;
}

// CRC-32 of n bytes at global address p, by the whole wave: four 256-entry
// slice tables are built in `tab` (4 KiB of LDS), lane l takes the l-th of
// 64 equal segments (slice-by-4), and the 64 segment CRCs are combined in
// order. Every lane returns the same value.
__device__ __forceinline__ uint32_t wave_crc32(const uint8_t *p, uint32_t n, uint32_t *tab) {
  const uint32_t lane = lane_id();
  // the bytes may have been stored by other lanes of this wave just before
  wave_sync();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  for (uint32_t i = lane; i < 256; i += kWave) {
    uint32_t c = i;
#pragma unroll
    for (int k = 0; k < 8; k++) c = (c & 1) ? (c >> 1) ^ kCrcPoly : c >> 1;
    tab[i] = c;
  }
  wave_sync();
#pragma unroll
  for (int t = 1; t < 4; t++) {
    for (uint32_t i = lane; i < 256; i += kWave) {
      const uint32_t c = tab[(t - 1) * 256 + i];
      tab[t * 256 + i] = (c >> 8) ^ tab[c & 0xff];
    }
    wave_sync();
  }
  const uint32_t seg = ((n + kWave - 1) / kWave + 3) & ~3u;
  const uint32_t lo = min(lane * seg, n), hi = min(lo + seg, n);
  uint32_t c = 0xFFFFFFFFu;
  uint32_t i = lo;
  for (; i + 4 <= hi; i += 4) {
    c ^= (uint32_t)p[i] | ((uint32_t)p[i + 1] << 8) | ((uint32_t)p[i + 2] << 16) |
         ((uint32_t)p[i + 3] << 24);
    c = tab[768 + (c & 0xff)] ^ tab[512 + ((c >> 8) & 0xff)] ^ tab[256 + ((c >> 16) & 0xff)] ^
        tab[c >> 24];
  }
  for (; i < hi; i++) c = tab[(c ^ p[i]) & 0xff] ^ (c >> 8);
  const uint32_t mine = ~c;
  // x^(8 * segment length) of this lane's segment, every lane at once
  const uint32_t shift = crc_x8n(hi - lo);
  uint32_t total = 0;  // CRC of the empty string
  for (uint32_t l = 0; l < (uint32_t)kWave; l++) {
    const uint32_t cl = rdlane(mine, l), xl = rdlane(shift, l);
    total = crc_mulmod(xl, total) ^ cl;
  }
  wave_sync();
  return uni(total);
}

// gzip member header (RFC 1952 2.3), checked as zlib's inflate() checks it:
// ID1 ID2 = 1f 8b, CM = 8, no reserved FLG bits, FEXTRA / FNAME / FCOMMENT
// skipped, FHCRC compared with the low 16 bits of the header's CRC-32.
// Returns the header length, or -kTrunc / -kCorrupt.
__device__ __forceinline__ int64_t gzip_header(const uint8_t *src, uint64_t len, uint32_t *tab) {
  const uint32_t lane = lane_id();
  if (len < 10) {
    // zlib reports a short header as truncated unless a byte it has already
    // read is wrong
    const uint32_t b0 = len > 0 ? uni(src[0]) : 0x1f, b1 = len > 1 ? uni(src[1]) : 0x8b;
    const uint32_t b2 = len > 2 ? uni(src[2]) : 8, b3 = len > 3 ? uni(src[3]) : 0;
    if (b0 != 0x1f || b1 != 0x8b || b2 != 8 || (b3 & 0xe0)) return -kCorrupt;
    return -kTrunc;
  }
  if (uni(src[0]) != 0x1f || uni(src[1]) != 0x8b || uni(src[2]) != 8) return -kCorrupt;
  const uint32_t flg = uni(src[3]);
  if (flg & 0xe0) return -kCorrupt;
  uint64_t at = 10;
  if (flg & 0x04) {  // FEXTRA
    if (at + 2 > len) return -kTrunc;
    const uint32_t xlen = uni(src[at]) | (uni(src[at + 1]) << 8);
    at += 2 + xlen;
    if (at > len) return -kTrunc;
  }
  for (uint32_t f = 0x08; f <= 0x10; f <<= 1) {  // FNAME, FCOMMENT: zero-terminated
    if (!(flg & f)) continue;
    for (;;) {
      if (at >= len) return -kTrunc;
      const uint64_t q = at + lane;
      const bool z = q < len && src[q] == 0;
      const uint64_t m = __ballot(z);
      if (m) {
        at += (uint64_t)__builtin_ctzll(m) + 1;
        break;
      }
      at += kWave;
    }
  }
  if (flg & 0x02) {  // FHCRC
    if (at + 2 > len) return -kTrunc;
    const uint32_t want = uni(src[at]) | (uni(src[at + 1]) << 8);
    if ((wave_crc32(src, (uint32_t)at, tab) & 0xffff) != want) return -kCorrupt;
    at += 2;
  }
  return (int64_t)at;
}

// Decodes one zlib / raw / gzip stream; `out` takes the decoded batches
// (SinkOne: this wave writes them; SinkQueue: a second wave does).
template <class Sink, class LD>
__device__ __forceinline__ Result inflate_stream_t(const uint8_t *src, uint64_t len, uint8_t *dst,
                                                   uint64_t cap, uint32_t flags, LD &S, Sink &out) {
  const uint32_t lane = lane_id();
  // the code lengths are dead while a block's symbols are decoded: their
  // bytes hold the window decoder's 64 token slots
  uint32_t *slot = reinterpret_cast<uint32_t *>(S.lens);
  Result res{kOk, 0, 1};
  Bits b;
  // 32-bit bit positions bound the stream to < 2^28 bytes; outputs are
  // capped below 2^31 bytes.
  if (len >= (1ull << 28)) {
    res.status = kCorrupt;
    return res;
  }
  if (cap > 0x7fffffffull) cap = 0x7fffffffull;
  const uintptr_t a = reinterpret_cast<uintptr_t>(src);
  const uint32_t lead = (uint32_t)(a & 3);
  b.w = reinterpret_cast<const uint32_t *>(a - lead);
  b.limit = (uint32_t)((lead + len) * 8);
  b.nw = (uint32_t)((lead + len + 3) / 4);
  const bool gzip = flags & CODAG_INFLATE_GZIP;
  const bool raw = (flags & CODAG_INFLATE_RAW) && !gzip;
  uint64_t start = 0;
  if (gzip) {
    const int64_t h = gzip_header(src, len, reinterpret_cast<uint32_t *>(S.ring));
    if (h < 0) {
      res.status = (int32_t)(-h);
      return res;
    }
    start = (uint64_t)h;
  } else if (!raw) {
    if (len < 2) {
      res.status = kTrunc;
      return res;
    }
    const uint32_t cmf = uni(src[0]), flg = uni(src[1]);
    // zlib inflate.c: method 8, window <= 32K, header check, no preset dictionary
    if ((cmf & 15) != 8 || (cmf >> 4) > 7 || ((cmf << 8) | flg) % 31 != 0 || (flg & 0x20)) {
      res.status = kCorrupt;
      return res;
    }
    start = 2;
  }
  bits_seek(b, lead + start);

  int32_t st = kOk;
  bool last = false;
  uint32_t tables = 0;  // 1 = fixed tables loaded
  while (!last && st == kOk) {
    bits_refill(b);
    const uint32_t h = (uint32_t)(b.buf & 7);
    bits_drop(b, 3);
    if (bits_past_end(b)) { st = kTrunc; break; }
    last = h & 1;
    const uint32_t type = h >> 1;
    if (type == 0) {  // stored (RFC 1951 3.2.4)
      bits_drop(b, (uint32_t)((8 - (bits_pos(b) & 7)) & 7));
      bits_refill(b);
      const uint32_t ln = (uint32_t)(b.buf & 0xffff);
      const uint32_t nln = (uint32_t)((b.buf >> 16) & 0xffff);
      bits_drop(b, 32);
      if (bits_past_end(b)) { st = kTrunc; break; }
      if (ln != (~nln & 0xffff)) { st = kCorrupt; break; }
      const uint64_t at = bits_pos(b) / 8 - lead;  // stream byte offset of the data
      if (at + ln > len) { st = kTrunc; break; }
      st = out.stored(src, at, ln, lane);
      if (st != kOk) break;
      bits_seek(b, lead + at + ln);
    } else if (type == 1) {  // fixed Huffman
      if (tables != 1) {
        load_fixed(S);
        tables = 1;
      }
      st = decode_block(b, out, S, slot, lane);
    } else if (type == 2) {  // dynamic Huffman
      uint32_t nlen = 0;
      tables = 0;
      st = read_dynamic(b, S, &nlen);
      if (st == kOk) st = decode_block(b, out, S, slot, lane);
    } else {
      st = kCorrupt;  // zlib: "invalid block type"
    }
  }
  // the last flush, the bytes written and their Adler-32; the status becomes
  // the first error in stream order of the decode and the output
  st = out.finish(lane, st, res);
  if (gzip) {
    // RFC 1952 trailer on the next byte boundary: CRC-32 and ISIZE, both
    // little-endian; zlib checks the CRC first ("incorrect data check"),
    // then the length ("incorrect length check")
    res.adler = wave_crc32(dst, (uint32_t)res.bytes, reinterpret_cast<uint32_t *>(S.ring));
    if (st == kOk) {
      bits_drop(b, (uint32_t)((8 - (bits_pos(b) & 7)) & 7));
      const uint64_t at = bits_pos(b) / 8 - lead;
      if (at + 8 > len) {
        st = kTrunc;
      } else {
        const uint32_t crc = uni(src[at]) | (uni(src[at + 1]) << 8) | (uni(src[at + 2]) << 16) |
                             (uni(src[at + 3]) << 24);
        const uint32_t isize = uni(src[at + 4]) | (uni(src[at + 5]) << 8) |
                               (uni(src[at + 6]) << 16) | (uni(src[at + 7]) << 24);
        if (!(flags & CODAG_INFLATE_SKIP_ADLER) && crc != res.adler) st = kCorrupt;
        else if (isize != (uint32_t)res.bytes) st = kCorrupt;
      }
    }
  } else if (st == kOk && !raw) {
    // RFC 1950 trailer: Adler-32, big-endian, on the next byte boundary
    bits_drop(b, (uint32_t)((8 - (bits_pos(b) & 7)) & 7));
    const uint64_t at = bits_pos(b) / 8 - lead;
    if (at + 4 > len) {
      st = kTrunc;
    } else if (!(flags & CODAG_INFLATE_SKIP_ADLER)) {
      const uint32_t t = (uni(src[at]) << 24) | (uni(src[at + 1]) << 16) |
                         (uni(src[at + 2]) << 8) | uni(src[at + 3]);
      if (t != res.adler) st = kCorrupt;
    }
  }
  res.status = st;
  return res;
}

__device__ __forceinline__ Result inflate_stream(const uint8_t *src, uint64_t len, uint8_t *dst,
                                                 uint64_t cap, uint32_t flags, Lds &S) {
  SinkOne<kRingOne> out;
  out.init(dst, S.ring, cap > 0x7fffffffull ? 0x7fffffffull : cap);
  return inflate_stream_t(src, len, dst, cap, flags, S, out);
}

__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(kWavesOne)))
inflate_chunks_kernel(const uint8_t *__restrict__ comp, uint64_t comp_bytes,
                      const uint64_t *__restrict__ off, uint64_t n, uint64_t chunk_bytes, uint8_t *__restrict__ out,
                      uint64_t *__restrict__ out_len, int32_t *__restrict__ status, uint32_t flags) {
  __shared__ Lds S;
  const uint64_t c = blockIdx.x;
  const uint64_t o0 = off[c], o1 = off[c + 1];
  Result r;
  if (o1 < o0 || o1 > comp_bytes) {
    r = Result{kCorrupt, 0, 0};
  } else {
    r = inflate_stream(comp + o0, o1 - o0, out + c * chunk_bytes, chunk_bytes, flags, S);
  }
  if (lane_id() == 0) {
    if (status) status[c] = status_code(r.status, flags);
    if (out_len) out_len[c] = r.bytes;
  }
}

__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(kWavesOne)))
inflate_batch_kernel(const codag_stream_desc *__restrict__ descs, uint64_t n,
                     codag_stream_result *__restrict__ results, uint32_t flags,
                     uint32_t *__restrict__ counter) {
  __shared__ Lds S;
  for (;;) {
    uint32_t i = 0;
    if (lane_id() == 0) i = atomicAdd(counter, 1u);
    i = __builtin_amdgcn_readfirstlane(__shfl(i, 0, kWave));
    if ((uint64_t)i >= n) break;
    const codag_stream_desc d = descs[i];
    // A stream decodes serially on one wave, and the waves of a CU share its
    // scalar pipe: in an uneven batch the largest streams set the finish
    // time. Their waves issue first (wave priority by output size), the
    // small streams fill the remaining issue slots.
    const uint64_t big = d.dst_cap;
    if (big >= (2u << 20)) __builtin_amdgcn_s_setprio(3);
    else if (big >= (1u << 20)) __builtin_amdgcn_s_setprio(2);
    else if (big >= (512u << 10)) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
    const Result r = inflate_stream(d.src, d.src_bytes, d.dst, d.dst_cap, flags, S);
    if (lane_id() == 0) {
      results[i].bytes_written = r.bytes;
      results[i].status = status_code(r.status, flags);
      results[i].adler32 = r.adler;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------ two-wave batch

// A workgroup of two waves decodes one stream at a time: wave 0 decodes
// (header, tables, symbols) and hands each batch of tokens, each stored
// block and the end of the stream to wave 1 through a four-deep LDS queue;
// wave 1 writes the output (ring copies, far gathers, flush, Adler-32). The
// decoding wave's serial chain no longer waits for the output phase, which
// bounds a large stream's time in an uneven batch. The two waves meet only
// through the queue's head / tail / done words (release / acquire on LDS)
// and at the top of each stream; every wait is bounded (kSpinLimit sleeps,
// between half a second and two), and one that runs out ends the stream as
// CORRUPT.
constexpr uint32_t kQueue = 4;
constexpr uint32_t kSpinLimit = 1u << 24;
enum : uint32_t { kMsgTokens = 0, kMsgStored = 1, kMsgEnd = 2 };

struct Msg {
  uint32_t kind, nt, ln, pad;
  uint64_t at;
};

struct Queue {
  uint32_t tok[kQueue][kWave];
  Msg msg[kQueue];
  uint32_t head;    // messages published by the decoding wave
  uint32_t tail;    // messages the writing wave has finished with
  int32_t wstat;    // the writing wave's first error
  uint32_t done;    // the writing wave has flushed the stream
  uint32_t bytes, adler;
  uint32_t stream;  // the workgroup's current stream index
};

__device__ __forceinline__ uint32_t lds_acquire(uint32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_release(uint32_t *p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The decoding wave's side: batches go into the queue.
struct SinkQueue {
  Queue *q;
  uint32_t head;
  uint32_t tail;  // the writer's last tail this wave read: slots below it are free
  bool ended, stuck;
  // Hands a message to the writer. The writer's tail is read again only when
  // the last one read leaves no free slot, and the writer's status is not
  // polled per batch (finish() returns the writer's first error before any
  // later decode error), so a batch costs one LDS round trip, the release.
  __device__ __forceinline__ int32_t push(uint32_t lane, uint32_t kind, uint32_t tok, uint32_t nt,
                                          uint64_t at, uint32_t ln) {
    for (uint32_t k = 0;; k++) {  // a free slot
      if (head - tail < kQueue) break;
      tail = uni(lds_acquire(&q->tail));
      if (head - tail < kQueue) break;
      if (k >= kSpinLimit) {
        stuck = true;
        return kCorrupt;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    const uint32_t sl = head % kQueue;
    q->tok[sl][lane] = tok;
    if (lane == 0) {
      q->msg[sl].kind = kind;
      q->msg[sl].nt = nt;
      q->msg[sl].ln = ln;
      q->msg[sl].at = at;
    }
    head++;
    lds_release(&q->head, head);
    return kOk;
  }
  __device__ __forceinline__ int32_t tokens(uint32_t lane, uint32_t tok, uint32_t nt) {
    if (nt == 0) return (int32_t)uni(lds_acquire(reinterpret_cast<uint32_t *>(&q->wstat)));
    return push(lane, kMsgTokens, tok, nt, 0, 0);
  }
  __device__ __forceinline__ int32_t stored(const uint8_t *, uint64_t at, uint32_t ln, uint32_t lane) {
    return push(lane, kMsgStored, 0, 0, at, ln);
  }
  // hands over the end of the stream and waits for the writing wave; the
  // writer's first error precedes any later error of the decode
  __device__ __forceinline__ int32_t finish(uint32_t lane, int32_t st, Result &res) {
    ended = true;
    res.bytes = 0;
    res.adler = 1;
    if (!stuck && push(lane, kMsgEnd, 0, 0, 0, 0) == kCorrupt && stuck) return kCorrupt;
    if (stuck) return kCorrupt;
    for (uint32_t k = 0;; k++) {
      if (uni(lds_acquire(&q->done))) break;
      if (k >= kSpinLimit) return kCorrupt;
      __builtin_amdgcn_s_sleep(1);
    }
    res.bytes = uni(q->bytes);
    res.adler = uni(q->adler);
    const int32_t ws = (int32_t)uni((uint32_t)q->wstat);
    return ws != kOk ? ws : st;
  }
};

// The writing wave's side: takes messages in order until the end of the
// stream, then publishes the bytes written, their Adler-32 and its first
// error. After an error it only drains the queue.
__device__ __forceinline__ void inflate_write(Queue &q, uint8_t *ring, uint8_t *dst, uint64_t cap,
                                              const uint8_t *src) {
  const uint32_t lane = lane_id();
  SinkOne<kRingPair> w;
  w.init(dst, ring, cap > 0x7fffffffull ? 0x7fffffffull : cap);
  int32_t wst = kOk;
  for (uint32_t t = 0;; t++) {
    bool ready = false;
    for (uint32_t k = 0; k < kSpinLimit; k++) {
      if (uni(lds_acquire(&q.head)) > t) {
        ready = true;
        break;
      }
      // an idle writer polls every ~200 cycles: a batch takes the decoder
      // thousands, and fewer polls leave the CU's issue slots to decoders
      __builtin_amdgcn_s_sleep(3);
    }
    if (!ready) {
      if (wst == kOk) wst = kCorrupt;
      break;
    }
    const uint32_t sl = t % kQueue;
    const uint32_t kind = uni(q.msg[sl].kind);
    if (kind == kMsgEnd) break;
    if (wst == kOk) {
      int32_t e;
      if (kind == kMsgTokens) e = w.tokens(lane, q.tok[sl][lane], uni(q.msg[sl].nt));
      else e = w.stored(src, q.msg[sl].at, uni(q.msg[sl].ln), lane);
      if (e != kOk) {
        wst = e;
        lds_release(reinterpret_cast<uint32_t *>(&q.wstat), (uint32_t)e);
      }
    }
    lds_release(&q.tail, t + 1);
  }
  Result r;
  w.finish(lane, kOk, r);
  q.bytes = (uint32_t)r.bytes;
  q.adler = r.adler;
  q.wstat = wst;
  // the output is in memory before the decoding wave reads it back (gzip
  // CRC-32)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  lds_release(&q.done, 1u);
}

struct Lds2 {
  LdsPair dec;                   // the decoding wave's tables, slots and CRC space
  alignas(16) uint8_t ring[kRingPair];  // the writing wave's output ring
  Queue q;
};

__global__ void __launch_bounds__(2 * kWave) __attribute__((amdgpu_waves_per_eu(kWavesPair)))
inflate_batch2_kernel(const codag_stream_desc *__restrict__ descs, uint64_t n,
                      codag_stream_result *__restrict__ results, uint32_t flags,
                      uint32_t *__restrict__ counter) {
  __shared__ Lds2 S;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x) / kWave;
  const uint32_t lane = lane_id();
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) {
      S.q.stream = atomicAdd(counter, 1u);
      S.q.head = 0;
      S.q.tail = 0;
      S.q.wstat = kOk;
      S.q.done = 0;
    }
    __syncthreads();
    const uint32_t i = uni(S.q.stream);
    if ((uint64_t)i >= n) break;
    const codag_stream_desc d = descs[i];
    const uint64_t big = d.dst_cap;
    if (big >= (2u << 20)) __builtin_amdgcn_s_setprio(3);
    else if (big >= (1u << 20)) __builtin_amdgcn_s_setprio(2);
    else if (big >= (512u << 10)) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
    if (wave == 0) {
      SinkQueue out;
      out.q = &S.q;
      out.head = 0;
      out.tail = 0;
      out.ended = false;
      out.stuck = false;
      Result r = inflate_stream_t(d.src, d.src_bytes, d.dst, d.dst_cap, flags, S.dec, out);
      if (!out.ended) {  // a header error before any output: release the writer
        Result t;
        const int32_t e = out.finish(lane, r.status, t);
        if (r.status == kOk) r.status = e;
      }
      if (lane == 0) {
        results[i].bytes_written = r.bytes;
        results[i].status = status_code(r.status, flags);
        results[i].adler32 = r.adler;
      }
    } else {
      inflate_write(S.q, S.ring, d.dst, d.dst_cap, d.src);
    }
  }
}

// The chunk container decoded by two-wave workgroups, one chunk each.
__global__ void __launch_bounds__(2 * kWave) __attribute__((amdgpu_waves_per_eu(kWavesPair)))
inflate_chunks2_kernel(const uint8_t *__restrict__ comp, uint64_t comp_bytes,
                       const uint64_t *__restrict__ off, uint64_t n, uint64_t chunk_bytes,
                       uint8_t *__restrict__ out, uint64_t *__restrict__ out_len,
                       int32_t *__restrict__ status, uint32_t flags) {
  __shared__ Lds2 S;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x) / kWave;
  const uint32_t lane = lane_id();
  const uint64_t c = blockIdx.x;
  const uint64_t o0 = off[c], o1 = off[c + 1];
  if (o1 < o0 || o1 > comp_bytes) {  // no stream: neither wave decodes
    if (threadIdx.x == 0) {
      if (status) status[c] = status_code(kCorrupt, flags);
      if (out_len) out_len[c] = 0;
    }
    return;
  }
  if (threadIdx.x == 0) {
    S.q.head = 0;
    S.q.tail = 0;
    S.q.wstat = kOk;
    S.q.done = 0;
  }
  __syncthreads();
  if (wave == 0) {
    SinkQueue sink;
    sink.q = &S.q;
    sink.head = 0;
    sink.tail = 0;
    sink.ended = false;
    sink.stuck = false;
    Result r = inflate_stream_t(comp + o0, o1 - o0, out + c * chunk_bytes, chunk_bytes, flags, S.dec, sink);
    if (!sink.ended) {
      Result t;
      const int32_t e = sink.finish(lane, r.status, t);
      if (r.status == kOk) r.status = e;
    }
    if (lane == 0) {
      if (status) status[c] = status_code(r.status, flags);
      if (out_len) out_len[c] = r.bytes;
    }
  } else {
    inflate_write(S.q, S.ring, out + c * chunk_bytes, chunk_bytes, comp + o0);
  }
}

__global__ void zero_counter_kernel(uint32_t *c) { *c = 0; }

// Device memory of the allocation-free batch entry point: word 0 is its
// work-queue counter. Every GPU holds its own copy for the life of the
// process; all calls on one GPU share it.
constexpr uint32_t kScratchWords = 64;  // 256 bytes
static __device__ __align__(256) uint32_t scratch[kScratchWords];
constexpr int kScratchDevices = 64;

}  // namespace inflate

// Device address of `scratch` on the current device, looked up once per
// device with that device current (the host-side name is not a device
// address and must never reach a kernel).
hipError_t inflate_scratch_ptr(uint32_t **out) {
  static std::mutex mu;
  static uint32_t *ptrs[inflate::kScratchDevices] = {};
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= inflate::kScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(inflate::scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    ptrs[dev] = static_cast<uint32_t *>(p);
  }
  *out = ptrs[dev];
  return hipSuccess;
}

hipError_t launch_inflate_chunks(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *off,
                                 uint64_t n,
                                 uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                                 int32_t *status, uint32_t flags, hipStream_t stream,
                                 bool two_wave) {
  if (n == 0) return hipSuccess;
  if (two_wave)
    hipLaunchKernelGGL(inflate::inflate_chunks2_kernel, dim3((uint32_t)n), dim3(2 * kWaveSize), 0,
                       stream, comp, comp_bytes, off, n, chunk_bytes, out, out_len, status, flags);
  else
    hipLaunchKernelGGL(inflate::inflate_chunks_kernel, dim3((uint32_t)n), dim3(kWaveSize), 0,
                       stream, comp, comp_bytes, off, n, chunk_bytes, out, out_len, status, flags);
  return hipGetLastError();
}

hipError_t launch_inflate_batch(const codag_stream_desc *descs, uint64_t n,
                                codag_stream_result *results, uint32_t flags, uint32_t *counter,
                                uint32_t groups, hipStream_t stream, bool two_wave) {
  hipLaunchKernelGGL(inflate::zero_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
  if (n == 0) return hipGetLastError();
  if (two_wave)
    hipLaunchKernelGGL(inflate::inflate_batch2_kernel, dim3(groups), dim3(2 * kWaveSize), 0, stream,
                       descs, n, results, flags, counter);
  else
    hipLaunchKernelGGL(inflate::inflate_batch_kernel, dim3(groups), dim3(kWaveSize), 0, stream,
                       descs, n, results, flags, counter);
  return hipGetLastError();
}

}  // namespace codag
