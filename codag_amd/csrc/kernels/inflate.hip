// inflate.hip -- gfx950 zlib/Deflate (RFC 1950/1951) decoder, one wave64 per stream.
//
// Mapping onto CDNA:
//  * Bit reader: wave-uniform (SGPR) 64-bit bit buffer. The compressed input
//    is held in two VGPR windows -- lane i holds dword i of a 256-byte window
//    -- and refilled into the bit buffer with v_readlane, so a refill costs no
//    memory round trip; the next window is prefetched with one coalesced load.
//  * Huffman decode: per-wave LDS lookup tables built per block by all 64
//    lanes (10-bit literal/length table, 8-bit distance table, canonical
//    count/symbol arrays for the rare longer codes). Table entries carry the
//    RFC 1951 length/distance base and extra-bit count directly.
//  * Literals: accumulated one byte per lane (lane k holds the k-th) and written
//    with one coalesced 64-lane store per 64 literals (or before a match).
//  * LZ77 matches: lanes copy 64 bytes per step from the stream's own output
//    in global memory (the window is the chunk itself); a workgroup fence
//    (s_waitcnt) is issued only when a match reads bytes whose stores may not
//    have retired yet.
//  * Adler-32: each lane accumulates sum(b) and sum(pos*b) over the bytes it
//    writes; the wave reduces them at the end, so the trailer is verified
//    without a second pass over the output.
//
// Reference counterpart: include/zlib/zlib.h:807-866 (full_warp_shared_lut),
// 531-707 (dynamic header / table build), 712-805 (symbol loop), 188-347
// (warp copy). Differences: stored blocks are decoded (the reference's branch
// is empty, zlib.h:840-842), BTYPE 11 is rejected (zlib.h:843 treats it as
// dynamic), the zlib header and Adler-32 trailer are checked as zlib does.
#include "device.hpp"
#include "kernels.hpp"

namespace codag {
namespace inflate {

using namespace codag::dev;

constexpr uint32_t kLitBits = 10;
constexpr uint32_t kDistBits = 8;
constexpr uint32_t kCodeBits = 7;

enum : int32_t { kOk = 0, kTrunc = 1, kOverflow = 2, kCorrupt = 3 };
enum : uint32_t { kKindLit = 0, kKindLen = 1, kKindEob = 2, kKindBad = 3 };
enum : uint32_t { kTabLit = 0, kTabDist = 1, kTabCodes = 2 };

// RFC 1951 section 3.2.5 base/extra tables (also include/zlib/zlib.h:35-51).
__constant__ uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                      31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
__constant__ uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2,
                                      2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
__constant__ uint16_t kDistBase[30] = {1,    2,    3,    4,    5,    7,     9,     13,    17,  25,
                                       33,   49,   65,   97,   129,  193,   257,   385,   513, 769,
                                       1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
__constant__ uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2,  3,  3,  4,  4,  5,  5,  6,
                                       6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
// Order of the code-length code lengths (RFC 1951 3.2.7).
__constant__ uint8_t kClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

template <int NSYM> struct Huff {
  uint16_t count[16];
  uint16_t first[16];
  uint16_t index[16];
  uint16_t sym[NSYM];
};

constexpr uint32_t kRingBytes = 4096;  // output ring (see the output side below)

struct Lds {
  alignas(16) uint8_t ring[kRingBytes];
  uint32_t lit[1u << kLitBits];
  uint32_t dist[1u << kDistBits];
  Huff<288> hl;
  Huff<32> hd;
  uint8_t lens[320];
};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// ------------------------------------------------------------- bit reader

struct Bits {
  const uint32_t *w;  // 4-aligned address at or below the stream start
  uint32_t limit;     // bit index (from w) one past the stream's last bit
  uint32_t nw;        // dwords that hold stream bytes
  uint32_t next;      // next dword to append to buf
  uint64_t buf;
  uint32_t cnt;
  uint32_t wbase;     // dword index held by lane 0 of win
  uint32_t win, win2; // dwords [wbase, wbase+64) and [wbase+64, wbase+128)
};

__device__ __forceinline__ void bits_windows(Bits &b) {
  const uint32_t lane = lane_id();
  uint32_t i = b.wbase + lane;
  b.win = i < b.nw ? b.w[i] : 0u;
  i += kWave;
  b.win2 = i < b.nw ? b.w[i] : 0u;
}

__device__ __forceinline__ void bits_refill(Bits &b) {
  if (b.cnt <= 32) {
    uint32_t k = b.next - b.wbase;
    if (k >= (uint32_t)kWave) {
      b.win = b.win2;
      b.wbase += kWave;
      const uint32_t i = b.wbase + kWave + lane_id();
      b.win2 = i < b.nw ? b.w[i] : 0u;
      k -= kWave;
    }
    const uint32_t d = __builtin_amdgcn_readlane(b.win, k);
    b.buf |= (uint64_t)d << b.cnt;
    b.cnt += 32;
    b.next++;
  }
}

__device__ __forceinline__ void bits_drop(Bits &b, uint32_t n) {
  b.buf >>= n;
  b.cnt -= n;
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

// The most recent kRing output bytes live in an LDS ring. Byte x of the
// stream sits at ring[(x + mis) & (kRing-1)], where mis = dst & 15, so a
// 16-byte aligned output address is also 16-byte aligned in the ring.
//  * literals are buffered one per lane and written to the ring in one
//    64-lane LDS store;
//  * matches with distance <= kNear copy ring -> ring (LDS only);
//  * farther matches read bytes that were flushed to HBM long ago; they are
//    queued as pieces of <= 64 bytes (lane j holds piece j) and resolved in
//    groups of 16 pieces per memory round trip;
//  * every kFlush bytes the ring is written to HBM with 16-byte stores, and
//    Adler-32 partial sums (sum b, sum pos*b) are accumulated on the way.
// A global load waits for all earlier stores of the wave (one in-order
// counter), so the design keeps global reads rare and batched.
constexpr uint32_t kRing = kRingBytes;
constexpr uint32_t kRingMask = kRing - 1;
constexpr uint32_t kFlush = 1024;
constexpr uint32_t kNear = kRing - 258 - 64 - kFlush;  // ring copies never read overwritten bytes
constexpr uint32_t kPieceGroup = 8;
// Latency of the serial symbol loop is hidden by other waves on the SIMD:
// keep register use low enough for this many waves per SIMD.
constexpr int kMinWavesPerEu = 6;

// Positions are 32-bit (streams and outputs below 2^31 bytes) so the
// wave-uniform state fits in scalar registers without spilling.
struct Out {
  uint8_t *dst;      // global output
  uint8_t *ring;     // LDS ring
  uint32_t cap;
  uint32_t pos;      // bytes produced (including buffered literals)
  uint32_t flushed;  // bytes [0, flushed) have been stored to dst
  uint32_t retired;  // bytes [0, retired) are known to have reached memory
  uint32_t mis;      // dst & 15
  uint32_t litbuf;   // pending literal byte of this lane
  uint32_t nlit;     // pending literals (uniform)
  uint32_t np;       // queued far pieces (uniform)
  uint32_t pdst, psrc, plen;  // piece `lane` (valid for lane < np)
  uint32_t hole_lo, hole_hi;  // output range covered by queued pieces
  uint32_t src_hi;            // highest source byte + 1 of queued pieces
  uint64_t s1, s2;            // per-lane Adler-32 partial sums
};

__device__ __forceinline__ uint32_t ridx(const Out &o, uint32_t x) {
  return (x + o.mis) & kRingMask;
}

__device__ __forceinline__ void out_lits_to_ring(Out &o) {
  if (o.nlit) {
    const uint32_t lane = lane_id();
    const uint64_t p0 = o.pos - o.nlit;
    if (lane < o.nlit) o.ring[ridx(o, p0 + lane)] = (uint8_t)o.litbuf;
    o.nlit = 0;
  }
}

__device__ __forceinline__ void out_maybe_flush(Out &o);

__device__ __forceinline__ void out_literal(Out &o, uint32_t byte) {
  if (lane_id() == o.nlit) o.litbuf = byte;
  o.nlit++;
  o.pos++;
  if (o.nlit == (uint32_t)kWave) {
    out_lits_to_ring(o);
    out_maybe_flush(o);
  }
}

// Copies the queued far pieces from HBM into the ring.
__device__ __forceinline__ void out_resolve(Out &o) {
  if (!o.np) return;
  if (o.src_hi > o.retired) {  // their sources must have reached memory
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    o.retired = o.flushed;
  }
  const uint32_t lane = lane_id();
  for (uint32_t g = 0; g < o.np; g += kPieceGroup) {
    uint32_t v[kPieceGroup];
#pragma unroll
    for (uint32_t k = 0; k < kPieceGroup; k++) {
      const uint32_t j = g + k;
      v[k] = 0;
      if (j < o.np) {
        const uint32_t s = __builtin_amdgcn_readlane(o.psrc, j);
        const uint32_t l = __builtin_amdgcn_readlane(o.plen, j);
        if (lane < l) v[k] = o.dst[s + lane];
      }
    }
#pragma unroll
    for (uint32_t k = 0; k < kPieceGroup; k++) {
      const uint32_t j = g + k;
      if (j < o.np) {
        const uint32_t d = __builtin_amdgcn_readlane(o.pdst, j);
        const uint32_t l = __builtin_amdgcn_readlane(o.plen, j);
        if (lane < l) o.ring[ridx(o, d + lane)] = (uint8_t)v[k];
      }
    }
  }
  o.np = 0;
  o.hole_lo = 0xffffffffu;
  o.hole_hi = 0;
  o.src_hi = 0;
}

__device__ __forceinline__ void adler_add(Out &o, uint32_t at, uint32_t b) {
  o.s1 += b;
  o.s2 += (uint64_t)at * b;
}

// Stores ring bytes [flushed, end) to HBM; `last` also writes a partial
// 16-byte tail. Pieces must have been resolved.
__device__ __forceinline__ void out_store(Out &o, uint32_t end, bool last) {
  const uint32_t lane = lane_id();
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
    // sum b and sum i*b over the 16 bytes, then s2 += at*sum + sum(i*b)
    const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
    uint32_t sum = 0, wsum = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const uint32_t b = (w4[i >> 2] >> (8 * (i & 3))) & 0xff;
      sum += b;
      wsum += (uint32_t)i * b;
    }
    o.s1 += sum;
    o.s2 += (uint64_t)at * sum + wsum;
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
}

__device__ __forceinline__ void out_maybe_flush(Out &o) {
  if (o.pos - o.flushed >= kFlush) {
    out_lits_to_ring(o);
    out_resolve(o);
    out_store(o, o.pos, false);
  }
}

// Appends the len bytes that start dist bytes back (periodic when dist < len).
__device__ __forceinline__ void out_match(Out &o, uint32_t dist, uint32_t len) {
  out_lits_to_ring(o);
  const uint32_t lane = lane_id();
  const uint32_t src0 = o.pos - dist;
  if (dist <= kNear) {
    const uint32_t span = dist < len ? dist : len;
    if (src0 < o.hole_hi && src0 + span > o.hole_lo) out_resolve(o);
    __syncthreads();  // ring writes above are visible to the reads below
    if (dist >= len) {
      for (uint32_t i = lane; i < len; i += kWave)
        o.ring[ridx(o, o.pos + i)] = o.ring[ridx(o, src0 + i)];
    } else {
      for (uint32_t i = lane; i < len; i += kWave)
        o.ring[ridx(o, o.pos + i)] = o.ring[ridx(o, src0 + (i % dist))];
    }
  } else {  // far: queue <= 64-byte pieces, sources are already in HBM
    const uint32_t pieces = (len + kWave - 1) / kWave;
    if (o.np + pieces > (uint32_t)kWave) out_resolve(o);
    for (uint32_t k = 0; k < pieces; k++) {
      if (lane == o.np) {
        o.pdst = o.pos + k * kWave;
        o.psrc = src0 + k * kWave;
        o.plen = (len - k * kWave) < (uint32_t)kWave ? (len - k * kWave) : (uint32_t)kWave;
      }
      o.np++;
    }
    if (o.pos < o.hole_lo) o.hole_lo = o.pos;
    o.hole_hi = o.pos + len;
    if (src0 + len > o.src_hi) o.src_hi = src0 + len;
  }
  o.pos += len;
  out_maybe_flush(o);
}

__device__ __forceinline__ void out_reduce_mod(Out &o) {
  o.s1 %= 65521u;
  o.s2 %= 65521u;
}

// ------------------------------------------------------------- table build

__device__ __forceinline__ uint32_t make_entry(uint32_t tab, uint32_t sym, uint32_t len) {
  if (tab == kTabCodes) return len | (sym << 8);
  if (tab == kTabDist) {
    if (sym >= 30) return len | (kKindBad << 4);
    return len | ((uint32_t)kDistBase[sym] << 8) | ((uint32_t)kDistExtra[sym] << 24);
  }
  if (sym < 256) return len | (kKindLit << 4) | (sym << 8);
  if (sym == 256) return len | (kKindEob << 4);
  if (sym <= 285)
    return len | (kKindLen << 4) | ((uint32_t)kLenBase[sym - 257] << 8) |
           ((uint32_t)kLenExtra[sym - 257] << 20);
  return len | (kKindBad << 4);
}

// Builds the canonical tables for n code lengths in lens[] and the LUT of
// 2^bits entries. Applies zlib's acceptance rules (inftrees.c): an
// over-subscribed set is rejected; an incomplete set is rejected unless it
// has a single code of length 1 (literal/distance tables) or no codes at all.
// Out of line: it runs once per block, and keeping its unrolled code out of
// the symbol loop leaves that loop's uniform state in scalar registers.
template <int NSYM>
__device__ __noinline__ int32_t build_table(const uint8_t *lens, uint32_t n, uint32_t tab, Huff<NSYM> &h,
                               uint32_t *lut, uint32_t bits) {
  const uint32_t lane = lane_id();
  if (lane < 16) h.count[lane] = 0;
  for (uint32_t i = lane; i < (1u << bits); i += kWave) lut[i] = 0;
  __syncthreads();
  for (uint32_t s = lane; s < n; s += kWave) {
    const uint32_t L = lens[s];
    if (L) atomicAdd(reinterpret_cast<uint32_t *>(&h.count[0]) + (L >> 1), 1u << (16 * (L & 1)));
  }
  __syncthreads();
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
  __syncthreads();
  if (lane >= 1 && lane < 16) {
    h.first[lane] = (uint16_t)first_l;
    h.index[lane] = (uint16_t)index_l;
  }
  __syncthreads();
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
  __syncthreads();
  return kOk;
}

// Canonical decode of codes longer than the LUT (or invalid ones), bit by bit
// (the count/first walk of RFC 1951 canonical codes). Returns the symbol or
// -1; *len gets the code length.
template <int NSYM>
__device__ __forceinline__ int32_t slow_decode(const Huff<NSYM> &h, uint64_t buf, uint32_t *len) {
  uint32_t code = 0, first = 0, index = 0;
  for (uint32_t L = 1; L < 16; L++) {
    code |= (uint32_t)(buf & 1);
    buf >>= 1;
    const uint32_t count = uni(h.count[L]);
    if (code - first < count) {
      *len = L;
      return (int32_t)uni(h.sym[index + code - first]);
    }
    index += count;
    first = (first + count) << 1;
    code <<= 1;
  }
  return -1;
}

// ------------------------------------------------------------- stream decode

struct Result {
  int32_t status;
  uint64_t bytes;
  uint32_t adler;
};

__device__ __forceinline__ int32_t read_dynamic(Bits &b, Lds &S, uint32_t *nlen_out) {
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
  __syncthreads();
  int32_t st = build_table(S.lens, 19, kTabCodes, S.hl, S.lit, kCodeBits);
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
      __syncthreads();
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
  __syncthreads();
  if (uni(S.lens[256]) == 0) return kCorrupt;  // zlib: "missing end-of-block"
  st = build_table(S.lens, nlen, kTabLit, S.hl, S.lit, kLitBits);
  if (st != kOk) return st;
  st = build_table(S.lens + nlen, ndist, kTabDist, S.hd, S.dist, kDistBits);
  *nlen_out = nlen;
  return st;
}

__device__ __forceinline__ void load_fixed(Lds &S) {
  const uint32_t lane = lane_id();
  for (uint32_t s = lane; s < 288; s += kWave)
    S.lens[s] = (uint8_t)(s < 144 ? 8 : (s < 256 ? 9 : (s < 280 ? 7 : 8)));
  __syncthreads();
  build_table(S.lens, 288, kTabLit, S.hl, S.lit, kLitBits);
  for (uint32_t s = lane; s < 32; s += kWave) S.lens[s] = 5;
  __syncthreads();
  build_table(S.lens, 32, kTabDist, S.hd, S.dist, kDistBits);
}

// Decodes symbols until end-of-block. Returns a status.
__device__ __forceinline__ int32_t decode_block(Bits &b, Out &o, Lds &S) {
  for (;;) {
    bits_refill(b);
    uint32_t e = uni(S.lit[b.buf & ((1u << kLitBits) - 1)]);
    uint32_t L = e & 15;
    if (L == 0) {
      int32_t sym = slow_decode(S.hl, b.buf, &L);
      if (sym < 0) return bits_past_end(b) ? kTrunc : kCorrupt;
      e = make_entry(kTabLit, (uint32_t)sym, L);
    }
    bits_drop(b, L);
    const uint32_t kind = (e >> 4) & 15;
    if (kind == kKindLit) {
      if (bits_past_end(b)) return kTrunc;
      if (o.pos >= o.cap) return kOverflow;
      out_literal(o, (e >> 8) & 0xff);
      continue;
    }
    if (kind == kKindEob) return bits_past_end(b) ? kTrunc : kOk;
    if (kind == kKindBad) return kCorrupt;
    // length
    const uint32_t lx = (e >> 20) & 15;
    const uint32_t len = ((e >> 8) & 0x1ff) + (uint32_t)(b.buf & ((1u << lx) - 1));
    bits_drop(b, lx);
    bits_refill(b);
    uint32_t d = uni(S.dist[b.buf & ((1u << kDistBits) - 1)]);
    uint32_t dl = d & 15;
    if (dl == 0) {
      int32_t sym = slow_decode(S.hd, b.buf, &dl);
      if (sym < 0) return bits_past_end(b) ? kTrunc : kCorrupt;
      d = make_entry(kTabDist, (uint32_t)sym, dl);
    }
    if ((d >> 4) & 15) return kCorrupt;  // distance code 30/31
    bits_drop(b, dl);
    const uint32_t dx = (d >> 24) & 15;
    const uint32_t dist = ((d >> 8) & 0xffff) + (uint32_t)(b.buf & ((1u << dx) - 1));
    bits_drop(b, dx);
    if (bits_past_end(b)) return kTrunc;
    if (dist > o.pos) return kCorrupt;  // zlib: "invalid distance too far back"
    if (o.pos + len > o.cap) return kOverflow;
    out_match(o, dist, len);
  }
}

__device__ __forceinline__ Result inflate_stream(const uint8_t *src, uint64_t len, uint8_t *dst, uint64_t cap,
                                 uint32_t flags, Lds &S) {
  const uint32_t lane = lane_id();
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
  const bool raw = flags & CODAG_INFLATE_RAW;
  uint64_t start = 0;
  if (!raw) {
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

  Out o;
  o.dst = dst;
  o.ring = S.ring;
  o.cap = (uint32_t)cap;
  o.pos = 0;
  o.flushed = 0;
  o.retired = 0;
  o.mis = (uint32_t)(reinterpret_cast<uintptr_t>(dst) & 15);
  o.litbuf = 0;
  o.nlit = 0;
  o.np = 0;
  o.pdst = o.psrc = o.plen = 0;
  o.hole_lo = 0xffffffffu;
  o.hole_hi = 0;
  o.src_hi = 0;
  o.s1 = 0;
  o.s2 = 0;

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
      if (o.pos + ln > o.cap) { st = kOverflow; break; }
      out_lits_to_ring(o);
      for (uint32_t done = 0; done < ln;) {  // through the ring, <= 256 bytes per step
        const uint32_t step = (ln - done) < 256u ? (ln - done) : 256u;
        for (uint32_t i = lane; i < step; i += kWave)
          o.ring[ridx(o, o.pos + i)] = src[at + done + i];
        o.pos += step;
        done += step;
        out_maybe_flush(o);
      }
      bits_seek(b, lead + at + ln);
    } else if (type == 1) {  // fixed Huffman
      if (tables != 1) {
        load_fixed(S);
        tables = 1;
      }
      st = decode_block(b, o, S);
    } else if (type == 2) {  // dynamic Huffman
      uint32_t nlen = 0;
      tables = 0;
      st = read_dynamic(b, S, &nlen);
      if (st == kOk) st = decode_block(b, o, S);
    } else {
      st = kCorrupt;  // zlib: "invalid block type"
    }
    out_reduce_mod(o);
  }
  // drain: pending literals and far pieces, then the last partial flush
  out_lits_to_ring(o);
  out_resolve(o);
  __syncthreads();
  out_store(o, o.pos, true);
  // wave reduction of the Adler-32 partial sums
  uint64_t s1 = o.s1 % 65521u, s2 = o.s2 % 65521u;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    s1 += __shfl_xor(s1, d, kWave);
    s2 += __shfl_xor(s2, d, kWave);
  }
  s1 %= 65521u;
  s2 %= 65521u;
  const uint64_t n = o.pos % 65521u;
  const uint32_t A = (uint32_t)((1 + s1) % 65521u);
  const uint32_t B = (uint32_t)((n * ((1 + s1) % 65521u) + 65521u * 65521u - s2) % 65521u);
  res.adler = uni((B << 16) | A);
  if (st == kOk && !raw) {
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
  res.bytes = o.pos;
  return res;
}

__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(kMinWavesPerEu)))
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
    if (status) status[c] = r.status;
    if (out_len) out_len[c] = r.bytes;
  }
}

__global__ void __launch_bounds__(kWave) __attribute__((amdgpu_waves_per_eu(kMinWavesPerEu)))
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
    const Result r = inflate_stream(d.src, d.src_bytes, d.dst, d.dst_cap, flags, S);
    if (lane_id() == 0) {
      results[i].bytes_written = r.bytes;
      results[i].status = r.status;
      results[i].adler32 = r.adler;
    }
    __syncthreads();
  }
}

__global__ void zero_counter_kernel(uint32_t *c) { *c = 0; }

}  // namespace inflate

hipError_t launch_inflate_chunks(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *off,
                                 uint64_t n,
                                 uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                                 int32_t *status, uint32_t flags, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(inflate::inflate_chunks_kernel, dim3((uint32_t)n), dim3(kWaveSize), 0, stream,
                     comp, comp_bytes, off, n, chunk_bytes, out, out_len, status, flags);
  return hipGetLastError();
}

hipError_t launch_inflate_batch(const codag_stream_desc *descs, uint64_t n,
                                codag_stream_result *results, uint32_t flags, uint32_t *counter,
                                uint32_t groups, hipStream_t stream) {
  hipLaunchKernelGGL(inflate::zero_counter_kernel, dim3(1), dim3(1), 0, stream, counter);
  if (n == 0) return hipGetLastError();
  hipLaunchKernelGGL(inflate::inflate_batch_kernel, dim3(groups), dim3(kWaveSize), 0, stream, descs,
                     n, results, flags, counter);
  return hipGetLastError();
}

}  // namespace codag
