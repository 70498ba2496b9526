// rlev1_pipe.hip -- ORC RLE v1 decode: one workgroup per chunk running a
// software pipeline over fixed 2 KiB windows of the chunk's stream.
//
// Segment chain in terminator numbers. Every RLE v1 segment ends on a varint
// terminator (a byte < 0x80). Number the stream's terminators 1, 2, ... and
// let 0 stand for the byte before the stream. The segment whose header
// follows terminator r ends on terminator
//   r + k                      literal group of k varints (header >= 0x80),
//   r + 2 + [delta byte < 0x80] run: the header itself (< 0x80) is r + 1,
// so the walk from header to header needs, per step, only the two bytes
// after terminator r (ent[r]). The walker holds its window's entries in its
// own registers and reads one per segment with register indexing and
// v_readlane -- no memory access on the walker's critical path.
//
// Pipeline (phase k, one workgroup barrier per phase):
//   helpers (waves 1-7; waves 1-2 load, stage and index)
//     load   : window k+2, 16 bytes per lane, into registers (HBM latency is
//              hidden behind two phases);
//     stage  : window k into an LDS ring of four windows; per 16-byte block
//              the terminator count;
//     index  : window k-1: block prefix sums number the terminators; for
//              each, sel = its ring position and ent = the two bytes after;
//     expand : window k-3, one lane per output value;
//   walker (wave 0): window k-2's part of the chain in scalar registers,
//              one descriptor per segment.
// Windows are fixed slices of the address range, so loading and indexing
// never wait for the walk. A segment may run into the next window, which is
// indexed before the segment is expanded (a well-formed segment spans at
// most 1 + 128 * 10 bytes < one window).
//
// Semantics and status follow include/rlev1/decoder.h:671-791 as restated in
// the oracle (first error in stream order): the walker reports a missing
// header byte or delta byte (TRUNCATED) and a segment that overflows the
// chunk (OVERFLOW at value index cap); the expanders report varints longer
// than 10 bytes (CORRUPT) and varints with no terminator among the indexed
// windows (TRUNCATED when the stream ends there; otherwise the varint is
// overlong and ends CORRUPT if any terminator follows, TRUNCATED if none).
#include "device.hpp"
#include "kernels.hpp"

namespace codag {
namespace rle1p {

using namespace codag::dev;

enum : int32_t { kOk = 0, kTrunc = 1, kOverflow = 2, kCorrupt = 3 };

constexpr uint32_t kWin = 2048;            // window bytes
constexpr uint32_t kBlocks = kWin / 16;    // 16-byte blocks per window
constexpr uint32_t kHelpers = 7;
constexpr uint32_t kThreads = kWave * (1 + kHelpers);
static_assert(kHelpers == 7, "expansion passes are dealt in rounds of 1 + 1 + 5 * 2");
constexpr uint32_t kRing = 4 * kWin;       // LDS byte ring
constexpr uint32_t kMirror = 32;           // copy of the ring's first bytes past its end
constexpr uint32_t kTerms = kWin + 2;      // terminators of one window (<= kWin: window 0's byte
                                           // before the stream stands for terminator 0)
constexpr uint32_t kDescs = kWin / 2 + 4;  // segments starting in one window (each spans >= 2 bytes)
constexpr uint32_t kTabRegs = kWin / 2 / kWave;  // walker registers holding a window's ent
constexpr uint32_t kMaxCap = 1u << 21;      // values per chunk (descriptor field)
constexpr uint32_t kSpare = kTerms - 1;    // sel / ent entry written for non-terminator bytes
constexpr uint32_t kNone = 0xffffffffu;

// Error keys: (value index << 2 | priority) << 32 | window. The smallest key
// is the first error in stream order; at one value index a varint read error
// comes before the overflow it would otherwise have hit.
enum : uint32_t { kPrCorrupt = 0, kPrTrunc = 1, kPrPending = 2, kPrOverflow = 3 };
constexpr uint64_t kNoKey = ~0ull;

struct Lds {
  uint8_t sbuf[kRing + kMirror];
  uint16_t sel[3][kTerms];     // ring byte offset of each terminator
  uint16_t ent[2][kTerms];     // the two stream bytes after each terminator
  uint32_t desc[2][kDescs];    // segment: r - T_w (11 bits) | first value index << 11
  uint8_t bcnt[2][kBlocks];    // terminators per 16-byte block
  uint32_t T[8];               // T[j & 7]: number of window j's first terminator
  uint32_t wnd[2][4];          // walked window: segments, first value, end value
  uint8_t mark[kHelpers][kWave];
  uint32_t last_w;             // window in which the walker stopped
  uint32_t lastterm;           // the stream's last byte is a terminator (or the stream is empty)
  uint64_t key;
};

// Bit j set where byte j of the 16-byte block is < 0x80.
__device__ __forceinline__ uint32_t term_mask16(const uint4 &x) {
  const uint32_t q[4] = {x.x, x.y, x.z, x.w};
  uint32_t m = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const uint32_t n = ~q[i] & 0x80808080u;
    m |= (((n >> 7) & 1) | ((n >> 14) & 2) | ((n >> 21) & 4) | ((n >> 28) & 8)) << (4 * i);
  }
  return m;
}

// 16 bytes at window block b (offset a0 + w*kWin + 16*b) if the block holds
// a stream byte, else zeros; the last block also fetches the first dword of
// the next window. Only blocks that hold stream bytes are read, so no load
// leaves the pages the stream occupies.
__device__ __forceinline__ void load_block(uintptr_t a0, uint32_t w, uint32_t b, uintptr_t s,
                                           uintptr_t e, uint4 &d, uint32_t &dn) {
  const uintptr_t a = a0 + (uintptr_t)w * kWin + 16u * b;
  d = (a + 16 > s && a < e) ? *reinterpret_cast<const uint4 *>(a) : make_uint4(0, 0, 0, 0);
  dn = 0;
  if (b == kBlocks - 1 && a + 16 < e) dn = *reinterpret_cast<const uint32_t *>(a + 16);
}

// Value of the LEB128 varint at ring bytes [s, s + n), n in 1..10 (bits past
// 64 dropped). Reads four LDS dwords from s & ~3 (the mirror covers the wrap).
template <uint32_t W>
__device__ __forceinline__ uint64_t ring_varint(const uint8_t *sbuf, uint32_t s, uint32_t n) {
  const uint32_t *d = reinterpret_cast<const uint32_t *>(sbuf + (s & ~3u));
  const uint32_t d0 = d[0], d1 = d[1], d2 = d[2], d3 = d[3];
  const uint32_t sh = s & 3;
  const uint32_t b0 = __builtin_amdgcn_alignbyte(d1, d0, sh);  // bytes s .. s+3
  const uint32_t b1 = __builtin_amdgcn_alignbyte(d2, d1, sh);  // bytes s+4 .. s+7
  // 7-bit groups of bytes s .. s+7: pairs into 14 bits, then into 28 bits
  const uint32_t p0 = (b0 & 0x007f007fu) | ((b0 >> 1) & 0x3f803f80u);
  const uint32_t p1 = (b1 & 0x007f007fu) | ((b1 >> 1) & 0x3f803f80u);
  const uint32_t q0 = (p0 & 0x3fffu) | ((p0 >> 2) & 0x0fffc000u);
  const uint32_t q1 = (p1 & 0x3fffu) | ((p1 >> 2) & 0x0fffc000u);
  uint32_t lo = q0 | (q1 << 28);
  if (W <= 4) {  // the column keeps the low 32 bits: 5 bytes decide them
    if (n <= 4) lo &= (1u << (7 * n)) - 1;
    return lo;
  }
  uint32_t hi = q1 >> 4;
  if (n > 8) {
    const uint32_t b2 = __builtin_amdgcn_alignbyte(d3, d2, sh);  // bytes s+8 .. s+11
    hi |= (b2 & 0x7fu) << 24;
    if (n > 9) hi |= (b2 & 0x100u) << 23;
  } else if (n <= 4) {
    lo &= (1u << (7 * n)) - 1;
    hi = 0;
  } else {
    hi &= (1u << (7 * n - 32)) - 1;
  }
  return ((uint64_t)hi << 32) | lo;
}

template <uint32_t W> __device__ __forceinline__ void store_val(uint8_t *dst, uint32_t i, uint64_t v) {
  if (W == 1) dst[i] = (uint8_t)v;
  else if (W == 2) reinterpret_cast<uint16_t *>(dst)[i] = (uint16_t)v;
  else if (W == 4) reinterpret_cast<uint32_t *>(dst)[i] = (uint32_t)v;
  else reinterpret_cast<uint64_t *>(dst)[i] = v;
}

__device__ __forceinline__ void report(Lds &L, uint32_t idx, uint32_t pr, uint32_t w) {
  atomicMin(reinterpret_cast<unsigned long long *>(&L.key),
            (unsigned long long)((((uint64_t)idx << 2) | pr) << 32 | w));
}

// Waits for this wave's LDS traffic, then the workgroup barrier. Global
// loads stay in flight across it (the prefetched windows).
__device__ __forceinline__ void phase_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Four workgroups (32 waves) per CU: every chunk of a 1024-chunk column is
// resident at once, which needs <= 64 VGPRs per lane.
template <uint32_t W>
__global__ void __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(8)))
rlev1_pipe_kernel(const uint8_t *__restrict__ comp, uint64_t comp_bytes,
                  const uint64_t *__restrict__ blk_off, const uint64_t *__restrict__ col_len,
                  uint64_t chunk_bytes, uint8_t *__restrict__ out, int32_t *__restrict__ status) {
  __shared__ Lds L;
  const uint64_t chunk = blockIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wave = uni(threadIdx.x / kWave);
  const uint64_t boff = blk_off[chunk];
  const uint64_t len64 = col_len[chunk];
  if (boff > comp_bytes || len64 > comp_bytes - boff || len64 >= (1ull << 30)) {
    if (threadIdx.x == 0 && status) status[chunk] = kCorrupt;  // uniform over the workgroup
    return;
  }
  const uint32_t len = (uint32_t)len64;
  const uint32_t cap = (uint32_t)(chunk_bytes / W);
  uint8_t *dst = out + chunk * chunk_bytes;
  const uintptr_t s_addr = reinterpret_cast<uintptr_t>(comp + boff);
  const uintptr_t e_addr = s_addr + len;
  const uintptr_t a0 = (s_addr - 1) & ~uintptr_t(15);   // window 0 holds the byte before the stream
  const uint32_t lead = (uint32_t)(s_addr - a0);        // 1..16
  const uint32_t nwin = (lead + len + kWin - 1) / kWin;  // >= 1

  if (threadIdx.x == 0) {
    L.T[0] = 0;
    L.last_w = kNone;
    L.lastterm = 0;
    L.key = kNoKey;
  }

  // helper lane gl = (wave - 1) * 64 + lane loads / stages / indexes block gl
  const uint32_t gl = (wave - 1) * kWave + lane;
  const bool loader = wave >= 1 && gl < kBlocks;
  uint4 pa = make_uint4(0, 0, 0, 0), pb = pa;  // prefetched windows k and k+1
  uint32_t pna = 0, pnb = 0;
  uint4 y = pa;                                // window k-1 (indexed this phase)
  uint32_t yn = 0, ym = 0;
  if (loader) {
    load_block(a0, 0, gl, s_addr, e_addr, pa, pna);
    if (nwin > 1) load_block(a0, 1, gl, s_addr, e_addr, pb, pnb);
  }

  // walker state (wave 0, scalar); the walker is the pipeline's critical
  // path, so its instructions issue ahead of the helpers'
  uint32_t r = 0, o = 0, wdone = 0;
  if (wave == 0) __builtin_amdgcn_s_setprio(3);
  else if (wave <= 2) __builtin_amdgcn_s_setprio(2);  // the waves that also load and index

  phase_barrier();
  for (uint32_t k = 0;; k++) {
    if (wave == 0) {
      // ------------------------------------------------------------ walk window k-2
      const int32_t w = (int32_t)k - 2;
      if (w >= 0 && !wdone) {
        const uint32_t Tw = uni(L.T[w & 7]);
        const uint32_t Tn = uni(L.T[(w + 1) & 7]);
        const bool last = (uint32_t)w == nwin - 1;
        const uint32_t lastterm = last ? uni(L.lastterm) : 0u;
        // In the last window, when the stream ends on a terminator, the last
        // two candidates may lack a header or delta byte (checked from Tfast on).
        const uint32_t Tfast = lastterm ? (Tn >= 2 ? Tn - 2 : 0u) : Tn;
        // the window's step entries in registers: lane l of tab[v] holds
        // entries 2 (64 v + l) and 2 (64 v + l) + 1
        typedef uint32_t tab_t __attribute__((ext_vector_type(kTabRegs)));
        tab_t tab;
        const uint32_t *ent32 = reinterpret_cast<const uint32_t *>(L.ent[w & 1]);
#pragma unroll
        for (uint32_t v = 0; v < kTabRegs; v++) tab[v] = ent32[v * kWave + lane];
        uint32_t *desc = L.desc[w & 1];
        const uint32_t o_first = o;
        uint32_t nd = 0, stop = 0, ovf = 0;
        // Step entry of local terminator rl, from the walker's registers.
#define CODAG_ENT(rl) \
  ((__builtin_amdgcn_readlane(tab[(rl) >> 7], ((rl) >> 1) & (kWave - 1)) >> (((rl) & 1) << 4)) & 0xffff)
        // One segment: run (header h < 0x80, decoder.h:736-790) of h + 3
        // values spanning 2 + [delta byte < 0x80] terminators, or literal
        // group (decoder.h:686-733) of 256 - h values and terminators;
        // descriptor (every lane writes the same word), chunk-full /
        // overflow check, advance.
#define CODAG_STEP(e)                                                                  \
  {                                                                                    \
    const uint32_t h_ = (e) & 0xff;                                                    \
    const uint32_t cnt = h_ < 0x80 ? h_ + 3 : 256 - h_;                                \
    const uint32_t adv = h_ < 0x80 ? 2 + ((e) < 0x8000) : cnt;                         \
    desc[nd++] = (r - Tw) | (o << 11);                                                 \
    if (o + cnt >= cap) {                                                              \
      if (o + cnt > cap) { /* values past the chunk: overflow at index cap */          \
        ovf = 1;                                                                       \
        o = cap + 1; /* the expanders still read the varint at index cap */            \
      } else {                                                                         \
        o = cap;                                                                       \
      }                                                                                \
      break;                                                                           \
    }                                                                                  \
    o += cnt;                                                                          \
    r += adv;                                                                          \
  }
        r = uni(r);
        o = uni(o);
        if (r < Tfast && o < cap) {
#pragma clang loop unroll(disable)
          do {  // the step ends the loop once the chunk is full
            const uint32_t rl = r - Tw;
            const uint32_t e = CODAG_ENT(rl);
            CODAG_STEP(e)
          } while (r < Tfast);
        }
#pragma clang loop unroll(disable)
        while (!ovf && r < Tn && o < cap) {  // the stream's last two terminators
          const uint32_t rl = r - Tw;
          if (r + 1 == Tn) {  // the header would follow the stream's last byte
            if (lane == 0) report(L, o, kPrTrunc, w);
            stop = 1;
            break;
          }
          const uint32_t e = CODAG_ENT(rl);
          if ((e & 0x80) == 0 && r + 2 == Tn) {  // a run header as the last byte: no delta byte
            if (lane == 0) report(L, o, kPrTrunc, w);
            stop = 1;
            break;
          }
          CODAG_STEP(e)
        }
#undef CODAG_STEP
#undef CODAG_ENT
        if (ovf) {
          if (lane == 0) report(L, cap, kPrOverflow, w);
          stop = 1;
        }
        if (o >= cap) stop = 1;
        if (lane == 0) {
          L.wnd[w & 1][0] = nd;
          L.wnd[w & 1][1] = o_first;
          L.wnd[w & 1][2] = o;
        }
        if (stop || last) {
          wdone = 1;
          if (lane == 0) L.last_w = (uint32_t)w;
        }
      }
    } else {
      const uint32_t h = wave - 1;
      // ------------------------------------------------------------ load / stage window k
      uint4 x = pa;
      uint32_t xn = pna;
      pa = pb;
      pna = pnb;
      if (loader && k + 2 < nwin) load_block(a0, k + 2, gl, s_addr, e_addr, pb, pnb);
      // ------------------------------------------------------------ index window k-1
      {
        const int32_t w = (int32_t)k - 1;
        if (w >= 0 && (uint32_t)w < nwin) {
          if (loader) {
            // block b's first terminator: prefix over the window's block counts
            const uint32_t c_lo = L.bcnt[w & 1][lane], c_hi = L.bcnt[w & 1][kWave + lane];
            const uint32_t s_lo = wave_scan_dpp(c_lo), s_hi = wave_scan_dpp(c_hi);
            const uint32_t tot_lo = __builtin_amdgcn_readlane(s_lo, kWave - 1);
            const uint32_t tot_hi = __builtin_amdgcn_readlane(s_hi, kWave - 1);
            if (gl == 0) L.T[(w + 1) & 7] = uni(L.T[w & 7]) + tot_lo + tot_hi;
            const uint32_t slot = w & 3;
            const uint32_t b = gl;
            const uint32_t base = b < kWave ? s_lo - c_lo : tot_lo + s_hi - c_hi;
            uint32_t dn = __shfl_down(y.x, 1, kWave);
            if (lane == kWave - 1)
              dn = b == kBlocks - 1
                       ? yn
                       : *reinterpret_cast<const uint32_t *>(L.sbuf + slot * kWin + 16 * (b + 1));
            const uint32_t gx = (uint32_t)w * kWin + 16 * b;
            const uint32_t lp = lead + len - 1 - gx;  // block byte of the stream's last byte
            if (lp < 16) L.lastterm = (ym >> lp) & 1;
            uint16_t *sel = L.sel[w % 3];
            uint16_t *ent = L.ent[w & 1];
            const uint32_t xo = slot * kWin + 16 * b;
            const uint32_t by[5] = {y.x, y.y, y.z, y.w, dn};
            // branch-free: bytes that are no terminator write the spare entry
#pragma unroll
            for (uint32_t j = 0; j < 16; j++) {
              const uint32_t ix = (ym >> j) & 1 ? base + __popc(ym & ((1u << j) - 1)) : kSpare;
              const uint32_t q = (j + 1) >> 2, sh = (j + 1) & 3;
              sel[ix] = (uint16_t)(xo + j);
              ent[ix] = (uint16_t)(sh == 0 ? by[q] : __builtin_amdgcn_alignbyte(by[q + 1], by[q], sh));
            }
          }
        } else if (w >= 0 && h == 0 && lane == 0) {
          L.T[(w + 1) & 7] = L.T[w & 7];  // past the stream: no terminators
        }
      }
      // ------------------------------------------------------------ stage window k
      if (loader && k < nwin) {
        const uint32_t slot = k & 3;
        reinterpret_cast<uint4 *>(L.sbuf + slot * kWin)[gl] = x;
        if (slot == 0 && gl < kMirror / 16) reinterpret_cast<uint4 *>(L.sbuf + kRing)[gl] = x;
        uint32_t m = term_mask16(x);
        // keep stream positions [0, len): block bytes [lead - gx, lead + len - gx)
        const int32_t gx = (int32_t)(k * kWin + 16 * gl);
        const int32_t lo = (int32_t)lead - gx, hi = (int32_t)(lead + len) - gx;
        const int32_t a = lo < 0 ? 0 : (lo > 16 ? 16 : lo);
        const int32_t z = hi < 0 ? 0 : (hi > 16 ? 16 : hi);
        const uint32_t valid = z > a ? ((0xffffu >> (16 - (z - a))) << a) : 0u;
        m &= valid;
        if (k == 0 && gl == 0) m |= 1u << (lead - 1);  // terminator 0: the byte before the stream
        L.bcnt[k & 1][gl] = (uint8_t)__popc(m);
        y = x;
        yn = xn;
        ym = m;
      }
      // ------------------------------------------------------------ expand window k-3
      const int32_t w = (int32_t)k - 3;
      if (w >= 0) {
        const uint32_t nd = uni(L.wnd[w & 1][0]);
        const uint32_t o_first = uni(L.wnd[w & 1][1]);
        const uint32_t o_end = uni(L.wnd[w & 1][2]);
        const uint32_t Tw = uni(L.T[w & 7]);
        const uint32_t n0 = uni(L.T[(w + 1) & 7]) - Tw;  // terminators of window w
        const uint32_t nk = uni(L.T[(w + 2) & 7]) - Tw;  // ... of windows w and w+1 (indexed)
        const uint32_t bad_pr = (uint32_t)w + 2 >= nwin ? kPrTrunc : kPrPending;
        const uint16_t *sel0 = L.sel[w % 3];
        // ring position of local terminator q: window w's entries, then w+1's
        const uint16_t *sel1m = L.sel[(w + 1) % 3] - n0;
        auto sel = [&](uint32_t q) -> uint32_t { return (q < n0 ? sel0 : sel1m)[q]; };
        const uint32_t *desc = L.desc[w & 1];
        // lanes exchange segment starts through these bytes (one wave's LDS
        // operations complete in order; a compiler barrier keeps the read a
        // real load after the marks)
        uint8_t *mk = L.mark[h];
        const uint32_t npass = (o_end - o_first + kWave - 1) / kWave;
        uint32_t cur = 0;
        // passes of 64 consecutive values: of every 12, helpers 0 and 1
        // (which also load, stage and index) take one each, the others two
        const uint32_t per = h < 2 ? 1u : 2u;
        const uint32_t first = h < 2 ? h : 2 * h - 2;
#pragma clang loop unroll(disable)
        for (uint32_t t = 0;; t++) {
          const uint32_t p = 12 * (t / per) + first + (t % per);
          if (p >= npass) break;
          const uint32_t ob = o_first + kWave * p;
          // cur: the last segment starting at or before value ob; lane i
          // holds segment cur + i
          uint32_t di, oi;
          bool iv;
          for (;;) {
            const uint32_t i = cur + lane;
            iv = i < nd;
            di = iv ? desc[i] : 0u;
            oi = di >> 11;
            const uint32_t c = (uint32_t)__popcll(__ballot(iv && lane >= 1 && oi <= ob));
            if (c == 0) break;
            cur += c;
          }
          // segment of value ob + lane = cur + (segments starting in (ob, ob + lane])
          mk[lane] = 0;
          if (lane >= 1 && iv && oi - ob < kWave) mk[oi - ob] = 1;
          asm volatile("" ::: "memory");
          const uint64_t starts = __ballot(mk[lane] != 0);
          const uint32_t j = __builtin_amdgcn_mbcnt_hi((uint32_t)(starts >> 32),
                                                       __builtin_amdgcn_mbcnt_lo((uint32_t)starts, 0)) +
                             (uint32_t)((starts >> lane) & 1);
          const uint32_t dsc = __shfl(di, (int)j, kWave);
          const uint32_t v = ob + lane;
          const bool act = v < o_end;
          const uint32_t rl = dsc & 0x7ff;
          const uint32_t os = dsc >> 11;
          const uint32_t ii = act ? v - os : 0u;  // lanes past the window read in range
          const uint32_t s0 = sel0[rl];  // terminator before the header
          const uint32_t hb = L.sbuf[s0 + 1], db = L.sbuf[s0 + 2];
          const bool run = hb < 0x80;
          // run (decoder.h:736-790): the base starts after the delta byte,
          // ends on terminator rl + 2 + [delta byte < 0x80] and belongs to the
          // run's first value; literal group (decoder.h:686-733): varint ii
          // ends on terminator rl + 1 + ii and starts after terminator rl + ii
          // (after the header for ii = 0)
          const uint32_t qb = run ? rl + 2 + (db < 0x80) : rl + 1 + ii;
          const bool complete = qb < nk;
          const uint32_t qa = (run || ii == 0) ? rl : qb - 1;
          const uint32_t start = sel(qa) + 1 + (run ? 2u : (ii == 0 ? 1u : 0u));
          const uint32_t n = ((sel(complete ? qb : qa) - start) & (kRing - 1)) + 1;
          if (act && complete && n <= 10) {
            if (v < cap) {
              uint64_t val = ring_varint<W>(L.sbuf, start, n);
              if (run) val += (uint64_t)((int64_t)ii * (int32_t)(int8_t)db);  // delta byte
              store_val<W>(dst, v, val);
            }
          } else if (act) {  // incomplete or longer than any 64-bit varint
            report(L, run ? os : v, complete ? kPrCorrupt : bad_pr, (uint32_t)w);
          }
        }
      }
    }
    phase_barrier();
    const uint32_t lw = uni(L.last_w);
    if ((lw != kNone && k >= lw + 3) || k >= nwin + 3) break;
  }

  if (threadIdx.x == 0) {
    const uint64_t key = L.key;
    int32_t st = kOk;
    uint32_t pend_w = kNone;
    if (key != kNoKey) {
      const uint32_t pr = (uint32_t)(key >> 32) & 3;
      st = pr == kPrCorrupt ? kCorrupt : pr == kPrTrunc ? kTrunc : pr == kPrOverflow ? kOverflow : kOk;
      if (pr == kPrPending) pend_w = (uint32_t)key;
    }
    L.last_w = pend_w;
    if (pend_w == kNone && status) status[chunk] = st;
  }
  __syncthreads();
  const uint32_t pend_w = uni(L.last_w);
  if (pend_w != kNone && wave == 0) {
    // An overlong varint runs past the indexed windows: it ends CORRUPT if
    // any terminator follows window pend_w + 1, TRUNCATED if none does.
    const uintptr_t from = a0 + (uintptr_t)(pend_w + 2) * kWin;  // 16-byte aligned
    uint32_t found = 0;
    for (uintptr_t a = from + 16u * lane; a < e_addr; a += 16u * kWave) {
      const uint4 d = *reinterpret_cast<const uint4 *>(a);
      uint32_t m = term_mask16(d);
      const uintptr_t rem = e_addr - a;
      if (rem < 16) m &= (1u << rem) - 1;
      found |= m;
    }
    if (lane == 0 && status) status[chunk] = __ballot(found != 0) ? kCorrupt : kTrunc;
  }
}

}  // namespace rle1p

hipError_t launch_rlev1_pipe(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                             const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                             uint64_t chunk_bytes, uint8_t *out, int32_t *status,
                             hipStream_t stream) {
  if (n_chunks == 0) return hipSuccess;
  if (chunk_bytes / w > rle1p::kMaxCap) return hipErrorInvalidValue;  // descriptor value field
  const dim3 grid((uint32_t)n_chunks), block(rle1p::kThreads);
#define CODAG_V1P(W)                                                                         \
  hipLaunchKernelGGL((rle1p::rlev1_pipe_kernel<W>), grid, block, 0, stream, comp, comp_bytes, \
                     blk_off, col_len, chunk_bytes, out, status)
  switch (w) {
  case 1: CODAG_V1P(1); break;
  case 2: CODAG_V1P(2); break;
  case 4: CODAG_V1P(4); break;
  case 8: CODAG_V1P(8); break;
  default: return hipErrorInvalidValue;
  }
#undef CODAG_V1P
  return hipGetLastError();
}

}  // namespace codag
