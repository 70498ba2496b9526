// rlev1_pc.hip -- ORC RLE v1 decode: one walker wave and kExp expander waves
// per chunk.
//
// Window bookkeeping (after each 4 KiB stage into LDS), built by the walker's
// 64 lanes: lane l holds the 64-bit mask of varint terminators (bytes < 0x80)
// among window bytes [64 l, 64 l + 64) and the exclusive prefix count of
// terminators before them, so rank(q) = terminators before byte q costs
// three v_readlane and a popcount; sel[r] (LDS) = position of terminator r.
//
// The walker follows the segment chain in scalar registers: a header byte
// comes from a 256-byte VGPR mirror of the window; a run's base varint ends
// at sel[rank(p + 2)], a literal group of k varints at sel[rank(p + 1) + k - 1],
// so every segment costs one LDS round trip. Each batch of up to 64 segment
// descriptors goes to an LDS double buffer; the expanders write batch k while
// the walker walks batch k + 1 (the barrier protocol of the RLE v2 walker +
// expander kernel). Expansion: a run's values base + i*delta one lane per
// value; a literal group's varints one lane per varint, each decoded from
// three LDS dwords.
//
// Errors follow the stream order of the CPU decoder (include/rlev1/
// decoder.h:671-791 restated in the oracle): the walker reports the value
// index of a truncation, overflow or overlong run base; the expanders report
// the smallest value index of an overlong literal varint; the earlier wins.
#include "device.hpp"
#include "kernels.hpp"

namespace codag {
namespace rle1 {

using namespace codag::dev;

enum : int32_t { kOk = 0, kTrunc = 1, kOverflow = 2, kCorrupt = 3 };

constexpr uint32_t kWin = 4096;      // window bytes: 64 lanes x 64-byte mask words
constexpr uint32_t kPosBits = 13;    // descriptor: p | out_idx << 13 (p < 8192)
constexpr uint32_t kBatchLast = 1u << 30;
constexpr uint32_t kBatchRestage = 1u << 29;
constexpr uint32_t kNoError = 0xffffffffu;

struct PcLds {
  uint8_t sbuf[kWin + 288];  // window + slack for the 256-byte VGPR mirror
  uint16_t sel[kWin];
  uint32_t desc[2][kWave];  // p | out_idx << kPosBits
  uint16_t rank[2][kWave];  // rank of the segment's first varint terminator
  uint16_t cnt[2][kWave];   // values to decode from the segment (checks) ...
  uint16_t keep[2][kWave];  // ... of which the first `keep` are stored
  uint32_t info[2];
  uint32_t bad_idx;         // smallest value index of an overlong literal varint
};

// Walker state. Stream offsets are 32-bit (a chunk stream is far below 4 GiB).
struct Walk {
  const uint8_t *src;
  uint32_t len, cap;
  uint32_t wpos, lead, P;  // window: sbuf[lead, P) = stream bytes [wpos, wpos + P - lead)
  bool at_end;
  uint32_t nterm;          // terminators in the window
  uint32_t mlo, mhi, cum;  // lane l: terminator mask of bytes [64l, 64l+64), prefix count
  uint32_t vbase, vw;      // VGPR mirror of sbuf[vbase, vbase + 256)
  uint32_t p, o;
};

__device__ __forceinline__ uint32_t rdl(uint32_t v, uint32_t l) {
  return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

// Stages stream bytes [pos, ...) and rebuilds the terminator masks, prefix
// counts and sel[] (walker wave only).
__device__ __forceinline__ void stage(Walk &s, PcLds &L, uint32_t pos, uint32_t lane) {
  const uint32_t n = (s.len - pos) < (kWin - 16) ? (s.len - pos) : (kWin - 16);
  s.lead = 0;
  if (n) stage_to_lds(L.sbuf, s.src, pos, n, lane, kWave, &s.lead);
  s.P = s.lead + n;
  s.at_end = (pos + n == s.len);
  s.wpos = pos;
  // lane l: bytes [64l, 64l + 64) restricted to [lead, P)
  const uint32_t *d = reinterpret_cast<const uint32_t *>(L.sbuf) + lane * 16;
  uint64_t m = 0;
#pragma unroll
  for (uint32_t k = 0; k < 16; k++) {
    const uint32_t t = ~d[k] & 0x80808080u;
    const uint32_t nib = ((t >> 7) & 1) | ((t >> 14) & 2) | ((t >> 21) & 4) | ((t >> 28) & 8);
    m |= (uint64_t)nib << (4 * k);
  }
  const uint32_t q0 = lane * 64;
  const uint32_t lo = s.lead > q0 ? min(s.lead - q0, 64u) : 0u;
  const uint32_t hi = s.P > q0 ? min(s.P - q0, 64u) : 0u;
  const uint64_t valid = hi > lo ? ((hi - lo == 64 ? ~0ull : ((1ull << (hi - lo)) - 1)) << lo) : 0ull;
  m &= valid;
  const uint32_t c = (uint32_t)__popcll(m);
  const uint32_t incl = wave_incl_scan(c);
  s.cum = incl - c;
  s.nterm = rdl(incl, kWave - 1);
  s.mlo = (uint32_t)m;
  s.mhi = (uint32_t)(m >> 32);
  uint32_t r = s.cum;
  while (m) {
    L.sel[r++] = (uint16_t)(q0 + (uint32_t)__builtin_ctzll(m));
    m &= m - 1;
  }
  s.p = s.lead;
  s.vbase = s.p & ~3u;
  s.vw = reinterpret_cast<const uint32_t *>(L.sbuf)[s.vbase / 4 + lane];
}

// Terminators before window byte q (q <= kWin).
__device__ __forceinline__ uint32_t rank_of(const Walk &s, uint32_t q) {
  if (q >= kWin) return s.nterm;
  const uint32_t l = q >> 6, b = q & 63;
  const uint64_t m = (uint64_t)rdl(s.mlo, l) | ((uint64_t)rdl(s.mhi, l) << 32);
  const uint64_t below = b ? (m & ((1ull << b) - 1)) : 0ull;
  return rdl(s.cum, l) + (uint32_t)__popcll(below);
}

__device__ __forceinline__ uint32_t byte_at(Walk &s, const PcLds &L, uint32_t p, uint32_t lane) {
  if (p - s.vbase > 248) {
    s.vbase = p & ~3u;
    s.vw = reinterpret_cast<const uint32_t *>(L.sbuf)[s.vbase / 4 + lane];
  }
  return (rdl(s.vw, (p - s.vbase) >> 2) >> ((p & 3) * 8)) & 0xff;
}

// Decodes the LEB128 varint at window bytes [start, end] (end - start < 10)
// from three LDS dwords; bits past 64 are dropped.
__device__ __forceinline__ uint64_t varint_at(const uint8_t *sbuf, uint32_t start, uint32_t end) {
  const uint32_t *d = reinterpret_cast<const uint32_t *>(sbuf) + (start >> 2);
  const uint32_t sh = (start & 3) * 8;
  const uint64_t w0 = (uint64_t)d[0] | ((uint64_t)d[1] << 32);
  const uint64_t w1 = (uint64_t)d[2] | ((uint64_t)d[3] << 32);
  const uint64_t lo = sh ? ((w0 >> sh) | (w1 << (64 - sh))) : w0;  // bytes start .. start+7
  const uint64_t hi = sh ? (w1 >> sh) : w1;                          // bytes start+8 .. start+9
  const uint32_t n = end - start + 1;
  uint64_t v = 0;
#pragma unroll
  for (uint32_t k = 0; k < 8; k++)
    if (k < n) v |= ((lo >> (8 * k)) & 0x7f) << (7 * k);
  if (n > 8) v |= (hi & 0x7f) << 56;
  if (n > 9) v |= ((hi >> 8) & 0x7f) << 63;
  return v;
}

template <uint32_t kExp>
__global__ void __launch_bounds__(kWave *(1 + kExp))
rlev1_pc_kernel(const uint8_t *__restrict__ comp, uint64_t comp_bytes,
                const uint64_t *__restrict__ blk_off, const uint64_t *__restrict__ col_len,
                uint64_t n_chunks, uint32_t w, uint64_t chunk_bytes, uint8_t *__restrict__ out,
                int32_t *__restrict__ status) {
  __shared__ PcLds L;
  const uint64_t chunk = blockIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const uint64_t boff = blk_off[chunk];
  const uint64_t len = col_len[chunk];
  if (boff > comp_bytes || len > comp_bytes - boff || len >= 0xffffffffull) {
    if (threadIdx.x == 0 && status) status[chunk] = kCorrupt;  // uniform over the workgroup
    return;
  }
  const uint32_t cap = (uint32_t)(chunk_bytes / w);
  const uint64_t mask = width_mask(w);
  uint8_t *dst = out + chunk * chunk_bytes;
  if (threadIdx.x == 0) L.bad_idx = kNoError;

  if (wave == 0) {  // ---- walker
    Walk s;
    s.src = comp + boff;
    s.len = (uint32_t)len;
    s.cap = cap;
    s.o = 0;
    stage(s, L, 0, lane);
    int32_t err = kOk;
    uint32_t err_idx = kNoError;
    bool done = false;
    for (uint32_t k = 0;; k++) {
      uint32_t nd = 0, dsc = 0, drk = 0, dcnt = 0, dkeep = 0;
      bool restage = false;
      const uint32_t p0 = s.p;
      while (!done && err == kOk && nd < (uint32_t)kWave) {
        if (s.o >= cap) {  // decoder.h:671 -- the chunk is full
          done = true;
          break;
        }
        const uint32_t p = s.p;
        if (p >= s.P) {
          if (s.at_end) {
            err = kTrunc;
            err_idx = s.o;
          } else {
            restage = true;
          }
          break;
        }
        const int32_t h = (int32_t)(int8_t)byte_at(s, L, p, lane);
        uint32_t r, cnt, keep, next;
        if (h < 0) {  // literal group of -h varints (decoder.h:686-733)
          cnt = (uint32_t)(-h);
          r = rank_of(s, p + 1);
          const uint32_t room = cap - s.o;  // values that fit
          const uint32_t need = cnt <= room ? cnt : room + 1;  // varints read before the outcome
          if (r + need > s.nterm) {  // a varint ends past the window
            if (!s.at_end) {
              restage = true;
              break;
            }
            // truncated at the first varint without a terminator: publish the
            // complete ones so their lengths are checked
            const uint32_t have = s.nterm - r;
            err = kTrunc;
            err_idx = s.o + have;
            cnt = have;
            keep = have;
            next = p;
          } else {
            keep = cnt <= room ? cnt : room;
            if (cnt > room) {
              err = kOverflow;
              err_idx = cap;
              cnt = need;
            }
            next = (uint32_t)L.sel[r + cnt - 1] + 1;
          }
        } else {  // run: delta byte + varint base (decoder.h:736-790)
          cnt = (uint32_t)h + 3;
          if (p + 2 > s.P) {  // no delta byte / base in the window
            if (!s.at_end) {
              restage = true;
              break;
            }
            err = kTrunc;
            err_idx = s.o;
            break;
          }
          r = rank_of(s, p + 2);
          if (r >= s.nterm) {
            if (!s.at_end) {
              restage = true;
              break;
            }
            err = kTrunc;
            err_idx = s.o;
            break;
          }
          const uint32_t end = L.sel[r];
          if (end - (p + 2) >= 10) {  // longer than any 64-bit varint
            err = kCorrupt;
            err_idx = s.o;
            break;
          }
          keep = cnt;
          if (s.o + cnt > cap) {
            err = kOverflow;
            err_idx = cap;
            keep = cap - s.o;
          }
          next = end + 1;
        }
        if (lane == nd) {
          dsc = p | (s.o << kPosBits);
          drk = r;
          dcnt = cnt;
          dkeep = keep;
        }
        nd++;
        s.o += keep;
        s.p = next;
      }
      // A segment never exceeds a fresh window unless it is malformed.
      if (restage && nd == 0 && p0 == s.lead && err == kOk) {
        err = kCorrupt;
        err_idx = s.o;
      }
      const bool last = done || err != kOk;
      if (last) restage = false;
      L.desc[k & 1][lane] = dsc;
      L.rank[k & 1][lane] = (uint16_t)drk;
      L.cnt[k & 1][lane] = (uint16_t)dcnt;
      L.keep[k & 1][lane] = (uint16_t)dkeep;
      if (lane == 0) L.info[k & 1] = nd | (last ? kBatchLast : 0u) | (restage ? kBatchRestage : 0u);
      __syncthreads();  // B_k
      if (last) {
        __syncthreads();  // F: expanders are done
        const uint32_t bad = L.bad_idx;
        int32_t st = err;
        if (bad != kNoError && bad <= err_idx) st = kCorrupt;
        if (lane == 0 && status) status[chunk] = st;
        return;
      }
      if (restage) {
        __syncthreads();  // R_k: the expanders have finished batch k
        stage(s, L, s.wpos + (s.p - s.lead), lane);
      }
    }
  } else {  // ---- expanders: wave e takes segments j with j % kExp == e - 1
    uint32_t bad = kNoError;
    for (uint32_t k = 0;; k++) {
      __syncthreads();  // B_k
      const uint32_t info = L.info[k & 1];
      const uint32_t nd = info & 0xffff;
      for (uint32_t j = wave - 1; j < nd; j += kExp) {
        const uint32_t e = L.desc[k & 1][j];
        const uint32_t p = e & ((1u << kPosBits) - 1);
        const uint32_t o = e >> kPosBits;
        const uint32_t r0 = L.rank[k & 1][j];
        const uint32_t cnt = L.cnt[k & 1][j];
        const uint32_t keep = L.keep[k & 1][j];
        const int32_t h = (int32_t)(int8_t)L.sbuf[p];
        if (h < 0) {
          for (uint32_t i = lane; i < cnt; i += kWave) {
            const uint32_t end = L.sel[r0 + i];
            const uint32_t start = i ? (uint32_t)L.sel[r0 + i - 1] + 1 : p + 1;
            if (end - start >= 10) {
              if (o + i < bad) bad = o + i;
            } else if (i < keep) {
              store_elem(dst + (uint64_t)(o + i) * w, w, varint_at(L.sbuf, start, end) & mask);
            }
          }
        } else {
          const int64_t delta = (int8_t)L.sbuf[p + 1];
          const uint64_t base = varint_at(L.sbuf, p + 2, L.sel[r0]);
          for (uint32_t i = lane; i < keep; i += kWave)
            store_elem(dst + (uint64_t)(o + i) * w, w, (base + (uint64_t)((int64_t)i * delta)) & mask);
        }
      }
      if (info & kBatchLast) {
        if (bad != kNoError) atomicMin(&L.bad_idx, bad);
        __syncthreads();  // F
        return;
      }
      if (info & kBatchRestage) __syncthreads();  // R_k
    }
  }
}

}  // namespace rle1

hipError_t launch_rlev1_pc(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                           const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                           uint64_t chunk_bytes, uint8_t *out, int32_t *status, uint32_t nexp,
                           hipStream_t stream) {
  if (n_chunks == 0) return hipSuccess;
  if (nexp == 1)
    hipLaunchKernelGGL((rle1::rlev1_pc_kernel<1>), dim3((uint32_t)n_chunks), dim3(2 * kWaveSize), 0,
                       stream, comp, comp_bytes, blk_off, col_len, n_chunks, w, chunk_bytes, out,
                       status);
  else if (nexp == 3)
    hipLaunchKernelGGL((rle1::rlev1_pc_kernel<3>), dim3((uint32_t)n_chunks), dim3(4 * kWaveSize), 0,
                       stream, comp, comp_bytes, blk_off, col_len, n_chunks, w, chunk_bytes, out,
                       status);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace codag
