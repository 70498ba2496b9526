// rlev2_wave.hip -- RLE v2 decode with one self-contained wave64 per chunk.
//
// The chunk's stream is staged into a per-wave LDS window (kWin bytes). The
// header walk runs in scalar registers: the 256 bytes around the walk
// position are mirrored in one VGPR (lane l holds dword l) and header bytes
// are read with v_readlane, so walking a block costs no memory round trip.
// The walk collects up to 64 blocks (lane j <- block j), then:
//   * SHORT_REPEAT blocks (3..10 values) expand one lane per block;
//   * DIRECT / DELTA / PATCHED_BASE blocks expand across the wave
//     (expand_big, one lane per value, wave prefix sums for DELTA, a wave
//     scan of the patch gaps for PATCHED_BASE).
// A block that does not fit in the window restarts the window at that
// block. There are no workgroup barriers: the chunk's progress depends only
// on its own wave, and several chunks share each SIMD to hide latency.
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

template <uint32_t kWin> struct WaveLds {
  uint8_t sbuf[kWin + 288];  // window + slack for the 256-byte VGPR mirror
};

template <uint32_t kWin>
__global__ void __launch_bounds__(kWave)
rlev2_wave_kernel(const uint8_t *__restrict__ comp, uint64_t comp_bytes,
                  const uint64_t *__restrict__ blk_off, const uint64_t *__restrict__ col_len,
                  uint64_t n_chunks, uint32_t w, uint64_t chunk_bytes, uint8_t *__restrict__ out,
                  int32_t *__restrict__ status) {
  __shared__ WaveLds<kWin> L;
  const uint64_t chunk = blockIdx.x;
  const uint32_t lane = lane_id();
  const uint64_t boff = blk_off[chunk];
  const uint64_t len = col_len[chunk];
  if (boff > comp_bytes || len > comp_bytes - boff) {
    if (lane == 0 && status) status[chunk] = kCorrupt;
    return;
  }
  const uint8_t *src = comp + boff;
  const uint32_t cap = (uint32_t)(chunk_bytes / w);
  const uint64_t mask = width_mask(w);
  const uint32_t nbits = 8 * w;
  uint8_t *dst = out + chunk * chunk_bytes;
  const uint32_t *sdw = reinterpret_cast<const uint32_t *>(L.sbuf);

  // Stream offsets are 32-bit here: a chunk's stream is far below 4 GiB.
  uint32_t wpos = 0;  // stream offset held at sbuf[lead]
  uint32_t lead = 0, P = 0;
  bool at_end = true;
  // The walk stops at p >= pdone (stream consumed: P when the window holds the
  // stream's end) and restages at p > pstop (header bytes might lie past the
  // window: P - 32 when it does not).
  uint32_t pdone = 0, pstop = 0;
  // SHORT_REPEAT fast path bounds (a block is at most 10 bytes and 10 values):
  // p < pfast keeps the header and the whole block inside the staged stream,
  // o < ofast keeps the values inside the chunk.
  uint32_t pfast = 0;
  const uint32_t ofast = cap >= 10 ? cap - 9 : 0;
  auto stage = [&](uint32_t pos) {
    __syncthreads();  // earlier LDS reads of this wave are done before overwrite
    const uint32_t n = ((uint32_t)len - pos) < (kWin - 16) ? ((uint32_t)len - pos) : (kWin - 16);
    lead = 0;
    if (n) stage_to_lds(L.sbuf, src, pos, n, lane, kWave, &lead);
    P = lead + n;
    at_end = (pos + n == (uint32_t)len);
    wpos = pos;
    pdone = at_end ? P : 0xffffffffu;
    pstop = at_end ? 0xffffffffu : P - 32;
    pfast = at_end ? (P >= 10 ? P - 9 : 0) : P - 31;
    __syncthreads();
  };

  if (len >= 0xffffffffull) {  // outside the 32-bit stream offsets used below
    if (lane == 0 && status) status[chunk] = kCorrupt;
    return;
  }
  stage(0);
  uint32_t p = lead, o = 0;
  uint32_t vbase = p & ~3u;
  uint32_t vw = sdw[vbase / 4 + lane];
  int32_t st = kOk;
  uint32_t bad = 0;
  bool done = (len == 0);
  while (!done && st == kOk) {
    // ---- walk: collect up to 64 blocks
    uint32_t nd = 0, dpos = 0, dout = 0;
    bool restage = false;
    while (nd < (uint32_t)kWave) {
      // Fast path for runs of SHORT_REPEAT blocks (the most frequent header):
      // while the position and output index are far enough from the window
      // end and the chunk end that no bound can be hit, a block costs one
      // header read, two bit-field extracts and the lane record.
      for (;;) {
        if (nd >= (uint32_t)kWave || p >= pfast || o >= ofast) break;
        if (p - vbase > 224) {
          vbase = p & ~3u;
          vw = sdw[vbase / 4 + lane];
        }
        const uint32_t k = (p - vbase) >> 2;
        const uint32_t first = (uint32_t)(((uint64_t)(uint32_t)__builtin_amdgcn_readlane(vw, k) |
                                           ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(vw, k + 1)
                                            << 32)) >> ((p & 3) * 8)) & 0xff;
        if (first >= 64) break;
        if (lane == nd) {
          dpos = p;
          dout = o;
        }
        nd++;
        o += (first & 7) + 3;
        p += 2 + ((first >> 3) & 7);
      }
      if (nd >= (uint32_t)kWave) break;
      if (p >= pdone) {  // decoder_new.h:1000 -- stream consumed
        done = true;
        break;
      }
      if (p > pstop) {  // header bytes might lie past the window
        restage = true;
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
          const uint32_t l1 = varint_len(L.sbuf, vw, vbase, q, P);
          const uint32_t l2 = q + l1 < P ? varint_len(L.sbuf, vw, vbase, q + l1, P) : 1;
          if (q + l1 + l2 <= P && (l1 > 10 || l2 > 10)) {  // longer than any 64-bit varint
            st = kCorrupt;
            break;
          }
          q += l1 + l2;
          size = (q - p) + (code ? (nd8 * fbw + 7) / 8 : 0);
          cnt = 2 + nd8;
        } else {  // PATCHED_BASE
          const uint32_t third = (uint32_t)((x >> 16) & 0xff), fourth = (uint32_t)((x >> 24) & 0xff);
          const uint32_t pw = decode_width(third & 0x1f);
          const uint32_t pgw = (fourth >> 5) + 1;
          if (pw + pgw > 64) {  // beyond the 65-entry closest-width table
            st = kCorrupt;
            break;
          }
          size = 4 + (third >> 5) + 1 + (len8 * fbw + 7) / 8 +
                 ((fourth & 0x1f) * closest_fixed(pw + pgw) + 7) / 8;
          cnt = len8;
        }
      }
      if (p + size > P) {
        if (at_end) st = kTrunc;
        else restage = true;
        break;
      }
      if (o + cnt > cap) {
        st = kOverflow;
        break;
      }
      if (lane == nd) {
        dpos = p;
        dout = o;
      }
      nd++;
      o += cnt;
      p += size;
    }
    // ---- expand the collected blocks
    const uint32_t myfirst = lane < nd ? L.sbuf[dpos] : 0u;
    const bool is_sr = lane < nd && (myfirst >> 6) == 0;
    if (is_sr) {  // SHORT_REPEAT: big-endian value, 3..10 copies (decoder_new.h:1003-1015)
      const uint32_t nb = ((myfirst >> 3) & 7) + 1;
      const uint32_t c = (myfirst & 7) + 3;
      const uint64_t v = lds_be(L.sbuf + dpos + 1, nb) & mask;
      for (uint32_t k = 0; k < c; k++) store_elem(dst + (uint64_t)(dout + k) * w, w, v);
    }
    uint64_t big = __ballot(lane < nd && !is_sr);
    while (big) {
      const uint32_t j = (uint32_t)__builtin_ctzll(big);
      big &= big - 1;
      const uint32_t bp = (uint32_t)__builtin_amdgcn_readlane(dpos, j);
      const uint32_t bo = (uint32_t)__builtin_amdgcn_readlane(dout, j);
      expand_big(L.sbuf, bp, bo, dst, w, mask, nbits, lane, &bad);
    }
    // A block never exceeds ~2.3 KB, so a fresh window always makes progress;
    // no progress would mean a malformed stream.
    if (restage && nd == 0 && p == lead) st = kCorrupt;
    if (restage && st == kOk) {
      stage(wpos + (p - lead));
      p = lead;
      vbase = p & ~3u;
      vw = sdw[vbase / 4 + lane];
    }
  }
  if (__ballot(bad) && st == kOk) st = kCorrupt;
  // positions the stream did not produce are zero (decoder_trans.h:81)
  if (st == kOk)
    for (uint32_t i = o + lane; i < cap; i += kWave) store_elem(dst + (uint64_t)i * w, w, 0);
  if (lane == 0 && status) status[chunk] = st;
}

}  // namespace rle2

hipError_t launch_rlev2_wave(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                             const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                             uint64_t chunk_bytes, uint8_t *out, int32_t *status, uint32_t win,
                             hipStream_t stream) {
  if (n_chunks == 0) return hipSuccess;
  if (win == 4096)
    hipLaunchKernelGGL(rle2::rlev2_wave_kernel<4096>, dim3((uint32_t)n_chunks), dim3(kWaveSize), 0,
                       stream, comp, comp_bytes, blk_off, col_len, n_chunks, w, chunk_bytes, out,
                       status);
  else if (win == 8192)
    hipLaunchKernelGGL(rle2::rlev2_wave_kernel<8192>, dim3((uint32_t)n_chunks), dim3(kWaveSize), 0,
                       stream, comp, comp_bytes, blk_off, col_len, n_chunks, w, chunk_bytes, out,
                       status);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace codag
