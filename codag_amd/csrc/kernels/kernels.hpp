// kernels.hpp -- host-callable launchers of the gfx950 decode kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "codag.h"

namespace codag {

constexpr int kWaveSize = 64;

// RLE launch shapes: (threads per group, LDS window) pairs that are compiled in.
bool rle_shape_supported(uint32_t nt, uint32_t stage);

hipError_t launch_rlev1_decode(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                               const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                               uint64_t chunk_bytes, uint8_t *out, int32_t *status,
                               uint32_t nt, uint32_t stage, hipStream_t stream);

hipError_t launch_rlev2_decode(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                               const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                               uint64_t chunk_bytes, uint8_t *out, int32_t *status,
                               uint32_t nt, uint32_t stage, hipStream_t stream);

// RLE v2, one wave64 per chunk; win = LDS window bytes (4096 or 8192).
hipError_t launch_rlev2_wave(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                             const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                             uint64_t chunk_bytes, uint8_t *out, int32_t *status, uint32_t win,
                             hipStream_t stream);

// two_wave: a decoding and a writing wave per chunk; else one wave.
hipError_t launch_inflate_chunks(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *off,
                                 uint64_t n,
                                 uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                                 int32_t *status, uint32_t flags, hipStream_t stream,
                                 bool two_wave = false);

// Persistent batch launch: `groups` resident workgroups pull streams from
// *counter (zeroed by a kernel launched first on the same stream). two_wave:
// workgroups of a decoding and a writing wave per stream; else one wave.
hipError_t launch_inflate_batch(const codag_stream_desc *descs, uint64_t n,
                                codag_stream_result *results, uint32_t flags, uint32_t *counter,
                                uint32_t groups, hipStream_t stream, bool two_wave = false);

// Device address (on the current device) of the inflate kernels' static
// 256-byte scratch array; its first word is the work-queue counter of the
// allocation-free batch entry point.
hipError_t inflate_scratch_ptr(uint32_t **out);

// RLE v1, one walker wave and nexp (1 or 3) expander waves per chunk.
hipError_t launch_rlev1_pc(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                           const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                           uint64_t chunk_bytes, uint8_t *out, int32_t *status, uint32_t nexp,
                           hipStream_t stream);

// RLE v1, software pipeline over 2 KiB windows: one walker wave and three
// helper waves per chunk (chunk_bytes / w <= 2^18).
hipError_t launch_rlev1_pipe(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                             const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                             uint64_t chunk_bytes, uint8_t *out, int32_t *status,
                             hipStream_t stream);

// RLE v2, one walker wave and nexp (1, 2 or 3) expander waves per chunk.
// poll_shift nonzero (a test) launches the variant whose bounded waits have
// no polls left, so every wait not met at once ends its chunk with kCorrupt.
hipError_t launch_rlev2_pc(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *blk_off,
                           const uint64_t *col_len, uint64_t n_chunks, uint32_t w,
                           uint64_t chunk_bytes, uint8_t *out, int32_t *status, uint32_t nexp,
                           uint32_t poll_shift, hipStream_t stream);

// BRLE / LZSS, one wave64 per chunk (chunk_bytes < 2^31).
hipError_t launch_brle_chunks(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *ptr,
                              uint64_t n, uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                              int32_t *status, hipStream_t stream);
hipError_t launch_lzss_chunks(const uint8_t *comp, uint64_t comp_bytes, const uint64_t *ends,
                              uint64_t n, uint64_t chunk_bytes, uint8_t *out, uint64_t *out_len,
                              int32_t *status, hipStream_t stream);

// Plain device copy (16-byte loads/stores): the reference point for HBM
// traffic counters.
hipError_t launch_copy(uint8_t *dst, const uint8_t *src, uint64_t bytes, hipStream_t stream);

}  // namespace codag
