// copy.hip -- plain device-to-device copy with 16-byte loads and stores.
// It moves exactly `bytes` bytes each way, so the HBM traffic counters of
// one launch calibrate the counters of the decode kernels.
#include "device.hpp"
#include "kernels.hpp"

namespace codag {
namespace {

__global__ void __launch_bounds__(256) copy_kernel(uint8_t *__restrict__ dst,
                                                   const uint8_t *__restrict__ src, uint64_t bytes) {
  const uint64_t nvec = bytes / 16;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint4 *s = reinterpret_cast<const uint4 *>(src);
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  for (; i < nvec; i += stride) d[i] = s[i];
  if (blockIdx.x == 0 && threadIdx.x < bytes % 16) dst[nvec * 16 + threadIdx.x] = src[nvec * 16 + threadIdx.x];
}

}  // namespace

hipError_t launch_copy(uint8_t *dst, const uint8_t *src, uint64_t bytes, hipStream_t stream) {
  if (bytes == 0) return hipSuccess;
  if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) & 15)
    return hipErrorInvalidValue;
  const uint64_t nvec = bytes / 16;
  uint64_t blocks = (nvec + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  hipLaunchKernelGGL(copy_kernel, dim3((uint32_t)blocks), dim3(256), 0, stream, dst, src, bytes);
  return hipGetLastError();
}

}  // namespace codag
