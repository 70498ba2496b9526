// zlib -- drop-in for the reference's build/exe/src/zlib/zlib.cu.exe.
//
// Usage: zlib <X_comp.bin> <out> <X_blk_offset.bin> <chunk_size>   (src/zlib/zlib.cu:28-38)
//    or: zlib <X_comp.bin> <X_blk_offset.bin> <chunk_size>         (no output file)
// Decodes the chunked zlib container through the reference's host API
// (deflate::decompress_gpu, codag_compat.hpp), which prints "\t<seconds>"
// like include/zlib/zlib.h:1683, and writes the decoded bytes to <out>.
// Exit status: 0 on success, 1 on an error.
#include <cstring>

#include "cli_common.hpp"
#include "codag_compat.hpp"

int main(int argc, char **argv) {
  const char *comp_f = nullptr, *out_f = nullptr, *blk_f = nullptr, *cs = nullptr;
  if (argc >= 5) {
    comp_f = argv[1];
    out_f = argv[2];
    blk_f = argv[3];
    cs = argv[4];
  } else if (argc == 4) {
    comp_f = argv[1];
    blk_f = argv[2];
    cs = argv[3];
  } else {
    std::fprintf(stderr, "Please provide arguments  input output blk_offset chunk_size\n");
    return 1;
  }
  const uint64_t chunk = (uint64_t)std::strtoull(cs, nullptr, 10);
  MappedFile comp, blk;
  if (!comp.open(comp_f)) {
    std::printf("Fatal Error: INPUT File open error\n");
    return 1;
  }
  if (!blk.open(blk_f)) {
    std::printf("Fatal Error: BLK_OFFSET File open error\n");
    return 1;
  }
  if (blk.n < 16 || chunk == 0) {
    std::fprintf(stderr, "Fatal Error: blk_offset file or chunk size invalid\n");
    return 1;
  }
  uint8_t *out = nullptr;
  uint64_t out_n = 0;
  deflate::decompress_gpu<uint32_t, 262144 / 2, 64, 8, 1>(
      comp.bytes(), &out, comp.n, &out_n,
      const_cast<uint64_t *>(reinterpret_cast<const uint64_t *>(blk.p)), blk.n, chunk);
  if (!out) return 1;
  int rc = 0;
  if (out_f && !write_file(out_f, out, out_n)) {
    std::printf("Fatal Error: OUTPUT File open error\n");
    rc = 1;
  }
  std::free(out);
  return rc;
}
