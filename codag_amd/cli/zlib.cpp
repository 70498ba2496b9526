// zlib -- drop-in for the reference's build/exe/src/zlib/zlib.cu.exe.
//
// Usage: zlib <X_comp.bin> <out> <X_blk_offset.bin> <chunk_size>   (src/zlib/zlib.cu:28-38)
//    or: zlib <X_comp.bin> <X_blk_offset.bin> <chunk_size>         (no output file)
// Decodes the chunked zlib container on the GPU, prints "\t<secs>" like
// include/zlib/zlib.h:1683 and writes the decoded bytes to <out>.
#include <cstring>
#include <vector>

#include "cli_common.hpp"

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
  const uint64_t *b = reinterpret_cast<const uint64_t *>(blk.p);
  const uint64_t orig = b[0];
  codag_ctx *ctx = nullptr;
  codag_status s = codag_ctx_create(0, &ctx);
  if (s) die("GPU context", s);
  std::vector<uint8_t> out(orig ? orig : 1);
  uint64_t out_n = 0;
  double ksec = 0, tsec = 0;
  s = codag_inflate_chunks_host(ctx, comp.bytes(), comp.n, b, blk.n / 8, chunk, 0, out.data(),
                                out.size(), &out_n, &ksec, &tsec);
  if (s) die("decode", s);
  std::printf("\t%g\n", ksec);
  if (out_f && !write_file(out_f, out.data(), out_n)) {
    std::printf("Fatal Error: OUTPUT File open error\n");
    return 1;
  }
  codag_ctx_destroy(ctx);
  return 0;
}
