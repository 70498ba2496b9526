// rlev2 -- drop-in for the reference's build/exe/src/rlev2/rlev2.cu.exe.
//
// Usage: rlev2 <raw column> <1|2|4|8>      (positional, src/rlev2/rlev2.cu:118-126)
//    or: rlev2 -f <raw column> -t <1|2|4|8> (README.md:17)
// Encodes with 128 KiB chunks (rlev2.cu:142), decodes on the GPU, checks the
// result (check_diff, rlev2.cu:18-43) and prints the reference's lines
// (rlev2.cu:145, 85, decoder_trans.h:124).
#include <cstring>
#include <string>
#include <vector>

#include "cli_common.hpp"

int main(int argc, char **argv) {
  std::string fname;
  int width = 0;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "-f") && i + 1 < argc) fname = argv[++i];
    else if (!std::strcmp(argv[i], "-t") && i + 1 < argc) width = std::atoi(argv[++i]);
    else if (fname.empty()) fname = argv[i];
    else if (!width) width = std::atoi(argv[i]);
  }
  if (fname.empty() || !width) {
    std::fprintf(stderr, "please provide arguments INPUT_FILE SIZE_OF_DATATYPE\n");
    return 1;
  }
  if (width != 1 && width != 2 && width != 4 && width != 8) {
    std::fprintf(stderr, "SIZE_OF_DATATYPE must be 1, 2, 4 or 8\n");
    return 1;
  }
  MappedFile in;
  if (!in.open(fname.c_str())) {
    std::printf("fatal error: input file open error\n");
    return 1;
  }
  const uint64_t chunk = 1024 * 128;
  std::printf("CHUNK_SIZE\t BUFFER SIZE \t COMPRESSED SIZE \t META DATA SIZE \t DECOMPRESSION TIME\n");
  uint8_t *comp = nullptr;
  uint64_t comp_bytes = 0, *blk = nullptr, *col = nullptr, n = 0;
  codag_status s = codag_rlev2_encode(in.bytes(), in.n, (uint32_t)width, chunk, &comp, &comp_bytes,
                                      &blk, &col, &n);
  if (s) die("encode", s);
  // meta_data_size as encoder_trans.h:505 counts it
  const uint64_t meta = 8 * n * 32 + 8 * (n + 1);
  codag_ctx *ctx = nullptr;
  if ((s = codag_ctx_create(0, &ctx))) die("GPU context", s);
  std::vector<uint8_t> out(in.n ? in.n : 1);
  uint64_t out_n = 0;
  double ksec = 0, tsec = 0;
  std::printf("%llu \t%d\t%llu\t%llu\t", (unsigned long long)chunk, 48,
              (unsigned long long)comp_bytes, (unsigned long long)meta);
  s = codag_rlev2_decode_host(ctx, comp, comp_bytes, blk, col, n, (uint32_t)width, chunk,
                              out.data(), out.size(), &out_n, &ksec, &tsec);
  if (s) die("decode", s);
  std::printf("%g secs\n", ksec);
  int rc = 0;
  if (out_n != in.n) {
    std::printf("fail: decoded %llu bytes, expected %llu\n", (unsigned long long)out_n,
                (unsigned long long)in.n);
    rc = 2;
  }
  int64_t d = first_diff(in.bytes(), out.data(), out_n < in.n ? out_n : in.n);
  if (d >= 0) {
    std::printf("fail at %lld %x(%x)\n", (long long)d, in.bytes()[d], out[d]);
    rc = 2;
  }
  codag_host_free(comp);
  codag_host_free(blk);
  codag_host_free(col);
  codag_ctx_destroy(ctx);
  return rc;
}
