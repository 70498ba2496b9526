// rlev1 -- drop-in for the reference's build/exe/src/rlev1/rlev1.cu.exe.
//
// Usage: rlev1 -f <raw column> -t <1|2|4|8> [-n <num_warp>]
// Behaviour follows src/rlev1/rlev1.cu:46-453: encode the column into the
// RLE v1 container with 128 KiB chunks, write ./input_data/rle_v1_col_len.bin
// and ./input_data/rle_v1_blk_offset.bin (rlev1.h:178-186), read them back,
// decode on the GPU and byte-compare against the input, printing the same
// tab-separated lines (rlev1.cu:156-159, rlev1.h:507,576-579).
#include <getopt.h>
#include <sys/stat.h>

#include <cstring>
#include <vector>

#include "cli_common.hpp"

static void usage() {
  std::printf("  %-35s Input binary filename (required).\n", "-f, --file");
  std::printf("  %-35s Data type size in bytes (required).\n", "-t, --data_byte");
  std::printf("  %-35s Accepted for compatibility (decode uses one launch shape).\n",
              "-n, --num_warp");
  std::exit(1);
}

int main(int argc, char **argv) {
  std::string fname;
  int width = 0;
  static struct option long_options[] = {{"file", required_argument, 0, 'f'},
                                         {"data_byte", required_argument, 0, 't'},
                                         {"num_warp", required_argument, 0, 'n'},
                                         {0, 0, 0, 0}};
  for (;;) {
    int idx = 0;
    int c = getopt_long(argc, argv, "f:o:r:t:d:l:n:?", long_options, &idx);
    if (c == -1) break;
    switch (c) {
    case 'f': fname = optarg; break;
    case 't': width = std::atoi(optarg); break;
    case 'o': case 'r': case 'd': case 'l': case 'n': break;  // parsed, unused (rlev1.cu:57-109)
    default: usage();
    }
  }
  if (fname.empty() || (width != 1 && width != 2 && width != 4 && width != 8)) usage();
  MappedFile in;
  if (!in.open(fname.c_str())) {
    std::printf("Fatal Error: INPUT File open error\n");
    return 1;
  }
  const uint64_t chunk = 131072;  // rlev1.cu:265
  std::printf("CHUNK_SIZE\tCOLUMN WIDTH\t BUFFER SIZE \t COMPRESSED SIZE \t META DATA SIZE \t "
              "DECOMPRESSION TIME\n");
  uint8_t *container = nullptr;
  uint64_t container_bytes = 0, *blk = nullptr, *col = nullptr, nblk = 0, ncol = 0, n = 0;
  codag_status s = codag_rlev1_encode(in.bytes(), in.n, (uint32_t)width, chunk, &container,
                                      &container_bytes, &blk, &nblk, &col, &ncol, &n);
  if (s) die("encode", s);
  mkdir("./input_data", 0755);
  if (!write_file("./input_data/rle_v1_col_len.bin", col, ncol * 8) ||
      !write_file("./input_data/rle_v1_blk_offset.bin", blk, nblk * 8)) {
    std::printf("Fatal Error: Col Len read error\n");
    return 1;
  }
  codag_host_free(blk);
  codag_host_free(col);
  // decode side: read the side files back (rlev1.h:436-467)
  MappedFile fcol, fblk;
  if (!fcol.open("./input_data/rle_v1_col_len.bin")) {
    std::printf("Fatal Error: Col Len read error\n");
    return 1;
  }
  if (!fblk.open("./input_data/rle_v1_blk_offset.bin")) {
    std::printf("Fatal Error: Block off read error\n");
    return 1;
  }
  codag_ctx *ctx = nullptr;
  if ((s = codag_ctx_create(0, &ctx))) die("GPU context", s);
  std::vector<uint8_t> out(in.n ? in.n : 1);
  uint64_t out_n = 0;
  double ksec = 0, tsec = 0;
  std::printf("NUM WARP: %i\n", 1);
  s = codag_rlev1_decode_host(ctx, container, container_bytes,
                              reinterpret_cast<const uint64_t *>(fblk.p), fblk.n / 8,
                              reinterpret_cast<const uint64_t *>(fcol.p), fcol.n / 8,
                              (uint32_t)width, chunk, out.data(), out.size(), &out_n, &ksec, &tsec);
  if (s) die("decode", s);
  const uint64_t meta = fcol.n + fblk.n;
  const float ratio = (float)container_bytes / (float)(out_n ? out_n : 1);
  std::printf("%llu\t%llu\t%g\t%g\n", (unsigned long long)container_bytes, (unsigned long long)meta,
              ratio, ksec);
  int64_t d = first_diff(in.bytes(), out.data(), out_n);
  if (d >= 0) std::printf("File Differ at %lld\n", (long long)d);
  codag_host_free(container);
  codag_ctx_destroy(ctx);
  return d >= 0 ? 2 : 0;
}
