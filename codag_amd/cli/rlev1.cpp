// rlev1 -- drop-in for the reference's build/exe/src/rlev1/rlev1.cu.exe.
//
// Usage: rlev1 -f <raw column> -t <1|2|4|8> [-n <num_warp>]
// Behaviour follows src/rlev1/rlev1.cu:46-453 through the reference's own
// host API (codag_compat.hpp): rle_v1::compress_gpu_orig encodes the column
// with 128 KiB chunks and writes ./input_data/rle_v1_col_len.bin and
// ./input_data/rle_v1_blk_offset.bin (rlev1.h:178-186); then
// rle_v1::decompress_gpu_orig decodes it five times, for the five NUM_CHUNK
// variants 1, 2, 4, 8 and 16 the reference instantiates (rlev1.cu:168-193),
// each printing "NUM WARP: 1" and "compressed\tmeta\tratio\tseconds"
// (rlev1.h:507,576-579), and each result is byte-compared with the input
// ("File Differ at i", rlev1.cu:21-28). Exit status: 0 when every decode
// matches, 2 on a mismatch, 1 on an error.
#include <getopt.h>
#include <sys/stat.h>

#include <cstring>
#include <string>

#include "cli_common.hpp"
#include "codag_compat.hpp"

static void usage() {
  std::printf("  %-35s Input binary filename (required).\n", "-f, --file");
  std::printf("  %-35s Data type size in bytes (required).\n", "-t, --data_byte");
  std::printf("  %-35s Accepted for compatibility (the decode launch shape is tuned).\n",
              "-n, --num_warp");
  std::exit(1);
}

// One decode variant: decode, compare, release as rlev1.cu does.
template <typename T, int NUM_CHUNK>
static bool variant(const uint8_t *in, uint64_t in_n, const uint8_t *cont, uint64_t cont_n,
                    uint64_t chunk) {
  uint8_t *out = nullptr;
  uint64_t out_n = 0;
  rle_v1::decompress_gpu_orig<uint32_t, T, uint32_t, 64, NUM_CHUNK>(cont, &out, cont_n, &out_n, 8,
                                                                     chunk);
  if (!out) return false;
  bool ok = out_n == in_n;
  const int64_t d = first_diff(in, out, out_n < in_n ? out_n : in_n);
  if (d >= 0) {
    std::printf("File Differ at %lld\n", (long long)d);
    ok = false;
  }
  std::free(out);
  return ok;
}

template <typename T> static int run(const uint8_t *in, uint64_t in_n) {
  const uint64_t chunk = 131072;  // rlev1.cu:166
  uint8_t *cont = nullptr;
  uint64_t cont_n = 0;
  rle_v1::compress_gpu_orig<uint64_t, T, uint32_t>(in, &cont, in_n, &cont_n, 8, chunk);
  if (!cont) return 1;
  bool ok = variant<T, 1>(in, in_n, cont, cont_n, chunk);
  ok = variant<T, 2>(in, in_n, cont, cont_n, chunk) && ok;
  ok = variant<T, 4>(in, in_n, cont, cont_n, chunk) && ok;
  ok = variant<T, 8>(in, in_n, cont, cont_n, chunk) && ok;
  ok = variant<T, 16>(in, in_n, cont, cont_n, chunk) && ok;
  codag_host_free(cont);
  return ok ? 0 : 2;
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
  if (fname.empty()) usage();
  MappedFile in;
  if (!in.open(fname.c_str())) {
    std::printf("Fatal Error: INPUT File open error\n");
    return 1;
  }
  std::printf("CHUNK_SIZE\tCOLUMN WIDTH\t BUFFER SIZE \t COMPRESSED SIZE \t META DATA SIZE \t "
              "DECOMPRESSION TIME\n");
  mkdir("./input_data", 0755);
  switch (width) {
  case 1: return run<uint8_t>(in.bytes(), in.n);
  case 2: return run<uint16_t>(in.bytes(), in.n);
  case 4: return run<uint32_t>(in.bytes(), in.n);
  case 8: return run<uint64_t>(in.bytes(), in.n);
  default: std::printf("Not supported input bytes\n"); return 1;
  }
}
