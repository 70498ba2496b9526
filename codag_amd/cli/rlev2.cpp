// rlev2 -- drop-in for the reference's build/exe/src/rlev2/rlev2.cu.exe.
//
// Usage: rlev2 <raw column> <1|2|4|8>      (positional, src/rlev2/rlev2.cu:118-126)
//    or: rlev2 -f <raw column> -t <1|2|4|8> (README.md:17)
// Behaviour follows rlev2.cu:44-115 through the reference's host API
// (codag_compat.hpp): rlev2::compress_gpu_orig encodes with 128 KiB chunks,
// then two groups of five rlev2::decompress_gpu_orig calls (queue sizes 48
// and 64, NUM_CHUNKS 1..16) each print "<seconds> secs" after the group's
// "CHUNK \t48\tencoded\tmeta\t" prefix, and every result is compared with the
// input (check_diff, rlev2.cu:18-43: "fail at i in(out)"). Exit status: 0
// when every decode matches, 2 on a mismatch, 1 on an error.
#include <cstring>
#include <string>

#include "cli_common.hpp"
#include "codag_compat.hpp"

template <typename T, int QS, int NW, int NC>
static bool variant(const T *in, uint64_t size, const uint8_t *enc, uint64_t enc_n, uint64_t n,
                    rlev2::blk_off_t *blk, rlev2::col_len_t *col, uint64_t chunk) {
  T *dec = nullptr;
  uint64_t dec_n = 0;
  rlev2::decompress_gpu_orig<1, QS, T, NW, NC>(enc, enc_n, n, blk, col, dec, dec_n, chunk);
  if (!dec) return false;
  bool ok = dec_n == size;
  const uint8_t *a = reinterpret_cast<const uint8_t *>(in);
  const uint8_t *b = reinterpret_cast<const uint8_t *>(dec);
  const int64_t d = first_diff(a, b, dec_n < size ? dec_n : size);
  if (d >= 0) {
    std::printf("fail at %llu %x(%x)\n", (unsigned long long)d, a[d], b[d]);
    ok = false;
  }
  delete[] dec;
  return ok;
}

template <typename T> static int run(const T *in, uint64_t size) {
  const uint64_t chunk = 1024 * 128;  // rlev2.cu:142
  uint8_t *enc = nullptr;
  uint64_t enc_n = 0, meta = 0, n = 0;
  rlev2::blk_off_t *blk = nullptr;
  rlev2::col_len_t *col = nullptr;
  rlev2::compress_gpu_orig<1, T>(in, size, enc, enc_n, meta, n, blk, col, chunk);
  if (!enc) return 1;
  bool ok = true;
  std::printf("%llu \t%d\t%llu\t%llu\t", (unsigned long long)chunk, 48,
              (unsigned long long)enc_n, (unsigned long long)meta);
  ok = variant<T, 48, 2, 1>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  ok = variant<T, 48, 2, 2>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  ok = variant<T, 48, 2, 4>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  ok = variant<T, 48, 2, 8>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  ok = variant<T, 48, 2, 16>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  std::printf("%llu \t%d\t%llu\t%llu\t", (unsigned long long)chunk, 48,
              (unsigned long long)enc_n, (unsigned long long)meta);
  ok = variant<T, 64, 1, 1>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  ok = variant<T, 64, 1, 2>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  ok = variant<T, 64, 1, 4>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  ok = variant<T, 64, 1, 8>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  ok = variant<T, 64, 1, 16>(in, size, enc, enc_n, n, blk, col, chunk) && ok;
  delete[] blk;
  delete[] col;
  delete[] enc;
  return ok ? 0 : 2;
}

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
  MappedFile in;
  if (!in.open(fname.c_str())) {
    std::printf("fatal error: input file open error\n");
    return 1;
  }
  std::printf("CHUNK_SIZE\t BUFFER SIZE \t COMPRESSED SIZE \t META DATA SIZE \t DECOMPRESSION TIME\n");
  switch (width) {
  case 1: return run(reinterpret_cast<const uint8_t *>(in.bytes()), in.n);
  case 2: return run(reinterpret_cast<const uint16_t *>(in.bytes()), in.n);
  case 4: return run(reinterpret_cast<const uint32_t *>(in.bytes()), in.n);
  case 8: return run(reinterpret_cast<const uint64_t *>(in.bytes()), in.n);
  default:
    std::fprintf(stderr, "SIZE_OF_DATATYPE must be 1, 2, 4 or 8\n");
    return 1;
  }
}
