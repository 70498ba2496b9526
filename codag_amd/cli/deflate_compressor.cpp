// deflate_compressor -- writes the chunked zlib container the zlib tool reads,
// as the reference's data-prep script does (deflate_compressor.py:44-126).
//
// Usage: deflate_compressor <input file> [<output dir>] [-l level] [-c chunk_bytes] [-j threads]
// For input "<dir>/<name>.<ext>" it writes, in <output dir> (default "."):
//   <name>_comp.bin        the chunk streams, each padded with zeros to 8 bytes
//   <name>_blk_offset.bin  [u64 input length][0][cumulative padded ends...]
//   <name>_col_len.bin     the padded stream length of each chunk (u64)
// <name> is the file name up to its first '.', as deflate_compressor.py:139
// takes it. The input is zero-padded to whole chunks (128 KiB by default) and
// every chunk is compressed as one zlib stream at level 9 (the script's
// zlib.compress(level=9), deflate_compressor.py:87). Prints "n chunks: N".
#include <cstring>
#include <string>

#include "cli_common.hpp"

int main(int argc, char **argv) {
  std::string in_path, out_dir = ".";
  int level = 9, threads = 0;
  uint64_t chunk = 131072;
  int pos = 0;
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "-l") && i + 1 < argc) level = std::atoi(argv[++i]);
    else if (!std::strcmp(argv[i], "-c") && i + 1 < argc) chunk = std::strtoull(argv[++i], nullptr, 10);
    else if (!std::strcmp(argv[i], "-j") && i + 1 < argc) threads = std::atoi(argv[++i]);
    else if (pos == 0 && ++pos) in_path = argv[i];
    else if (pos == 1 && ++pos) out_dir = argv[i];
    else pos = 99;
  }
  if (in_path.empty() || pos > 2 || level < 0 || level > 9 || chunk == 0) {
    std::fprintf(stderr,
                 "usage: deflate_compressor <input file> [<output dir>] [-l level] "
                 "[-c chunk_bytes] [-j threads]\n");
    return 1;
  }
  MappedFile in;
  if (!in.open(in_path.c_str())) {
    std::fprintf(stderr, "Fatal Error: cannot open %s\n", in_path.c_str());
    return 1;
  }
  std::string base = in_path.substr(in_path.find_last_of('/') + 1);
  base = base.substr(0, base.find('.'));
  uint8_t *comp = nullptr;
  uint64_t comp_n = 0, *blk = nullptr, *col = nullptr, n = 0;
  codag_status s = codag_deflate_compress_chunks(in.bytes(), in.n, chunk, level, threads, &comp,
                                                 &comp_n, &blk, &col, &n);
  if (s) die("compress", s);
  std::printf("n chunks:  %llu\n", (unsigned long long)n);
  const std::string stem = out_dir + "/" + base;
  int rc = 0;
  if (!write_file(stem + "_comp.bin", comp, comp_n) ||
      !write_file(stem + "_blk_offset.bin", blk, (n + 2) * 8) ||
      !write_file(stem + "_col_len.bin", col, n * 8)) {
    std::fprintf(stderr, "Fatal Error: cannot write %s_*.bin\n", stem.c_str());
    rc = 1;
  }
  codag_host_free(comp);
  codag_host_free(blk);
  codag_host_free(col);
  return rc;
}
