// cli_common.hpp -- small helpers shared by the drop-in command-line tools.
#pragma once

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "codag.h"

// Read-only memory map of a whole file (the reference CLIs mmap their inputs,
// e.g. src/rlev1/rlev1.cu:134-152, src/zlib/zlib.cu:60-89).
struct MappedFile {
  void *p = nullptr;
  uint64_t n = 0;
  int fd = -1;
  bool open(const char *path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat sb;
    if (fstat(fd, &sb) != 0) return false;
    n = (uint64_t)sb.st_size;
    if (n == 0) return true;
    p = mmap(nullptr, n, PROT_READ, MAP_PRIVATE, fd, 0);
    if (p == MAP_FAILED) {
      p = nullptr;
      return false;
    }
    return true;
  }
  const uint8_t *bytes() const { return static_cast<const uint8_t *>(p); }
  ~MappedFile() {
    if (p) munmap(p, n);
    if (fd >= 0) close(fd);
  }
};

inline bool write_file(const std::string &path, const void *data, uint64_t n) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  bool ok = n == 0 || std::fwrite(data, 1, n, f) == n;
  return std::fclose(f) == 0 && ok;
}

[[noreturn]] inline void die(const char *what, codag_status s) {
  std::fprintf(stderr, "Fatal Error: %s: %s (%s)\n", what, codag_status_string(s), codag_last_error());
  std::exit(1);
}

// First differing byte, as the reference's comp_test / check_diff report it
// (src/rlev1/rlev1.cu:21-28, src/rlev2/rlev2.cu:18-43). Returns -1 if equal.
inline int64_t first_diff(const uint8_t *a, const uint8_t *b, uint64_t n) {
  for (uint64_t i = 0; i < n; i++)
    if (a[i] != b[i]) return (int64_t)i;
  return -1;
}
