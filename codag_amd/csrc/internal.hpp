// internal.hpp -- shared host-side helpers of libcodag (not part of the ABI).
#pragma once

#include "codag.h"

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

extern "C" codag_status codag_test_poll_shift(codag_ctx *ctx, uint32_t shift);
extern "C" codag_status codag_test_chunks_two_wave(codag_ctx *ctx, uint32_t on);

namespace codag {

// Records the message returned by codag_last_error() on this thread.
void set_error(const std::string &msg);
void clear_error();

inline codag_status fail(codag_status s, const std::string &msg) {
  set_error(msg);
  return s;
}

template <typename T> T *host_alloc(uint64_t n) {
  void *p = std::malloc(n ? n * sizeof(T) : 1);
  if (p) std::memset(p, 0, n ? n * sizeof(T) : 1);
  return static_cast<T *>(p);
}

inline uint64_t width_mask(uint32_t w) {
  return w >= 8 ? ~0ull : ((1ull << (8 * w)) - 1);
}

inline bool valid_width(uint32_t w) { return w == 1 || w == 2 || w == 4 || w == 8; }

inline uint64_t load_le(const uint8_t *p, uint32_t w) {
  uint64_t v = 0;
  std::memcpy(&v, p, w);
  return v;
}

inline void store_le(uint8_t *p, uint32_t w, uint64_t v) { std::memcpy(p, &v, w); }

// Appends an unsigned LEB128 varint (7 bits per byte, LSB group first).
inline void put_uvarint(std::string &s, uint64_t v) {
  while (v >= 0x80) {
    s.push_back(static_cast<char>((v & 0x7f) | 0x80));
    v >>= 7;
  }
  s.push_back(static_cast<char>(v));
}

// Small deterministic PRNG (splitmix64 seeding + xoshiro256**).
struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    for (int i = 0; i < 4; i++) {
      seed += 0x9e3779b97f4a7c15ull;
      uint64_t z = seed;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
      s[i] = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    uint64_t r = rotl(s[1] * 5, 7) * 9;
    uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // Uniform in [lo, hi] (inclusive).
  uint64_t range(uint64_t lo, uint64_t hi) {
    uint64_t span = hi - lo + 1;
    return span == 0 ? next() : lo + next() % span;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  bool chance(double p) { return uniform() < p; }
};

// Runs fn(i) for i in [0, n) on `threads` host threads (< 1: all cores),
// items handed out in order; returns the thread count used.
template <typename Fn> int run_threads(uint64_t n, int threads, const Fn &fn) {
  if (threads < 1) threads = (int)std::max(1u, std::thread::hardware_concurrency());
  if ((uint64_t)threads > n) threads = (int)std::max<uint64_t>(1, n);
  std::vector<std::thread> ts;
  std::atomic<uint64_t> next{0};
  auto body = [&]() {
    for (;;) {
      const uint64_t i = next.fetch_add(1);
      if (i >= n) break;
      fn(i);
    }
  };
  for (int t = 1; t < threads; t++) ts.emplace_back(body);
  body();
  for (auto &t : ts) t.join();
  return threads;
}

// RLE v2 width tables (include/rlev2/utils.h:28-59).
extern const uint8_t kV2DecodeWidth[32];
uint32_t v2_closest_fixed(uint32_t bits);
// Smallest encoded width code whose decoded width >= bits (bits 1..64).
uint32_t v2_width_code(uint32_t bits);

}  // namespace codag
