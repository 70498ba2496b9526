// byte_producers.cpp -- CPU encoders and container writers of the
// reference's two byte-oriented codecs (BRLE, LZSS), and the generator of
// their benchmark inputs.
//
//   BRLE  stream: brle::encode, include/brle/brle.h:129-230, on its direct
//         path (brle.h:23 defines USE_LOOKAHEADd, so USE_LOOKAHEAD is off);
//         container: brle::compress_gpu, brle.h:251-303.
//   LZSS  stream: compress_func + find_match, include/lzss/lzss.h:75-113,
//         310-419; container: compress_gpu, lzss.h:1060-1124.
// Both encoders make the reference's choices, so their streams are the
// reference's bytes.
#include "../internal.hpp"

#include <cmath>
#include <string>
#include <vector>

namespace codag {

// ------------------------------------------------------------------ BRLE

// brle::encode (brle.h:129-230): literals accumulate up to 128 bytes; a third
// equal byte in a row turns the last two literals and it into a run, which
// grows to 130 bytes. A literal group is written as [h = n + 127][n bytes],
// a run as [h = n - 3][byte] (the WRITE_VALUES macro, brle.h:130-143).
static std::string brle_encode_chunk(const uint8_t *in, uint64_t n) {
  std::string out;
  uint8_t lit[128];
  uint32_t nl = 0, tail = 0;
  bool rep = false;
  auto flush = [&]() {
    if (nl == 0) return;
    if (rep) {
      out.push_back((char)(uint8_t)(nl - 3));
      out.push_back((char)lit[0]);
    } else {
      out.push_back((char)(uint8_t)(nl + 127));
      out.append(reinterpret_cast<const char *>(lit), nl);
    }
    rep = false;
    tail = 0;
    nl = 0;
  };
  for (uint64_t i = 0; i < n; i++) {
    const uint8_t c = in[i];
    if (nl == 0) {  // brle.h:174-176
      lit[nl++] = c;
      tail = 1;
    } else if (rep) {  // brle.h:177-194
      if (c == lit[0]) {
        if (++nl == 130) flush();
      } else {
        flush();
        lit[nl++] = c;
        tail = 1;
      }
    } else {  // brle.h:195-222
      tail = (c == lit[nl - 1]) ? tail + 1 : 1;
      if (tail == 3) {
        if (nl + 1 == 3) {
          nl += 1;
          rep = true;
        } else {
          nl -= 2;
          flush();
          lit[0] = c;
          nl = 3;
          rep = true;
        }
      } else {
        lit[nl++] = c;
        if (nl == 128) flush();
      }
    }
  }
  flush();
  return out;
}

// ------------------------------------------------------------------ LZSS

// compress_func (lzss.h:310-419) with find_match (lzss.h:75-113). The
// reference scans the history oldest first and keeps a strictly longer
// match, stopping at the first 34-byte one, so it chooses the longest match
// and, of equal lengths, the oldest start. Its candidates start 1..2047
// bytes back but not at the oldest history byte (hist_offset starts at 1,
// lzss.h:77), which is the chunk's first byte while fewer than 2048 bytes
// precede; a match may not run into the current position (lzss.h:86) and
// is bounded by the 512-byte lookahead (the bytes left in the chunk, at
// most). Matches shorter than 3 bytes are dropped (lzss.h:108-109), so only
// starts whose next three bytes agree can matter: this finds them through
// hash chains on three bytes instead of the reference's scan of all 2047.
static std::string lzss_encode_chunk(const uint8_t *in, uint32_t n) {
  constexpr uint32_t kHash = 1u << 16;
  std::vector<int32_t> head(kHash, -1), prev(n, -1), cand;
  cand.reserve(2048);
  auto hash3 = [&](uint32_t p) {
    const uint32_t v = (uint32_t)in[p] | ((uint32_t)in[p + 1] << 8) | ((uint32_t)in[p + 2] << 16);
    return (v * 2654435761u) >> 16;
  };
  auto insert = [&](uint32_t p) {
    if (p + 2 >= n) return;
    const uint32_t h = hash3(p);
    prev[p] = head[h];
    head[h] = (int32_t)p;
  };
  std::string out(1, '\0');  // out_bytes starts at 1: the first flag byte (lzss.h:326)
  size_t flag_at = 0;
  uint8_t flags = 0;
  uint32_t items = 0, cur = 0;
  while (cur < n) {
    const uint32_t look = std::min<uint32_t>(512, n - cur);
    uint32_t best = 0, best_p = 0;
    if (look >= 3 && cur >= 4) {
      const uint32_t lo = cur >= 2048 ? cur - 2047 : 1;
      cand.clear();
      for (int32_t p = head[hash3(cur)]; p >= (int32_t)lo; p = prev[p])
        if ((uint32_t)p + 3 <= cur) cand.push_back(p);
      for (size_t k = cand.size(); k-- > 0;) {  // oldest first
        const uint32_t p = (uint32_t)cand[k];
        const uint32_t lim = std::min<uint32_t>({34u, look, cur - p});
        uint32_t len = 0;
        while (len < lim && in[p + len] == in[cur + len]) len++;
        if (len > best) {
          best = len;
          best_p = p;
          if (best == 34) break;
        }
      }
    }
    if (best < 3) {  // literal (lzss.h:353-377)
      out.push_back((char)in[cur]);
      flags |= (uint8_t)(1u << items);
      insert(cur);
      cur += 1;
    } else {  // reference (lzss.h:379-405)
      const uint32_t v = ((cur - best_p) << 5) | (best - 3);
      out.push_back((char)(uint8_t)(v & 0xff));
      out.push_back((char)(uint8_t)(v >> 8));
      for (uint32_t k = 0; k < best; k++) insert(cur + k);
      cur += best;
    }
    if (++items == 8) {  // lzss.h:406-411: the next flag byte is reserved at once
      out[flag_at] = (char)flags;
      flags = 0;
      items = 0;
      flag_at = out.size();
      out.push_back('\0');
    }
  }
  if (items) out[flag_at] = (char)flags;  // lzss.h:415-417
  return out;
}

}  // namespace codag

using namespace codag;

extern "C" {

codag_status codag_brle_encode(const uint8_t *in, uint64_t in_bytes, uint64_t chunk_bytes,
                               int threads, uint8_t **container, uint64_t *container_bytes) {
  if ((!in && in_bytes) || !container || !container_bytes)
    return fail(CODAG_ERR_INVALID, "brle_encode: null argument");
  if (chunk_bytes == 0) chunk_bytes = 262144;
  const uint64_t n = (in_bytes + chunk_bytes - 1) / chunk_bytes;
  if (n + 1 > 0xffffffffull) return fail(CODAG_ERR_INVALID, "brle_encode: too many chunks");
  std::vector<std::string> s(n);
  run_threads(n, threads, [&](uint64_t i) {
    s[i] = brle_encode_chunk(in + i * chunk_bytes, std::min(chunk_bytes, in_bytes - i * chunk_bytes));
  });
  uint64_t data = 0;
  for (auto &x : s) data += x.size();
  const uint64_t head = 4 + 8 * (n + 1) + 8;
  uint8_t *c = host_alloc<uint8_t>(head + data);
  if (!c) return fail(CODAG_ERR_NOMEM, "brle_encode");
  const uint32_t n_ptr = (uint32_t)(n + 1);
  std::memcpy(c, &n_ptr, 4);
  uint64_t at = 0;
  for (uint64_t i = 0; i <= n; i++) {
    std::memcpy(c + 4 + 8 * i, &at, 8);
    if (i < n) {
      std::memcpy(c + head + at, s[i].data(), s[i].size());
      at += s[i].size();
    }
  }
  std::memcpy(c + 4 + 8 * (n + 1), &in_bytes, 8);
  *container = c;
  *container_bytes = head + data;
  return CODAG_OK;
}

codag_status codag_lzss_encode(const uint8_t *in, uint64_t in_bytes, uint64_t chunk_bytes,
                               int threads, uint8_t **container, uint64_t *container_bytes) {
  if ((!in && in_bytes) || !container || !container_bytes)
    return fail(CODAG_ERR_INVALID, "lzss_encode: null argument");
  if (chunk_bytes == 0) chunk_bytes = 131072;
  if (chunk_bytes > 0xffffffffull) return fail(CODAG_ERR_INVALID, "lzss_encode: chunk too large");
  const uint64_t n = (in_bytes + chunk_bytes - 1) / chunk_bytes;
  if (n > 0xffffffffull) return fail(CODAG_ERR_INVALID, "lzss_encode: too many chunks");
  std::vector<std::string> s(n);
  run_threads(n, threads, [&](uint64_t i) {
    s[i] = lzss_encode_chunk(in + i * chunk_bytes,
                             (uint32_t)std::min(chunk_bytes, in_bytes - i * chunk_bytes));
  });
  uint64_t data = 0;
  for (auto &x : s) data += x.size();
  const uint64_t head = 28 + 8 * n;  // HEAD_INTS * 4 + lens (lzss.h:1072-1073)
  uint8_t *c = host_alloc<uint8_t>(head + data);
  if (!c) return fail(CODAG_ERR_NOMEM, "lzss_encode");
  const uint32_t h[5] = {(uint32_t)chunk_bytes, (uint32_t)n, 2048u, 3u, 34u};
  std::memcpy(c, &in_bytes, 8);
  std::memcpy(c + 8, h, 20);
  uint64_t at = 0;
  for (uint64_t i = 0; i < n; i++) {
    std::memcpy(c + head + at, s[i].data(), s[i].size());
    at += s[i].size();
    std::memcpy(c + 28 + 8 * i, &at, 8);  // inclusive scan of the lengths (lzss.h:1089)
  }
  *container = c;
  *container_bytes = head + data;
  return CODAG_OK;
}

codag_status codag_synth_byte_payload(uint64_t n_bytes, uint32_t kind, uint64_t seed,
                                      uint8_t **raw) {
  if (!raw || kind > 1) return fail(CODAG_ERR_INVALID, "synth_byte_payload: bad argument");
  uint8_t *r = host_alloc<uint8_t>(n_bytes);
  if (!r) return fail(CODAG_ERR_NOMEM, "synth_byte_payload");
  // order-1 Markov over 96 printable bytes (each favours 4 successors, 75%)
  // for BRLE's literal stretches; a 512-word vocabulary of 2-9 letter words
  // for the LZSS text
  std::vector<uint8_t> fav(256 * 4);
  std::vector<std::string> words(512);
  {
    Rng t(seed ^ 0xb17eULL);
    for (auto &f : fav) f = (uint8_t)t.range(32, 127);
    for (auto &w : words) {
      const uint64_t len = t.range(2, 9);
      for (uint64_t k = 0; k < len; k++) w.push_back((char)t.range('a', 'z'));
    }
  }
  constexpr uint64_t kPart = 1u << 20;
  const uint64_t parts = (n_bytes + kPart - 1) / kPart;
  run_threads(parts, 0, [&](uint64_t p) {
    Rng rng(seed * 0x9e3779b97f4a7c15ull + p);
    uint8_t *d = r + p * kPart;
    const uint64_t n = std::min(kPart, n_bytes - p * kPart);
    uint64_t i = 0;
    if (kind == 0) {
      uint8_t prev = 'a';
      bool run = (p & 1) != 0;
      while (i < n) {
        if (run) {
          // 3 + Geometric(mean 30), at most 130
          const double u = std::max(rng.uniform(), 1e-12);
          uint64_t len = 3 + (uint64_t)(std::log(u) / std::log(1.0 - 1.0 / 31.0));
          len = std::min<uint64_t>({len, 130, n - i});
          const uint8_t v = (uint8_t)rng.range(0, 255);
          std::memset(d + i, v, len);
          i += len;
          prev = v;
        } else {
          const uint64_t len = std::min<uint64_t>(rng.range(1, 64), n - i);
          for (uint64_t k = 0; k < len; k++) {
            const uint8_t c =
                rng.chance(0.75) ? fav[prev * 4 + rng.range(0, 3)] : (uint8_t)rng.range(32, 127);
            d[i + k] = c;
            prev = c;
          }
          i += len;
        }
        run = !run;
      }
    } else {
      // words drawn from a skewed choice (the square of a uniform index),
      // separated by spaces, a number every 12 words on average, a newline
      // every 16
      uint64_t wcount = 0;
      while (i < n) {
        std::string tok;
        if (rng.chance(1.0 / 12)) {
          tok = std::to_string(rng.range(0, 99999));
        } else {
          const double u = rng.uniform();
          tok = words[(size_t)(u * u * 511.0)];
        }
        tok.push_back(++wcount % 16 == 0 ? '\n' : ' ');
        const uint64_t len = std::min<uint64_t>(tok.size(), n - i);
        std::memcpy(d + i, tok.data(), len);
        i += len;
      }
    }
  });
  *raw = r;
  return CODAG_OK;
}

}  // extern "C"
