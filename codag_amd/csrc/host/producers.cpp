// producers.cpp -- CPU encoders and container writers (the ★(producer) side).
//
// These write the byte formats the reference decoders read:
//   RLE v1 stream/container : include/rlev1/decoder.h:671-791, rlev1.h:103-192,
//                             encoder.h:780-783 (4-byte stream padding)
//   RLE v2 stream/container : include/rlev2/decoder_new.h:909-1187,
//                             encoder_trans.h:482-564 (128-byte alignment)
//   Deflate container       : deflate_compressor.py:44-126
// codag_rlev1_encode restates the reference's RLE v1 encoder choices, so its
// streams are the reference's byte for byte; the synthetic column generators
// emit their own controlled segment mixes.
#include "../internal.hpp"

#include <zlib.h>

#include <algorithm>
#include <memory>
#include <atomic>
#include <cmath>
#include <string>
#include <thread>
#include <vector>

namespace codag {

const uint8_t kV2DecodeWidth[32] = {1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11,
                                    12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22,
                                    23, 24, 26, 28, 30, 32, 40, 48, 56, 64};

uint32_t v2_closest_fixed(uint32_t bits) {
  for (uint32_t c = 0; c < 32; c++)
    if (kV2DecodeWidth[c] >= bits) return kV2DecodeWidth[c];
  return 64;
}

uint32_t v2_width_code(uint32_t bits) {
  if (bits < 1) bits = 1;
  for (uint32_t c = 0; c < 32; c++)
    if (kV2DecodeWidth[c] >= bits) return c;
  return 31;
}

static uint32_t bit_length(uint64_t v) { return v ? 64 - __builtin_clzll(v) : 0; }

// ------------------------------------------------------------------ RLE v1

// Greedy ORC v1 chunk encoder: runs of 3..130 values with a constant signed
// 8-bit delta (in the column's unsigned arithmetic), literal groups of up to
// 128 values otherwise.
static std::string rlev1_encode_chunk(const uint8_t *in, uint64_t n, uint32_t w) {
  const uint64_t mask = width_mask(w);
  const uint32_t nb = 8 * w;
  std::string out;
  std::vector<uint64_t> lits;
  auto flush = [&]() {
    size_t i = 0;
    while (i < lits.size()) {
      size_t g = std::min<size_t>(128, lits.size() - i);
      out.push_back(static_cast<char>(static_cast<int8_t>(-(int)g)));
      for (size_t k = 0; k < g; k++) put_uvarint(out, lits[i + k]);
      i += g;
    }
    lits.clear();
  };
  auto val = [&](uint64_t i) { return load_le(in + i * w, w); };
  auto sdelta = [&](uint64_t d) -> int64_t {
    d &= mask;
    if (nb < 64 && (d >> (nb - 1))) return (int64_t)(d | ~mask);
    return (int64_t)d;
  };
  uint64_t i = 0;
  while (i < n) {
    if (i + 2 < n) {
      uint64_t d = (val(i + 1) - val(i)) & mask;
      int64_t sd = sdelta(d);
      if (sd >= -128 && sd <= 127 && ((val(i + 2) - val(i + 1)) & mask) == d) {
        uint64_t j = i + 2;
        while (j + 1 < n && j + 1 - i < 130 && ((val(j + 1) - val(j)) & mask) == d) j++;
        flush();
        uint64_t len = j - i + 1;
        out.push_back(static_cast<char>(len - 3));
        out.push_back(static_cast<char>(static_cast<int8_t>(sd)));
        put_uvarint(out, val(i));
        i = j + 1;
        continue;
      }
    }
    lits.push_back(val(i));
    i++;
  }
  flush();
  return out;
}

// The reference's RLE v1 chunk encoder, restated for byte-identical streams:
// the writer state machine of compression_warp_orig (include/rlev1/
// encoder.h:1352-1695), one chunk of exactly n values (the reference encodes
// the zero-padded column a full chunk at a time, encoder.h:60-74). Its
// choices, all kept here:
//  * a run needs three values with one signed 8-bit step; the step and the
//    run test are computed in the column type T before widening, with C++'s
//    promotions (encoder.h:1416, 1459, 1491): uint8/uint16 steps never wrap,
//    uint32 steps are unsigned (a decrease is never a run), uint64 steps
//    wrap to a signed 64-bit difference;
//  * runs hold 3..129 values: the 130th value closes the run (header 126)
//    and starts the next sequence (encoder.h:1509-1523);
//  * literal groups hold at most 127 values (encoder.h:1447, 1573); the
//    group header is a placeholder byte (0x01 or 0x03, as the reference
//    writes them) overwritten with -count when the group closes;
//  * values and run bases are unsigned LEB128 of the T value
//    (write_varint_op_orig, encoder.h:839-853).
template <typename T> static std::string rlev1_ref_encode_chunk(const T *v, uint64_t n) {
  std::string out;
  auto byte = [&](uint8_t b) { out.push_back(static_cast<char>(b)); };
  auto varint = [&](T val) {
    do {
      uint8_t b = (uint8_t)(val & 0x7F);
      if ((val & ~(T)0x7F) != 0) b |= 0x80;
      byte(b);
      val = (T)(val >> 7);
    } while (val != 0);
  };
  auto in_range = [](int64_t d) { return !(d > 127 || d < -128); };
  T buf[2] = {0, 0};
  uint32_t head = 0, count = 0;
  T first = 0, prev = 0;
  uint32_t dcount = 0;
  int8_t cur = 0;
  bool dflag = false;
  uint64_t lit_idx = 0;
  uint32_t lit = 0;
  auto push = [&](T x) {  // buffer x as the newest value
    buf[head] = x;
    head ^= 1u;
  };
  auto close_group = [&]() { out[lit_idx] = static_cast<char>(static_cast<int8_t>(-(int)lit)); };
  if (n == 0) return out;
  // first value
  T x = v[0];
  dcount = 1;
  prev = x;
  push(x);
  count = 1;
  uint64_t i = 1;
  if (n > 1) {  // second value (the reference always reads two)
    x = v[1];
    i = 2;
    const int64_t d = (int64_t)(x - prev);
    if (!in_range(d)) {
      dflag = false;
      lit_idx = out.size();
      varint((T)1);
      lit = 1;
      varint(buf[0]);
      count--;
    } else {
      dflag = true;
      cur = (int8_t)d;
      dcount++;
    }
    prev = x;
    push(x);
    count++;
  }
  for (; i < n; i++) {
    if (lit == 127) {
      close_group();
      lit = 0;
    }
    x = v[i];
    if (dcount == 1) {
      const int64_t d = (int64_t)(x - prev);
      if (!in_range(d)) {
        dflag = false;
        if (lit == 0) {
          lit_idx = out.size();
          varint((T)1);
        }
        varint(buf[head ^ 1u]);
        lit++;
        count--;
      } else {
        dflag = true;
        cur = (int8_t)d;
        dcount++;
      }
      prev = x;
      push(x);
      count++;
      continue;
    }
    if (prev + cur == x && dflag) {  // in T with C++ promotions (encoder.h:1491)
      dcount++;
      if (dcount == 3) {
        first = buf[head];
        if (lit != 0) {
          close_group();
          lit = 0;
        }
      } else if (dcount == 130) {
        byte((uint8_t)(dcount - 4));
        byte((uint8_t)cur);
        varint(first);
        dcount = 1;
        count = 0;
        lit = 0;
      }
      push(x);
      count = std::min(count + 1, 2u);
      prev = x;
    } else if (dcount >= 3) {
      byte((uint8_t)(dcount - 3));
      byte((uint8_t)cur);
      varint(first);
      dcount = 1;
      lit = 0;
      push(x);
      count = 1;
      prev = x;
    } else {
      if (lit == 0) {
        lit_idx = out.size();
        byte(3);
      }
      lit++;
      varint(buf[head]);
      if (lit == 127) {
        close_group();
        lit = 0;
      }
      const int64_t d = (int64_t)(x - prev);
      if (!in_range(d)) {
        if (lit == 0) {
          lit_idx = out.size();
          byte(3);
        }
        dflag = false;
        count = 0;
        varint(buf[head ^ 1u]);
        lit++;
        dcount = 1;
      } else {
        count = 1;
        dflag = true;
        cur = (int8_t)d;
        dcount = 2;
      }
      prev = x;
      push(x);
      count = std::min(count + 1, 2u);
    }
  }
  if (dcount >= 3 && dflag) {
    byte((uint8_t)(dcount - 3));
    byte((uint8_t)cur);
    varint(first);
    return out;
  }
  if (count == 1) {
    if (lit == 127) {
      close_group();
      lit = 0;
    }
    if (lit == 0) {
      lit_idx = out.size();
      varint((T)1);
    }
    varint(buf[head ^ 1u]);
    lit++;
  }
  if (count == 2) {
    if (lit == 127) {
      close_group();
      lit = 0;
    }
    if (lit == 0) {
      lit_idx = out.size();
      byte(1);
    }
    varint(buf[head]);
    lit++;
    if (lit == 127) {
      close_group();
      lit = 0;
    }
    if (lit == 0) {
      lit_idx = out.size();
      byte(1);
    }
    head ^= 1u;
    varint(buf[head]);
    lit++;
  }
  close_group();
  return out;
}

static std::string rlev1_ref_encode(const uint8_t *chunk, uint64_t n, uint32_t w) {
  switch (w) {
  case 1: return rlev1_ref_encode_chunk(chunk, n);
  case 2: return rlev1_ref_encode_chunk(reinterpret_cast<const uint16_t *>(chunk), n);
  case 4: return rlev1_ref_encode_chunk(reinterpret_cast<const uint32_t *>(chunk), n);
  default: return rlev1_ref_encode_chunk(reinterpret_cast<const uint64_t *>(chunk), n);
  }
}

// Assembles the RLE v1 container + side files from per-chunk streams.
static codag_status rlev1_assemble(const std::vector<std::string> &streams, uint64_t in_bytes,
                                   uint8_t **container, uint64_t *container_bytes,
                                   uint64_t **blk_off, uint64_t *blk_off_entries,
                                   uint64_t **col_len, uint64_t *col_len_entries) {
  const uint64_t nc = streams.size();
  uint64_t total = 0;
  for (auto &s : streams) total += (s.size() + 3) / 4 * 4;  // encoder.h:780-783
  uint8_t *c = host_alloc<uint8_t>(8 + total);
  const uint64_t n_off = nc * 32 + 1;    // rlev1.h:117 blk_n_bytes
  const uint64_t n_len = nc * 32 * 32;   // rlev1.h:181 file length
  uint64_t *bo = host_alloc<uint64_t>(n_off);
  uint64_t *cl = host_alloc<uint64_t>(blk_off_entries ? n_len : nc);
  if (!c || !bo || !cl) {
    std::free(c); std::free(bo); std::free(cl);
    return fail(CODAG_ERR_NOMEM, "rlev1 container allocation");
  }
  store_le(c, 8, in_bytes);  // rlev1.h:163
  uint64_t off = 0;
  for (uint64_t i = 0; i < nc; i++) {
    bo[i] = off;
    cl[i] = streams[i].size();
    std::memcpy(c + 8 + off, streams[i].data(), streams[i].size());
    off += (streams[i].size() + 3) / 4 * 4;
  }
  for (uint64_t i = nc; i < n_off; i++) bo[i] = off;
  *container = c;
  *container_bytes = 8 + total;
  *blk_off = bo;
  if (blk_off_entries) *blk_off_entries = n_off;
  *col_len = cl;
  if (col_len_entries) *col_len_entries = n_len;
  return CODAG_OK;
}

// ------------------------------------------------------------------ RLE v2

struct BitWriter {  // MSB-first, flushed to a byte boundary by finish()
  std::string *out;
  uint64_t acc = 0;
  uint32_t n = 0;
  void put(uint64_t v, uint32_t bits) {
    // feed in pieces of <= 32 bits to keep acc from overflowing
    while (bits > 32) {
      bits -= 32;
      put32((v >> bits) & 0xffffffffull, 32);
    }
    put32(bits == 64 ? v : (v & ((1ull << bits) - 1)), bits);
  }
  void put32(uint64_t v, uint32_t bits) {
    acc = (acc << bits) | v;
    n += bits;
    while (n >= 8) {
      n -= 8;
      out->push_back(static_cast<char>((acc >> n) & 0xff));
    }
    acc &= (n ? ((1ull << n) - 1) : 0);
  }
  void finish() {
    if (n) out->push_back(static_cast<char>((acc << (8 - n)) & 0xff));
    acc = 0;
    n = 0;
  }
};

static void v2_put_be(std::string &s, uint64_t v, uint32_t nbytes) {
  for (int k = (int)nbytes - 1; k >= 0; k--) s.push_back(static_cast<char>((v >> (8 * k)) & 0xff));
}

static void v2_short_repeat(std::string &s, uint64_t v, uint32_t count) {
  uint32_t nbytes = std::max(1u, (bit_length(v) + 7) / 8);
  s.push_back(static_cast<char>(((nbytes - 1) << 3) | (count - 3)));
  v2_put_be(s, v, nbytes);
}

static void v2_header2(std::string &s, uint8_t type, uint32_t code, uint32_t len) {
  s.push_back(static_cast<char>(type | (code << 1) | (((len - 1) >> 8) & 1)));
  s.push_back(static_cast<char>((len - 1) & 0xff));
}

static void v2_direct(std::string &s, const uint64_t *v, uint32_t len, uint32_t code) {
  v2_header2(s, 0x40, code, len);
  BitWriter bw{&s};
  for (uint32_t i = 0; i < len; i++) bw.put(v[i], kV2DecodeWidth[code]);
  bw.finish();
}

static uint64_t zigzag(int64_t x) { return ((uint64_t)x << 1) ^ (uint64_t)(x >> 63); }

// DELTA with packed deltas (code >= 1) or a fixed delta (code 0).
static void v2_delta(std::string &s, uint64_t base, int64_t base_delta, const uint64_t *deltas,
                     uint32_t len, uint32_t code) {
  v2_header2(s, 0xc0, code, len);
  put_uvarint(s, base);
  put_uvarint(s, zigzag(base_delta));
  if (code != 0) {
    BitWriter bw{&s};
    for (uint32_t i = 0; i + 2 < len; i++) bw.put(deltas[i], kV2DecodeWidth[code]);
    bw.finish();
  }
}

struct Patch {
  uint32_t gap;
  uint64_t value;
};

static void v2_patched_base(std::string &s, const uint64_t *low, uint32_t len, uint32_t code,
                            uint64_t base, const std::vector<Patch> &patches, uint32_t pw_code,
                            uint32_t pgw) {
  uint32_t bw = std::max(1u, (bit_length(base) + 7) / 8);
  v2_header2(s, 0x80, code, len);
  s.push_back(static_cast<char>(((bw - 1) << 5) | pw_code));
  s.push_back(static_cast<char>(((pgw - 1) << 5) | (uint32_t)patches.size()));
  v2_put_be(s, base, bw);
  BitWriter w{&s};
  for (uint32_t i = 0; i < len; i++) w.put(low[i], kV2DecodeWidth[code]);
  w.finish();
  uint32_t pw = kV2DecodeWidth[pw_code];
  uint32_t cfb = v2_closest_fixed(pw + pgw);
  for (auto &p : patches) w.put(((uint64_t)p.gap << pw) | p.value, cfb);
  w.finish();
}

static int64_t sign_extend(uint64_t v, uint32_t nb) {
  if (nb >= 64) return (int64_t)v;
  uint64_t m = 1ull << (nb - 1);
  return (int64_t)((v ^ m) - m);
}

// Encodes one block of values (1..128) with the smallest of DIRECT, DELTA
// and PATCHED_BASE.
static void v2_encode_block(std::string &s, const uint64_t *v, uint32_t len, uint32_t w) {
  const uint64_t mask = width_mask(w);
  const uint32_t nb = 8 * w;
  // DIRECT
  uint64_t vmax = 0;
  for (uint32_t i = 0; i < len; i++) vmax |= v[i];
  uint32_t dcode = v2_width_code(std::max(1u, bit_length(vmax)));
  uint64_t best = 2 + ((uint64_t)len * kV2DecodeWidth[dcode] + 7) / 8;
  int choice = 0;
  // DELTA
  uint32_t tcode = 0;
  uint64_t d0 = 0, tbytes = ~0ull;
  int64_t sd0 = 0;
  std::vector<uint64_t> deltas;
  if (len >= 2) {
    d0 = (v[1] - v[0]) & mask;
    bool sub = (d0 == 0);
    uint64_t dmax = 0;
    bool fixed = true;
    deltas.resize(len > 2 ? len - 2 : 0);
    for (uint32_t i = 2; i < len; i++) {
      uint64_t d = sub ? ((v[i - 1] - v[i]) & mask) : ((v[i] - v[i - 1]) & mask);
      deltas[i - 2] = d;
      dmax |= d;
      if (sub ? d != 0 : d != d0) fixed = false;
    }
    sd0 = sign_extend(d0, nb);
    uint64_t zz = zigzag(sd0);
    if (zz < (1ull << 63) && (fixed || bit_length(dmax) <= 32)) {
      std::string tmp;
      put_uvarint(tmp, v[0]);
      put_uvarint(tmp, zz);
      if (fixed) {
        tcode = 0;
        tbytes = 2 + tmp.size();
      } else {
        tcode = v2_width_code(std::max(2u, bit_length(dmax)));  // width 1 -> 2 (code 0 = fixed)
        tbytes = 2 + tmp.size() + ((uint64_t)(len - 2) * kV2DecodeWidth[tcode] + 7) / 8;
      }
      if (tbytes < best) {
        best = tbytes;
        choice = 1;
      }
    }
  }
  // PATCHED_BASE: base = min, low bits at the ~95th percentile width
  std::vector<uint64_t> low(len);
  std::vector<Patch> patches;
  uint32_t pcode = 0, pwcode = 0, pgw = 1;
  uint64_t base = 0;
  if (len >= 8) {
    base = v[0];
    for (uint32_t i = 1; i < len; i++) base = std::min(base, v[i]);
    std::vector<uint64_t> red(len);
    for (uint32_t i = 0; i < len; i++) red[i] = (v[i] - base) & mask;
    std::vector<uint64_t> sorted(red);
    std::sort(sorted.begin(), sorted.end());
    uint64_t p95 = sorted[(len * 95) / 100 == len ? len - 1 : (len * 95) / 100];
    pcode = v2_width_code(std::max(1u, bit_length(p95)));
    uint32_t fbw = kV2DecodeWidth[pcode];
    uint64_t pmax = 0, gmax = 0, prev = 0;
    bool ok = fbw < nb;
    for (uint32_t i = 0; i < len && ok; i++) {
      uint64_t hi = fbw >= 64 ? 0 : red[i] >> fbw;
      low[i] = fbw >= 64 ? red[i] : red[i] & ((1ull << fbw) - 1);
      if (hi) {
        patches.push_back({(uint32_t)(i - prev), hi});
        gmax = std::max<uint64_t>(gmax, i - prev);
        pmax |= hi;
        prev = i;
      }
    }
    if (ok && !patches.empty() && patches.size() <= 31) {
      pwcode = v2_width_code(std::max(1u, bit_length(pmax)));
      pgw = std::max(1u, bit_length(gmax));
      uint32_t pw = kV2DecodeWidth[pwcode];
      if (pw < 32 && pw + pgw <= 64 && pgw <= 8) {
        uint32_t bw = std::max(1u, (bit_length(base) + 7) / 8);
        uint64_t pb = 4 + bw + ((uint64_t)len * fbw + 7) / 8 +
                      ((uint64_t)patches.size() * v2_closest_fixed(pw + pgw) + 7) / 8;
        if (pb < best) {
          best = pb;
          choice = 2;
        }
      }
    }
  }
  if (choice == 0) {
    v2_direct(s, v, len, dcode);
  } else if (choice == 1) {
    v2_delta(s, v[0], sd0, deltas.data(), len, tcode);
  } else {
    v2_patched_base(s, low.data(), len, pcode, base, patches, pwcode, pgw);
  }
}

static std::string rlev2_encode_chunk(const uint8_t *in, uint64_t n, uint32_t w) {
  std::string s;
  std::vector<uint64_t> v(n);
  for (uint64_t i = 0; i < n; i++) v[i] = load_le(in + i * w, w);
  uint64_t i = 0;
  while (i < n) {
    uint64_t r = 1;
    while (i + r < n && r < 10 && v[i + r] == v[i]) r++;
    if (r >= 3) {
      v2_short_repeat(s, v[i], (uint32_t)r);
      i += r;
      continue;
    }
    uint64_t end = i + 1;
    while (end < n && end - i < 128) {
      if (end + 2 < n && v[end] == v[end + 1] && v[end] == v[end + 2]) break;
      end++;
    }
    v2_encode_block(s, &v[i], (uint32_t)(end - i), w);
    i = end;
  }
  return s;
}

// The reference's RLE v2 chunk encoder, restated for byte-identical streams:
// block_encode_orig (include/rlev2/encoder_trans.h:184-357) driving the
// encoder of include/rlev2/encoder.h:83-591 (a port of ORC's RleEncoderV2)
// with its dialect, all kept here: values are held as uint64 (READ_TYPE is
// int64_t, utils.h:4-8), so nothing is ever negative; DELTA deltas are the
// unsigned 64-bit differences (no absolute value); a block whose minimum is 0
// and maximum at most 128 fails is_safe_subtract and is written DIRECT
// (encoder.h:64-67); DIRECT widths are rounded to the aligned widths
// 1, 2, 4, 8, 16, ...; a varying DELTA of width 1 is written with width 2;
// a patch width of 64 becomes 56 with 8-bit base-reduced values
// (encoder.h:205-211); histograms use uint8 counters (MAX_LITERAL_SIZE 128
// never overflows them); the DIRECT/DELTA/PATCHED_BASE length field is
// var_runlen - 1. One deviation, at the 128-zero run (see encode()).
namespace v2ref {

constexpr uint32_t kMaxLit = 128;  // utils.h:65 MAX_LITERAL_SIZE
constexpr uint32_t kMinRep = 3;    // utils.h:66 MINIMUM_REPEAT
constexpr uint32_t kMaxShort = 10; // utils.h:69 MAX_SHORT_REPEAT_LENGTH

// utils.h:30-43 __BIT_WIDTH_ENCODE_MAP: width w -> 5-bit code
inline uint8_t enc_width(uint32_t w) {
  if (w <= 1) return 0;
  if (w <= 24) return (uint8_t)(w - 1);
  if (w <= 32) return (uint8_t)(24 + (w - 25) / 2);
  return (uint8_t)(28 + (w - 33) / 8);
}
// utils.h:45-52 __CLOSEST_FIXED_BIT_MAP
inline uint8_t closest(uint32_t b) {
  if (b <= 1) return 1;
  if (b <= 24) return (uint8_t)b;
  if (b <= 32) return (uint8_t)((b + 1) & ~1u);
  return (uint8_t)((b + 7) & ~7u);
}
// utils.h:54-62 __CLOSEST_ALIGNED_FIXED_BIT_MAP
inline uint8_t closest_aligned(uint32_t b) {
  if (b <= 1) return 1;
  if (b == 2) return 2;
  if (b <= 4) return 4;
  if (b <= 8) return 8;
  return (uint8_t)((b + 7) & ~7u);
}
inline uint8_t nbits(uint64_t v) { return closest(v ? 64 - __builtin_clzll(v) : 0); }

struct Patch {
  uint32_t len = 0, width = 0, gap_width = 0;
  uint8_t bits95 = 0, bits100 = 0;
  uint64_t min = 0;
  uint64_t red[kMaxLit];
  uint64_t gap_patch[kMaxLit];
};

struct Enc {
  std::string out;
  uint64_t deltas[kMaxLit] = {};
  uint64_t lit[kMaxLit] = {};
  uint32_t n = 0, fix = 0, var = 0;
  uint8_t delta_bits = 0;
  bool fixed_delta = false;

  void put(uint8_t b) { out.push_back(static_cast<char>(b)); }

  // encoder.h:83-120
  void aligned(const uint64_t *in, uint32_t len, uint32_t bits) {
    if (bits < 8) {
      const uint64_t m = (1u << bits) - 1;
      const uint32_t hops = 8 / bits, rem = len % hops, end = len - rem;
      for (uint32_t i = 0; i < end; i += hops) {
        uint8_t b = 0;
        for (uint32_t j = 0; j < hops; j++) b |= (uint8_t)((in[i + j] & m) << (8 - (j + 1) * bits));
        put(b);
      }
      if (rem) {
        uint32_t sh = 8 - bits;
        uint8_t b = 0;
        for (uint32_t i = end; i < len; i++) {
          b |= (uint8_t)((in[i] & m) << sh);
          sh -= bits;
        }
        put(b);
      }
    } else {
      const int32_t bytes = (int32_t)(bits / 8);
      for (uint32_t i = 0; i < len; i++)
        for (int32_t j = bytes - 1; j >= 0; j--) put((uint8_t)((in[i] >> (j * 8)) & 0xff));
    }
  }

  // encoder.h:122-156
  void unaligned(const uint64_t *in, uint32_t len, uint32_t bits) {
    uint32_t left = 8;
    uint8_t cur = 0;
    for (uint32_t i = 0; i < len; i++) {
      uint64_t v = in[i];
      uint32_t todo = bits;
      while (todo > left) {
        cur |= (uint8_t)(v >> (todo - left));
        todo -= left;
        v &= (todo >= 64 ? ~0ull : ((1ull << todo) - 1));
        put(cur);
        cur = 0;
        left = 8;
      }
      left -= todo;
      cur |= (uint8_t)(v << left);
      if (left == 0) {
        put(cur);
        cur = 0;
        left = 8;
      }
    }
    if (left != 8) put(cur);
  }

  // encoder.h:158-178: hist of width codes; width covering p percent
  static void hist_of(const uint64_t *d, uint32_t len, uint8_t *h) {
    for (int i = 0; i < 32; i++) h[i] = 0;
    for (uint32_t i = 0; i < len; i++) h[enc_width(nbits(d[i]))]++;
  }
  static uint8_t pct_width(const uint8_t *h, uint32_t len, uint32_t p) {
    p = std::min<uint32_t>(100 - p, 100);
    int32_t plen = p == 0 ? 0 : (int32_t)((len * p - 1) / 100 + 1);
    for (int i = 31; i >= 0; i--) {
      plen -= h[i];
      if (plen < 0) return kV2DecodeWidth[i];
    }
    return 0;
  }

  void short_repeat() {  // encoder.h:396-412
    const uint64_t v = lit[0];
    const uint32_t vb = (nbits(v) - 1) / 8 + 1;
    put((uint8_t)(((vb - 1) << 3) | (fix - kMinRep)));
    for (int32_t i = (int32_t)vb - 1; i >= 0; i--) put((uint8_t)((v >> (i * 8)) & 0xff));
    fix = 0;
    n = 0;
    var = 0;
  }

  void direct() {  // encoder.h:424-440
    uint8_t h[32];
    hist_of(lit, n, h);
    const uint32_t bits = closest_aligned(pct_width(h, n, 100));
    const uint32_t len = var - 1;
    put((uint8_t)(0x40 | (enc_width(bits) << 1) | ((len & 0x100) >> 8)));
    put((uint8_t)(len & 0xff));
    aligned(lit, n, bits);
    var = 0;
    n = 0;
  }

  void vulong(uint64_t v) {  // encoder.h:443-454
    while (v & ~0x7full) {
      put((uint8_t)(0x80 | (v & 0x7f)));
      v >>= 7;
    }
    put((uint8_t)v);
  }

  void delta() {  // encoder.h:468-520
    uint32_t len = 0;
    uint32_t enc = 0;
    uint32_t bits = closest_aligned(delta_bits);
    if (fixed_delta) {
      if (fix > kMinRep) {
        len = fix - 1;
        fix = 0;
      } else {
        len = var - 1;
        var = 0;
      }
    } else {
      if (bits == 1) bits = 2;
      enc = enc_width(bits);
      len = var - 1;
      var = 0;
    }
    put((uint8_t)(0xC0 | (enc << 1) | ((len & 0x100) >> 8)));
    put((uint8_t)(len & 0xff));
    vulong(lit[0]);
    const int64_t d0 = (int64_t)deltas[0];
    vulong((uint64_t)((d0 << 1) ^ (d0 >> 63)));
    if (!fixed_delta) aligned(deltas + 1, n - 2, bits);
    for (uint32_t i = 0; i < n; i++) deltas[i] = 0;
    n = 0;
    var = 0;
  }

  void prepare_patch(Patch &pb) {  // encoder.h:196-266
    uint64_t mask = (uint64_t)((int64_t)(1ull << pb.bits95) - 1);
    pb.len = (n + 19) / 20;
    pb.width = closest(pb.bits100 - pb.bits95);
    if (pb.width == 64) {
      pb.width = 56;
      pb.bits95 = 8;
      mask = (1ull << pb.bits95) - 1;
    }
    uint32_t gi = 0;
    uint64_t gaps[kMaxLit], patches[kMaxLit];
    uint64_t prev = 0, max_gap = 0;
    for (uint32_t i = 0; i < n; i++) {
      if (pb.red[i] > mask) {
        const uint64_t gap = i - prev;
        if (gap > max_gap) max_gap = gap;
        prev = i;
        gaps[gi] = gap;
        patches[gi++] = pb.red[i] >> pb.bits95;
        pb.red[i] &= mask;
      }
    }
    pb.len = gi;
    if (max_gap == 0 && pb.len != 0) pb.gap_width = 1;
    else pb.gap_width = nbits(max_gap);
    if (pb.gap_width > 8) {  // unreachable for 128-value blocks (gaps < 128)
      pb.gap_width = 8;
      pb.len += max_gap == 511 ? 2 : 1;
    }
    uint32_t pos = 0, k = 0;
    for (uint32_t i = 0; i < pb.len; i++) {
      uint64_t g = gaps[k], pv = patches[k];
      k++;
      while (g > 255) {
        pb.gap_patch[pos++] = 255ull << pb.width;
        i++;
        g -= 255;
      }
      pb.gap_patch[pos++] = (g << pb.width) | pv;
    }
  }

  void patched(Patch &pb) {  // encoder.h:522-580 (no value is negative)
    var -= 1;
    put((uint8_t)(0x80 | (enc_width(pb.bits95) << 1) | ((var & 0x100) >> 8)));
    put((uint8_t)(var & 0xff));
    const uint32_t vb = ((uint32_t)nbits(pb.min) + 1 - 1) / 8 + 1;
    put((uint8_t)(((vb - 1) << 5) | enc_width(pb.width)));
    put((uint8_t)(((pb.gap_width - 1) << 5) | pb.len));
    for (int32_t i = (int32_t)vb - 1; i >= 0; i--) put((uint8_t)((pb.min >> (i * 8)) & 0xff));
    unaligned(pb.red, n, closest(pb.bits95));
    unaligned(pb.gap_patch, pb.len, closest(pb.gap_width + pb.width));
    var = 0;
    n = 0;
  }

  void determine() {  // encoder.h:268-394
    if (n <= kMinRep) {
      direct();
      return;
    }
    bool inc = true, dec = true;
    fixed_delta = true;
    uint64_t lo = lit[0], hi = lit[0];
    const uint64_t init = lit[1] - lit[0];
    uint64_t maxd = 0;
    for (uint32_t i = 1; i < n; i++) {
      const uint64_t l1 = lit[i], l0 = lit[i - 1];
      const uint64_t d = l1 - l0;
      lo = std::min(lo, l1);
      hi = std::max(hi, l1);
      inc &= (l0 <= l1);
      dec &= (l0 >= l1);
      fixed_delta &= (d == init);
      deltas[i - 1] = d;
      maxd = std::max(maxd, d);
    }
    deltas[0] = lit[1] - lit[0];
    // is_safe_subtract(hi, lo): lo > 0, or hi > lo - INT8_MIN (unsigned)
    if (!(lo > 0 || hi > lo + 128)) {
      direct();
      return;
    }
    if (lo == hi || fixed_delta) {
      delta();
      return;
    }
    if (init != 0) {
      delta_bits = nbits(maxd);
      if (inc || dec) {
        delta();
        return;
      }
    }
    uint8_t h[32];
    hist_of(lit, n, h);
    const uint32_t b100 = pct_width(h, n, 100), b90 = pct_width(h, n, 90);
    if (b100 - b90 > 1) {
      Patch pb;
      pb.min = lo;
      for (uint32_t i = 0; i < n; i++) pb.red[i] = lit[i] - lo;
      hist_of(pb.red, n, h);
      pb.bits95 = pct_width(h, n, 95);
      pb.bits100 = pct_width(h, n, 100);
      if (pb.bits100 - pb.bits95 != 0) {
        prepare_patch(pb);
        patched(pb);
      } else {
        direct();
      }
    } else {
      direct();
    }
  }

  // encoder_trans.h:229-339: one chunk's values
  void encode(const uint64_t *vals, uint64_t count) {
    uint64_t prev_delta = 0;
    for (uint64_t r = 0; r < count; r++) {
      const uint64_t val = vals[r];
      if (n == 0) {
        lit[n++] = val;
        fix = 1;
        var = 1;
        continue;
      }
      if (n == 1) {
        lit[n++] = val;
        prev_delta = val - lit[0];
        if (val == lit[0]) {
          fix = 2;
          var = 0;
        } else {
          fix = 0;
          var = 2;
        }
        continue;
      }
      const uint64_t cur = val - lit[n - 1];
      if (prev_delta == 0 && cur == 0) {
        lit[n++] = val;
        if (var > 0) fix = 2;
        fix++;
        if (fix >= kMinRep && var > 0) {
          n -= kMinRep;
          var -= kMinRep - 1;
          determine();
          prev_delta = 0;
          for (uint32_t ii = 0; ii < kMinRep; ii++) {
            lit[ii] = val;
            deltas[ii] = 0;
          }
          n = kMinRep;
        } else if (fix >= kMinRep) {
          deltas[0] = 0;
        }
        if (fix == kMaxLit) {
          // encoder_trans.h:289-291 calls determineEncoding here with
          // var_runlen 0. For a run of 128 zeros is_safe_subtract(0, 0)
          // fails, so the reference writes DIRECT with length field
          // var_runlen - 1 = 511 ahead of 128 values: a stream its own
          // decoder cannot read. The one deviation from the reference's
          // bytes: that run is written as the fixed DELTA block determine()
          // writes for a run of any other value.
          if (lit[0] == 0) {
            fixed_delta = true;
            delta();
          } else {
            determine();
          }
        }
        continue;
      }
      if (fix >= kMinRep) {
        if (fix <= kMaxShort) {
          short_repeat();
        } else {
          fixed_delta = true;
          delta();
        }
      }
      if (fix > 0 && fix < kMinRep && val != lit[n - 1]) {
        var = fix;
        fix = 0;
      }
      if (n == 0) {
        lit[n++] = val;
        fix = 1;
        var = 1;
      } else {
        prev_delta = val - lit[n - 1];
        lit[n++] = val;
        var++;
        if (var == kMaxLit) determine();
      }
    }
    if (n != 0) {
      if (var != 0) {
        determine();
      } else if (fix != 0) {
        if (fix < kMinRep) {
          var = fix;
          fix = 0;
          determine();
        } else if (fix <= kMaxShort) {
          short_repeat();
        } else {
          fixed_delta = true;
          delta();
        }
      }
    }
  }
};

}  // namespace v2ref

// One chunk of a w-byte column (the values [0, count)), reference choices.
static std::string rlev2_ref_encode(const uint8_t *in, uint64_t count, uint32_t w) {
  std::vector<uint64_t> v(count);
  for (uint64_t i = 0; i < count; i++) v[i] = load_le(in + i * w, w);
  auto e = std::make_unique<v2ref::Enc>();
  e->encode(v.data(), count);
  return std::move(e->out);
}

static codag_status rlev2_assemble(const std::vector<std::string> &streams, uint64_t raw_bytes,
                                   uint8_t **comp, uint64_t *comp_bytes, uint64_t **blk_off,
                                   uint64_t **col_len) {
  const uint64_t nc = streams.size();
  uint64_t total = 0;
  for (auto &s : streams) total += (s.size() + 127) / 128 * 128;  // encoder_trans.h:352
  uint8_t *c = host_alloc<uint8_t>(total);
  uint64_t *bo = host_alloc<uint64_t>(nc + 1);
  uint64_t *cl = host_alloc<uint64_t>(nc);
  if (!c || !bo || !cl) {
    std::free(c); std::free(bo); std::free(cl);
    return fail(CODAG_ERR_NOMEM, "rlev2 container allocation");
  }
  uint64_t off = 0;
  for (uint64_t i = 0; i < nc; i++) {
    bo[i] = off;
    cl[i] = streams[i].size();
    std::memcpy(c + off, streams[i].data(), streams[i].size());
    off += (streams[i].size() + 127) / 128 * 128;
  }
  bo[nc] = raw_bytes;  // encoder_trans.h:545
  *comp = c;
  *comp_bytes = total;
  *blk_off = bo;
  *col_len = cl;
  return CODAG_OK;
}

}  // namespace codag

using namespace codag;

extern "C" {

void codag_host_free(void *p) { std::free(p); }

codag_status codag_rlev1_encode(const uint8_t *in, uint64_t in_bytes, uint32_t w,
                                uint64_t chunk_bytes, uint8_t **container,
                                uint64_t *container_bytes, uint64_t **blk_off,
                                uint64_t *blk_off_entries, uint64_t **col_len,
                                uint64_t *col_len_entries, uint64_t *n_chunks) {
  if ((!in && in_bytes) || !container || !container_bytes || !blk_off || !col_len)
    return fail(CODAG_ERR_INVALID, "rlev1_encode: null argument");
  if (!valid_width(w) || chunk_bytes == 0 || chunk_bytes % w)
    return fail(CODAG_ERR_INVALID, "rlev1_encode: bad width or chunk size");
  // rlev1.h:111-112: chunk count rounded up to a multiple of 32
  uint64_t nc = (in_bytes + chunk_bytes * 32 - 1) / (chunk_bytes * 32) * 32;
  if (nc == 0) nc = 32;
  std::vector<std::string> streams(nc);
  run_threads(nc, 0, [&](uint64_t i) {
    // zero padding (rlev1.h:127); 8-byte aligned so the column reads as T
    std::vector<uint64_t> buf((chunk_bytes + 7) / 8, 0);
    uint8_t *b = reinterpret_cast<uint8_t *>(buf.data());
    uint64_t lo = i * chunk_bytes;
    if (lo < in_bytes) std::memcpy(b, in + lo, std::min(chunk_bytes, in_bytes - lo));
    streams[i] = rlev1_ref_encode(b, chunk_bytes / w, w);
  });
  if (n_chunks) *n_chunks = nc;
  return rlev1_assemble(streams, in_bytes, container, container_bytes, blk_off, blk_off_entries,
                        col_len, col_len_entries);
}

codag_status codag_rlev2_encode(const uint8_t *in, uint64_t in_bytes, uint32_t w,
                                uint64_t chunk_bytes, uint8_t **comp, uint64_t *comp_bytes,
                                uint64_t **blk_off, uint64_t **col_len, uint64_t *n_chunks) {
  if ((!in && in_bytes) || !comp || !comp_bytes || !blk_off || !col_len)
    return fail(CODAG_ERR_INVALID, "rlev2_encode: null argument");
  if (!valid_width(w) || chunk_bytes == 0 || chunk_bytes % w)
    return fail(CODAG_ERR_INVALID, "rlev2_encode: bad width or chunk size");
  uint64_t nc = (in_bytes + chunk_bytes * 32 - 1) / (chunk_bytes * 32) * 32;  // encoder_trans.h:490
  if (nc == 0) nc = 32;
  std::vector<std::string> streams(nc);
  run_threads(nc, 0, [&](uint64_t i) {
    // encoder_trans.h:207-209: chunk i encodes the whole values of
    // [i*chunk, min((i+1)*chunk, in_bytes)); no padding values
    const uint64_t lo = i * chunk_bytes;
    const uint64_t hi = std::min((i + 1) * chunk_bytes, in_bytes);
    const uint64_t count = hi > lo ? (hi - lo) / w : 0;
    streams[i] = rlev2_ref_encode(in + lo, count, w);
  });
  if (n_chunks) *n_chunks = nc;
  return rlev2_assemble(streams, in_bytes, comp, comp_bytes, blk_off, col_len);
}

codag_status codag_deflate_compress_chunks(const uint8_t *in, uint64_t in_bytes,
                                           uint64_t chunk_bytes, int level, int threads,
                                           uint8_t **comp, uint64_t *comp_bytes,
                                           uint64_t **blk_offset, uint64_t **col_len,
                                           uint64_t *n_chunks) {
  if ((!in && in_bytes) || !comp || !comp_bytes || !blk_offset || chunk_bytes == 0)
    return fail(CODAG_ERR_INVALID, "deflate_compress_chunks: bad argument");
  // deflate_compressor.py:62-66: zero-pad to whole chunks
  uint64_t nc = (in_bytes + chunk_bytes - 1) / chunk_bytes;
  std::vector<std::string> streams(nc);
  std::atomic<int> bad{0};
  run_threads(nc, threads, [&](uint64_t i) {
    std::vector<uint8_t> buf(chunk_bytes, 0);
    uint64_t lo = i * chunk_bytes;
    std::memcpy(buf.data(), in + lo, std::min(chunk_bytes, in_bytes - lo));
    uLongf cap = compressBound(chunk_bytes);
    std::string s(cap, '\0');
    if (compress2(reinterpret_cast<Bytef *>(&s[0]), &cap, buf.data(), chunk_bytes, level) != Z_OK)
      bad = 1;
    s.resize(cap);
    s.resize((s.size() + 7) / 8 * 8, '\0');  // deflate_compressor.py:90,104-112
    streams[i] = std::move(s);
  });
  if (bad) return fail(CODAG_ERR_INVALID, "zlib compress2 failed");
  uint64_t total = 0;
  for (auto &s : streams) total += s.size();
  uint8_t *c = host_alloc<uint8_t>(total);
  uint64_t *bo = host_alloc<uint64_t>(nc + 2);
  uint64_t *cl = host_alloc<uint64_t>(nc);
  if (!c || !bo || !cl) {
    std::free(c); std::free(bo); std::free(cl);
    return fail(CODAG_ERR_NOMEM, "deflate container allocation");
  }
  bo[0] = in_bytes;  // deflate_compressor.py:59-60
  bo[1] = 0;
  uint64_t off = 0;
  for (uint64_t i = 0; i < nc; i++) {
    std::memcpy(c + off, streams[i].data(), streams[i].size());
    off += streams[i].size();
    bo[i + 2] = off;  // :118-124
    cl[i] = streams[i].size();  // :125
  }
  *comp = c;
  *comp_bytes = total;
  *blk_offset = bo;
  if (col_len) *col_len = cl; else std::free(cl);
  if (n_chunks) *n_chunks = nc;
  return CODAG_OK;
}

// ----------------------------------------------- synthetic RLE v1 column

codag_status codag_synth_rlev1(uint64_t n_values, uint32_t w, uint64_t chunk_bytes,
                               uint64_t value_range, uint64_t seed, uint8_t **raw,
                               uint8_t **container, uint64_t *container_bytes, uint64_t **blk_off,
                               uint64_t **col_len, uint64_t *n_chunks) {
  if (!valid_width(w) || chunk_bytes == 0 || chunk_bytes % w || !raw || !container)
    return fail(CODAG_ERR_INVALID, "synth_rlev1: bad argument");
  const uint64_t per = chunk_bytes / w;
  const uint64_t in_bytes = n_values * w;
  uint64_t nc = (in_bytes + chunk_bytes * 32 - 1) / (chunk_bytes * 32) * 32;
  if (nc == 0) nc = 32;
  const uint64_t mask = width_mask(w);
  uint8_t *r = host_alloc<uint8_t>(nc * chunk_bytes);
  if (!r) return fail(CODAG_ERR_NOMEM, "synth_rlev1 raw");
  std::vector<std::string> streams(nc);
  run_threads(nc, 0, [&](uint64_t c) {
    Rng rng(seed * 0x100000001b3ull + c);
    uint8_t *dst = r + c * chunk_bytes;
    uint64_t have = std::min<uint64_t>(per, n_values > c * per ? n_values - c * per : 0);
    std::string s;
    uint64_t i = 0;
    bool run = rng.chance(0.5);
    while (i < have) {
      uint64_t left = have - i;
      if (run && left >= 3) {
        // length min(130, 3 + Geometric(mean 5))
        uint64_t g = 0;
        while (g < 127 && rng.chance(5.0 / 6.0)) g++;
        uint64_t len = std::min<uint64_t>({130, 3 + g, left});
        int64_t d = 0;
        if (rng.chance(0.3)) {
          do d = (int64_t)rng.range(0, 255) - 128; while (d == 0);
        }
        uint64_t base = rng.range(0, value_range - 1) & mask;
        s.push_back(static_cast<char>(len - 3));
        s.push_back(static_cast<char>(static_cast<int8_t>(d)));
        put_uvarint(s, base);
        for (uint64_t k = 0; k < len; k++)
          store_le(dst + (i + k) * w, w, (base + (uint64_t)((int64_t)k * d)) & mask);
        i += len;
      } else {
        uint64_t len = std::min<uint64_t>(rng.range(1, 128), left);
        s.push_back(static_cast<char>(static_cast<int8_t>(-(int)len)));
        for (uint64_t k = 0; k < len; k++) {
          uint64_t v = rng.range(0, value_range - 1) & mask;
          put_uvarint(s, v);
          store_le(dst + (i + k) * w, w, v);
        }
        i += len;
      }
      run = !run;
    }
    if (have < per) {  // zero padding of the column tail, as the reference pads
      std::string z = rlev1_encode_chunk(dst + have * w, per - have, w);
      // the padding zeros are already in dst (host_alloc zero-fills)
      s += z;
    }
    streams[c] = std::move(s);
  });
  *raw = r;
  if (n_chunks) *n_chunks = nc;
  return rlev1_assemble(streams, in_bytes, container, container_bytes, blk_off, nullptr, col_len,
                        nullptr);
}

// ----------------------------------------------- synthetic RLE v2 column

codag_status codag_synth_rlev2(uint64_t n_values, uint32_t w, uint64_t chunk_bytes,
                               uint64_t seed, uint8_t **raw, uint8_t **comp,
                               uint64_t *comp_bytes, uint64_t **blk_off, uint64_t **col_len,
                               uint64_t *n_chunks) {
  if (!valid_width(w) || chunk_bytes == 0 || chunk_bytes % w || !raw || !comp)
    return fail(CODAG_ERR_INVALID, "synth_rlev2: bad argument");
  const uint64_t per = chunk_bytes / w;
  const uint64_t in_bytes = n_values * w;
  uint64_t nc = (in_bytes + chunk_bytes * 32 - 1) / (chunk_bytes * 32) * 32;
  if (nc == 0) nc = 32;
  const uint64_t mask = width_mask(w);
  const uint32_t nb = 8 * w;
  uint8_t *r = host_alloc<uint8_t>(nc * chunk_bytes);
  if (!r) return fail(CODAG_ERR_NOMEM, "synth_rlev2 raw");
  // decode-map codes whose width fits the column
  std::vector<uint32_t> codes;
  for (uint32_t c = 0; c < 32; c++)
    if (kV2DecodeWidth[c] <= nb) codes.push_back(c);
  std::vector<std::string> streams(nc);
  run_threads(nc, 0, [&](uint64_t c) {
    Rng rng(seed * 0x9e3779b97f4a7c15ull + c);
    uint8_t *dst = r + c * chunk_bytes;
    uint64_t have = std::min<uint64_t>(per, n_values > c * per ? n_values - c * per : 0);
    std::string s;
    uint64_t cnt[4] = {0, 0, 0, 0};
    std::vector<uint64_t> v(128), low(128);
    uint64_t i = 0;
    while (i < have) {
      uint64_t left = have - i;
      int t = 0;  // least-served type so far: 25/25/25/25 by value count
      for (int k = 1; k < 4; k++)
        if (cnt[k] < cnt[t]) t = k;
      if (t == 0 && left < 3) t = 1;
      if ((t == 2 && left < 8) || (t == 3 && left < 3)) t = 1;
      uint32_t len = 0;
      if (t == 0) {  // SHORT_REPEAT, runs 3..10
        len = (uint32_t)std::min<uint64_t>(rng.range(3, 10), left);
        uint32_t vb = (uint32_t)rng.range(1, nb);
        uint64_t val = (vb >= 64 ? rng.next() : rng.next() & ((1ull << vb) - 1)) & mask;
        v2_short_repeat(s, val, len);
        for (uint32_t k = 0; k < len; k++) store_le(dst + (i + k) * w, w, val);
      } else if (t == 1) {  // DIRECT
        len = (uint32_t)std::min<uint64_t>(rng.range(16, 128), left);
        uint32_t code = codes[rng.range(0, codes.size() - 1)];
        uint32_t bits = kV2DecodeWidth[code];
        for (uint32_t k = 0; k < len; k++) {
          v[k] = bits >= 64 ? rng.next() : rng.next() & ((1ull << bits) - 1);
          store_le(dst + (i + k) * w, w, v[k] & mask);
        }
        v2_direct(s, v.data(), len, code);
      } else if (t == 2) {  // PATCHED_BASE, ~5% outliers
        len = (uint32_t)std::min<uint64_t>(rng.range(20, 128), left);
        uint32_t code = (uint32_t)rng.range(0, v2_width_code(std::min(nb / 2, 24u)));
        uint32_t fbw = kV2DecodeWidth[code];
        // patch width: a decode-map width with fbw + pw <= column bits and pw < 32
        uint32_t pwmax = std::min<uint32_t>(nb - fbw, 31);
        uint32_t ncodes = 0;
        while (ncodes < 32 && kV2DecodeWidth[ncodes] <= pwmax) ncodes++;
        uint32_t pcode = (uint32_t)rng.range(0, ncodes - 1);
        uint32_t pw = kV2DecodeWidth[pcode];
        uint64_t base = rng.next() & (nb >= 40 ? ((1ull << 32) - 1) : (mask >> 1));
        std::vector<Patch> patches;
        uint32_t prev = 0, gmax = 0;
        for (uint32_t k = 0; k < len; k++) {
          low[k] = rng.next() & ((1ull << fbw) - 1);
          uint64_t full = low[k];
          if (patches.size() < 31 && rng.chance(0.05)) {
            uint64_t p = (rng.next() & ((1ull << pw) - 1)) | 1;
            patches.push_back({k - prev, p});
            gmax = std::max(gmax, k - prev);
            prev = k;
            full |= p << fbw;
          }
          store_le(dst + (i + k) * w, w, (full + base) & mask);
        }
        if (patches.empty()) {  // force one outlier so the block is a real PATCHED_BASE
          uint32_t k = (uint32_t)rng.range(0, len - 1);
          uint64_t p = 1;
          patches.push_back({k, p});
          gmax = k;
          store_le(dst + (i + k) * w, w, ((low[k] | (p << fbw)) + base) & mask);
        }
        uint32_t pgw = std::max(1u, bit_length(gmax));
        v2_patched_base(s, low.data(), len, code, base, patches, pcode, pgw);
      } else {  // DELTA: monotone, deltas in [0,31], the first >= 1
        len = (uint32_t)std::min<uint64_t>(rng.range(16, 128), left);
        uint64_t base = rng.next() & (mask >> 8);
        uint64_t d0 = rng.range(1, 31);
        bool fixed = rng.chance(0.2);
        uint64_t dmax = 0;
        v[0] = base;
        v[1] = (base + d0) & mask;
        std::vector<uint64_t> deltas(len);
        for (uint32_t k = 2; k < len; k++) {
          uint64_t d = fixed ? d0 : rng.range(0, 31);
          deltas[k - 2] = d;
          dmax |= d;
          v[k] = (v[k - 1] + d) & mask;
        }
        for (uint32_t k = 0; k < len; k++) store_le(dst + (i + k) * w, w, v[k]);
        uint32_t code = fixed ? 0 : v2_width_code(std::max(2u, bit_length(dmax)));
        v2_delta(s, base, sign_extend(d0, nb), deltas.data(), len, code);
      }
      cnt[t] += len;
      i += len;
    }
    if (have < per) s += rlev2_encode_chunk(dst + have * w, per - have, w);
    streams[c] = std::move(s);
  });
  *raw = r;
  if (n_chunks) *n_chunks = nc;
  return rlev2_assemble(streams, in_bytes, comp, comp_bytes, blk_off, col_len);
}

// ------------------------------------------------------- synthetic Deflate

codag_status codag_synth_deflate_payload(uint64_t n_bytes, uint64_t part_bytes, uint64_t seed,
                                         uint8_t **raw) {
  if (!raw || part_bytes == 0) return fail(CODAG_ERR_INVALID, "synth_deflate_payload");
  uint8_t *r = host_alloc<uint8_t>(n_bytes);
  if (!r) return fail(CODAG_ERR_NOMEM, "synth_deflate_payload");
  // Shared order-2 transition table over a 64-symbol alphabet: each context
  // favours 6 successors (70%) and otherwise picks uniformly.
  static const char kAlpha[65] =
      "etaoinshrdlucmfwypvbgkjqxzETAOINSHRDLUCMFWYP .,;:!?-'\"()0123456";
  std::vector<uint8_t> fav(64 * 64 * 6);
  {
    Rng t(seed ^ 0x5eedULL);
    for (auto &f : fav) f = (uint8_t)t.range(0, 63);
  }
  // Zipf(1.1) over 1..2^20, sampled by rejection-inversion (Hormann and
  // Derflinger, "Rejection-inversion to generate variates from monotone
  // discrete distributions", ACM TOMACS 6(3), 1996).
  const double s = 1.1, K = 1048576.0;
  auto Hf = [&](double x) { return (std::pow(x, 1.0 - s) - 1.0) / (1.0 - s); };
  auto Hinv = [&](double u) { return std::pow(1.0 + u * (1.0 - s), 1.0 / (1.0 - s)); };
  auto hf = [&](double x) { return std::pow(x, -s); };
  const double hx1 = Hf(1.5) - 1.0, hN = Hf(K + 0.5);
  const double sq = 2.0 - Hinv(Hf(2.5) - hf(2.0));
  const uint64_t parts = (n_bytes + part_bytes - 1) / part_bytes;
  run_threads(parts, 0, [&](uint64_t p) {
    Rng rng(seed * 0x2545f4914f6cdd1dull + p);
    uint8_t *d = r + p * part_bytes;
    uint64_t n = std::min(part_bytes, n_bytes - p * part_bytes);
    if (p % 2 == 0) {  // text
      uint32_t c1 = 0, c2 = 0;
      uint64_t i = 0;
      while (i < n) {
        if (i > 64 && rng.chance(0.012)) {
          // repeat of length 3..258 (log-uniform) from up to 32 KiB back
          uint64_t len = (uint64_t)std::exp(std::log(3.0) + rng.uniform() * (std::log(258.0) - std::log(3.0)));
          len = std::min<uint64_t>({len, 258, n - i});
          uint64_t maxd = std::min<uint64_t>(i, 32768);
          uint64_t dist = rng.range(1, maxd);
          for (uint64_t k = 0; k < len; k++) d[i + k] = d[i + k - dist];
          i += len;
          if (i >= 2) {
            c2 = c1;
            c1 = 0;
          }
          continue;
        }
        uint32_t sym = rng.chance(0.7) ? fav[(c2 * 64 + c1) * 6 + rng.range(0, 5)]
                                       : (uint32_t)rng.range(0, 63);
        d[i++] = (uint8_t)kAlpha[sym];
        c2 = c1;
        c1 = sym;
      }
    } else {  // little-endian int32 column, Zipf(1.1)
      uint64_t i = 0;
      while (i + 4 <= n) {
        int32_t v;
        for (;;) {
          const double u = hN + rng.uniform() * (hx1 - hN);
          const double x = Hinv(u);
          double k = std::floor(x + 0.5);
          if (k < 1) k = 1;
          if (k > K) k = K;
          if (k - x <= sq || u >= Hf(k + 0.5) - hf(k)) {
            v = (int32_t)k;
            break;
          }
        }
        std::memcpy(d + i, &v, 4);
        i += 4;
      }
      while (i < n) d[i++] = 0;
    }
  });
  *raw = r;
  return CODAG_OK;
}

codag_status codag_deflate_compress_mixed(const uint8_t *in, uint64_t in_bytes,
                                          uint32_t pct_fixed, uint32_t pct_stored,
                                          uint64_t block_bytes, uint64_t seed, uint8_t **stream,
                                          uint64_t *stream_bytes, uint64_t counts[3]) {
  if ((!in && in_bytes) || !stream || !stream_bytes || block_bytes == 0 ||
      pct_fixed + pct_stored > 100)
    return fail(CODAG_ERR_INVALID, "deflate_compress_mixed: bad argument");
  z_stream z;
  std::memset(&z, 0, sizeof z);
  if (deflateInit(&z, 6) != Z_OK) return fail(CODAG_ERR_NOMEM, "deflateInit");
  std::string out;
  out.resize(compressBound(in_bytes) + 64 * (in_bytes / block_bytes + 2) + 64);
  z.next_out = reinterpret_cast<Bytef *>(&out[0]);
  z.avail_out = (uInt)out.size();
  Rng rng(seed);
  uint64_t c[3] = {0, 0, 0};
  uint64_t pos = 0;
  int cur_level = 6, cur_strategy = Z_DEFAULT_STRATEGY;
  do {
    uint32_t roll = (uint32_t)rng.range(0, 99);
    int kind = roll < pct_fixed ? 0 : (roll < pct_fixed + pct_stored ? 2 : 1);
    int level = kind == 2 ? 0 : 6;
    int strategy = kind == 0 ? Z_FIXED : Z_DEFAULT_STRATEGY;
    if (level != cur_level || strategy != cur_strategy) {
      if (deflateParams(&z, level, strategy) != Z_OK) {
        deflateEnd(&z);
        return fail(CODAG_ERR_INVALID, "deflateParams");
      }
      cur_level = level;
      cur_strategy = strategy;
    }
    uint64_t n = std::min(block_bytes, in_bytes - pos);
    z.next_in = const_cast<Bytef *>(in + pos);
    z.avail_in = (uInt)n;
    pos += n;
    int flush = pos >= in_bytes ? Z_FINISH : Z_BLOCK;
    int zr = deflate(&z, flush);
    if (zr == Z_STREAM_ERROR || z.avail_in != 0) {
      deflateEnd(&z);
      return fail(CODAG_ERR_INVALID, "deflate (mixed) failed");
    }
    c[kind]++;
    if (flush == Z_FINISH && zr != Z_STREAM_END) {
      deflateEnd(&z);
      return fail(CODAG_ERR_INVALID, "deflate (mixed) did not finish");
    }
  } while (pos < in_bytes);
  if (in_bytes == 0) deflate(&z, Z_FINISH);
  uint64_t len = z.total_out;
  deflateEnd(&z);
  uint8_t *s = host_alloc<uint8_t>(len);
  if (!s) return fail(CODAG_ERR_NOMEM, "deflate (mixed) output");
  std::memcpy(s, out.data(), len);
  *stream = s;
  *stream_bytes = len;
  if (counts) {
    counts[0] = c[0];
    counts[1] = c[1];
    counts[2] = c[2];
  }
  return CODAG_OK;
}

}  // extern "C"
