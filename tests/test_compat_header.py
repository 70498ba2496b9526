"""CPU checks of the reference-shaped C++ layer (include/codag_compat.hpp) and
of the container writer CLI.

* A translation unit that makes the reference CLIs' own calls (template
  arguments and argument types of src/rlev1/rlev1.cu:168-193,
  src/rlev2/rlev2.cu:55-110 and src/zlib/zlib.cu:102-104), and the BRLE /
  LZSS host functions with the reference's signatures (brle.h:251,305;
  lzss.h:1060,1126), compiles and links against libcodag.so.
* deflate_compressor writes the files deflate_compressor.py:44-126 writes:
  zero-padded 128 KiB chunks, each zlib.compress(level=9)-ed and padded to 8
  bytes, [length][0][cumulative ends] offsets and padded lengths.
"""
import os
import subprocess
import zlib

import numpy as np

import codag_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CALLS = r"""
#include "codag_compat.hpp"
int main() {
  const uint8_t *in = nullptr; uint8_t *out = nullptr, *out2 = nullptr;
  uint64_t n = 0, n2 = 0;
  rle_v1::compress_gpu_orig<uint64_t, uint16_t, uint32_t>(in, &out, 0, &n, 8, 131072);
  rle_v1::decompress_gpu_orig<uint32_t, uint16_t, uint32_t, 64, 4, 1>(out, &out2, n, &n2, 8, 131072);
  rle_v1::decompress_gpu_orig<uint32_t, uint64_t, uint32_t, 64, 16>(out, &out2, n, &n2, 8, 131072);
  uint64_t *col = nullptr;
  uint32_t *in32 = nullptr, *dec = nullptr;
  uint8_t *enc = nullptr; uint64_t enc_n = 0, meta = 0, nch = 0, dec_n = 0;
  rlev2::blk_off_t *blk_off = nullptr; rlev2::col_len_t *col_len = nullptr;
  rlev2::compress_gpu_orig<8, uint32_t>(in32, 0, enc, enc_n, meta, nch, blk_off, col_len, 131072);
  rlev2::decompress_gpu_orig<8, 48, uint32_t, 2, 4>(enc, enc_n, nch, blk_off, col_len, dec, dec_n, 131072);
  rlev2::decompress_gpu_orig<8, 64, uint32_t, 1, 16>(enc, enc_n, nch, blk_off, col_len, dec, dec_n, 131072);
  deflate::decompress_gpu<uint32_t, 262144 / 2, 64, 8, 1>(in, &out, 0, &n, col, 16, 131072);
  brle::compress_gpu(in, &out, 0, &n);          // brle.h:251
  brle::decompress_gpu(out, &out2, n, &n2);     // brle.h:305
  lzss::compress_gpu(in, &out, 0, &n);          // lzss.h:1060
  lzss::decompress_gpu_base(out, &out2, n, &n2);  // lzss.h:1126
  return 0;
}
"""


def test_compat_header_compiles_with_the_reference_calls(tmp_path):
    src = tmp_path / "calls.cpp"
    src.write_text(CALLS)
    exe = tmp_path / "calls"
    p = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        str(src), "-o", str(exe), "-L", os.path.join(ROOT, "codag_amd", "lib"),
                        "-lcodag"], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr


def test_deflate_compressor_matches_the_script_layout(tmp_path):
    rng = np.random.default_rng(3)
    raw = (rng.integers(0, 7, 300000) * (np.arange(300000) % 5 == 0)).astype(np.uint8).tobytes()
    src = tmp_path / "col.v1.bin"  # the stem is the name up to its first dot
    src.write_bytes(raw)
    p = subprocess.run([os.path.join(ROOT, "codag_amd", "bin", "deflate_compressor"), str(src),
                        str(tmp_path)], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    k = 131072
    padded = raw + b"\0" * ((-len(raw)) % k)
    exp, ends = b"", [0]
    for i in range(len(padded) // k):
        z = zlib.compress(padded[i * k:(i + 1) * k], 9)
        z += b"\0" * ((-len(z)) % 8)
        exp += z
        ends.append(ends[-1] + len(z))
    assert (tmp_path / "col_comp.bin").read_bytes() == exp
    blk = np.frombuffer((tmp_path / "col_blk_offset.bin").read_bytes(), np.uint64)
    col = np.frombuffer((tmp_path / "col_col_len.bin").read_bytes(), np.uint64)
    assert blk.tolist() == [len(raw)] + ends
    assert col.tolist() == [ends[i + 1] - ends[i] for i in range(len(ends) - 1)]
    # the in-memory writer gives the same bytes
    comp, b2, c2, n = ca.deflate_compress_chunks(np.frombuffer(raw, np.uint8), k, level=9)
    assert comp.tobytes() == exp and b2.tolist() == blk.tolist()
