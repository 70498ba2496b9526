#!/usr/bin/env python3
"""Writes tests/golden/kats.json: known-answer vectors for the three decoders.

Every vector is derived by hand from the stream formats (ORC RLE v1/v2 as the
reference decoders read them, RFC 1950/1951 for zlib); nothing here runs
reference code. Sources of each expected value:

  RLE v1 (include/rlev1/decoder.h:671-791):
    header h >= 0: run of h+3 values, signed delta byte, unsigned LEB128 base;
    header h < 0: -h LEB128 literals.
  RLE v2 (include/rlev2/decoder_new.h:909-1187, utils.h:28-59):
    the ORC specification's unsigned examples, re-checked against the
    reference semantics (big-endian SHORT_REPEAT value, DELTA sign rule,
    running-sum PATCHED_BASE gaps).
  zlib: the zlib stream of empty input and a hand-built stored block.

    python3 tests/golden/make_kats.py      # rewrites kats.json
"""
import json
import os
import zlib

M64 = (1 << 64) - 1


def uleb(x: int) -> bytes:
    out = bytearray()
    while True:
        b = x & 0x7F
        x >>= 7
        if x:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def v1_vectors():
    vs = []
    # run of 3 (header 0), delta +1, base 5
    vs.append(dict(name="v1_run", hex="000105", width=8, values=[5, 6, 7]))
    # literal group of 2 (header -2 = 0xFE): 129 = 81 01, 127 = 7F
    vs.append(dict(name="v1_literals", hex="fe81017f", width=8, values=[129, 127]))
    # run of 130 (header 127), delta -1 (0xFF), base 10: 10, 9, ..., -119 mod 2^64
    vs.append(dict(name="v1_max_run_negative_delta", hex="7fff0a", width=8,
                   values=[(10 - i) & M64 for i in range(130)]))
    # header -128 (0x80): 128 literals (never emitted by the reference encoder)
    lits = [(i * 1000003) % (1 << 40) for i in range(128)]
    vs.append(dict(name="v1_header_minus_128", hex=(bytes([0x80]) + b"".join(uleb(x) for x in lits)).hex(),
                   width=8, values=lits))
    # 4-byte elements keep the low 32 bits of a wide literal
    vs.append(dict(name="v1_width4_truncates", hex=(bytes([0xFF]) + uleb(0x1_2345_6789)).hex(),
                   width=4, values=[0x2345_6789]))
    return vs


def v2_vectors():
    vs = []
    # SHORT_REPEAT: W=2 bytes (bits 5..3 = 1), count 3+2 = 5, value 0x2710 = 10000
    vs.append(dict(name="v2_short_repeat", hex="0a2710", width=8, values=[10000] * 5))
    # DIRECT: width code 15 -> 16 bits, length 4, big-endian 16-bit values
    vs.append(dict(name="v2_direct", hex="5e035ca1ab1edeadbeef", width=4,
                   values=[0x5CA1, 0xAB1E, 0xDEAD, 0xBEEF]))
    # DELTA: width code 3 (4 bits), length 10, base 2, delta base 1, then the
    # 4-bit deltas 2,2,4,2,4,2,4,6
    primes = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]
    vs.append(dict(name="v2_delta", hex="c609020222424246", width=8, values=primes))
    # PATCHED_BASE: base 2000 (2 bytes), width 5 bits, 20 values, one patch
    # (gap 3, 12-bit patch value 3898) on top of value index 3.
    pb = [2030, 2000, 2020, 1000000, 2040, 2050, 2060, 2070, 2080, 2090,
          2100, 2110, 2120, 2130, 2140, 2150, 2160, 2170, 2180, 2190]
    vs.append(dict(name="v2_patched_base",
                   hex="8e132b2107d01e00147028323c46505a646e78828c96a0aab4befce8",
                   width=8, values=pb))
    # DELTA with width code 0 (fixed delta): base 7, delta base -2 (zigzag 3),
    # length 4 -> 7, 5, 3, 1 (the delta sign comes from the delta base)
    vs.append(dict(name="v2_delta_fixed_negative", hex="c003" + uleb(7).hex() + "03",
                   width=8, values=[7, 5, 3, 1]))
    return vs


def zlib_vectors():
    vs = []
    vs.append(dict(name="zlib_empty", hex="789c030000000001", out_hex=""))
    body = b"hello"
    stored = bytes([0x78, 0x01, 0x01, 0x05, 0x00, 0xFA, 0xFF]) + body + zlib.adler32(body).to_bytes(4, "big")
    vs.append(dict(name="zlib_stored_block", hex=stored.hex(), out_hex=body.hex()))
    # fixed-Huffman block: literal 'a' (0x61 -> code 0x91, 8 bits), a match of
    # length 3 at distance 1, end of block; written by hand bit by bit below.
    bits = []

    def put(v, n, rev=False):
        # Huffman codes are sent MSB first; other fields LSB first
        for i in range(n):
            bits.append((v >> (n - 1 - i)) & 1 if rev else (v >> i) & 1)

    put(1, 1)          # BFINAL
    put(1, 2)          # BTYPE = fixed
    put(0x30 + 0x61, 8, rev=True)  # literal 0x61: codes 0..143 are 0x30 + sym, 8 bits
    put(1, 7, rev=True)            # length symbol 257 (length 3): 7-bit code 0000001
    put(0, 5, rev=True)            # distance code 0 (distance 1)
    put(0, 7, rev=True)            # end of block (256): 0000000
    while len(bits) % 8:
        bits.append(0)
    data = bytes(sum(bits[i + k] << k for k in range(8)) for i in range(0, len(bits), 8))
    out = b"aaaa"
    stream = bytes([0x78, 0x01]) + data + zlib.adler32(out).to_bytes(4, "big")
    vs.append(dict(name="zlib_fixed_match", hex=stream.hex(), out_hex=out.hex()))
    return vs


def main():
    kats = dict(rlev1=v1_vectors(), rlev2=v2_vectors(), zlib=zlib_vectors())
    # every zlib vector must also be what zlib itself decodes
    for v in kats["zlib"]:
        assert zlib.decompress(bytes.fromhex(v["hex"])).hex() == v["out_hex"], v["name"]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kats.json")
    with open(path, "w") as f:
        json.dump(kats, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
