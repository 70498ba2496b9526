"""Pins the CPU oracle (oracle/) before anything is checked against it.

Every vector here is either hand-derived from the reference decoder text (the
file:line is given), a published ORC / zlib example, or a closed-form round
trip through an independent encoder in tests/pycodecs.py. Plausible decoder
mistakes (a dropped term, a wrong sign or shift, a transposed table) fail at
least one of them.
"""
import random
import zlib

import numpy as np
import pytest

import oracle
from tests import pycodecs as pc

U64 = (1 << 64) - 1


# ---------------------------------------------------------------- RLE v1 KATs
# Hand-derived from include/rlev1/decoder.h:671-791.

def test_v1_run_kat():
    rc, v = oracle.rlev1_decode_chunk(bytes([0x00, 0x01, 0x05]), 8, 3 * 8)
    assert rc == oracle.OK and v.tolist() == [5, 6, 7]


def test_v1_literal_kat():
    rc, v = oracle.rlev1_decode_chunk(bytes([0xFE, 0x81, 0x01, 0x7F]), 8, 2 * 8)
    assert rc == oracle.OK and v.tolist() == [129, 127]


def test_v1_negative_delta_max_run_kat():
    rc, v = oracle.rlev1_decode_chunk(bytes([0x7F, 0xFF, 0x0A]), 8, 130 * 8)
    assert rc == oracle.OK
    assert v.tolist() == [(10 - i) & U64 for i in range(130)]


def test_v1_header_minus_128_kat():
    vals = list(range(1000, 1128))
    stream = bytes([0x80]) + b"".join(pc.uleb(x) for x in vals)
    rc, v = oracle.rlev1_decode_chunk(stream, 8, 128 * 8)
    assert rc == oracle.OK and v.tolist() == vals


def test_v1_width_truncation_kat():
    # run base 300 in a uint8 column: (300 + i*1) mod 256 (decoder.h:649-652 casts)
    stream = bytes([0x00, 0x01]) + pc.uleb(300)
    rc, v = oracle.rlev1_decode_chunk(stream, 1, 3)
    assert rc == oracle.OK and v.tolist() == [44, 45, 46]
    # a 10-byte varint carrying 2^63 in a uint64 column
    rc, v = oracle.rlev1_decode_chunk(bytes([0xFF]) + pc.uleb(1 << 63), 8, 8)
    assert rc == oracle.OK and v.tolist() == [1 << 63]


def test_v1_stops_at_chunk_bytes():
    # decoder.h:671 stops once CHUNK_SIZE bytes are produced; trailing bytes ignored
    stream = bytes([0x00, 0x00, 0x07, 0x00, 0x00, 0x09])
    rc, v = oracle.rlev1_decode_chunk(stream, 4, 3 * 4)
    assert rc == oracle.OK and v.tolist() == [7, 7, 7]


def test_v1_errors():
    rc, _ = oracle.rlev1_decode_chunk(bytes([0x00, 0x01]), 8, 24)  # missing base
    assert rc == oracle.TRUNCATED
    rc, _ = oracle.rlev1_decode_chunk(bytes([0x05, 0x00, 0x01]), 8, 24)  # 8 values into 3
    assert rc == oracle.OVERFLOW
    rc, _ = oracle.rlev1_decode_chunk(b"", 8, 8)
    assert rc == oracle.TRUNCATED


@pytest.mark.parametrize("width", [1, 2, 4, 8])
def test_v1_round_trip(width):
    rng = random.Random(width)
    vals = []
    while len(vals) < 3000:
        if rng.random() < 0.5:
            d = rng.randint(-128, 127)
            b = rng.getrandbits(8 * width)
            vals += [(b + i * d) for i in range(rng.randint(3, 130))]
        else:
            vals += [rng.getrandbits(rng.choice([7, 14, 8 * width])) for _ in
                     range(rng.randint(1, 200))]
    vals = pc.as_np(vals, width)
    stream = pc.rlev1_encode(vals, width)
    rc, out = oracle.rlev1_decode_chunk(stream, width, len(vals) * width)
    assert rc == oracle.OK
    assert np.array_equal(out, vals)


# ---------------------------------------------------------------- RLE v2 KATs
# ORC spec examples, hand-checked against include/rlev2/decoder_new.h:909-1187.

def test_v2_short_repeat_kat():
    rc, v = oracle.rlev2_decode_chunk(bytes.fromhex("0a2710"), 8, 64)
    assert rc == oracle.OK and v.tolist() == [10000] * 5


def test_v2_direct_kat():
    rc, v = oracle.rlev2_decode_chunk(bytes.fromhex("5e035ca1ab1edeadbeef"), 4, 64)
    assert rc == oracle.OK and v.tolist() == [23713, 43806, 57005, 48879]


def test_v2_delta_kat():
    rc, v = oracle.rlev2_decode_chunk(bytes.fromhex("c609020222424246"), 8, 64)
    assert rc == oracle.OK and v.tolist() == [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]


def test_v2_patched_base_kat():
    s = bytes.fromhex("8e132b2107d01e00147028323c46505a646e78828c96a0aab4befce8")
    rc, v = oracle.rlev2_decode_chunk(s, 8, 64)
    assert rc == oracle.OK
    assert v.tolist() == [2030, 2000, 2020, 1000000] + list(range(2040, 2200, 10))


def test_v2_delta_zero_base_delta_subtracts():
    # decoder_new.h:1057: multiplier = base_delta > 0 ? 1 : -1 with unsigned
    # DATA_TYPE -> a zero base_delta makes every packed delta subtract.
    s = pc.v2_delta(100, 0, [1, 2, 3], code=3)
    rc, v = oracle.rlev2_decode_chunk(s, 8, 64)
    assert rc == oracle.OK and v.tolist() == [100, 100, 99, 97, 94]
    # negative base_delta wraps to a large unsigned value -> adds (mod 2^64)
    s = pc.v2_delta(100, -1, [1, 2, 3], code=3)
    rc, v = oracle.rlev2_decode_chunk(s, 8, 64)
    assert v.tolist() == [100, 99, 100, 102, 105]


def test_v2_delta_fixed():
    s = pc.v2_delta(7, 5, [0] * 6, code=None)
    rc, v = oracle.rlev2_decode_chunk(s, 4, 64)
    assert rc == oracle.OK and v.tolist() == [7 + 5 * i for i in range(8)]


def test_v2_uint8_length_wraps():
    # decoder_new.h:1019 stores len in a uint8_t: a length field of 255 -> 256 -> 0.
    rc, v = oracle.rlev2_decode_chunk(bytes([0x40 | (7 << 1), 0xFF]), 8, 600)
    assert rc == oracle.OK and len(v) == 0


def test_v2_svarint_arithmetic_shift():
    # decoder_new.h:993-997: ret is int64_t, so ret >> 1 is arithmetic. The
    # uvarint 2^64-1 decodes to 0 (a logical shift would give INT64_MIN).
    s = bytes([0xC0, 0x02]) + pc.uleb(5) + pc.uleb(U64)
    rc, v = oracle.rlev2_decode_chunk(s, 8, 64)
    assert rc == oracle.OK and v.tolist() == [5, 5, 5]


@pytest.mark.parametrize("code", range(32))
def test_v2_direct_every_width(code):
    bits = pc.DECODE_WIDTHS[code]
    rng = random.Random(code)
    vals = [rng.getrandbits(bits) for _ in range(97)]
    vals[0] = (1 << bits) - 1
    s = pc.v2_direct(vals, code)
    rc, v = oracle.rlev2_decode_chunk(s, 8, 200)
    assert rc == oracle.OK and v.tolist() == vals
    # in a uint32 column values are reduced mod 2^32
    rc, v = oracle.rlev2_decode_chunk(s, 4, 200)
    assert v.tolist() == [x & 0xFFFFFFFF for x in vals]


def test_v2_short_repeat_wide_bytes_truncate():
    s = pc.v2_short_repeat(0x0102030405, 3, 5)
    rc, v = oracle.rlev2_decode_chunk(s, 4, 8)
    assert v.tolist() == [0x02030405] * 3


def test_v2_patched_base_round_trip():
    rng = random.Random(5)
    for trial in range(40):
        n = rng.randint(1, 128)
        code = rng.choice([3, 7, 11, 15])
        fbw = pc.DECODE_WIDTHS[code]
        reduced = [rng.getrandbits(fbw) for _ in range(n)]
        pw_code = rng.choice([0, 3, 7, 11, 15, 20])
        pw = pc.DECODE_WIDTHS[pw_code]
        pgw = rng.randint(1, 8)
        idx = sorted(rng.sample(range(n), min(n, rng.randint(0, 6))))
        patches, prev, full = [], 0, list(reduced)
        for i in idx:
            g = i - prev
            if g >= (1 << pgw):
                continue
            p = rng.getrandbits(pw)
            patches.append((g, p))
            full[i] |= p << fbw
            prev = i
        base = rng.getrandbits(30)
        s = pc.v2_patched_base(reduced, base, code, 4, patches, pw_code, pgw)
        rc, v = oracle.rlev2_decode_chunk(s, 8, 256)
        assert rc == oracle.OK
        assert v.tolist() == [(x + base) & U64 for x in full], trial


def test_v2_errors():
    rc, _ = oracle.rlev2_decode_chunk(bytes.fromhex("5e035ca1"), 8, 64)
    assert rc == oracle.TRUNCATED
    rc, _ = oracle.rlev2_decode_chunk(bytes.fromhex("0a2710"), 8, 3)
    assert rc == oracle.OVERFLOW
    # patch gap pointing past the block
    s = pc.v2_patched_base([1, 2], 0, 3, 1, [(5, 1)], 3, 3)
    rc, _ = oracle.rlev2_decode_chunk(s, 8, 64)
    assert rc == oracle.CORRUPT


def test_v2_width_tables():
    # utils.h:28 and utils.h:42-48, transcribed as data
    assert [oracle.rlev2_decode_width(i) for i in range(32)] == pc.DECODE_WIDTHS
    closest = [1, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22,
               23, 24, 26, 26, 28, 28, 30, 30, 32, 32] + [40] * 8 + [48] * 8 + [56] * 8 + [64] * 8
    assert [oracle.rlev2_closest_width(i) for i in range(65)] == closest


# ---------------------------------------------------------------- zlib / Deflate

def test_zlib_empty_stream_kat():
    rc, out = oracle.zlib_inflate(bytes.fromhex("789c030000000001"), 16)
    assert rc == oracle.OK and len(out) == 0


def test_zlib_stored_block_kat():
    body = b"hello"
    s = bytes([0x78, 0x01, 0x01, 0x05, 0x00, 0xFA, 0xFF]) + body + (0x062C0215).to_bytes(4, "big")
    rc, out = oracle.zlib_inflate(s, 64)
    assert rc == oracle.OK and out.tobytes() == body


def test_checksum_kats():
    assert oracle.adler32(b"Wikipedia") == 0x11E60398
    assert oracle.crc32(b"123456789") == 0xCBF43926
    assert oracle.adler32(b"") == 1


@pytest.mark.parametrize("level,strategy", [(6, zlib.Z_DEFAULT_STRATEGY), (9, zlib.Z_DEFAULT_STRATEGY),
                                            (6, zlib.Z_FIXED), (0, zlib.Z_DEFAULT_STRATEGY),
                                            (1, zlib.Z_HUFFMAN_ONLY)])
def test_zlib_round_trip(level, strategy):
    rng = np.random.default_rng(level)
    data = rng.integers(0, 8, 50000, dtype=np.uint8).tobytes() + bytes(range(256)) * 40
    c = zlib.compressobj(level, zlib.DEFLATED, 15, 8, strategy)
    s = c.compress(data) + c.flush()
    rc, out = oracle.zlib_inflate(s, len(data))
    assert rc == oracle.OK and out.tobytes() == data


def test_zlib_corrupt_and_truncated():
    s = zlib.compress(b"abc" * 1000, 6)
    rc, _ = oracle.zlib_inflate(s[: len(s) // 2], 4000)
    assert rc == oracle.TRUNCATED
    bad = bytearray(s)
    bad[2] |= 0x06  # BTYPE 11 in the first block header
    rc, _ = oracle.zlib_inflate(bytes(bad), 4000)
    assert rc == oracle.CORRUPT


def test_inflate_chunks_container():
    chunk = 4096
    rng = np.random.default_rng(1)
    raw = rng.integers(0, 4, chunk * 5, dtype=np.uint8).tobytes()
    streams = [zlib.compress(raw[i * chunk:(i + 1) * chunk], 6) for i in range(5)]
    padded = [s + b"\0" * ((-len(s)) % 8) for s in streams]
    off = np.cumsum([0] + [len(p) for p in padded]).astype(np.uint64)
    st, out = oracle.inflate_chunks(b"".join(padded), off, chunk, threads=3)
    assert st.tolist() == [0] * 5 and out.tobytes() == raw


def test_v1_overlong_varint_is_corrupt():
    # 11-byte varint: beyond what any column width needs
    stream = bytes([0xFF]) + bytes([0x80] * 10) + bytes([0x01])
    rc, _ = oracle.rlev1_decode_chunk(stream, 8, 8)
    assert rc == oracle.CORRUPT


def test_v2_overlong_varint_is_corrupt():
    # DELTA base varint of 11 bytes
    s = bytes([0xC0, 0x02]) + bytes([0x80] * 10) + bytes([0x01]) + pc.uleb(2)
    rc, _ = oracle.rlev2_decode_chunk(s, 8, 64)
    assert rc == oracle.CORRUPT
    # truncated inside the varint: TRUNCATED wins
    rc, _ = oracle.rlev2_decode_chunk(bytes([0xC0, 0x02]) + bytes([0x80] * 12), 8, 64)
    assert rc == oracle.TRUNCATED


def test_periodic_copy_reciprocal_is_exact():
    # the inflate kernel's periodic match copy takes i // d as
    # (i * (2^16 // d + 1)) >> 16 for i < 512 (match bytes), d < 64 (distance)
    for d in range(1, 64):
        m = 0x10000 // d + 1
        assert all(((i * m) >> 16) == i // d for i in range(512))
