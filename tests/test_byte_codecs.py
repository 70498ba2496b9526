"""CPU checks of the BRLE and LZSS side: the oracle decoders against vectors
derived by hand from the reference's decoders (include/brle/brle.h:76-90,
include/lzss/lzss.h:264-303), the product's encoders against vectors derived
by hand from the reference's encoders (brle.h:129-230, lzss.h:75-113,
310-419) and against the oracle's statement-by-statement restatements of
them, round trips, and the container layouts (brle.h:281-296,
lzss.h:1100-1119)."""
import numpy as np
import pytest

import codag_amd as ca
import oracle

B = bytes.fromhex


# ------------------------------------------------------------- BRLE decoder

@pytest.mark.parametrize("stream,want", [
    ("0241", b"A" * 5),                  # h <= 127: h + 3 copies of the next byte
    ("8042", b"B"),                      # h > 127: h - 127 literal bytes
    ("814344", b"CD"),
    ("7f45", b"E" * 130),                # longest run
    ("ff" + bytes(range(128)).hex(), bytes(range(128))),  # longest literal group
    ("0046" + "8047", b"FFFG"),
    ("", b""),
])
def test_brle_decode_kats(stream, want):
    st, out = oracle.brle_decode_chunk(B(stream), 1 << 12)
    assert st == oracle.OK
    assert out.tobytes() == want


@pytest.mark.parametrize("stream,cap,st,produced", [
    ("8241", 100, oracle.TRUNCATED, b""),       # three literals announced, one present
    ("05", 100, oracle.TRUNCATED, b""),         # a run without its byte
    ("024103", 100, oracle.TRUNCATED, b"AAAAA"),
    ("0241", 4, oracle.OVERFLOW, b""),          # five bytes do not fit four
    ("0241814243", 6, oracle.OVERFLOW, b"AAAAA"),
    ("02418142", 6, oracle.TRUNCATED, b"AAAAA"),  # the truncation is found first
])
def test_brle_decode_errors(stream, cap, st, produced):
    got, out = oracle.brle_decode_chunk(B(stream), cap)
    assert got == st
    assert out.tobytes() == produced


# ------------------------------------------------------------- LZSS decoder

@pytest.mark.parametrize("stream,want", [
    # flags 0x07: three literals, then v = 0x0060: length 0 + 3, offset 3
    ("07616263" + "6000", b"abcabc"),
    # a literal, then v = 0x0027: length 7 + 3 = 10 from 1 byte back (copied
    # forward one byte at a time, so the match repeats its own output)
    ("0161" + "2700", b"a" * 11),
    # v = 0x005f: length 34 from 2 back
    ("036162" + "5f00", b"ab" * 18),
    # eight literals and the flag byte the encoder reserves after them
    ("ff" + b"abcdefgh".hex() + "00", b"abcdefgh"),
    ("", b""),
])
def test_lzss_decode_kats(stream, want):
    st, out = oracle.lzss_decode_chunk(B(stream), 1 << 12)
    assert st == oracle.OK
    assert out.tobytes() == want


@pytest.mark.parametrize("stream,cap,st,produced", [
    ("002000", 100, oracle.CORRUPT, b""),     # offset 1 with nothing produced
    ("000000", 100, oracle.CORRUPT, b""),     # offset 0
    ("016127", 100, oracle.TRUNCATED, b"a"),  # the reference's second byte is missing
    ("01612700", 5, oracle.OVERFLOW, b"a"),
    ("036162" + "4000", 4, oracle.OVERFLOW, b"ab"),
    ("036162" + "0001", 100, oracle.CORRUPT, b"ab"),  # offset 8 with 2 bytes produced
])
def test_lzss_decode_errors(stream, cap, st, produced):
    got, out = oracle.lzss_decode_chunk(B(stream), cap)
    assert got == st
    assert out.tobytes() == produced


# ------------------------------------------------------------- encoders

def _brle_stream(raw):
    data, ptr, orig = ca.brle_container(ca.brle_encode(raw))
    assert orig == len(raw)
    return data[int(ptr[0]):int(ptr[1])].tobytes() if len(ptr) > 1 else b""


def _lzss_stream(raw):
    data, ends, chunk, orig = ca.lzss_container(ca.lzss_encode(raw))
    assert orig == len(raw)
    return data[: int(ends[0])].tobytes() if len(ends) else b""


# Hand-derived from brle::encode (brle.h:159-226): a third equal byte turns
# the last two literals into a run (the literals before them are written
# first); runs close at 130 bytes, literal groups at 128.
BRLE_ENC = [
    (b"a", "8061"),
    (b"aaa", "0061"),
    (b"aaab" + b"x" * 10 + bytes(range(20)),
     "0061" + "8062" + "0778" + "93" + bytes(range(20)).hex()),
    (b"abccc", "816162" + "0063"),
    (b"a" * 131, "7f61" + "8061"),
    (b"a" * 132, "7f61" + "816161"),
    (b"a" * 133, "7f61" + "0061"),
    (bytes(range(128)) + bytes(range(5)), "ff" + bytes(range(128)).hex() + "84" + bytes(range(5)).hex()),
]


@pytest.mark.parametrize("raw,want", BRLE_ENC)
def test_brle_encoder_kats(raw, want):
    assert _brle_stream(raw).hex() == want
    assert oracle.brle_encode_chunk(raw).hex() == want


# Hand-derived from compress_func / find_match (lzss.h:75-113, 326-417).
# "aaaa": no match can be 3 long before 4 bytes are in the history (the
# history's oldest byte is never a start and a match ends before the current
# position). "a" * 40: then 3 from 3 back, 6 from 6, 12 from 12 and the 15
# bytes left from 24 back (the oldest of the longest); 8 items fill a group
# and the next flag byte is reserved and stays 0.
LZSS_ENC = [
    (b"aaaa", "0f61616161"),
    (b"a" * 40, "0f61616161" + "6000" + "c300" + "8901" + "0c03" + "00"),
    (b"abcabcabcabcabcabc hello hello hello",
     "8f61626361" + "6000" + "c300" + "8201" + "20" + "1f68656c6c6f" + "c300" + "8301"),
]


@pytest.mark.parametrize("raw,want", LZSS_ENC)
def test_lzss_encoder_kats(raw, want):
    assert _lzss_stream(raw).hex() == want
    assert oracle.lzss_encode_chunk(raw).hex() == want


def test_empty_inputs():
    # no chunks (the reference divides by a chunk count of 0 here,
    # lzss.h:1066-1067, and brle.h:252 underflows)
    data, ends, chunk, orig = ca.lzss_container(ca.lzss_encode(b""))
    assert (len(data), len(ends), orig) == (0, 0, 0)
    data, ptr, orig = ca.brle_container(ca.brle_encode(b""))
    assert (len(data), list(ptr), orig) == (0, [0], 0)


def _inputs(seed, n):
    rng = np.random.default_rng(seed)
    text = ca.synth_byte_payload(n, 1, seed=seed)
    runs = ca.synth_byte_payload(n, 0, seed=seed)
    rand = rng.integers(0, 256, n, dtype=np.uint8)
    small = rng.integers(0, 3, n, dtype=np.uint8)          # many ties between candidates
    period = np.tile(rng.integers(0, 256, 7, dtype=np.uint8), n // 7 + 1)[:n]
    return [text, runs, rand, small, period]


@pytest.mark.parametrize("seed", [1, 2])
def test_lzss_encoder_matches_reference_restatement(seed):
    # the product's hash-chain search against the oracle's exhaustive
    # oldest-first scan over the 2048-byte ring, chunks of 5000 bytes (the
    # history wraps within a chunk and restarts at each)
    for raw in _inputs(seed, 14000):
        cont = ca.lzss_encode(raw, chunk_bytes=5000)
        data, ends, chunk, orig = ca.lzss_container(cont)
        assert (chunk, orig, len(ends)) == (5000, 14000, 3)
        lo = 0
        for i, e in enumerate(ends):
            want = oracle.lzss_encode_chunk(raw[i * 5000:(i + 1) * 5000])
            assert data[lo:int(e)].tobytes() == want, (i, raw[:16])
            lo = int(e)


@pytest.mark.parametrize("seed", [1, 2])
def test_brle_encoder_matches_reference_restatement(seed):
    for raw in _inputs(seed, 20000):
        cont = ca.brle_encode(raw, chunk_bytes=6000)
        data, ptr, orig = ca.brle_container(cont)
        assert orig == 20000 and len(ptr) == 5
        for i in range(4):
            want = oracle.brle_encode_chunk(raw[i * 6000:(i + 1) * 6000])
            assert data[int(ptr[i]):int(ptr[i + 1])].tobytes() == want


def test_brle_container_layout_and_round_trip():
    raw = ca.synth_byte_payload(3 * ca.BRLE_CHUNK_BYTES + 12345, 0, seed=5)
    cont = ca.brle_encode(raw)
    n_ptr = int(cont[:4].view(np.uint32)[0])
    assert n_ptr == 5
    data, ptr, orig = ca.brle_container(cont)
    assert ptr[0] == 0 and np.all(np.diff(ptr.astype(np.int64)) > 0)
    assert int(ptr[-1]) == len(data) and orig == len(raw)
    st, out_len, out = oracle.brle_decode_chunks(data, ptr, ca.BRLE_CHUNK_BYTES)
    assert np.all(st == 0)
    assert list(out_len) == [ca.BRLE_CHUNK_BYTES] * 3 + [12345]
    assert np.array_equal(out[:orig], raw)


def test_lzss_container_layout_and_round_trip():
    raw = ca.synth_byte_payload(2 * ca.LZSS_CHUNK_BYTES + 777, 1, seed=6)
    cont = ca.lzss_encode(raw)
    hdr = cont[8:28].view(np.uint32)
    assert list(hdr) == [ca.LZSS_CHUNK_BYTES, 3, 2048, 3, 34]
    data, ends, chunk, orig = ca.lzss_container(cont)
    assert int(ends[-1]) == len(data) and orig == len(raw)
    st, out_len, out = oracle.lzss_decode_chunks(data, ends, chunk)
    assert np.all(st == 0)
    assert list(out_len) == [chunk, chunk, 777]
    assert np.array_equal(out[:orig], raw)


def test_lzss_matches_are_longest_and_oldest():
    # property of the encoder's choice, checked by brute force over the
    # stream: every match is the longest available and the oldest of those;
    # every literal has no 3-byte match available
    raw = ca.synth_byte_payload(6000, 1, seed=9)
    s = _lzss_stream(raw)
    r = raw.tobytes()
    i = cur = 0
    while i < len(s):
        flags = s[i]
        i += 1
        for k in range(8):
            if i >= len(s) or cur >= len(r):
                break
            lo = max(cur - 2047, 1)
            cands = []
            for p in range(lo, cur):
                lim = min(34, len(r) - cur, cur - p)
                n = 0
                while n < lim and r[p + n] == r[cur + n]:
                    n += 1
                cands.append((n, -p))
            best = max(cands) if cands else (0, 0)
            if flags >> k & 1:
                assert best[0] < 3
                cur += 1
                i += 1
            else:
                v = s[i] | s[i + 1] << 8
                i += 2
                length, off = (v & 31) + 3, v >> 5
                assert (length, cur - off) == (best[0], -best[1])
                cur += length
    assert cur == len(r)
