"""GPU parity of the BRLE and LZSS decoders (codag_brle_decode,
codag_lzss_decode) against the oracle's restatements of
include/brle/brle.h:76-90 and include/lzss/lzss.h:264-303: byte-identical
output, the same byte counts and the same status codes, on the reference
encoders' streams, on hand-built and random streams that reach every batch
limit of the kernels, and on broken streams."""
import numpy as np
import pytest

import codag_amd as ca
import oracle
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu_required):
    c = ca.Context(0)
    yield c
    c.close()


def run(ctx, lzss, data, idx, chunk, shift=0, comp_bytes=None):
    """Decode on the GPU into an output that starts `shift` bytes past a
    16-byte boundary. Returns (status, out_len, out)."""
    data = np.ascontiguousarray(data, dtype=np.uint8)
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    n = len(idx) if lzss else len(idx) - 1
    d_comp = ctx.upload(data if len(data) else np.zeros(1, np.uint8))
    d_idx = ctx.upload(idx if len(idx) else np.zeros(1, np.uint64))
    d_out = ctx.alloc(n * chunk + 32)
    d_out.fill(0xA5)
    d_len = ctx.alloc(8 * max(n, 1))
    d_st = ctx.alloc(4 * max(n, 1))
    d_st.fill(0x7F)
    fn = ctx.lzss_decode if lzss else ctx.brle_decode
    fn(d_comp, len(data) if comp_bytes is None else comp_bytes, d_idx, n, chunk,
       d_out.offset(shift), d_len, d_st)
    ctx.sync()
    out = d_out.download(n * chunk + 32)[shift: shift + n * chunk]
    return d_st.download(4 * n, np.int32), d_len.download(8 * n, np.uint64), out


def compare(ctx, lzss, data, idx, chunk, shift=0):
    st, ln, out = run(ctx, lzss, data, idx, chunk, shift)
    fn = oracle.lzss_decode_chunks if lzss else oracle.brle_decode_chunks
    ost, oln, oout = fn(data, idx, chunk)
    assert st.tolist() == ost.tolist()
    assert ln.tolist() == oln.tolist()
    for i in range(len(st)):
        a = out[i * chunk: i * chunk + int(oln[i])]
        b = oout[i * chunk: i * chunk + int(oln[i])]
        if not np.array_equal(a, b):
            bad = int(np.flatnonzero(a != b)[0])
            raise AssertionError(f"chunk {i}: first difference at byte {bad} of {int(oln[i])}")
    return st, ln


def brle_container(streams):
    ptr = np.cumsum([0] + [len(s) for s in streams]).astype(np.uint64)
    return np.frombuffer(b"".join(streams) or b"\0", np.uint8)[: int(ptr[-1])], ptr


def lzss_container(streams):
    ends = np.cumsum([len(s) for s in streams]).astype(np.uint64)
    return np.frombuffer(b"".join(streams) or b"\0", np.uint8)[: int(ends[-1]) if len(ends) else 0], ends


# ----------------------------------------------------------------- KATs

B = bytes.fromhex
BRLE_KATS = [("0241", 64), ("8042", 64), ("814344", 64), ("7f45", 256),
             ("ff" + bytes(range(128)).hex(), 256), ("00468047", 64), ("", 64),
             ("8241", 64), ("05", 64), ("024103", 64), ("0241", 4), ("0241814243", 6),
             ("02418142", 6)]
LZSS_KATS = [("076162636000", 64), ("01612700", 64), ("0361625f00", 64),
             ("ff" + b"abcdefgh".hex() + "00", 64), ("", 64), ("002000", 64), ("000000", 64),
             ("016127", 64), ("01612700", 5), ("0361624000", 4), ("0361620001", 64)]


@pytest.mark.parametrize("shift", [0, 9])
def test_kats(ctx, shift):
    say(f"byte codec KATs, output shift {shift}")
    for s, cap in BRLE_KATS:
        data, ptr = brle_container([B(s)])
        compare(ctx, False, data, ptr, cap, shift)
    for s, cap in LZSS_KATS:
        data, ends = lzss_container([B(s)])
        compare(ctx, True, data, ends, cap, shift)


# ------------------------------------------------- reference encoders' streams

@pytest.mark.parametrize("shift", [0, 1, 15])
def test_brle_reference_streams(ctx, shift):
    say(f"brle reference streams, shift {shift}")
    raw = ca.synth_byte_payload(5 * ca.BRLE_CHUNK_BYTES + 4321, 0, seed=11 + shift)
    data, ptr, orig = ca.brle_container(ca.brle_encode(raw))
    st, ln = compare(ctx, False, data, ptr, ca.BRLE_CHUNK_BYTES, shift)
    assert np.all(st == 0) and int(ln.sum()) == len(raw)


@pytest.mark.parametrize("shift", [0, 1, 15])
def test_lzss_reference_streams(ctx, shift):
    say(f"lzss reference streams, shift {shift}")
    parts = [ca.synth_byte_payload(3 * ca.LZSS_CHUNK_BYTES, 1, seed=12 + shift),
             ca.synth_byte_payload(2 * ca.LZSS_CHUNK_BYTES, 0, seed=13),
             np.tile(np.arange(7, dtype=np.uint8), 20000)[: ca.LZSS_CHUNK_BYTES],
             np.zeros(ca.LZSS_CHUNK_BYTES - 999, np.uint8)]
    raw = np.concatenate(parts)
    data, ends, chunk, orig = ca.lzss_container(ca.lzss_encode(raw))
    st, ln = compare(ctx, True, data, ends, chunk, shift)
    assert np.all(st == 0) and int(ln.sum()) == len(raw)


def test_host_wrappers_round_trip(ctx):
    say("brle / lzss host wrappers")
    raw = ca.synth_byte_payload(3 * ca.BRLE_CHUNK_BYTES + 17, 0, seed=21)
    out, ks = ca.brle.decompress_gpu(ctx, ca.brle.compress_gpu(raw))
    assert np.array_equal(out, raw) and ks > 0
    raw = ca.synth_byte_payload(4 * ca.LZSS_CHUNK_BYTES + 17, 1, seed=22)
    out, ks = ca.lzss.decompress_gpu_base(ctx, ca.lzss.compress_gpu(raw))
    assert np.array_equal(out, raw) and ks > 0


# ------------------------------------------------------ random valid streams

def random_brle_stream(rng, out_bytes, mix):
    """Items until out_bytes: runs of 3..130, literal groups of 1..128;
    mix = probability of a run. Short items stress the 64-item batch limit,
    long literal groups the header window."""
    s = bytearray()
    made = 0
    while made < out_bytes:
        if rng.random() < mix:
            k = int(rng.integers(3, 131)) if rng.random() < 0.5 else 3
            k = min(k, max(out_bytes - made, 3))
            s += bytes([k - 3, int(rng.integers(0, 256))])
        else:
            k = int(rng.integers(1, 129)) if rng.random() < 0.7 else 128
            s += bytes([k + 127]) + rng.integers(0, 256, k, dtype=np.uint8).tobytes()
        made += k
    return bytes(s), made


def random_lzss_stream(rng, out_bytes, near):
    """Items until out_bytes: literals and matches of 3..34 from 1..2047 back
    (never before the chunk start); `near` = probability that a match comes
    from at most 40 bytes back, so it overlaps itself or reads the batch's
    own output."""
    s = bytearray()
    made = 0
    while made < out_bytes:
        flag_at = len(s)
        s.append(0)
        flags = 0
        for k in range(8):
            if made >= out_bytes:
                break
            if made < 3 or rng.random() < 0.35:
                flags |= 1 << k
                s.append(int(rng.integers(0, 256)))
                made += 1
            else:
                ln = int(rng.integers(3, 35))
                hi = min(made, 40 if rng.random() < near else 2047)
                off = int(rng.integers(1, hi + 1))
                v = (off << 5) | (ln - 3)
                s += bytes([v & 255, v >> 8])
                made += ln
        s[flag_at] = flags
    return bytes(s), made


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_brle_random_streams(ctx, seed):
    say(f"brle random streams seed {seed}")
    rng = np.random.default_rng(seed)
    chunk = 65536
    streams = []
    for mix in (0.0, 0.2, 0.5, 0.9, 1.0):
        s, made = random_brle_stream(rng, chunk - int(rng.integers(0, 300)), mix)
        streams.append(s)
    # one stream whose output runs past the chunk (OVERFLOW after the last
    # item that fits) and one cut inside an item (TRUNCATED)
    s, _ = random_brle_stream(rng, chunk + 500, 0.5)
    streams.append(s)
    s, _ = random_brle_stream(rng, 3000, 0.3)
    streams.append(s[:-1])
    streams.append(b"")
    data, ptr = brle_container(streams)
    st, ln = compare(ctx, False, data, ptr, chunk, shift=seed)
    assert st.tolist()[-3:] == [oracle.OVERFLOW, oracle.TRUNCATED, oracle.OK]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_lzss_random_streams(ctx, seed):
    say(f"lzss random streams seed {seed}")
    rng = np.random.default_rng(seed)
    chunk = 65536
    streams = []
    for near in (0.0, 0.3, 0.8, 1.0):
        s, made = random_lzss_stream(rng, chunk - int(rng.integers(0, 300)), near)
        streams.append(s)
    s, _ = random_lzss_stream(rng, chunk + 500, 0.3)
    streams.append(s)                                   # OVERFLOW
    s, _ = random_lzss_stream(rng, 3000, 0.3)
    streams.append(s[:-1])                              # cut: TRUNCATED or a shorter stream
    bad = bytearray(random_lzss_stream(rng, 5000, 0.3)[0])
    bad[1000:1010] = b"\x00" * 10                       # a flag byte of 0 and offsets of 0
    streams.append(bytes(bad))
    streams.append(rng.integers(0, 256, 4000, dtype=np.uint8).tobytes())  # noise
    data, ends = lzss_container(streams)
    st, ln = compare(ctx, True, data, ends, chunk, shift=seed)
    assert st.tolist()[4] == oracle.OVERFLOW
    assert oracle.CORRUPT in st.tolist()[6:]


def test_chunks_outside_the_input_are_corrupt(ctx):
    say("byte codec chunk bounds")
    data, ptr = brle_container([B("0241"), B("0342")])
    st, ln, _ = run(ctx, False, data, ptr, 64, comp_bytes=3)
    assert st.tolist() == [oracle.OK, oracle.CORRUPT] and ln.tolist() == [5, 0]
    st, ln, _ = run(ctx, False, data, np.array([2, 0, 4], np.uint64), 64)
    assert st.tolist() == [oracle.CORRUPT, oracle.OK]
    data, ends = lzss_container([B("076162636000"), B("0161")])
    st, ln, _ = run(ctx, True, data, ends, 64, comp_bytes=7)
    assert st.tolist() == [oracle.OK, oracle.CORRUPT]


def test_full_chunks_round_trip(ctx):
    # 64 MiB of each through the device entry points: every chunk full,
    # output equal to the encoder's input
    say("brle / lzss 64 MiB round trip")
    raw = ca.synth_byte_payload(64 << 20, 0, seed=31)
    data, ptr, orig = ca.brle_container(ca.brle_encode(raw))
    st, ln, out = run(ctx, False, data, ptr, ca.BRLE_CHUNK_BYTES)
    assert np.all(st == 0) and np.all(ln == ca.BRLE_CHUNK_BYTES)
    assert np.array_equal(out, raw)
    raw = ca.synth_byte_payload(64 << 20, 1, seed=32)
    data, ends, chunk, orig = ca.lzss_container(ca.lzss_encode(raw))
    st, ln, out = run(ctx, True, data, ends, chunk)
    assert np.all(st == 0) and np.all(ln == chunk)
    assert np.array_equal(out, raw)
