"""GPU parity of the Deflate decoder (codag_inflate_chunks / codag_inflate_batch)
against zlib's inflate (the oracle): byte-identical output, same byte counts,
same status codes, and the Adler-32 the kernel computes equals zlib's."""
import ctypes
import zlib

import numpy as np
import pytest

import codag_amd as ca
import oracle
from tests.gpu_helpers import inflate, say

pytestmark = pytest.mark.gpu
CH = 131072


@pytest.fixture(scope="module")
def ctx(gpu_required):
    c = ca.Context(0)
    yield c
    c.close()


def container(chunks):
    padded = [s + b"\0" * ((-len(s)) % 8) for s in chunks]
    off = np.cumsum([0] + [len(p) for p in padded]).astype(np.uint64)
    return np.frombuffer(b"".join(padded) or b"\0", np.uint8), off


def check(ctx, streams, cap, flags=0):
    comp, off = container(streams)
    st, ln, out = inflate(ctx, comp, off, cap, flags)
    for i in range(len(streams)):
        # the kernel sees the 8-byte-padded stream, so zlib gets the same bytes
        s = comp[int(off[i]): int(off[i + 1])].tobytes()
        rc, ref = oracle.zlib_inflate(s, cap)
        assert st[i] == rc, (i, st[i], rc)
        if rc == 0:
            assert int(ln[i]) == len(ref)
            assert out[i * cap: i * cap + len(ref)].tobytes() == ref.tobytes(), i
    return st


def z(data, level=6, strategy=zlib.Z_DEFAULT_STRATEGY, wbits=15):
    c = zlib.compressobj(level, zlib.DEFLATED, wbits, 8, strategy)
    return c.compress(data) + c.flush()


def test_kats(ctx):
    say("inflate KATs")
    empty = bytes.fromhex("789c030000000001")
    stored = bytes([0x78, 0x01, 0x01, 0x05, 0x00, 0xFA, 0xFF]) + b"hello" + \
        (0x062C0215).to_bytes(4, "big")
    st = check(ctx, [empty, stored], 64)
    assert st.tolist() == [0, 0]


@pytest.mark.parametrize("level,strategy", [(0, 0), (1, 0), (6, 0), (9, 0), (6, zlib.Z_FIXED),
                                            (6, zlib.Z_HUFFMAN_ONLY), (6, zlib.Z_RLE),
                                            (6, zlib.Z_FILTERED)])
def test_levels_and_strategies(ctx, level, strategy):
    say(f"inflate level={level} strategy={strategy}")
    payload = ca.synth_deflate_payload(8 * CH, CH, seed=level * 10 + strategy)
    rng = np.random.default_rng(level)
    parts = [payload[i * CH:(i + 1) * CH].tobytes() for i in range(8)]
    parts.append(rng.integers(0, 256, CH, dtype=np.uint8).tobytes())  # incompressible
    parts.append(b"\0" * CH)                                           # one long run
    parts.append(bytes(range(256)) * (CH // 256))                     # period-256 copies
    parts.append(b"ab" * (CH // 2))                                    # dist-2 overlap
    streams = [z(p, level, strategy) for p in parts]
    st = check(ctx, streams, CH)
    assert np.all(st == 0)


def test_short_and_odd_sizes(ctx):
    say("inflate odd sizes")
    rng = np.random.default_rng(3)
    streams = []
    for n in [0, 1, 2, 3, 7, 63, 64, 65, 255, 256, 1000, 4095, 4097, 70000]:
        streams.append(z(rng.integers(0, 6, n, dtype=np.uint8).tobytes(), 6))
    check(ctx, streams, CH)


def test_unaligned_stream_starts(ctx):
    # streams packed without 8-byte padding start at arbitrary byte offsets
    say("inflate unaligned")
    rng = np.random.default_rng(4)
    streams = [z(rng.integers(0, 20, int(rng.integers(100, 30000)), dtype=np.uint8).tobytes(), 6)
               for _ in range(40)]
    off = np.cumsum([0] + [len(s) for s in streams]).astype(np.uint64)
    comp = np.frombuffer(b"".join(streams), np.uint8)
    st, ln, out = inflate(ctx, comp, off, CH)
    for i, s in enumerate(streams):
        ref = zlib.decompress(s)
        assert st[i] == 0 and int(ln[i]) == len(ref)
        assert out[i * CH: i * CH + len(ref)].tobytes() == ref


def test_raw_and_skip_adler_flags(ctx):
    say("inflate flags")
    data = ca.synth_deflate_payload(CH, CH, seed=5).tobytes()
    raw = z(data, 6, wbits=-15)
    comp, off = container([raw])
    st, ln, out = inflate(ctx, comp, off, CH, ca.INFLATE_RAW)
    assert st.tolist() == [0] and out[:len(data)].tobytes() == data
    s = bytearray(z(data, 6))
    s[-1] ^= 0xFF  # wrong Adler-32
    comp, off = container([bytes(s)])
    st, _, _ = inflate(ctx, comp, off, CH, 0)
    assert st.tolist() == [3]
    st, _, out = inflate(ctx, comp, off, CH, ca.INFLATE_SKIP_ADLER)
    assert st.tolist() == [0] and out[:len(data)].tobytes() == data


def test_corrupt_and_truncated(ctx):
    say("inflate corrupt/truncated")
    data = ca.synth_deflate_payload(CH, CH, seed=6).tobytes()
    good = z(data, 6)
    cases = [good[: len(good) // 2], good[:1], good[:-2]]
    bad = bytearray(good)
    bad[2] |= 0x06  # BTYPE 11
    cases.append(bytes(bad))
    hdr = bytearray(good)
    hdr[1] ^= 1  # header check
    cases.append(bytes(hdr))
    st = check(ctx, cases, CH)
    assert np.all(st != 0)


def test_random_corruption_matches_zlib(ctx):
    # flip bytes inside dynamic-block headers / data: status must match zlib's
    say("inflate fuzz")
    rng = np.random.default_rng(7)
    data = ca.synth_deflate_payload(CH, CH, seed=7).tobytes()[:20000]
    good = z(data, 6)
    streams = []
    for t in range(300):
        s = bytearray(good)
        for _ in range(int(rng.integers(1, 4))):
            s[int(rng.integers(2, min(len(s), 200)))] ^= int(rng.integers(1, 256))
        streams.append(bytes(s))
    comp, off = container(streams)
    st, ln, out = inflate(ctx, comp, off, CH)
    agree = 0
    for i, s in enumerate(streams):
        rc, ref = oracle.zlib_inflate(s, CH)
        if rc == 0:
            assert st[i] == 0 and out[i * CH: i * CH + len(ref)].tobytes() == ref.tobytes()
        else:
            assert st[i] != 0, i
        agree += int(st[i] == rc)
    assert agree >= 0.9 * len(streams)


def test_symbol_window_edges(ctx):
    # the window decoder hands these to the symbol loop: codes longer than the
    # 10-bit literal/length and 8-bit distance LUTs, end-of-block at every
    # position of a window (frequent flushes), and batches cut by the 64-token
    # and 512-byte limits at every token of a window
    say("inflate window edges")
    rng = np.random.default_rng(11)
    parts = []
    # skewed bytes: Huffman-only codes up to 15 bits
    p = 0.5 ** np.arange(1, 257)
    parts.append(z(rng.choice(256, CH, p=p / p.sum()).astype(np.uint8).tobytes(), 6, zlib.Z_HUFFMAN_ONLY))
    # random far copies: many distance codes, some longer than 8 bits
    base = rng.integers(0, 256, 4096, dtype=np.uint8).tobytes()
    buf = bytearray(base)
    while len(buf) < CH:
        n = int(rng.integers(3, 259))
        d = int(rng.integers(1, min(len(buf), 32768)))
        buf += buf[len(buf) - d: len(buf) - d + n] if d >= n else (buf[-d:] * (n // d + 1))[:n]
        buf += rng.integers(0, 256, int(rng.integers(0, 4)), dtype=np.uint8).tobytes()
    parts.append(z(bytes(buf[:CH]), 9))
    # many small blocks: end-of-block and empty stored blocks (sync flushes)
    for step in (37, 200, 1001):
        c = zlib.compressobj(6, zlib.DEFLATED, 15)
        data = ca.synth_deflate_payload(CH // 4, CH, seed=step).tobytes()
        out = b""
        for i in range(0, len(data), step):
            out += c.compress(data[i:i + step]) + c.flush(zlib.Z_SYNC_FLUSH if i % 3 else zlib.Z_FULL_FLUSH)
        parts.append(out + c.flush())
    # long matches between literals: the 512-byte batch cut at varied tokens
    for k in range(4):
        s = bytearray()
        while len(s) < CH:
            s += rng.integers(0, 256, int(rng.integers(1, 6 + 8 * k)), dtype=np.uint8).tobytes()
            s += bytes([int(rng.integers(0, 256))]) * int(rng.integers(3, 300))
        parts.append(z(bytes(s[:CH]), 6))
    st = check(ctx, parts, CH)
    assert np.all(st == 0)


def test_corruption_inside_blocks_matches_zlib(ctx):
    # flipped bytes anywhere in the stream, so the window decoder meets invalid
    # and truncated codes mid-block
    say("inflate fuzz inside blocks")
    rng = np.random.default_rng(12)
    data = ca.synth_deflate_payload(CH, CH, seed=12).tobytes()[:60000]
    good = z(data, 6)
    streams = []
    for t in range(400):
        s = bytearray(good)
        for _ in range(int(rng.integers(1, 3))):
            s[int(rng.integers(2, len(s)))] ^= int(rng.integers(1, 256))
        streams.append(bytes(s))
    comp, off = container(streams)
    st, ln, out = inflate(ctx, comp, off, CH)
    agree = 0
    for i, s in enumerate(streams):
        rc, ref = oracle.zlib_inflate(s, CH)
        if rc == 0:
            assert st[i] == 0 and out[i * CH: i * CH + len(ref)].tobytes() == ref.tobytes()
        else:
            assert st[i] != 0, i
        agree += int(st[i] == rc)
    assert agree >= 0.9 * len(streams)


def test_chunked_container_parity(ctx):
    say("inflate container 64 chunks")
    raw = ca.synth_deflate_payload(64 * CH - 999, CH, seed=8)
    comp, blk, cl, n = ca.deflate_compress_chunks(raw, CH, level=6, threads=8)
    st, ln, out = inflate(ctx, comp, blk[1:], CH)
    assert np.all(st == 0) and np.all(ln == CH)
    assert out[: len(raw)].tobytes() == raw.tobytes()
    assert oracle.crc32(out[: len(raw)]) == oracle.crc32(raw)


def test_host_wrapper_reference_shape(ctx):
    say("inflate host wrapper")
    raw = ca.synth_deflate_payload(5 * CH + 17, CH, seed=9)
    comp, blk, cl, n = ca.deflate.compress_chunks(raw, CH, level=9)
    out, secs = ca.deflate.decompress_gpu(ctx, comp, blk, CH)
    assert out.tobytes() == raw.tobytes() and secs > 0


class Batch:
    """Device-resident batch: streams, descriptors, outputs and results are
    uploaded/allocated once, so repeated launches use the same addresses."""

    def __init__(self, ctx, streams, caps):
        self.ctx, self.n = ctx, len(streams)
        src = np.frombuffer(b"".join(streams) or b"\0", np.uint8)
        soff = np.cumsum([0] + [len(s) for s in streams]).astype(np.uint64)
        self.doff = np.cumsum([0] + list(caps)).astype(np.uint64)
        self.d_src = ctx.upload(src)
        self.d_dst = ctx.alloc(int(self.doff[-1]) or 1)
        self.descs = (ca.StreamDesc * self.n)()
        for i in range(self.n):
            self.descs[i].src = self.d_src.offset(int(soff[i]))
            self.descs[i].src_bytes = len(streams[i])
            self.descs[i].dst = self.d_dst.offset(int(self.doff[i]))
            self.descs[i].dst_cap = int(caps[i])
        self.d_desc = ctx.upload(np.frombuffer(bytes(self.descs), np.uint8))
        self.d_res = ctx.alloc(16 * max(self.n, 1))

    def launch(self, flags=0, stream=None, clear=True):
        if clear:
            self.d_dst.fill(0)
            self.d_res.fill(0x7F)
        self.ctx.inflate_batch(self.d_desc, self.n, self.d_res, flags, stream)

    def results(self):
        res = self.d_res.download(16 * self.n).view(np.uint8).reshape(self.n, 16)
        written = res[:, :8].copy().view(np.uint64).reshape(-1)
        status = res[:, 8:12].copy().view(np.int32).reshape(-1)
        adler = res[:, 12:16].copy().view(np.uint32).reshape(-1)
        return status, written, adler, self.d_dst.download(), self.doff


def _batch(ctx, streams, caps, flags=0):
    b = Batch(ctx, streams, caps)
    b.launch(flags)
    ctx.sync()
    return b.results()


def _uneven(seed, count, hi, big=()):
    """`count` zlib streams log-uniform in 4 KiB..hi plus streams of the sizes
    in `big`, with the fixed/dynamic/stored block mix 20/75/5."""
    rng = np.random.default_rng(seed)
    sizes = list(np.exp(rng.uniform(np.log(4096), np.log(hi), count)).astype(int)) + list(big)
    payload = ca.synth_deflate_payload(int(sum(sizes)), CH, seed=seed)
    streams, raws, pos = [], [], 0
    for i, n in enumerate(sizes):
        r = payload[pos: pos + n].tobytes()
        pos += n
        raws.append(r)
        s, counts = ca.deflate_compress_mixed(np.frombuffer(r, np.uint8), 20, 5, 32768, seed=i)
        streams.append(s.tobytes())
    return streams, raws


def _check_batch(res, raws):
    status, written, adler, dst, doff = res
    for i, r in enumerate(raws):
        assert status[i] == 0 and int(written[i]) == len(r), (i, status[i], written[i], len(r))
        assert dst[int(doff[i]): int(doff[i]) + len(r)].tobytes() == r, i
        assert int(adler[i]) == zlib.adler32(r), i


def test_batch_uneven_streams(ctx):
    say("inflate batch uneven, streams up to 4 MiB")
    # two 4 MiB streams (the largest size of the uneven-batch workload) and
    # 60 smaller ones, largest first as the work queue prefers
    streams, raws = _uneven(10, 60, 600000, big=(4 << 20, (4 << 20) - 12345))
    order = sorted(range(len(raws)), key=lambda i: -len(raws[i]))
    streams = [streams[i] for i in order]
    raws = [raws[i] for i in order]
    _check_batch(_batch(ctx, streams, [len(r) for r in raws]), raws)


def test_batch_graph_replay(ctx):
    say("inflate batch graph replay on fixed buffers")
    streams, raws = _uneven(12, 40, 300000)
    b = Batch(ctx, streams, [len(r) for r in raws])
    for rep in range(3):  # the same (descs, n, results) replays the captured graph
        b.launch()
        ctx.sync()
        _check_batch(b.results(), raws)
    # new descriptor contents at the same address: the replay reads them
    rev = list(range(b.n))[::-1]
    moved = (ca.StreamDesc * b.n)()
    for k, i in enumerate(rev):
        moved[k] = b.descs[i]
    ca.check(ca.lib().codag_copy_h2d(ctx.h, b.d_desc.ptr, ctypes.cast(moved, ctypes.c_void_p),
                                     ctypes.sizeof(moved)), "copy_h2d")
    b.launch()
    ctx.sync()
    status, written, adler, dst, doff = b.results()
    for k, i in enumerate(rev):  # result k now describes stream i
        assert status[k] == 0 and int(written[k]) == len(raws[i])
        assert int(adler[k]) == zlib.adler32(raws[i])
        assert dst[int(doff[i]): int(doff[i]) + len(raws[i])].tobytes() == raws[i]


def test_batch_two_streams_one_context(ctx):
    say("inflate batch on two HIP streams of one context")
    sa, ra = _uneven(13, 50, 400000)
    sb, rb = _uneven(14, 70, 200000)
    a = Batch(ctx, sa, [len(r) for r in ra])
    b = Batch(ctx, sb, [len(r) for r in rb])
    s1, s2 = ctx.stream_create(), ctx.stream_create()
    try:
        for rep in range(2):
            for x in (a, b):
                x.d_dst.fill(0)
                x.d_res.fill(0x7F)
            ctx.sync()
            a.launch(stream=s1, clear=False)
            b.launch(stream=s2, clear=False)
            ctx.stream_sync(s1)
            ctx.stream_sync(s2)
            _check_batch(a.results(), ra)
            _check_batch(b.results(), rb)
    finally:
        ctx.stream_destroy(s1)
        ctx.stream_destroy(s2)


def test_batch_errors_and_overflow(ctx):
    say("inflate batch errors")
    data = ca.synth_deflate_payload(50000, CH, seed=11).tobytes()
    good = z(data, 6)
    streams = [good, good[:100], good, b"", good]
    caps = [len(data), len(data), 1000, 16, len(data)]
    status, written, adler, dst, doff = _batch(ctx, streams, caps)
    exp = [oracle.zlib_inflate(s, c)[0] for s, c in zip(streams, caps)]
    assert status.tolist() == exp


def gz(data, level=6, name=None, comment=None, extra=None, hcrc=False, mtime=1234567):
    """A gzip member (RFC 1952) with the optional header fields."""
    flg = (0x08 if name is not None else 0) | (0x10 if comment is not None else 0) | \
        (0x04 if extra is not None else 0) | (0x02 if hcrc else 0)
    h = bytes([0x1F, 0x8B, 8, flg]) + mtime.to_bytes(4, "little") + bytes([0, 3])
    if extra is not None:
        h += len(extra).to_bytes(2, "little") + extra
    if name is not None:
        h += name + b"\0"
    if comment is not None:
        h += comment + b"\0"
    if hcrc:
        h += (zlib.crc32(h) & 0xFFFF).to_bytes(2, "little")
    body = z(data, level, wbits=-15)
    return h + body + zlib.crc32(data).to_bytes(4, "little") + \
        (len(data) & 0xFFFFFFFF).to_bytes(4, "little")


def _gz_cases():
    data = ca.synth_deflate_payload(300000, CH, seed=77).tobytes()
    good = gz(data, name=b"col.bin", comment=b"x" * 150, extra=b"ab" * 40, hcrc=True)
    bad_hcrc = bytearray(good)
    bad_hcrc[10 + 2 + 80 + 8 + 151] ^= 1  # the FHCRC bytes follow the comment's terminator
    streams = [gz(data), gz(data[:1000], 9), gz(b""), good, bytes(bad_hcrc),
               good[:5], good[:-3],                       # truncated header / trailer
               good[:3] + bytes([good[3] | 0x20]) + good[4:],  # reserved flag bit
               good[:-8] + (zlib.crc32(data) ^ 1).to_bytes(4, "little") + good[-4:],  # CRC
               good[:-4] + (len(data) + 1).to_bytes(4, "little"),                     # ISIZE
               bytes([0x1F, 0x8C]) + good[2:]]             # magic
    # the stream cut inside its trailer gets one spare output byte: with the
    # output exactly full, zlib cannot tell a short input from a full buffer
    caps = [len(data), 1000, 16, len(data), len(data), 16, len(data) + 1, len(data), len(data),
            len(data), len(data)]
    return data, streams, caps


def test_gzip_members_match_zlib(ctx):
    say("gzip members against zlib (windowBits 31)")
    data, streams, caps = _gz_cases()
    status, written, crc, dst, doff = _batch(ctx, streams, caps, ca.INFLATE_GZIP)
    for i, (s, c) in enumerate(zip(streams, caps)):
        rc, ref = oracle.zlib_inflate(s, c, gzip=True)
        assert status[i] == rc, (i, status[i], rc)
        if rc == 0:
            assert int(written[i]) == len(ref)
            assert dst[int(doff[i]): int(doff[i]) + len(ref)].tobytes() == ref.tobytes()
            assert int(crc[i]) == zlib.crc32(ref.tobytes())
    assert status.tolist()[:4] == [0, 0, 0, 0]
    # the CRC check can be skipped; ISIZE is still checked
    status, written, crc, dst, doff = _batch(ctx, streams[8:10], caps[8:10],
                                             ca.INFLATE_GZIP | ca.INFLATE_SKIP_ADLER)
    assert status.tolist() == [0, ca.CHUNK_CORRUPT]


def test_gzip_chunks_entry_point(ctx):
    say("gzip through codag_inflate_chunks")
    raw = ca.synth_deflate_payload(3 * CH, CH, seed=78).tobytes()
    chunks = [gz(raw[i * CH:(i + 1) * CH], name=b"c%d" % i) for i in range(3)]
    comp, off = container(chunks)
    st, ln, out = inflate(ctx, comp, off, CH, ca.INFLATE_GZIP)
    assert st.tolist() == [0, 0, 0] and out[: 3 * CH].tobytes() == raw


def test_cudf_status_codes(ctx):
    say("cuDF status convention")
    data = ca.synth_deflate_payload(50000, CH, seed=79).tobytes()
    good = z(data, 6)
    streams = [good, good, good[:100], b"\x78\x9c\xff\xff\xff"]
    caps = [len(data), 1000, len(data), 100]
    plain, *_ = _batch(ctx, streams, caps)
    assert plain.tolist() == [ca.CHUNK_OK, ca.CHUNK_OVERFLOW, ca.CHUNK_TRUNCATED, ca.CHUNK_CORRUPT]
    cudf, *_ = _batch(ctx, streams, caps, ca.INFLATE_CUDF_STATUS)
    assert cudf.tolist() == [0, 1, 2, -1]  # cudf_deflate.cuh:1243-1343
