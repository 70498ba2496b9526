"""GPU parity of the two-wave batch kernel (inflate_batch2_kernel: a
decoding wave hands token batches, stored blocks and the stream end to a
writing wave through an LDS queue). Its results -- status, bytes written,
Adler-32 / CRC-32 and every output byte -- must equal the one-wave batch
kernel's, which the other inflate tests hold to zlib, on valid, damaged,
truncated, overflowing, stored-heavy, raw and gzip streams; valid streams
are also checked against their original bytes."""
import zlib

import numpy as np
import pytest

import codag_amd as ca
import oracle
from tests.gpu_helpers import say
from tests.test_gpu_inflate import CH, Batch, _check_batch, _gz_cases, _uneven, z

pytestmark = pytest.mark.gpu

ONE = ca.TuneParams(64, 0, 28, 1)
PAIR = ca.TuneParams(128, 0, 12, 1)


@pytest.fixture(scope="module")
def ctx(gpu_required):
    c = ca.Context(0)
    yield c
    c.tune_set(ca.DEFLATE, ONE)
    c.close()


def run(ctx, streams, caps, flags, shape):
    ctx.tune_set(ca.DEFLATE, shape)
    try:
        b = Batch(ctx, streams, caps)
        b.launch(flags)
        ctx.sync()
        return b.results()
    finally:
        ctx.tune_set(ca.DEFLATE, ONE)


def same(ctx, streams, caps, flags=0):
    a = run(ctx, streams, caps, flags, ONE)
    b = run(ctx, streams, caps, flags, PAIR)
    sa, wa, xa, da, oa = a
    sb, wb, xb, db, ob = b
    assert sa.tolist() == sb.tolist()
    assert wa.tolist() == wb.tolist()
    assert xa.tolist() == xb.tolist()
    for i in range(len(streams)):
        n = int(wa[i])
        assert np.array_equal(da[int(oa[i]): int(oa[i]) + n], db[int(ob[i]): int(ob[i]) + n]), i
    return b


def test_pair_uneven_streams(ctx):
    say("two-wave batch: uneven streams up to 4 MiB")
    streams, raws = _uneven(10, 60, 600000, big=(4 << 20, (4 << 20) - 12345))
    order = sorted(range(len(raws)), key=lambda i: -len(raws[i]))
    streams = [streams[i] for i in order]
    raws = [raws[i] for i in order]
    _check_batch(same(ctx, streams, [len(r) for r in raws]), raws)


def test_pair_damaged_streams(ctx):
    say("two-wave batch: damaged, truncated and overflowing streams")
    rng = np.random.default_rng(21)
    data = ca.synth_deflate_payload(200000, CH, seed=21).tobytes()
    good = z(data, 6)
    streams, caps = [], []
    for t in range(160):  # byte flips anywhere in the stream
        s = bytearray(good)
        for _ in range(int(rng.integers(1, 4))):
            s[int(rng.integers(2, len(s)))] ^= int(rng.integers(1, 256))
        streams.append(bytes(s))
        caps.append(len(data))
    for cut in (0, 1, 2, 7, 100, len(good) // 3, len(good) - 5, len(good) - 1):
        streams.append(good[:cut])
        caps.append(len(data))
    for cap in (0, 1, 1000, 65536, len(data) - 1):  # output buffer too small
        streams.append(good)
        caps.append(cap)
    stored = z(data[:150000], 0)  # stored blocks only
    streams += [stored, stored[:70000], stored]
    caps += [150000, 150000, 100000]
    mixed, _ = ca.deflate_compress_mixed(np.frombuffer(data, np.uint8), 20, 30, 8192, seed=3)
    streams += [mixed.tobytes(), mixed.tobytes()]
    caps += [len(data), 120000]
    status, written, adler, dst, doff = same(ctx, streams, caps)
    for i, (s, c) in enumerate(zip(streams, caps)):
        rc, ref = oracle.zlib_inflate(s, c)
        if rc == 0:
            assert status[i] == 0 and dst[int(doff[i]): int(doff[i]) + len(ref)].tobytes() == \
                ref.tobytes(), i
        else:
            assert status[i] != 0, i


def test_pair_raw_and_skip_adler(ctx):
    say("two-wave batch: raw deflate and skipped Adler-32")
    data = ca.synth_deflate_payload(300000, CH, seed=22).tobytes()
    raw = z(data, 6, wbits=-15)
    bad_adler = bytearray(z(data, 9))
    bad_adler[-1] ^= 1
    same(ctx, [raw, raw[:-10]], [len(data), len(data)], ca.INFLATE_RAW)
    st, *_ = same(ctx, [bytes(bad_adler)], [len(data)], ca.INFLATE_SKIP_ADLER)
    assert st.tolist() == [0]
    st, *_ = same(ctx, [bytes(bad_adler)], [len(data)])
    assert st.tolist() == [ca.CHUNK_CORRUPT]


def test_pair_gzip_and_cudf_status(ctx):
    say("two-wave batch: gzip members and cuDF status codes")
    data, streams, caps = _gz_cases()
    status, written, crc, dst, doff = same(ctx, streams, caps, ca.INFLATE_GZIP)
    for i, (s, c) in enumerate(zip(streams, caps)):
        rc, ref = oracle.zlib_inflate(s, c, gzip=True)
        assert status[i] == rc, (i, status[i], rc)
        if rc == 0:
            assert int(crc[i]) == zlib.crc32(ref.tobytes())
    d2 = ca.synth_deflate_payload(50000, CH, seed=79).tobytes()
    good = z(d2, 6)
    st, *_ = same(ctx, [good, good, good[:100], b"\x78\x9c\xff\xff\xff"],
                  [len(d2), 1000, len(d2), 100], ca.INFLATE_CUDF_STATUS)
    assert st.tolist() == [0, 1, 2, -1]


def test_pair_graph_replay(ctx):
    say("two-wave batch: graph replays on fixed buffers")
    streams, raws = _uneven(12, 40, 300000)
    ctx.tune_set(ca.DEFLATE, PAIR)
    try:
        b = Batch(ctx, streams, [len(r) for r in raws])
        for rep in range(3):
            b.launch()
            ctx.sync()
            _check_batch(b.results(), raws)
    finally:
        ctx.tune_set(ca.DEFLATE, ONE)


def test_pair_chunk_container(ctx):
    say("two-wave chunk kernel: chunk container against the one-wave kernel and zlib")
    from tests.gpu_helpers import inflate
    from tests.test_gpu_inflate import container
    raw = ca.synth_deflate_payload(24 * CH, CH, seed=23).tobytes()
    chunks = [z(raw[i * CH:(i + 1) * CH], 6) for i in range(24)]
    chunks[5] = chunks[5][:-9]                     # truncated
    chunks[9] = bytes([0x78, 0x9c, 0xff]) + chunks[9][3:]  # corrupt
    comp, off = container(chunks)
    try:
        ca.check(ca.lib().codag_test_chunks_two_wave(ctx.h, 2), "test_chunks_two_wave")
        a = inflate(ctx, comp, off, CH)  # one wave per chunk
        ca.check(ca.lib().codag_test_chunks_two_wave(ctx.h, 1), "test_chunks_two_wave")
        b = inflate(ctx, comp, off, CH)  # two waves per chunk
    finally:
        ca.check(ca.lib().codag_test_chunks_two_wave(ctx.h, 0), "test_chunks_two_wave")
    assert a[0].tolist() == b[0].tolist() and a[1].tolist() == b[1].tolist()
    for i in range(24):
        n = int(a[1][i])
        assert np.array_equal(a[2][i * CH: i * CH + n], b[2][i * CH: i * CH + n]), i
        rc, ref = oracle.zlib_inflate(comp[int(off[i]): int(off[i + 1])].tobytes(), CH)
        assert b[0][i] == rc, i
