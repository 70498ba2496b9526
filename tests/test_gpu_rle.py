"""GPU parity of the RLE v1 / RLE v2 decode kernels against the CPU oracle.

Every decode goes through the C ABI (codag_rlev1_decode / codag_rlev2_decode)
on device-resident buffers and must be byte-identical to the oracle, with the
same per-chunk status codes."""
import numpy as np
import pytest

import codag_amd as ca
import oracle
from tests import pycodecs as pc
from tests.gpu_helpers import rle_decode, say, single_chunk

pytestmark = pytest.mark.gpu
CH = 131072
U64 = (1 << 64) - 1


@pytest.fixture(scope="module")
def ctx(gpu_required):
    c = ca.Context(0)
    yield c
    c.close()


def _check_chunks(ctx, codec, comp, blk, cl, width, chunk):
    n = len(cl)
    st, out = rle_decode(ctx, codec, comp, blk[:n], cl, width, chunk)
    fn = oracle.rlev1_decode_chunks if codec == ca.RLEV1 else oracle.rlev2_decode_chunks
    ost, oout = fn(comp, blk[:n], cl, width, chunk, threads=8)
    assert st.tolist() == ost.tolist()
    ok = ost == 0
    o = out.reshape(n, chunk)[ok]
    e = oout.reshape(n, chunk)[ok]
    if not np.array_equal(o, e):
        bad = np.nonzero((o != e).any(axis=1))[0]
        raise AssertionError(f"{len(bad)} chunks differ, first {bad[:5]}")
    return st, out


# ------------------------------------------------------------------ RLE v1

V1_KATS = [
    (bytes([0x00, 0x01, 0x05]), 8, [5, 6, 7]),
    (bytes([0xFE, 0x81, 0x01, 0x7F]), 8, [129, 127]),
    (bytes([0x7F, 0xFF, 0x0A]), 8, [(10 - i) & U64 for i in range(130)]),
    (bytes([0x80]) + b"".join(pc.uleb(x) for x in range(1000, 1128)), 8, list(range(1000, 1128))),
    (bytes([0x00, 0x01]) + pc.uleb(300), 1, [44, 45, 46]),
    (bytes([0xFF]) + pc.uleb(1 << 63), 8, [1 << 63]),
]


@pytest.mark.parametrize("k", range(len(V1_KATS)))
def test_v1_kats(ctx, k):
    stream, w, expect = V1_KATS[k]
    say(f"rlev1 KAT {k}")
    comp, blk, cl = single_chunk(stream)
    st, out = rle_decode(ctx, ca.RLEV1, comp, blk, cl, w, len(expect) * w)
    dt = {1: np.uint8, 8: np.uint64}[w]
    assert st.tolist() == [0]
    assert out.view(dt).tolist() == expect


@pytest.mark.parametrize("width", [8, 4, 2, 1])
def test_v1_synth_parity(ctx, width):
    say(f"rlev1 synth width={width}")
    nvals = 40 * CH // width + 333
    raw, cont, blk, cl, n = ca.synth_rlev1(nvals, width, CH, 1 << 16, seed=11 + width)
    st, out = _check_chunks(ctx, ca.RLEV1, cont[8:], blk, cl, width, CH)
    assert np.all(st == 0)
    assert out[: len(raw)].tobytes() == raw.tobytes()


@pytest.mark.parametrize("width", [8, 4, 2, 1])
def test_v1_encoder_parity(ctx, width):
    say(f"rlev1 encoder width={width}")
    rng = np.random.default_rng(width)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[width]
    col = np.concatenate([rng.integers(0, 1 << 62, 20000).astype(dt),
                          np.repeat(rng.integers(0, 255, 3000), 50).astype(dt),
                          (np.arange(90000) * 3).astype(dt)])
    cont, blk, cl, n = ca.rlev1_encode(col, width)
    st, out = _check_chunks(ctx, ca.RLEV1, cont[8:], blk, cl[:n], width, CH)
    assert out[: col.nbytes].tobytes() == col.tobytes()


def test_v1_many_small_chunks_and_max_segments(ctx):
    # chunks of 64 values: runs of max length 130 overflow, literal groups of 128 do not
    say("rlev1 small chunks")
    rng = np.random.default_rng(5)
    streams, vals = [], []
    for i in range(300):
        if i % 3 == 0:
            v = [int(x) for x in rng.integers(0, 1 << 40, 64)]
            s = bytes([0xC0]) + b"".join(pc.uleb(x) for x in v)
        elif i % 3 == 1:
            v = [7 + j * 2 for j in range(64)]
            s = bytes([61, 2]) + pc.uleb(7)
        else:
            v = [int(x) for x in rng.integers(0, 1 << 62, 64)]
            s = pc.rlev1_encode(v, 8)
        streams.append(s)
        vals.append(v)
    off = np.cumsum([0] + [len(s) for s in streams])[:-1].astype(np.uint64)
    cl = np.array([len(s) for s in streams], np.uint64)
    comp = np.frombuffer(b"".join(streams), np.uint8)
    st, out = _check_chunks(ctx, ca.RLEV1, comp, off, cl, 8, 64 * 8)
    assert np.all(st == 0)
    assert out.view(np.uint64).tolist() == [x & U64 for v in vals for x in v]


V1_BAD = [
    (bytes([0x00, 0x01]), 24),                       # truncated run
    (bytes([0x05, 0x00, 0x01]), 24),                 # run overflows the chunk
    (b"", 8),                                        # empty stream
    (bytes([0xFF]) + bytes([0x80] * 10) + bytes([1]), 8),  # 11-byte varint
    (bytes([0xFE, 0x81]), 16),                       # truncated literal
]


@pytest.mark.parametrize("k", range(len(V1_BAD)))
def test_v1_error_status_matches_oracle(ctx, k):
    stream, chunk = V1_BAD[k]
    say(f"rlev1 error case {k}")
    comp, blk, cl = single_chunk(stream, pad=16)
    st, _ = rle_decode(ctx, ca.RLEV1, comp, blk, cl, 8, chunk)
    rc, _ = oracle.rlev1_decode_chunk(stream, 8, chunk)
    assert rc != 0 and st.tolist() == [rc]


@pytest.mark.parametrize("shape", [(512, 2048, 4), (256, 4096, 2), (128, 4096, 2)])
def test_v1_walker_and_expander_kernels(ctx, shape):
    # the walker + expander kernel on the synthetic columns, the known answers,
    # the malformed corpus, small chunks, and randomly damaged streams whose
    # status (first error in stream order) must equal the oracle's
    say(f"rlev1 kernel shape {shape}")
    saved = ctx.tune_get(ca.RLEV1)
    ctx.tune_set(ca.RLEV1, ca.TuneParams(shape[0], shape[1], 0, shape[2]))
    try:
        for width in (8, 4, 1):
            raw, cont, blk, cl, n = ca.synth_rlev1(24 * CH // width + 333, width, CH, 1 << 16,
                                                   seed=31 + width)
            st, out = _check_chunks(ctx, ca.RLEV1, cont[8:], blk, cl, width, CH)
            assert np.all(st == 0) and out[: len(raw)].tobytes() == raw.tobytes()
        for k in range(len(V1_KATS)):
            test_v1_kats(ctx, k)
        for k in range(len(V1_BAD)):
            test_v1_error_status_matches_oracle(ctx, k)
        test_v1_many_small_chunks_and_max_segments(ctx)
        rng = np.random.default_rng(17)
        base = []
        for b in range(6):
            v = [int(x) for x in rng.integers(0, 1 << (8 * (b + 1)), 200)]
            v[50:90] = [v[50]] * 40
            base.append(pc.rlev1_encode(v, 8))
        streams = []
        for t in range(600):
            s = bytearray(base[t % 6])
            for _ in range(int(rng.integers(1, 4))):
                s[int(rng.integers(0, len(s)))] = int(rng.integers(0, 256))
            if t % 3 == 0:
                s = s[: int(rng.integers(0, len(s) + 1))]
            streams.append(bytes(s))
        off = np.cumsum([0] + [len(s) for s in streams])[:-1].astype(np.uint64)
        cl = np.array([len(s) for s in streams], np.uint64)
        comp = np.frombuffer(b"".join(streams) + b"\0" * 16, np.uint8)
        st, _ = _check_chunks(ctx, ca.RLEV1, comp, off, cl, 8, 200 * 8)
        assert len(set(st.tolist())) > 2  # the damage produced several kinds of status
    finally:
        ctx.tune_set(ca.RLEV1, saved)


@pytest.mark.parametrize("width", [8, 4, 1])
def test_v1_pipeline_damaged_multiwindow_chunks(ctx, width):
    # 128 KiB chunks whose streams span many 2 KiB windows, damaged by single
    # bytes, by runs of continuation bytes (overlong varints that run past the
    # indexed windows) and by truncation; status must equal the oracle's and
    # undamaged chunks must decode exactly
    say(f"rlev1 pipeline damaged multi-window chunks w={width}")
    saved = ctx.tune_get(ca.RLEV1)
    ctx.tune_set(ca.RLEV1, ca.TuneParams(512, 2048, 0, 4))
    try:
        raw, cont, blk, cl, n = ca.synth_rlev1(6 * CH // width, width, CH, 1 << 16, seed=40 + width)
        comp = cont[8:]
        rng = np.random.default_rng(50 + width)
        streams = []
        for t in range(48):
            c = t % n
            s = bytearray(comp[int(blk[c]): int(blk[c]) + int(cl[c])])
            if t % 8 != 7:
                for _ in range(int(rng.integers(1, 3))):
                    k = int(rng.integers(0, len(s)))
                    if t % 4 == 0:
                        s[k:k + 3000] = b"\xff" * min(3000, len(s) - k)
                    else:
                        s[k] = int(rng.integers(0, 256))
                if t % 5 == 0:
                    s = s[: int(rng.integers(0, len(s) + 1))]
            # padding between streams: every alignment of the stream start
            streams.append((bytes(s), bytes(int(rng.integers(0, 16)))))
        off = np.cumsum([0] + [len(b) + len(p) for b, p in streams])[:-1].astype(np.uint64)
        cln = np.array([len(b) for b, _ in streams], np.uint64)
        comp2 = np.frombuffer(b"".join(b + p for b, p in streams) + bytes(16), np.uint8)
        st, _ = _check_chunks(ctx, ca.RLEV1, comp2, off, cln, width, CH)
        assert (st == 0).sum() >= 6 and len(set(st.tolist())) > 2
    finally:
        ctx.tune_set(ca.RLEV1, saved)


def test_v1_stream_outside_buffer_is_corrupt(ctx):
    say("rlev1 out-of-range blk_off")
    comp = np.zeros(64, np.uint8)
    st, _ = rle_decode(ctx, ca.RLEV1, comp, np.array([60], np.uint64),
                       np.array([100], np.uint64), 8, 64)
    assert st.tolist() == [3]


# ------------------------------------------------------------------ RLE v2

V2_KATS = [
    ("0a2710", 8, [10000] * 5),
    ("5e035ca1ab1edeadbeef", 4, [23713, 43806, 57005, 48879]),
    ("c609020222424246", 8, [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]),
    ("8e132b2107d01e00147028323c46505a646e78828c96a0aab4befce8", 8,
     [2030, 2000, 2020, 1000000] + list(range(2040, 2200, 10))),
]


@pytest.mark.parametrize("k", range(len(V2_KATS)))
def test_v2_kats(ctx, k):
    hexs, w, expect = V2_KATS[k]
    say(f"rlev2 KAT {k}")
    comp, blk, cl = single_chunk(bytes.fromhex(hexs), pad=16)
    st, out = rle_decode(ctx, ca.RLEV2, comp, blk, cl, w, 64 * w)
    dt = {4: np.uint32, 8: np.uint64}[w]
    assert st.tolist() == [0]
    v = out.view(dt)
    assert v[: len(expect)].tolist() == expect and not v[len(expect):].any()


def test_v2_semantic_corners(ctx):
    say("rlev2 corners")
    cases = [pc.v2_delta(100, 0, [1, 2, 3], code=3), pc.v2_delta(100, -1, [1, 2, 3], code=3),
             pc.v2_delta(7, 5, [0] * 6, code=None), bytes([0x40 | (7 << 1), 0xFF]),
             bytes([0xC0, 0x02]) + pc.uleb(5) + pc.uleb(U64),
             pc.v2_short_repeat(0x0102030405, 3, 5)]
    for w in (8, 4):
        streams = cases
        off = np.cumsum([0] + [len(s) for s in streams])[:-1].astype(np.uint64)
        cl = np.array([len(s) for s in streams], np.uint64)
        comp = np.frombuffer(b"".join(streams) + b"\0" * 16, np.uint8)
        st, out = _check_chunks(ctx, ca.RLEV2, comp, off, cl, w, 512 * w)
        assert np.all(st == 0)


@pytest.mark.parametrize("code", range(32))
def test_v2_every_width(ctx, code):
    bits = pc.DECODE_WIDTHS[code]
    rng = np.random.default_rng(code)
    vals = [int(rng.integers(0, 1 << min(bits, 62))) << max(0, bits - 62) for _ in range(255)]
    vals[0] = (1 << bits) - 1
    s = pc.v2_direct(vals[:128], code) + pc.v2_direct(vals[128:], code)
    comp, blk, cl = single_chunk(s, pad=16)
    for w in (8, 4, 2, 1):
        st, out = _check_chunks(ctx, ca.RLEV2, comp, blk, cl, w, 256 * w)
        assert st.tolist() == [0]


@pytest.mark.parametrize("code", range(1, 32))
def test_v2_delta_every_width(ctx, code):
    # DELTA blocks with packed deltas of every width, up to 255 values (four
    # 64-value groups), a 10-byte base varint and both signs of the delta
    # base (zero subtracts): the 32-bit and the 64-bit prefix sums and the
    # varint assembly, on the default kernel and the one-wave kernel
    bits = pc.DECODE_WIDTHS[code]
    rng = np.random.default_rng(1000 + code)
    s = b""
    for nvals, base, bdelta in ((255, (1 << 64) - 1, 3), (130, 1 << 63, 0), (66, 12345, -7),
                                (3, 0, 1)):
        deltas = [int(rng.integers(0, 1 << min(bits, 62))) << max(0, bits - 62)
                  for _ in range(nvals - 2)]
        deltas[0] = (1 << bits) - 1
        s += pc.v2_delta(base, bdelta, deltas, code)
    comp, blk, cl = single_chunk(s, pad=16)
    saved = ctx.tune_get(ca.RLEV2)
    try:
        for shape in (None, (64, 4096, 0)):
            if shape:
                ctx.tune_set(ca.RLEV2, ca.TuneParams(shape[0], shape[1], 0, shape[2]))
            for w in (8, 4, 2, 1):
                st, out = _check_chunks(ctx, ca.RLEV2, comp, blk, cl, w, 512 * w)
                assert st.tolist() == [0]
    finally:
        ctx.tune_set(ca.RLEV2, saved)


@pytest.mark.parametrize("width", [8, 4, 2, 1])
def test_v2_synth_parity(ctx, width):
    say(f"rlev2 synth width={width}")
    nvals = 40 * CH // width + 77
    raw, comp, blk, cl, n = ca.synth_rlev2(nvals, width, CH, seed=21 + width)
    st, out = _check_chunks(ctx, ca.RLEV2, comp, blk, cl, width, CH)
    assert np.all(st == 0)
    assert out[: len(raw)].tobytes() == raw.tobytes()


@pytest.mark.parametrize("width", [8, 4, 2, 1])
def test_v2_encoder_parity(ctx, width):
    say(f"rlev2 encoder width={width}")
    rng = np.random.default_rng(40 + width)
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[width]
    col = np.concatenate([rng.integers(0, 1 << 62, 20000).astype(dt),
                          np.repeat(rng.integers(0, 255, 3000), 7).astype(dt),
                          (np.arange(90000) * 3 + 5).astype(dt),
                          (rng.integers(0, 50, 30000) + (rng.random(30000) < 0.03) * 10**6).astype(dt)])
    comp, blk, cl, n = ca.rlev2_encode(col, width)
    st, out = _check_chunks(ctx, ca.RLEV2, comp, blk, cl, width, CH)
    assert out[: col.nbytes].tobytes() == col.tobytes()


def test_v2_patched_base_random(ctx):
    say("rlev2 patched base random")
    rng = np.random.default_rng(9)
    streams = []
    for t in range(200):
        n = int(rng.integers(1, 129))
        code = int(rng.choice([3, 7, 11, 15]))
        fbw = pc.DECODE_WIDTHS[code]
        reduced = [int(x) for x in rng.integers(0, 1 << fbw, n)]
        pw_code = int(rng.choice([0, 3, 7, 11, 15, 20, 27]))
        pw = pc.DECODE_WIDTHS[pw_code]
        pgw = int(rng.integers(1, 9))
        idx = sorted(rng.choice(n, min(n, int(rng.integers(0, 8))), replace=False).tolist())
        patches, prev = [], 0
        for i in idx:
            if i - prev < (1 << pgw):
                patches.append((i - prev, int(rng.integers(0, 1 << pw))))
                prev = i
        streams.append(pc.v2_patched_base(reduced, int(rng.integers(0, 1 << 30)), code, 4,
                                          patches, pw_code, pgw))
    # patches in the second half of a 128-value block (second lane pass)
    red = [int(x) for x in rng.integers(0, 8, 128)]
    streams.append(pc.v2_patched_base(red, 1000, 2, 2, [(70, 5), (57, 9)], 3, 8))
    streams.append(pc.v2_patched_base(red, 7, 2, 1, [(0, 1), (64, 3), (63, 2)], 7, 8))
    off = np.cumsum([0] + [len(s) for s in streams])[:-1].astype(np.uint64)
    cl = np.array([len(s) for s in streams], np.uint64)
    comp = np.frombuffer(b"".join(streams) + b"\0" * 16, np.uint8)
    st, _ = _check_chunks(ctx, ca.RLEV2, comp, off, cl, 8, 256 * 8)
    assert np.all(st == 0)


V2_BAD = [
    ("5e035ca1", 64),          # DIRECT truncated
    ("0a2710", 3),             # overflow
    ("c6", 64),                # DELTA header truncated
    ("c002" + "80" * 10 + "0102", 64),  # 11-byte varint
    ("c002" + "80" * 12, 64),  # varint runs off the stream
]


@pytest.mark.parametrize("k", range(len(V2_BAD)))
def test_v2_error_status_matches_oracle(ctx, k):
    hexs, cap = V2_BAD[k]
    say(f"rlev2 error case {k}")
    s = bytes.fromhex(hexs)
    comp, blk, cl = single_chunk(s, pad=16)
    st, _ = rle_decode(ctx, ca.RLEV2, comp, blk, cl, 8, cap * 8)
    rc, _ = oracle.rlev2_decode_chunk(s, 8, cap)
    assert rc != 0 and st.tolist() == [rc]


@pytest.mark.parametrize("shape", [(64, 4096, 0), (64, 8192, 0), (256, 4096, 0), (128, 4096, 0),
                                   (512, 8192, 0), (128, 4096, 2), (192, 4096, 2), (256, 4096, 2)])
def test_v2_every_launch_shape(ctx, shape):
    say(f"rlev2 launch shape {shape}")
    raw, comp, blk, cl, n = ca.synth_rlev2(12 * CH // 8 + 5, 8, CH, seed=77)
    saved = ctx.tune_get(ca.RLEV2)
    ctx.tune_set(ca.RLEV2, ca.TuneParams(shape[0], shape[1], 0, shape[2]))
    try:
        st, out = _check_chunks(ctx, ca.RLEV2, comp, blk, cl, 8, CH)
        assert np.all(st == 0) and out[: len(raw)].tobytes() == raw.tobytes()
    finally:
        ctx.tune_set(ca.RLEV2, saved)


@pytest.mark.parametrize("shape", [(128, 4096, 2), (192, 4096, 2), (256, 4096, 2), (64, 4096, 0)])
def test_v2_walker_and_expander_kernels(ctx, shape):
    # the one-wave kernel and the walker + expander kernels on the same cases:
    # synthetic columns of several widths (many windows and batches per
    # chunk), the malformed streams, an empty stream and random PATCHED_BASE
    say(f"rlev2 kernel shape {shape}")
    saved = ctx.tune_get(ca.RLEV2)
    ctx.tune_set(ca.RLEV2, ca.TuneParams(shape[0], shape[1], 0, shape[2]))
    try:
        for width in (8, 4, 1):
            raw, comp, blk, cl, n = ca.synth_rlev2(24 * CH // width + 77, width, CH,
                                                   seed=121 + width)
            st, out = _check_chunks(ctx, ca.RLEV2, comp, blk, cl, width, CH)
            assert np.all(st == 0) and out[: len(raw)].tobytes() == raw.tobytes()
        for hexs, cap in V2_BAD + [("", 16)]:
            s = bytes.fromhex(hexs)
            comp, blk, cl = single_chunk(s, pad=16)
            st, out = rle_decode(ctx, ca.RLEV2, comp, blk, cl, 8, cap * 8)
            rc, ref = oracle.rlev2_decode_chunk(s, 8, cap)
            assert st.tolist() == [rc], hexs
            if rc == 0:
                assert out.view(np.uint64).tolist() == ref.tolist() + [0] * (cap - len(ref))
        test_v2_patched_base_random(ctx)
    finally:
        ctx.tune_set(ca.RLEV2, saved)


def test_v2_queue_waits_are_bounded(ctx):
    # the walker + expander kernel's waits are bounded: with the bound cut to
    # nothing (a host-side test hook), every wait not met at once ends its
    # chunk with CORRUPT and the launch returns; chunks that finish are still
    # exact, and with the bound restored every chunk decodes
    import ctypes
    lib = ca.lib()
    lib.codag_test_poll_shift.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
    lib.codag_test_poll_shift.restype = ctypes.c_int
    raw, comp, blk, cl, n = ca.synth_rlev2(16 * CH // 4, 4, CH, seed=77)
    ost, oout = oracle.rlev2_decode_chunks(comp, blk[:n], cl, 4, CH, threads=8)
    assert np.all(ost == 0)
    saved = ctx.tune_get(ca.RLEV2)
    say("rlev2 queue kernel with zero poll limits")
    try:
        for shape in ((128, 4096, 2), (256, 4096, 2)):
            ctx.tune_set(ca.RLEV2, ca.TuneParams(shape[0], shape[1], 0, shape[2]))
            assert lib.codag_test_poll_shift(ctx.h, 31) == 0
            st, out = rle_decode(ctx, ca.RLEV2, comp, blk[:n], cl, 4, CH)
            assert set(st.tolist()) <= {0, 3} and 3 in st.tolist()
            ok = st == 0
            assert np.array_equal(out.reshape(n, CH)[ok], oout.reshape(n, CH)[ok])
            assert lib.codag_test_poll_shift(ctx.h, 0) == 0
            st, out = rle_decode(ctx, ca.RLEV2, comp, blk[:n], cl, 4, CH)
            assert np.all(st == 0) and np.array_equal(out, oout)
    finally:
        lib.codag_test_poll_shift(ctx.h, 0)
        ctx.tune_set(ca.RLEV2, saved)


def test_v2_patch_outside_block_is_corrupt(ctx):
    say("rlev2 bad patch")
    s = pc.v2_patched_base([1, 2], 0, 3, 1, [(5, 1)], 3, 3)
    comp, blk, cl = single_chunk(s, pad=16)
    st, _ = rle_decode(ctx, ca.RLEV2, comp, blk, cl, 8, 64 * 8)
    assert st.tolist() == [3]
