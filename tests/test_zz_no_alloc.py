"""The allocation-free batch entry point (codag_inflate_batch_noalloc) against
the existing batch entry point (codag_inflate_batch): byte-identical outputs,
byte counts, statuses and Adler-32 on two calls in a row, the second with a
smaller batch, and both checked against zlib."""
import ctypes
import zlib

import numpy as np
import pytest

import codag_amd as ca
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu
CH = 131072


def _streams(seed, count, hi):
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(4096), np.log(hi), count)).astype(int)
    payload = ca.synth_deflate_payload(int(sizes.sum()), CH, seed=seed)
    streams, raws, pos = [], [], 0
    for i, n in enumerate(sizes):
        r = payload[pos: pos + n].tobytes()
        pos += n
        raws.append(r)
        s, _ = ca.deflate_compress_mixed(np.frombuffer(r, np.uint8), 20, 5, 32768, seed=i)
        streams.append(s.tobytes())
    # one truncated stream and one that overflows its capacity: statuses too
    streams.append(streams[0][: len(streams[0]) // 2])
    raws.append(None)
    streams.append(streams[1])
    raws.append(None)
    return streams, raws


def _run(ctx, streams, raws, noalloc):
    caps = [len(r) if r is not None else 1000 for r in raws]
    caps[-2] = len(raws[0])
    n = len(streams)
    src = np.frombuffer(b"".join(streams), np.uint8)
    soff = np.cumsum([0] + [len(s) for s in streams]).astype(np.uint64)
    doff = np.cumsum([0] + caps).astype(np.uint64)
    d_src = ctx.upload(src)
    d_dst = ctx.alloc(int(doff[-1]))
    d_dst.fill(0)
    descs = (ca.StreamDesc * n)()
    for i in range(n):
        descs[i].src, descs[i].src_bytes = d_src.offset(int(soff[i])), len(streams[i])
        descs[i].dst, descs[i].dst_cap = d_dst.offset(int(doff[i])), caps[i]
    d_desc = ctx.upload(np.frombuffer(bytes(descs), np.uint8))
    d_res = ctx.alloc(16 * n)
    d_res.fill(0x7F)
    ctx.sync()
    if noalloc:
        # on the context's stream, with the context's device current
        ca.check(ca.lib().codag_inflate_batch_noalloc(d_desc.ptr, n, d_res.ptr, 0, 0,
                                                      ca.lib().codag_ctx_stream(ctx.h)),
                 "inflate_batch_noalloc")
    else:
        ctx.inflate_batch(d_desc, n, d_res)
    ctx.sync()
    res = d_res.download(16 * n).view(np.uint8).reshape(n, 16)
    return res.copy(), d_dst.download(int(doff[-1])), doff


def test_noalloc_batch_matches_batch(gpu_required):
    say("inflate_batch_noalloc vs inflate_batch, two calls (second smaller)")
    ctx = ca.Context(0)
    try:
        for seed, count, hi in ((21, 300, 1 << 20), (22, 40, 200000)):
            streams, raws = _streams(seed, count, hi)
            say(f"  batch seed {seed}: {len(streams)} streams")
            r_ref, out_ref, doff = _run(ctx, streams, raws, noalloc=False)
            r_new, out_new, _ = _run(ctx, streams, raws, noalloc=True)
            assert r_new.tobytes() == r_ref.tobytes()
            assert out_new.tobytes() == out_ref.tobytes()
            status = r_new[:, 8:12].copy().view(np.int32).reshape(-1)
            adler = r_new[:, 12:16].copy().view(np.uint32).reshape(-1)
            for i, r in enumerate(raws):
                if r is None:
                    continue
                assert status[i] == 0
                assert out_new[int(doff[i]): int(doff[i]) + len(r)].tobytes() == r
                assert int(adler[i]) == zlib.adler32(r)
            assert status[-2] == ca.CHUNK_TRUNCATED and status[-1] == ca.CHUNK_OVERFLOW
    finally:
        ctx.close()


def test_noalloc_rejects_oversized_batch():
    # argument checks run before any device call: no GPU needed
    st = ca.lib().codag_inflate_batch_noalloc(ctypes.c_void_p(16), 1 << 32, ctypes.c_void_p(16),
                                              0, 0, None)
    assert st == ca.ERR_INVALID
    assert b"2^32" in ca.lib().codag_last_error()
