"""GPU tests of the autotuner (codag_tune_run): for each codec it times its
launch shapes on a device-resident sample, keeps one whose output equals the
first candidate's and makes it the context's shape; a decode with that shape
still reproduces the column, and the C4 batch bench's use (a Deflate sample
decoded through the batch launch) returns a compiled shape."""
import numpy as np
import pytest

import codag_amd as ca
from tests.gpu_helpers import rle_decode, say

pytestmark = pytest.mark.gpu

CH = ca.CHUNK_BYTES


@pytest.fixture(scope="module")
def ctx(gpu_required):
    c = ca.Context(0)
    yield c
    c.close()


def fields(t):
    return (t.threads_per_group, t.window_bytes, t.groups_per_cu, t.queue_depth)


@pytest.mark.parametrize("codec", [ca.RLEV1, ca.RLEV2])
def test_tune_rle_keeps_a_shape_that_decodes(ctx, codec):
    say(f"autotuner, codec {codec}")
    if codec == ca.RLEV1:
        raw, cont, blk, cl, n = ca.synth_rlev1(1 << 19, 8, CH, 1 << 16, seed=5)
        comp = cont[8:]
    else:
        raw, comp, blk, cl, n = ca.synth_rlev2(1 << 19, 8, CH, seed=6)
    saved = ctx.tune_get(codec)
    d_comp = ctx.upload(comp)
    d_off = ctx.upload(np.ascontiguousarray(blk[:n], dtype=np.uint64))
    d_len = ctx.upload(np.ascontiguousarray(cl[:n], dtype=np.uint64))
    d_out = ctx.alloc(n * CH)
    try:
        best, ms = ctx.tune_run(codec, d_comp, len(comp), d_off, d_len, n, 8, CH, d_out, reps=1)
        assert ms > 0
        assert fields(ctx.tune_get(codec)) == fields(best)
        assert d_out.download(len(raw)).tobytes() == raw.tobytes()
        st, out = rle_decode(ctx, codec, comp, blk, cl[:n], 8, CH)  # with the kept shape
        assert np.all(st == 0) and out[: len(raw)].tobytes() == raw.tobytes()
    finally:
        ctx.tune_set(codec, saved)


def test_tune_deflate_batch_sample(ctx):
    say("autotuner, deflate batch launch")
    raw = ca.synth_deflate_payload(48 * CH, CH, seed=7)
    comp, blk, cl, n = ca.deflate_compress_chunks(raw, CH, level=6)
    off = blk[1:].astype(np.uint64)
    saved = ctx.tune_get(ca.DEFLATE)
    d_comp = ctx.upload(comp)
    d_off = ctx.upload(off)
    d_out = ctx.alloc(n * CH)
    try:
        best, ms = ctx.tune_run(ca.DEFLATE, d_comp, len(comp), d_off, None, n, 1, CH, d_out, reps=1)
        assert ms > 0
        assert best.threads_per_group in (64, 128) and 1 <= best.groups_per_cu <= 64
        assert fields(ctx.tune_get(ca.DEFLATE)) == fields(best)
        assert d_out.download(len(raw)).tobytes() == raw.tobytes()
    finally:
        ctx.tune_set(ca.DEFLATE, saved)
