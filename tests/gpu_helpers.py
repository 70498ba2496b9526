"""Helpers for the GPU parity tests: run a decode through the C ABI on
device-resident buffers and bring the result back."""
import sys

import numpy as np

import codag_amd as ca


def say(msg: str):
    print(msg, flush=True)
    sys.stdout.flush()


def rle_decode(ctx, codec, comp, blk_off, col_len, width, chunk_bytes, fill=0xA5):
    """Decode an RLE container body on the GPU. Returns (status[n], out bytes)."""
    comp = np.ascontiguousarray(comp, dtype=np.uint8)
    n = len(col_len)
    d_comp = ctx.upload(comp)
    d_off = ctx.upload(np.ascontiguousarray(blk_off[:n], dtype=np.uint64))
    d_len = ctx.upload(np.ascontiguousarray(col_len[:n], dtype=np.uint64))
    d_out = ctx.alloc(n * chunk_bytes)
    d_out.fill(fill)
    d_st = ctx.alloc(4 * max(n, 1))
    d_st.fill(0x7F)
    fn = ctx.rlev1_decode if codec == ca.RLEV1 else ctx.rlev2_decode
    fn(d_comp, len(comp), d_off, d_len, n, width, chunk_bytes, d_out, d_st)
    ctx.sync()
    st = d_st.download(4 * n, np.int32)
    out = d_out.download(n * chunk_bytes)
    return st, out


def single_chunk(stream: bytes, pad: int = 0):
    """(comp, blk_off, col_len) for one chunk holding `stream`."""
    comp = np.frombuffer(bytes(stream) + b"\0" * pad, dtype=np.uint8)
    return comp, np.zeros(1, np.uint64), np.array([len(stream)], np.uint64)


def inflate(ctx, comp, offsets, chunk_bytes, flags=0):
    comp = np.ascontiguousarray(comp, dtype=np.uint8)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = len(off) - 1
    d_comp = ctx.upload(comp)
    d_off = ctx.upload(off)
    d_out = ctx.alloc(n * chunk_bytes)
    d_out.fill(0xA5)
    d_len = ctx.alloc(8 * max(n, 1))
    d_st = ctx.alloc(4 * max(n, 1))
    d_st.fill(0x7F)
    ctx.inflate_chunks(d_comp, len(comp), d_off, n, chunk_bytes, d_out, d_len, d_st, flags)
    ctx.sync()
    return (d_st.download(4 * n, np.int32), d_len.download(8 * n, np.uint64),
            d_out.download(n * chunk_bytes))
