"""CPU oracle for the CODAG decoders -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package, and only as the checker. The product
(``codag_amd``, the C ABI in ``include/codag.h`` and the CLIs) never loads it.

It wraps ``liboracle.so`` (built from ``rle_oracle.c`` / ``zlib_oracle.c`` by
``oracle/Makefile``):

* RLE v1: restates ``include/rlev1/decoder.h:671-791``.
* RLE v2: restates ``include/rlev2/decoder_new.h:909-1187`` and the width
  tables of ``include/rlev2/utils.h:28-59``.
* Deflate: the system zlib ``inflate()`` (RFC 1950/1951), which is the
  arithmetic ``include/zlib/zlib.h:807-866`` implements.
* BRLE / LZSS (``byte_oracle.c``): restate ``include/brle/brle.h:26-100`` and
  ``include/lzss/lzss.h:234-305`` (decoders) and ``brle.h:129-230``,
  ``lzss.h:75-113, 310-419`` (encoders, statement by statement, with the
  reference's rings and exhaustive match scan).

Pinning (see tests/test_oracle.py): hand-derived known-answer vectors from the
reference decoder text, the public ORC v2 examples, closed-form round trips
through independent Python encoders, zlib's own streams, and brute-force
decodes of tiny inputs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OK, TRUNCATED, OVERFLOW, CORRUPT = 0, 1, 2, 3

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)


def build() -> str:
    """Compile liboracle.so (gcc; no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liboracle.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.oracle_rlev1_decode_chunk.argtypes = [_u8p, ctypes.c_uint64, ctypes.c_uint32, _u8p,
                                                ctypes.c_uint64, _u64p]
        L.oracle_rlev2_decode_chunk.argtypes = [_u8p, ctypes.c_uint64, ctypes.c_uint32, _u8p,
                                                ctypes.c_uint64, _u64p]
        for fn in (L.oracle_rlev1_decode_chunks, L.oracle_rlev2_decode_chunks):
            fn.argtypes = [_u8p, _u64p, _u64p, ctypes.c_uint64, ctypes.c_uint32,
                           ctypes.c_uint64, _u8p, _i32p, ctypes.c_int]
            fn.restype = ctypes.c_uint64
        L.oracle_rlev2_decode_width.argtypes = [ctypes.c_uint32]
        L.oracle_rlev2_decode_width.restype = ctypes.c_uint32
        L.oracle_rlev2_closest_width.argtypes = [ctypes.c_uint32]
        L.oracle_rlev2_closest_width.restype = ctypes.c_uint32
        L.oracle_zlib_version.restype = ctypes.c_char_p
        for fn in (L.oracle_zlib_inflate, L.oracle_raw_inflate, L.oracle_gzip_inflate):
            fn.argtypes = [_u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64, _u64p]
        for fn in (L.oracle_crc32, L.oracle_adler32):
            fn.argtypes = [_u8p, ctypes.c_uint64]
            fn.restype = ctypes.c_uint32
        L.oracle_inflate_chunks.argtypes = [_u8p, _u64p, ctypes.c_uint64, ctypes.c_uint64, _u8p,
                                            _i32p, ctypes.c_int]
        L.oracle_inflate_chunks.restype = ctypes.c_uint64
        L.oracle_inflate_batch.argtypes = [_u8p, _u64p, _u64p, ctypes.c_uint64, _u8p, _u64p,
                                           _u64p, _u64p, _i32p, ctypes.c_int]
        L.oracle_inflate_batch.restype = ctypes.c_uint64
        for fn in (L.oracle_brle_decode_chunk, L.oracle_lzss_decode_chunk):
            fn.argtypes = [_u8p, ctypes.c_uint64, _u8p, ctypes.c_uint64, _u64p]
        for fn in (L.oracle_brle_decode_chunks, L.oracle_lzss_decode_chunks):
            fn.argtypes = [_u8p, _u64p, ctypes.c_uint64, ctypes.c_uint64, _u8p, _u64p, _i32p]
            fn.restype = ctypes.c_uint64
        for fn in (L.oracle_brle_encode_chunk, L.oracle_lzss_encode_chunk):
            fn.argtypes = [_u8p, ctypes.c_uint64, _u8p]
            fn.restype = ctypes.c_uint64
        _LIB = L
    return _LIB


def _p(a: np.ndarray, t=_u8p):
    return a.ctypes.data_as(t)


def _bytes(b) -> np.ndarray:
    if isinstance(b, np.ndarray):
        return np.ascontiguousarray(b).view(np.uint8).reshape(-1)
    return np.frombuffer(bytes(b), dtype=np.uint8)


_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def rlev1_decode_chunk(stream, width: int, chunk_bytes: int):
    """Decode one RLE v1 chunk stream; returns (status, values ndarray)."""
    src = _bytes(stream)
    out = np.zeros(chunk_bytes, dtype=np.uint8)
    produced = ctypes.c_uint64(0)
    pad = np.concatenate([src, np.zeros(1, np.uint8)])  # keep a valid pointer for empty input
    rc = lib().oracle_rlev1_decode_chunk(_p(pad), len(src), width, _p(out), chunk_bytes,
                                         ctypes.byref(produced))
    return rc, out[: produced.value].view(_DT[width])


def rlev2_decode_chunk(stream, width: int, cap_values: int):
    """Decode one RLE v2 chunk stream (col_len = len(stream)); returns (status, values)."""
    src = _bytes(stream)
    out = np.zeros(cap_values * width, dtype=np.uint8)
    produced = ctypes.c_uint64(0)
    pad = np.concatenate([src, np.zeros(1, np.uint8)])
    rc = lib().oracle_rlev2_decode_chunk(_p(pad), len(src), width, _p(out), cap_values,
                                         ctypes.byref(produced))
    return rc, out.view(_DT[width])[: produced.value]


def _outbuf(out, nbytes):
    """A caller's preallocated output (timing runs) or a fresh zeroed one."""
    if out is None:
        return np.zeros(nbytes, dtype=np.uint8)
    assert out.dtype == np.uint8 and out.flags.c_contiguous and len(out) >= nbytes
    return out


def rlev1_decode_chunks(comp, blk_off, col_len, width, chunk_bytes, threads=1, out=None):
    """Whole RLE v1 container body; returns (status array, output bytes)."""
    comp = _bytes(comp)
    blk_off = np.ascontiguousarray(blk_off, dtype=np.uint64)
    col_len = np.ascontiguousarray(col_len, dtype=np.uint64)
    n = len(col_len)
    out = _outbuf(out, n * chunk_bytes)
    st = np.zeros(n, dtype=np.int32)
    lib().oracle_rlev1_decode_chunks(_p(comp), _p(blk_off, _u64p), _p(col_len, _u64p), n, width,
                                     chunk_bytes, _p(out), _p(st, _i32p), threads)
    return st, out


def rlev2_decode_chunks(comp, blk_off, col_len, width, chunk_bytes, threads=1, out=None):
    comp = _bytes(comp)
    blk_off = np.ascontiguousarray(blk_off, dtype=np.uint64)
    col_len = np.ascontiguousarray(col_len, dtype=np.uint64)
    n = len(col_len)
    out = _outbuf(out, n * chunk_bytes)
    st = np.zeros(n, dtype=np.int32)
    lib().oracle_rlev2_decode_chunks(_p(comp), _p(blk_off, _u64p), _p(col_len, _u64p), n, width,
                                     chunk_bytes, _p(out), _p(st, _i32p), threads)
    return st, out


def zlib_version() -> str:
    return lib().oracle_zlib_version().decode()


def zlib_inflate(stream, out_cap: int, raw: bool = False, gzip: bool = False):
    src = _bytes(stream)
    pad = np.concatenate([src, np.zeros(1, np.uint8)])
    out = np.zeros(max(out_cap, 1), dtype=np.uint8)
    got = ctypes.c_uint64(0)
    fn = lib().oracle_gzip_inflate if gzip else (
        lib().oracle_raw_inflate if raw else lib().oracle_zlib_inflate)
    rc = fn(_p(pad), len(src), _p(out), out_cap, ctypes.byref(got))
    return rc, out[: got.value]


def crc32(data) -> int:
    d = _bytes(data)
    pad = np.concatenate([d, np.zeros(1, np.uint8)])
    return int(lib().oracle_crc32(_p(pad), len(d)))


def adler32(data) -> int:
    d = _bytes(data)
    pad = np.concatenate([d, np.zeros(1, np.uint8)])
    return int(lib().oracle_adler32(_p(pad), len(d)))


def inflate_chunks(comp, offsets, chunk_bytes, threads=1, out=None):
    """offsets: n+1 entries (stream i = comp[off[i]:off[i+1]])."""
    comp = _bytes(comp)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = len(off) - 1
    out = _outbuf(out, n * chunk_bytes)
    st = np.zeros(n, dtype=np.int32)
    lib().oracle_inflate_chunks(_p(comp), _p(off, _u64p), n, chunk_bytes, _p(out),
                                _p(st, _i32p), threads)
    return st, out


def inflate_batch(src, src_off, src_len, dst_off, dst_cap, dst_total, threads=1, out=None):
    src = _bytes(src)
    src_off = np.ascontiguousarray(src_off, dtype=np.uint64)
    src_len = np.ascontiguousarray(src_len, dtype=np.uint64)
    dst_off = np.ascontiguousarray(dst_off, dtype=np.uint64)
    dst_cap = np.ascontiguousarray(dst_cap, dtype=np.uint64)
    n = len(src_off)
    dst = _outbuf(out, max(dst_total, 1))
    out_len = np.zeros(n, dtype=np.uint64)
    st = np.zeros(n, dtype=np.int32)
    lib().oracle_inflate_batch(_p(src), _p(src_off, _u64p), _p(src_len, _u64p), n, _p(dst),
                               _p(dst_off, _u64p), _p(dst_cap, _u64p), _p(out_len, _u64p),
                               _p(st, _i32p), threads)
    return st, out_len, dst


def rlev2_decode_width(code: int) -> int:
    return int(lib().oracle_rlev2_decode_width(code))


def rlev2_closest_width(bits: int) -> int:
    return int(lib().oracle_rlev2_closest_width(bits))


def _pad(b) -> np.ndarray:
    d = _bytes(b)
    return np.concatenate([d, np.zeros(1, np.uint8)])


def brle_decode_chunk(stream, cap: int):
    """(status, bytes) of one BRLE chunk stream decoded into at most cap bytes."""
    src = _pad(stream)
    out = np.zeros(max(cap, 1), dtype=np.uint8)
    got = ctypes.c_uint64(0)
    st = lib().oracle_brle_decode_chunk(_p(src), len(src) - 1, _p(out), cap, ctypes.byref(got))
    return st, out[: got.value]


def lzss_decode_chunk(stream, cap: int):
    src = _pad(stream)
    out = np.zeros(max(cap, 1), dtype=np.uint8)
    got = ctypes.c_uint64(0)
    st = lib().oracle_lzss_decode_chunk(_p(src), len(src) - 1, _p(out), cap, ctypes.byref(got))
    return st, out[: got.value]


def _byte_chunks(fn, data, idx, n, chunk_bytes, out):
    data = _bytes(data)
    if len(data) == 0:
        data = np.zeros(1, np.uint8)
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    out = _outbuf(out, max(n * chunk_bytes, 1))
    out_len = np.zeros(max(n, 1), dtype=np.uint64)
    st = np.zeros(max(n, 1), dtype=np.int32)
    fn(_p(data), _p(idx, _u64p), n, chunk_bytes, _p(out), _p(out_len, _u64p), _p(st, _i32p))
    return st[:n], out_len[:n], out


def brle_decode_chunks(data, ptr, chunk_bytes, out=None):
    """ptr: n+1 offsets. Returns (status[n], out_len[n], out)."""
    n = len(ptr) - 1
    return _byte_chunks(lib().oracle_brle_decode_chunks, data, ptr, n, chunk_bytes, out)


def lzss_decode_chunks(data, ends, chunk_bytes, out=None):
    """ends: n cumulative stream ends. Returns (status[n], out_len[n], out)."""
    return _byte_chunks(lib().oracle_lzss_decode_chunks, data, ends, len(ends), chunk_bytes, out)


def brle_encode_chunk(raw) -> bytes:
    src = _pad(raw)
    n = len(src) - 1
    out = np.zeros(n + n // 128 + 2, dtype=np.uint8)
    k = lib().oracle_brle_encode_chunk(_p(src), n, _p(out))
    return out[:k].tobytes()


def lzss_encode_chunk(raw) -> bytes:
    src = _pad(raw)
    n = len(src) - 1
    out = np.zeros(2 * n + n // 8 + 2, dtype=np.uint8)
    k = lib().oracle_lzss_encode_chunk(_p(src), n, _p(out))
    return out[:k].tobytes()
