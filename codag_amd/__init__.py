"""codag_amd -- MI355X-native CODAG decompressors (RLE v1, RLE v2, Deflate, BRLE, LZSS).

Python host mirror of the C ABI in ``include/codag.h`` (``lib/libcodag.so``),
loaded with ctypes. The GPU work happens in the library's gfx950 HIP kernels;
this module only moves buffers and calls the entry points. It never falls back
to a CPU decoder: if the library is missing, :func:`lib` raises.

The sub-namespaces mirror the reference's host interfaces:

* :class:`rle_v1`  -- ``rle_v1::compress_gpu_orig`` / ``decompress_gpu_orig``
  (include/rlev1/rlev1.h:103-192, 419-592)
* :class:`rlev2`   -- ``rlev2::compress_gpu_orig`` / ``decompress_gpu_orig``
  (include/rlev2/encoder_trans.h:482-564, decoder_trans.h:57-133)
* :class:`deflate` -- ``deflate::decompress_gpu`` (include/zlib/zlib.h:1583-1700)
  and the container writer of deflate_compressor.py:44-126
* :class:`brle`    -- ``brle::compress_gpu`` / ``decompress_gpu`` (include/brle/brle.h:251-342)
* :class:`lzss`    -- ``lzss::compress_gpu`` / ``decompress_gpu_base``
  (include/lzss/lzss.h:1060-1189)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcodag.so")
INCLUDE_DIR = os.path.join(os.path.dirname(_HERE), "include")

# codag_status
OK, ERR_INVALID, ERR_HIP, ERR_NOMEM, ERR_IO, ERR_FORMAT, ERR_DECODE, ERR_BUSY, \
    ERR_UNSUPPORTED, ERR_TIMEOUT = range(10)
# per-chunk status
CHUNK_OK, CHUNK_TRUNCATED, CHUNK_OVERFLOW, CHUNK_CORRUPT = range(4)
# inflate flags
INFLATE_SKIP_ADLER = 1
INFLATE_RAW = 2
INFLATE_GZIP = 4
INFLATE_CUDF_STATUS = 8
# codecs
RLEV1, RLEV2, DEFLATE = 1, 2, 3
# server request flags
REQ_DEVICE_IO = 1

CHUNK_BYTES = 131072
BRLE_CHUNK_BYTES = 262144  # include/brle/brle.h:8
LZSS_CHUNK_BYTES = 131072  # include/lzss/lzss.h:21

_LIB = None
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i32p = ctypes.POINTER(ctypes.c_int32)
_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32


class CodagError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        msg = lib().codag_last_error().decode(errors="replace")
        name = lib().codag_status_string(status).decode()
        super().__init__(f"{what}: {name} ({msg})")


def build(jobs: int = 8) -> str:
    """Compile libcodag.so and the CLIs with hipcc for gfx950 (no GPU needed)."""
    subprocess.run(["make", "-s", "-C", _HERE, f"-j{jobs}"], check=True)
    return LIB_PATH


class TuneParams(ctypes.Structure):
    _fields_ = [("threads_per_group", _u32), ("window_bytes", _u32), ("groups_per_cu", _u32),
                ("queue_depth", _u32)]


class StreamDesc(ctypes.Structure):
    _fields_ = [("src", _vp), ("src_bytes", _u64), ("dst", _vp), ("dst_cap", _u64)]


class StreamResult(ctypes.Structure):
    _fields_ = [("bytes_written", _u64), ("status", ctypes.c_int32), ("adler32", _u32)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget_bytes", _u64), ("max_inflight", _u32)]


class Request(ctypes.Structure):
    _fields_ = [("codec", ctypes.c_int), ("flags", _u32), ("elem_bytes", _u32), ("chunk_bytes", _u64),
                ("n_chunks", _u64), ("comp", _vp), ("comp_bytes", _u64), ("blk_off", _vp),
                ("col_len", _vp), ("out", _vp), ("user_tag", _u64)]


class Completion(ctypes.Structure):
    _fields_ = [("request_id", _u64), ("user_tag", _u64), ("status", ctypes.c_int),
                ("failed_chunks", _u32), ("queue_seconds", ctypes.c_double),
                ("latency_seconds", ctypes.c_double)]


def _sig(L, name, res, args):
    f = getattr(L, name)
    f.restype = res
    f.argtypes = args


def lib():
    """The loaded libcodag.so. Raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with codag_amd.build() "
                               "(the HIP decode path has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        st = ctypes.c_int
        pp = ctypes.POINTER(_vp)
        p64 = ctypes.POINTER(_u64)
        _sig(L, "codag_version", ctypes.c_char_p, [])
        _sig(L, "codag_status_string", ctypes.c_char_p, [st])
        _sig(L, "codag_last_error", ctypes.c_char_p, [])
        _sig(L, "codag_device_count", ctypes.c_int, [])
        _sig(L, "codag_ctx_create", st, [ctypes.c_int, pp])
        _sig(L, "codag_ctx_destroy", st, [_vp])
        _sig(L, "codag_ctx_stream", _vp, [_vp])
        dec = [_vp, _vp, _u64, _vp, _vp, _u64, _u32, _u64, _vp, _vp, _vp]
        _sig(L, "codag_rlev1_decode", st, dec)
        _sig(L, "codag_rlev2_decode", st, dec)
        _sig(L, "codag_inflate_chunks", st, [_vp, _vp, _u64, _vp, _u64, _u64, _vp, _vp, _vp, _u32, _vp])
        _sig(L, "codag_inflate_batch", st, [_vp, _vp, _u64, _vp, _u32, _vp])
        _sig(L, "codag_inflate_batch_noalloc", st, [_vp, _u64, _vp, _u32, _u32, _vp])
        _sig(L, "codag_test_chunks_two_wave", st, [_vp, _u32])
        _sig(L, "codag_brle_decode", st, [_vp, _vp, _u64, _vp, _u64, _u64, _vp, _vp, _vp, _vp])
        _sig(L, "codag_lzss_decode", st, [_vp, _vp, _u64, _vp, _u64, _u64, _vp, _vp, _vp, _vp])
        dd = ctypes.POINTER(ctypes.c_double)
        _sig(L, "codag_rlev1_decode_host", st, [_vp, _vp, _u64, _vp, _u64, _vp, _u64, _u32, _u64,
                                                _vp, _u64, p64, dd, dd])
        _sig(L, "codag_rlev2_decode_host", st, [_vp, _vp, _u64, _vp, _vp, _u64, _u32, _u64, _vp,
                                                _u64, p64, dd, dd])
        _sig(L, "codag_inflate_chunks_host", st, [_vp, _vp, _u64, _vp, _u64, _u64, _u32, _vp, _u64,
                                                  p64, dd, dd])
        _sig(L, "codag_brle_decode_host", st, [_vp, _vp, _u64, _u64, _vp, _u64, p64, dd, dd])
        _sig(L, "codag_lzss_decode_host", st, [_vp, _vp, _u64, _vp, _u64, p64, dd, dd])
        _sig(L, "codag_host_free", None, [_vp])
        _sig(L, "codag_brle_encode", st, [_vp, _u64, _u64, ctypes.c_int, pp, p64])
        _sig(L, "codag_lzss_encode", st, [_vp, _u64, _u64, ctypes.c_int, pp, p64])
        _sig(L, "codag_synth_byte_payload", st, [_u64, _u32, _u64, pp])
        _sig(L, "codag_rlev1_encode", st, [_vp, _u64, _u32, _u64, pp, p64, pp, p64, pp, p64, p64])
        _sig(L, "codag_rlev2_encode", st, [_vp, _u64, _u32, _u64, pp, p64, pp, pp, p64])
        _sig(L, "codag_deflate_compress_chunks", st, [_vp, _u64, _u64, ctypes.c_int, ctypes.c_int,
                                                      pp, p64, pp, pp, p64])
        _sig(L, "codag_synth_rlev1", st, [_u64, _u32, _u64, _u64, _u64, pp, pp, p64, pp, pp, p64])
        _sig(L, "codag_synth_rlev2", st, [_u64, _u32, _u64, _u64, pp, pp, p64, pp, pp, p64])
        _sig(L, "codag_synth_deflate_payload", st, [_u64, _u64, _u64, pp])
        _sig(L, "codag_deflate_compress_mixed", st, [_vp, _u64, _u32, _u32, _u64, _u64, pp, p64,
                                                     ctypes.POINTER(_u64 * 3)])
        _sig(L, "codag_tune_get", st, [_vp, ctypes.c_int, ctypes.POINTER(TuneParams)])
        _sig(L, "codag_tune_set", st, [_vp, ctypes.c_int, ctypes.POINTER(TuneParams)])
        _sig(L, "codag_tune_run", st, [_vp, ctypes.c_int, _vp, _u64, _vp, _vp, _u64, _u32, _u64,
                                       _vp, ctypes.c_int, ctypes.POINTER(TuneParams),
                                       ctypes.POINTER(ctypes.c_float)])
        _sig(L, "codag_server_create", st, [ctypes.POINTER(ServerConfig), pp])
        _sig(L, "codag_server_submit", st, [_vp, ctypes.POINTER(Request), p64])
        _sig(L, "codag_server_poll", st, [_vp, ctypes.POINTER(Completion), _u32,
                                          ctypes.POINTER(_u32), ctypes.c_int])
        _sig(L, "codag_server_stats", st, [_vp, p64, ctypes.POINTER(_u32), ctypes.POINTER(_u32)])
        _sig(L, "codag_server_destroy", st, [_vp])
        _sig(L, "codag_dev_alloc", st, [_vp, _u64, pp])
        _sig(L, "codag_dev_free", st, [_vp, _vp])
        _sig(L, "codag_copy_h2d", st, [_vp, _vp, _vp, _u64])
        _sig(L, "codag_copy_d2h", st, [_vp, _vp, _vp, _u64])
        _sig(L, "codag_memset_d", st, [_vp, _vp, ctypes.c_int, _u64])
        _sig(L, "codag_copy_d2d", st, [_vp, _vp, _vp, _u64])
        _sig(L, "codag_sync", st, [_vp])
        _sig(L, "codag_stream_create", st, [_vp, pp])
        _sig(L, "codag_stream_destroy", st, [_vp, _vp])
        _sig(L, "codag_stream_sync", st, [_vp, _vp])
        _sig(L, "codag_timer_start", st, [_vp])
        _sig(L, "codag_timer_stop", st, [_vp, ctypes.POINTER(ctypes.c_float)])
        _sig(L, "codag_device_info", st, [_vp, ctypes.c_char_p, _u32, ctypes.POINTER(_u32)])
        _sig(L, "codag_event_record", st, [_vp, _u32, _vp])
        _sig(L, "codag_event_elapsed", st, [_vp, _u32, _u32, ctypes.POINTER(ctypes.c_float)])
        _LIB = L
    return _LIB


def check(status: int, what: str):
    if status != OK:
        raise CodagError(status, what)


def _ptr(a: np.ndarray):
    return ctypes.c_void_p(a.ctypes.data)


def _as_u8(a) -> np.ndarray:
    if isinstance(a, np.ndarray):
        return np.ascontiguousarray(a).view(np.uint8).reshape(-1)
    return np.frombuffer(bytes(a), dtype=np.uint8)


def _take(p: ctypes.c_void_p, nbytes: int, dtype=np.uint8) -> np.ndarray:
    """Copy a library-allocated host buffer into numpy and free it."""
    if not p.value:
        return np.zeros(0, dtype=dtype)
    buf = (ctypes.c_uint8 * nbytes).from_address(p.value)
    out = np.frombuffer(bytearray(buf), dtype=np.uint8).view(dtype).copy()
    lib().codag_host_free(p)
    return out


# ---------------------------------------------------------------- device side

def device_count() -> int:
    return int(lib().codag_device_count())


class Context:
    """One device + one HIP stream + tuned launch parameters (codag_ctx)."""

    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        check(lib().codag_ctx_create(device, ctypes.byref(self.h)), "codag_ctx_create")

    def close(self):
        if self.h:
            lib().codag_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # plumbing
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def upload(self, arr) -> "DeviceBuffer":
        a = _as_u8(arr)
        b = DeviceBuffer(self, max(len(a), 1))
        b.nbytes = len(a)
        if len(a):
            check(lib().codag_copy_h2d(self.h, b.ptr, _ptr(a), len(a)), "copy_h2d")
        return b

    def sync(self):
        check(lib().codag_sync(self.h), "sync")

    def stream_create(self) -> ctypes.c_void_p:
        """A new HIP stream of this context's device (codag_stream_create)."""
        s = ctypes.c_void_p()
        check(lib().codag_stream_create(self.h, ctypes.byref(s)), "stream_create")
        return s

    def stream_destroy(self, stream):
        check(lib().codag_stream_destroy(self.h, stream), "stream_destroy")

    def stream_sync(self, stream=None):
        check(lib().codag_stream_sync(self.h, stream), "stream_sync")

    def copy(self, dst, src, nbytes: int):
        """Asynchronous device-to-device copy by the library's plain copy kernel."""
        check(lib().codag_copy_d2d(self.h, _dp(dst), _dp(src), nbytes), "copy_d2d")

    def timer_start(self):
        check(lib().codag_timer_start(self.h), "timer_start")

    def timer_stop(self) -> float:
        ms = ctypes.c_float()
        check(lib().codag_timer_stop(self.h, ctypes.byref(ms)), "timer_stop")
        return float(ms.value)

    def event(self, slot: int):
        check(lib().codag_event_record(self.h, slot, None), "event_record")

    def elapsed(self, a: int, b: int) -> float:
        ms = ctypes.c_float()
        check(lib().codag_event_elapsed(self.h, a, b, ctypes.byref(ms)), "event_elapsed")
        return float(ms.value)

    def info(self):
        name = ctypes.create_string_buffer(256)
        cu = _u32()
        check(lib().codag_device_info(self.h, name, 256, ctypes.byref(cu)), "device_info")
        return name.value.decode(), int(cu.value)

    # decode entry points (device pointers)
    def rlev1_decode(self, comp, comp_bytes, blk_off, col_len, n, w, chunk, out, status=None):
        check(lib().codag_rlev1_decode(self.h, _dp(comp), comp_bytes, _dp(blk_off), _dp(col_len),
                                       n, w, chunk, _dp(out), _dp(status), None), "rlev1_decode")

    def rlev2_decode(self, comp, comp_bytes, blk_off, col_len, n, w, chunk, out, status=None):
        check(lib().codag_rlev2_decode(self.h, _dp(comp), comp_bytes, _dp(blk_off), _dp(col_len),
                                       n, w, chunk, _dp(out), _dp(status), None), "rlev2_decode")

    def inflate_chunks(self, comp, comp_bytes, offsets, n, chunk, out, out_len=None, status=None,
                       flags=0):
        check(lib().codag_inflate_chunks(self.h, _dp(comp), comp_bytes, _dp(offsets), n, chunk,
                                         _dp(out), _dp(out_len), _dp(status), flags, None),
              "inflate_chunks")

    def inflate_batch(self, descs, n, results, flags=0, stream=None):
        check(lib().codag_inflate_batch(self.h, _dp(descs), n, _dp(results), flags, stream),
              "inflate_batch")

    def brle_decode(self, comp, comp_bytes, ptr, n, chunk, out, out_len=None, status=None,
                    stream=None):
        check(lib().codag_brle_decode(self.h, _dp(comp), comp_bytes, _dp(ptr), n, chunk, _dp(out),
                                      _dp(out_len), _dp(status), stream), "brle_decode")

    def lzss_decode(self, comp, comp_bytes, ends, n, chunk, out, out_len=None, status=None,
                    stream=None):
        check(lib().codag_lzss_decode(self.h, _dp(comp), comp_bytes, _dp(ends), n, chunk, _dp(out),
                                      _dp(out_len), _dp(status), stream), "lzss_decode")

    def tune_get(self, codec) -> TuneParams:
        p = TuneParams()
        check(lib().codag_tune_get(self.h, codec, ctypes.byref(p)), "tune_get")
        return p

    def tune_set(self, codec, p: TuneParams):
        check(lib().codag_tune_set(self.h, codec, ctypes.byref(p)), "tune_set")

    def tune_run(self, codec, comp, comp_bytes, blk_off, col_len, n, w, chunk, out, reps=3):
        best = TuneParams()
        ms = ctypes.c_float()
        check(lib().codag_tune_run(self.h, codec, _dp(comp), comp_bytes, _dp(blk_off),
                                   _dp(col_len), n, w, chunk, _dp(out), reps, ctypes.byref(best),
                                   ctypes.byref(ms)), "tune_run")
        return best, float(ms.value)


def _dp(b):
    if b is None:
        return None
    if isinstance(b, DeviceBuffer):
        return b.ptr
    return ctypes.c_void_p(int(b))


class DeviceBuffer:
    """A device allocation owned by a Context (hipMalloc through the C ABI)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = nbytes
        self.ptr = ctypes.c_void_p()
        check(lib().codag_dev_alloc(ctx.h, max(nbytes, 1), ctypes.byref(self.ptr)), "dev_alloc")

    def download(self, nbytes=None, dtype=np.uint8) -> np.ndarray:
        n = self.nbytes if nbytes is None else nbytes
        out = np.empty(n, dtype=np.uint8)
        if n:
            check(lib().codag_copy_d2h(self.ctx.h, _ptr(out), self.ptr, n), "copy_d2h")
        return out.view(dtype)

    def fill(self, value: int = 0):
        check(lib().codag_memset_d(self.ctx.h, self.ptr, value, max(self.nbytes, 1)), "memset_d")

    def offset(self, k: int) -> int:
        return self.ptr.value + k

    def free(self):
        if self.ptr and self.ctx.h:
            lib().codag_dev_free(self.ctx.h, self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


# ---------------------------------------------------------------- producers

def _c():
    return ctypes.c_void_p()


def rlev1_encode(raw, elem_bytes: int, chunk_bytes: int = CHUNK_BYTES):
    """Returns (container u8, blk_off file u64, col_len file u64, n_chunks)."""
    a = _as_u8(raw)
    c, b, l = _c(), _c(), _c()
    cn, bn, ln, n = _u64(), _u64(), _u64(), _u64()
    check(lib().codag_rlev1_encode(_ptr(a) if len(a) else None, len(a), elem_bytes, chunk_bytes,
                                   ctypes.byref(c), ctypes.byref(cn), ctypes.byref(b),
                                   ctypes.byref(bn), ctypes.byref(l), ctypes.byref(ln),
                                   ctypes.byref(n)), "rlev1_encode")
    return (_take(c, cn.value), _take(b, bn.value * 8, np.uint64), _take(l, ln.value * 8, np.uint64),
            n.value)


def rlev2_encode(raw, elem_bytes: int, chunk_bytes: int = CHUNK_BYTES):
    """Returns (comp u8, blk_off u64 [n+1, last = raw bytes], col_len u64 [n], n)."""
    a = _as_u8(raw)
    c, b, l = _c(), _c(), _c()
    cn, n = _u64(), _u64()
    check(lib().codag_rlev2_encode(_ptr(a) if len(a) else None, len(a), elem_bytes, chunk_bytes,
                                   ctypes.byref(c), ctypes.byref(cn), ctypes.byref(b),
                                   ctypes.byref(l), ctypes.byref(n)), "rlev2_encode")
    return (_take(c, cn.value), _take(b, (n.value + 1) * 8, np.uint64),
            _take(l, n.value * 8, np.uint64), n.value)


def deflate_compress_chunks(raw, chunk_bytes: int = CHUNK_BYTES, level: int = 6, threads: int = 0):
    """deflate_compressor.py container: (comp, blk_offset file [n+2], col_len [n], n)."""
    a = _as_u8(raw)
    c, b, l = _c(), _c(), _c()
    cn, n = _u64(), _u64()
    check(lib().codag_deflate_compress_chunks(_ptr(a) if len(a) else None, len(a), chunk_bytes,
                                              level, threads, ctypes.byref(c), ctypes.byref(cn),
                                              ctypes.byref(b), ctypes.byref(l), ctypes.byref(n)),
          "deflate_compress_chunks")
    return (_take(c, cn.value), _take(b, (n.value + 2) * 8, np.uint64),
            _take(l, n.value * 8, np.uint64), n.value)


def deflate_compress_mixed(raw, pct_fixed=20, pct_stored=5, block_bytes=32768, seed=1):
    a = _as_u8(raw)
    s, sn = _c(), _u64()
    counts = (_u64 * 3)()
    check(lib().codag_deflate_compress_mixed(_ptr(a) if len(a) else None, len(a), pct_fixed,
                                             pct_stored, block_bytes, seed, ctypes.byref(s),
                                             ctypes.byref(sn), ctypes.byref(counts)),
          "deflate_compress_mixed")
    return _take(s, sn.value), tuple(int(x) for x in counts)


def synth_rlev1(n_values, elem_bytes=8, chunk_bytes=CHUNK_BYTES, value_range=1 << 16, seed=0xC0DA7):
    r, c, b, l = _c(), _c(), _c(), _c()
    cn, n = _u64(), _u64()
    check(lib().codag_synth_rlev1(n_values, elem_bytes, chunk_bytes, value_range, seed,
                                  ctypes.byref(r), ctypes.byref(c), ctypes.byref(cn),
                                  ctypes.byref(b), ctypes.byref(l), ctypes.byref(n)), "synth_rlev1")
    nc = n.value
    raw = _take(r, nc * chunk_bytes)[: n_values * elem_bytes]
    return raw, _take(c, cn.value), _take(b, (nc * 32 + 1) * 8, np.uint64), \
        _take(l, nc * 8, np.uint64), nc


def synth_rlev2(n_values, elem_bytes=8, chunk_bytes=CHUNK_BYTES, seed=0xC0DA8):
    r, c, b, l = _c(), _c(), _c(), _c()
    cn, n = _u64(), _u64()
    check(lib().codag_synth_rlev2(n_values, elem_bytes, chunk_bytes, seed, ctypes.byref(r),
                                  ctypes.byref(c), ctypes.byref(cn), ctypes.byref(b),
                                  ctypes.byref(l), ctypes.byref(n)), "synth_rlev2")
    nc = n.value
    raw = _take(r, nc * chunk_bytes)[: n_values * elem_bytes]
    return raw, _take(c, cn.value), _take(b, (nc + 1) * 8, np.uint64), \
        _take(l, nc * 8, np.uint64), nc


def brle_encode(raw, chunk_bytes=BRLE_CHUNK_BYTES, threads=0) -> np.ndarray:
    """brle::compress_gpu's container (u8): [u32 n+1][u64 ptr[n+1]][u64 len][streams]."""
    a = _as_u8(raw)
    c, cn = _c(), _u64()
    check(lib().codag_brle_encode(_ptr(a) if len(a) else None, len(a), chunk_bytes, threads,
                                  ctypes.byref(c), ctypes.byref(cn)), "brle_encode")
    return _take(c, cn.value)


def lzss_encode(raw, chunk_bytes=LZSS_CHUNK_BYTES, threads=0) -> np.ndarray:
    """lzss::compress_gpu's container (u8): [u64 len][u32 x5][u64 ends[n]][streams]."""
    a = _as_u8(raw)
    c, cn = _c(), _u64()
    check(lib().codag_lzss_encode(_ptr(a) if len(a) else None, len(a), chunk_bytes, threads,
                                  ctypes.byref(c), ctypes.byref(cn)), "lzss_encode")
    return _take(c, cn.value)


def brle_container(container):
    """Splits a BRLE container into (data u8, ptr u64 [n+1], original length)."""
    c = _as_u8(container)
    n_ptr = int(c[:4].view(np.uint32)[0])
    ptr = c[4:4 + 8 * n_ptr].view(np.uint64)
    orig = int(c[4 + 8 * n_ptr:12 + 8 * n_ptr].view(np.uint64)[0])
    return c[12 + 8 * n_ptr:], ptr, orig


def lzss_container(container):
    """Splits an LZSS container into (data u8, ends u64 [n], chunk bytes, original length)."""
    c = _as_u8(container)
    orig = int(c[:8].view(np.uint64)[0])
    chunk, n = (int(x) for x in c[8:16].view(np.uint32))
    ends = c[28:28 + 8 * n].view(np.uint64)
    return c[28 + 8 * n:], ends, chunk, orig


def synth_byte_payload(n_bytes, kind, seed=0xC0DAB):
    """kind 0: BRLE benchmark column (runs and literal stretches); 1: LZSS text."""
    r = _c()
    check(lib().codag_synth_byte_payload(n_bytes, kind, seed, ctypes.byref(r)), "synth_byte_payload")
    return _take(r, n_bytes)


def synth_deflate_payload(n_bytes, part_bytes=CHUNK_BYTES, seed=0xC0DA9):
    r = _c()
    check(lib().codag_synth_deflate_payload(n_bytes, part_bytes, seed, ctypes.byref(r)),
          "synth_deflate_payload")
    return _take(r, n_bytes)


# ------------------------------------------------- reference-shaped host API

class rle_v1:
    """Mirror of namespace rle_v1 (include/rlev1/rlev1.h)."""

    @staticmethod
    def compress_gpu_orig(raw, elem_bytes, chunk_size=CHUNK_BYTES):
        return rlev1_encode(raw, elem_bytes, chunk_size)

    @staticmethod
    def decompress_gpu_orig(ctx: Context, container, blk_off_file, col_len_file, elem_bytes,
                            chunk_size=CHUNK_BYTES):
        """Returns (decoded bytes, kernel seconds). Raises CodagError on failure."""
        c = _as_u8(container)
        b = np.ascontiguousarray(blk_off_file, dtype=np.uint64)
        l = np.ascontiguousarray(col_len_file, dtype=np.uint64)
        orig = int(c[:8].view(np.uint64)[0]) if len(c) >= 8 else 0
        out = np.empty(max(orig, 1), dtype=np.uint8)
        on, ks, ts = _u64(), ctypes.c_double(), ctypes.c_double()
        check(lib().codag_rlev1_decode_host(ctx.h, _ptr(c), len(c), _ptr(b), len(b), _ptr(l), len(l),
                                            elem_bytes, chunk_size, _ptr(out), len(out),
                                            ctypes.byref(on), ctypes.byref(ks), ctypes.byref(ts)),
              "rlev1_decode_host")
        return out[: on.value], ks.value


class rlev2:
    """Mirror of namespace rlev2 (include/rlev2/encoder_trans.h, decoder_trans.h)."""

    @staticmethod
    def compress_gpu_orig(raw, elem_bytes, chunk_size=CHUNK_BYTES):
        return rlev2_encode(raw, elem_bytes, chunk_size)

    @staticmethod
    def decompress_gpu_orig(ctx: Context, comp, n_chunks, blk_off, col_len, elem_bytes,
                            chunk_size=CHUNK_BYTES):
        c = _as_u8(comp)
        b = np.ascontiguousarray(blk_off, dtype=np.uint64)
        l = np.ascontiguousarray(col_len, dtype=np.uint64)
        orig = int(b[n_chunks])
        out = np.empty(max(orig, 1), dtype=np.uint8)
        on, ks, ts = _u64(), ctypes.c_double(), ctypes.c_double()
        cp = _ptr(c) if len(c) else None
        check(lib().codag_rlev2_decode_host(ctx.h, cp, len(c), _ptr(b), _ptr(l), n_chunks,
                                            elem_bytes, chunk_size, _ptr(out), len(out),
                                            ctypes.byref(on), ctypes.byref(ks), ctypes.byref(ts)),
              "rlev2_decode_host")
        return out[: on.value], ks.value


class deflate:
    """Mirror of namespace deflate (include/zlib/zlib.h) + the container writer."""

    @staticmethod
    def compress_chunks(raw, chunk_size=CHUNK_BYTES, level=6, threads=0):
        return deflate_compress_chunks(raw, chunk_size, level, threads)

    @staticmethod
    def decompress_gpu(ctx: Context, comp, blk_offset_file, chunk_size=CHUNK_BYTES, flags=0):
        c = _as_u8(comp)
        b = np.ascontiguousarray(blk_offset_file, dtype=np.uint64)
        orig = int(b[0])
        out = np.empty(max(orig, 1), dtype=np.uint8)
        on, ks, ts = _u64(), ctypes.c_double(), ctypes.c_double()
        cp = _ptr(c) if len(c) else None
        check(lib().codag_inflate_chunks_host(ctx.h, cp, len(c), _ptr(b), len(b), chunk_size, flags,
                                              _ptr(out), len(out), ctypes.byref(on),
                                              ctypes.byref(ks), ctypes.byref(ts)),
              "inflate_chunks_host")
        return out[: on.value], ks.value


class brle:
    """Mirror of namespace brle (include/brle/brle.h)."""

    @staticmethod
    def compress_gpu(raw, chunk_size=BRLE_CHUNK_BYTES, threads=0):
        return brle_encode(raw, chunk_size, threads)

    @staticmethod
    def decompress_gpu(ctx: Context, container, chunk_size=BRLE_CHUNK_BYTES):
        """Returns (decoded bytes, kernel seconds). Raises CodagError on failure."""
        c = _as_u8(container)
        _, _, orig = brle_container(c)
        out = np.empty(max(orig, 1), dtype=np.uint8)
        on, ks, ts = _u64(), ctypes.c_double(), ctypes.c_double()
        check(lib().codag_brle_decode_host(ctx.h, _ptr(c), len(c), chunk_size, _ptr(out), len(out),
                                           ctypes.byref(on), ctypes.byref(ks), ctypes.byref(ts)),
              "brle_decode_host")
        return out[: on.value], ks.value


class lzss:
    """Mirror of namespace lzss (include/lzss/lzss.h)."""

    @staticmethod
    def compress_gpu(raw, chunk_size=LZSS_CHUNK_BYTES, threads=0):
        return lzss_encode(raw, chunk_size, threads)

    @staticmethod
    def decompress_gpu_base(ctx: Context, container):
        c = _as_u8(container)
        orig = int(c[:8].view(np.uint64)[0]) if len(c) >= 8 else 0
        out = np.empty(max(orig, 1), dtype=np.uint8)
        on, ks, ts = _u64(), ctypes.c_double(), ctypes.c_double()
        check(lib().codag_lzss_decode_host(ctx.h, _ptr(c), len(c), _ptr(out), len(out),
                                           ctypes.byref(on), ctypes.byref(ks), ctypes.byref(ts)),
              "lzss_decode_host")
        return out[: on.value], ks.value
