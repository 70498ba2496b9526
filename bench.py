#!/usr/bin/env python3
"""Benchmark of the MI355X CODAG decoders (one GPU).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
                    [--config all|c1|c2|c3|c4|c5|brle|lzss]

Prints ONE JSON line. With the default ``--config all`` the line is headlined
on the largest single-GPU workload -- the batched uneven Deflate decode of
10,000 zlib streams (4 KiB..4 MiB) -- and carries the other four workloads
under ``config.configs``:

* c1  ORC RLE v1 decode of a 2^24-value uint64 column (128 MiB), 128 KiB chunks
* c2  ORC RLE v2 decode of a uint32 and a uint64 column, 2^26 values each
* c3  Deflate decode of 1 GiB in 8192 zlib level-6 chunks of 128 KiB
* c4  batched uneven Deflate (the headline)
* c5  served mixed-codec requests, Poisson arrivals, p50/p99 latency
* brle  byte RLE decode of 1 GiB in 4096 chunks of 256 KiB (brle.h's format)
* lzss  LZSS decode of 1 GiB in 8192 chunks of 128 KiB (lzss.h's format)

A step is one pass of the hot path over the whole workload with the
compressed input already resident in HBM. Every workload's GPU output is
checked bit-exact (byte compare + CRC32) against the raw data before the
warm-up and again after the timed steps.

Roofline: the dominant kernel's algorithmic bytes per launch (compressed
bytes read + decoded bytes written + offset/length metadata) divided by its
mean launch time, measured with HIP events on the stream the kernel runs on,
against MI355X's 8 TB/s HBM peak. ``traffic`` is the HBM bytes per launch that
rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE measured for the same kernel and
workload (profiles/traffic.json, scaled by a plain copy kernel).

cpu_baseline: the CPU restatement of the same decoder (system zlib inflate
for Deflate), multithreaded over chunks/streams on every core this process
may run on, decoding RAM to preallocated RAM; also a single-thread figure.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import codag_amd as ca  # noqa: E402

HBM_PEAK_GBS = 8000.0
CH = 131072
METRIC = "decompressed GB/s + % HBM peak (bit-exact), 1 GPU; served p50/p99 latency"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_threads() -> int:
    """Every core this process may run on."""
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:
        return max(1, os.cpu_count() or 1)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def crc32(b: np.ndarray) -> int:
    import zlib
    return zlib.crc32(memoryview(np.ascontiguousarray(b).view(np.uint8))) & 0xFFFFFFFF


def traffic_from_profiles(kernel_key: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any.

    Keys are "<kernel>:<workload>:<workgroups>"; the bytes come from FETCH_SIZE
    and WRITE_SIZE scaled by the same counters of the plain copy kernel."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        return t.get(kernel_key)
    except (OSError, ValueError):
        return None


def roofline(kernel, kernel_ms, alg_bytes, traffic_key):
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    tr = traffic_from_profiles(traffic_key)
    r = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": tr, "kernel": kernel,
         "kernel_ms": round(kernel_ms, 4), "alg_bytes_per_launch": int(alg_bytes)}
    if tr:
        r["traffic_over_alg"] = round(tr / alg_bytes, 3)
    return r


def time_cpu(fn, min_runs=5, budget_s=2.0):
    """Median wall time of fn() after one warm-up run: at least min_runs runs,
    more while the budget lasts."""
    fn()
    times = []
    t0 = time.perf_counter()
    while len(times) < min_runs or (time.perf_counter() - t0 < budget_s and len(times) < 50):
        t1 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t1)
    return float(np.median(times)), len(times)


def cpu_leg(multi, multi_bytes, multi_desc, single, single_bytes, single_desc, what):
    """cpu_baseline object: `multi` on every core, `single` on one thread."""
    threads = host_threads()
    tm, nm = time_cpu(multi)
    ts, ns = time_cpu(single, min_runs=5, budget_s=1.0)
    return {"value": round(multi_bytes / tm / 1e9, 3), "unit": "GB/s", "cores": threads,
            "kind": "port", "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "single_thread_gbs": round(single_bytes / ts / 1e9, 4),
            "sample": f"{multi_desc}, {what} on {threads} threads (all cores of this process), "
                      f"median of {nm} runs after a warm-up; single thread: {single_desc}, "
                      f"median of {ns}"}


# ----------------------------------------------------------------- RLE (c1, c2)

class Column:
    """One device-resident RLE container."""

    def __init__(self, ctx, codec, width, raw, comp, blk, cl, n):
        self.codec, self.width, self.n = codec, width, n
        self.raw = raw
        self.comp_bytes = len(comp)
        self.d_comp = ctx.upload(comp)
        self.d_off = ctx.upload(blk[:n].astype(np.uint64))
        self.d_len = ctx.upload(cl[:n].astype(np.uint64))
        self.d_out = ctx.alloc(n * CH)
        self.d_st = ctx.alloc(4 * n)
        self.host = (comp, blk[:n].copy(), cl[:n].copy())
        self.alg_bytes = len(comp) + n * CH + 16 * n
        self.tune = None  # launch shape chosen by the autotuner for this column

    def autotune(self, ctx):
        self.tune, ms = ctx.tune_run(self.codec, self.d_comp, self.comp_bytes, self.d_off,
                                     self.d_len, self.n, self.width, CH, self.d_out)
        return ms

    def shape(self):
        t = self.tune
        return None if t is None else [t.threads_per_group, t.window_bytes, t.queue_depth]

    def kernel(self, ctx):
        """Name of the kernel the column's launch shape selects."""
        t = self.tune or ctx.tune_get(self.codec)
        if self.codec == ca.RLEV1:
            if t.queue_depth == 4 and t.window_bytes == 2048:
                return "rlev1_pipe_kernel"
            return "rlev1_pc_kernel" if t.queue_depth == 2 else "rlev1_decode_kernel"
        if t.queue_depth == 2 and t.threads_per_group in (128, 192, 256):
            return "rlev2_pc_kernel"
        return "rlev2_wave_kernel" if t.threads_per_group == 64 else "rlev2_decode_kernel"

    def launch(self, ctx):
        if self.tune is not None:  # host-side parameter switch, no device work
            ctx.tune_set(self.codec, self.tune)
        fn = ctx.rlev1_decode if self.codec == ca.RLEV1 else ctx.rlev2_decode
        fn(self.d_comp, self.comp_bytes, self.d_off, self.d_len, self.n, self.width, CH,
           self.d_out, self.d_st)

    def verify(self):
        st = self.d_st.download(4 * self.n, np.int32)
        out = self.d_out.download(len(self.raw))
        ok = bool(np.all(st == 0)) and out.tobytes() == self.raw.tobytes()
        return ok, crc32(out)


def make_rle(ctx, cfg):
    cols = []
    if cfg == "c1":
        raw, cont, blk, cl, n = ca.synth_rlev1(1 << 24, 8, CH, 1 << 16, seed=0xC0DA6 + 1)
        cols.append(Column(ctx, ca.RLEV1, 8, raw, cont[8:], blk, cl, n))
        desc = "rlev1 uint64, 2^24 values (128 MiB), 128 KiB chunks"
        dtype = "uint64"
    else:
        for w in (4, 8):
            raw, comp, blk, cl, n = ca.synth_rlev2(1 << 26, w, CH, seed=0xC0DA6 + 2 + w)
            cols.append(Column(ctx, ca.RLEV2, w, raw, comp, blk, cl, n))
        desc = "rlev2 uint32 + uint64 columns, 2^26 values each (256 + 512 MiB), 128 KiB chunks"
        dtype = "uint32+uint64"
    return cols, desc, dtype


def cpu_rle_baseline(cols):
    import oracle

    threads = host_threads()
    outs = [np.empty(c.n * CH, np.uint8) for c in cols]

    def run(th, k=None):
        for c, o in zip(cols, outs):
            comp, blk, cl = c.host
            fn = oracle.rlev1_decode_chunks if c.codec == ca.RLEV1 else oracle.rlev2_decode_chunks
            m = c.n if k is None else min(k, c.n)
            fn(comp, blk[:m], cl[:m], c.width, CH, th, out=o)

    total = sum(c.n * CH for c in cols)
    k1 = 64
    single = sum(min(k1, c.n) * CH for c in cols)
    return cpu_leg(lambda: run(threads), total, f"the whole workload ({total >> 20} MiB out)",
                   lambda: run(1, k1), single,
                   f"the first {k1} chunks of each column ({single >> 20} MiB out)",
                   "oracle C decoder")


def bench_rle(ctx, args, cfg):
    log(f"[bench] {cfg}: generating")
    cols, desc, dtype = make_rle(ctx, cfg)
    out_bytes = sum(len(c.raw) for c in cols)
    if not args.no_autotune:
        for c in cols:
            ms = c.autotune(ctx)
            log(f"[bench] {cfg}: autotuner: {c.width * 8}-bit column -> launch shape {c.shape()} "
                f"({ms:.3f} ms)")

    def check():
        crcs = []
        for c in cols:
            ok, crc = c.verify()
            if not ok or crc != crc32(c.raw):
                raise SystemExit(f"{cfg}: bit-exact check failed for codec {c.codec}/{c.width}")
            crcs.append(crc)
        return crcs

    for c in cols:
        c.launch(ctx)
    ctx.sync()
    crcs = check()
    log(f"[bench] {cfg}: bit-exact check passed")
    for _ in range(args.warmup):
        for c in cols:
            c.launch(ctx)
    ctx.sync()
    # timed region: events around every launch on the context stream
    dom = max(range(len(cols)), key=lambda i: cols[i].alg_bytes)
    per = [0.0] * len(cols)
    ctx.event(0)
    slot = 1
    marks = []
    for s in range(args.steps):
        for i, c in enumerate(cols):
            ctx.event(slot)
            c.launch(ctx)
            ctx.event(slot + 1)
            marks.append((i, slot, slot + 1))
            slot += 2
            if slot >= 250:  # drain event slots
                for (j, x, y) in marks:
                    per[j] += ctx.elapsed(x, y)
                marks.clear()
                slot = 1
    ctx.event(255)
    total_ms = ctx.elapsed(0, 255)
    ctx.sync()
    for (j, x, y) in marks:
        per[j] += ctx.elapsed(x, y)
    if check() != crcs:
        raise SystemExit(f"{cfg}: output changed during the timed steps")
    ms_step = total_ms / args.steps
    k_ms = per[dom] / args.steps
    kname = cols[dom].kernel(ctx)
    return {
        "workload": desc, "value": round(out_bytes / (ms_step / 1e3) / 1e9, 2), "unit": "GB/s",
        "ms_per_step": round(ms_step, 4), "dtype": dtype, "bit_exact": True,
        "chunks": [c.n for c in cols], "compressed_bytes": [c.comp_bytes for c in cols],
        "decompressed_bytes": out_bytes, "crc32": crcs,
        "launch_shapes": [c.shape() for c in cols],
        "kernel_ms_per_column": [round(p / args.steps, 4) for p in per],
        "roofline": roofline(kname, k_ms, cols[dom].alg_bytes, f"{kname}:{cfg}:{cols[dom].n}"),
        "cpu_baseline": None if args.no_cpu else cpu_rle_baseline(cols),
    }


# ---------------------------------------------------------------- Deflate (c3)

def bench_deflate(ctx, args):
    log("[bench] c3: generating 1 GiB Deflate workload")
    n_bytes = 1 << 30
    raw = ca.synth_deflate_payload(n_bytes, CH, seed=0xC0DA6 + 3)
    comp, blk, cl, n = ca.deflate_compress_chunks(raw, CH, level=6, threads=host_threads())
    off = blk[1:].astype(np.uint64)
    d_comp = ctx.upload(comp)
    d_off = ctx.upload(off)
    d_out = ctx.alloc(n * CH)
    d_len = ctx.alloc(8 * n)
    d_st = ctx.alloc(4 * n)

    if args.c3_waves:
        ca.check(ca.lib().codag_test_chunks_two_wave(ctx.h, 1 if args.c3_waves == 2 else 2),
                 "test_chunks_two_wave")

    def launch():
        ctx.inflate_chunks(d_comp, len(comp), d_off, n, CH, d_out, d_len, d_st, 0)

    def check():
        st = d_st.download(4 * n, np.int32)
        out = d_out.download(len(raw))
        if not (np.all(st == 0) and out.tobytes() == raw.tobytes()):
            raise SystemExit("c3: Deflate bit-exact check failed")
        return crc32(out)

    launch()
    ctx.sync()
    crc = check()
    log("[bench] c3: bit-exact check passed")
    for _ in range(args.warmup):
        launch()
    ctx.sync()
    ctx.event(0)
    for s in range(args.steps):
        launch()
    ctx.event(1)
    ms = ctx.elapsed(0, 1) / args.steps
    if check() != crc:
        raise SystemExit("c3: output changed during the timed steps")
    alg = len(comp) + n * CH + 8 * (n + 1)
    cpu = None
    if not args.no_cpu:
        import oracle
        out_h = np.empty(n * CH, np.uint8)
        k1 = 256
        cpu = cpu_leg(lambda: oracle.inflate_chunks(comp, off, CH, host_threads(), out=out_h),
                      n * CH, f"the whole workload ({n} chunks, {n * CH >> 20} MiB out)",
                      lambda: oracle.inflate_chunks(comp, off[: k1 + 1], CH, 1, out=out_h),
                      k1 * CH, f"the first {k1} chunks ({k1 * CH >> 20} MiB out)",
                      f"system zlib {oracle.zlib_version()} inflate")
    return {
        "workload": "deflate 1 GiB, 8192 x 128 KiB zlib level-6 chunks "
                    "(50% order-2 Markov text, 50% Zipf(1.1) int32)",
        "value": round(n_bytes / (ms / 1e3) / 1e9, 2), "unit": "GB/s", "ms_per_step": round(ms, 4),
        "dtype": "u8", "bit_exact": True, "chunks": int(n), "compressed_bytes": int(len(comp)),
        "decompressed_bytes": n_bytes, "crc32": crc,
        "roofline": roofline("inflate_chunks_kernel", ms, alg, f"inflate_chunks_kernel:c3:{n}"),
        "cpu_baseline": cpu,
    }


# ------------------------------------------------------ BRLE / LZSS

def bench_bytecodec(ctx, args, kind):
    """BRLE (kind "brle") or LZSS ("lzss") over 1 GiB of its benchmark bytes,
    encoded by the reference encoders' restatements."""
    lz = kind == "lzss"
    chunk = ca.LZSS_CHUNK_BYTES if lz else ca.BRLE_CHUNK_BYTES
    n_bytes = 1 << 30
    log(f"[bench] {kind}: generating and encoding 1 GiB")
    raw = ca.synth_byte_payload(n_bytes, 1 if lz else 0, seed=0xC0DA6 + (11 if lz else 10))
    if lz:
        data, idx, chunk, _ = ca.lzss_container(ca.lzss_encode(raw, chunk, host_threads()))
        n = len(idx)
    else:
        data, idx, _ = ca.brle_container(ca.brle_encode(raw, chunk, host_threads()))
        n = len(idx) - 1
    data = np.ascontiguousarray(data)
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    d_comp = ctx.upload(data)
    d_idx = ctx.upload(idx)
    d_out = ctx.alloc(n * chunk)
    d_len = ctx.alloc(8 * n)
    d_st = ctx.alloc(4 * n)
    fn = ctx.lzss_decode if lz else ctx.brle_decode
    kname = "lzss_chunks_kernel" if lz else "brle_chunks2_kernel"

    def launch():
        fn(d_comp, len(data), d_idx, n, chunk, d_out, d_len, d_st)

    def check():
        st = d_st.download(4 * n, np.int32)
        out = d_out.download(n_bytes)
        if not (np.all(st == 0) and np.array_equal(out, raw)):
            raise SystemExit(f"{kind}: bit-exact check failed")
        return crc32(out)

    launch()
    ctx.sync()
    crc = check()
    log(f"[bench] {kind}: bit-exact check passed")
    for _ in range(args.warmup):
        launch()
    ctx.sync()
    ctx.event(0)
    for _ in range(args.steps):
        launch()
    ctx.event(1)
    ms = ctx.elapsed(0, 1) / args.steps
    if check() != crc:
        raise SystemExit(f"{kind}: output changed during the timed steps")
    alg = len(data) + n_bytes + 8 * len(idx)
    cpu = None
    if not args.no_cpu:
        import oracle
        dec = oracle.lzss_decode_chunks if lz else oracle.brle_decode_chunks
        out_h = np.empty(n * chunk, np.uint8)
        parts = np.array_split(np.arange(n), host_threads())

        def part(ids):
            if len(ids) == 0:
                return
            a, b = int(ids[0]), int(ids[-1]) + 1
            if lz:
                lo = int(idx[a - 1]) if a else 0
                sub = idx[a:b] - np.uint64(lo)
                dec(data[lo:int(idx[b - 1])], sub, chunk, out=out_h[a * chunk:b * chunk])
            else:
                lo = int(idx[a])
                sub = idx[a:b + 1] - np.uint64(lo)
                dec(data[lo:int(idx[b])], sub, chunk, out=out_h[a * chunk:b * chunk])

        k1 = 256 if lz else 128
        cpu = cpu_leg(lambda: _parallel(part, parts), n_bytes,
                      f"the whole workload ({n} chunks, {n_bytes >> 20} MiB out)",
                      lambda: part(np.arange(k1)), k1 * chunk,
                      f"the first {k1} chunks ({k1 * chunk >> 20} MiB out)",
                      f"oracle C {kind} decoder")
    what = ("LZSS (2 KiB history, matches 3-34) of word text" if lz else
            "byte RLE of runs (3 + Geometric(mean 30)) and literal stretches (1-64)")
    return {
        "workload": f"{kind} 1 GiB, {n} x {chunk >> 10} KiB chunks, {what}",
        "value": round(n_bytes / (ms / 1e3) / 1e9, 2), "unit": "GB/s", "ms_per_step": round(ms, 4),
        "dtype": "u8", "bit_exact": True, "chunks": int(n), "compressed_bytes": int(len(data)),
        "decompressed_bytes": n_bytes, "crc32": crc,
        "roofline": roofline(kname, ms, alg, f"{kname}:{kind}:{n}"),
        "cpu_baseline": cpu,
    }


# ----------------------------------------------------------- batched (c4)

def _parallel(fn, items):
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(host_threads()) as ex:
        return list(ex.map(fn, items))


def bench_batch(ctx, args):
    """Uneven batch: n zlib streams, sizes log-uniform 4 KiB..4 MiB, blocks mixed
    fixed/dynamic/stored 20/75/5, decoded by one persistent graph-replayed launch."""
    rng = np.random.default_rng(0xC0DA6 + 4)
    n = args.streams
    sizes = np.exp(rng.uniform(np.log(4096), np.log(4 << 20), n)).astype(np.int64)
    total = int(sizes.sum())
    log(f"[bench] c4: generating {n} streams, {total >> 20} MiB")
    payload = ca.synth_deflate_payload(total, CH, seed=0xC0DA6 + 4)
    poff = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    res = _parallel(lambda i: ca.deflate_compress_mixed(payload[poff[i]:poff[i + 1]], 20, 5, 32768,
                                                        seed=i), range(n))
    mix = np.sum([r[1] for r in res], axis=0)
    order = np.argsort(-sizes, kind="stable")  # largest first: the queue balances the tail
    slen = np.array([len(res[i][0]) for i in order], np.int64)
    soff = np.concatenate([[0], np.cumsum(slen)]).astype(np.int64)
    cap = sizes[order]
    doff = np.concatenate([[0], np.cumsum(cap)]).astype(np.int64)
    src_h = np.concatenate([res[i][0] for i in order])
    del res
    d_src = ctx.upload(src_h)
    d_dst = ctx.alloc(int(doff[-1]))
    descs = (ca.StreamDesc * n)()
    for k in range(n):
        descs[k].src, descs[k].src_bytes = d_src.offset(int(soff[k])), int(slen[k])
        descs[k].dst, descs[k].dst_cap = d_dst.offset(int(doff[k])), int(cap[k])
    d_descs = ctx.upload(np.frombuffer(descs, np.uint8))
    rsz = 16
    d_res = ctx.alloc(n * rsz)
    tuned = None
    if args.c4_shape:
        tpg, per = (int(x) for x in args.c4_shape.split(","))
        ctx.tune_set(ca.DEFLATE, ca.TuneParams(tpg, 0, per, 1))
    elif not args.no_autotune:
        # the autotuner times its launch shapes (one or two waves per stream,
        # workgroups per CU) on the largest streams, which set the batch's
        # finish time; it keeps the fastest shape whose output matches
        k = min(n, 512)
        ch = int(cap[0])
        d_soff = ctx.upload(soff[: k + 1].astype(np.uint64))
        d_tout = ctx.alloc(k * ch)
        tbest, tms = ctx.tune_run(ca.DEFLATE, d_src, int(soff[k]), d_soff, None, k, 1, ch, d_tout,
                                  reps=2)
        del d_tout, d_soff
        tuned = {"sample": f"the {k} largest streams", "threads_per_group": tbest.threads_per_group,
                 "groups_per_cu": tbest.groups_per_cu, "sample_ms": round(tms, 3)}
        log(f"[bench] c4: autotuner on {tuned['sample']}: {tbest.threads_per_group} threads x "
            f"{tbest.groups_per_cu} workgroups per CU ({tms:.1f} ms)")
    best = ctx.tune_get(ca.DEFLATE)
    kname = "inflate_batch2_kernel" if best.threads_per_group == 128 else "inflate_batch_kernel"
    log(f"[bench] c4: persistent launch, {best.threads_per_group} threads x "
        f"{best.groups_per_cu} workgroups per CU")

    def check():
        rr = np.frombuffer(d_res.download(n * rsz).tobytes(), dtype=np.dtype(
            [("bytes", "<u8"), ("status", "<i4"), ("adler", "<u4")]))
        out = d_dst.download(int(doff[-1]))
        for k in range(n):
            i = order[k]
            if rr["status"][k] != 0 or int(rr["bytes"][k]) != sizes[i] or \
                    not np.array_equal(out[doff[k]:doff[k + 1]], payload[poff[i]:poff[i + 1]]):
                raise SystemExit(f"c4: batch bit-exact check failed at stream {i}")
        return crc32(out)

    ctx.inflate_batch(d_descs, n, d_res)
    ctx.sync()
    crc = check()
    log("[bench] c4: bit-exact check passed")
    for _ in range(args.warmup):
        ctx.inflate_batch(d_descs, n, d_res)
    ctx.sync()
    t_wall = time.perf_counter()
    ctx.event(0)
    for _ in range(args.steps):
        ctx.inflate_batch(d_descs, n, d_res)
    ctx.event(1)
    ms = ctx.elapsed(0, 1) / args.steps
    wall = time.perf_counter() - t_wall
    if check() != crc:
        raise SystemExit("c4: output changed during the timed steps")

    # plain per-file launches (one inflate call per stream) on a seeded sample
    samp = np.sort(np.random.default_rng(5).choice(n, size=min(n, 64), replace=False))
    pairs = np.zeros(2 * len(samp), np.uint64)
    pairs[1::2] = slen[samp]
    d_pairs = ctx.upload(pairs)
    d_st1 = ctx.alloc(4 * len(samp))

    def per_file():
        for j, k in enumerate(samp):
            ctx.inflate_chunks(d_src.offset(int(soff[k])), int(slen[k]), d_pairs.offset(16 * j), 1,
                               int(cap[k]), d_dst.offset(int(doff[k])), None, d_st1.offset(4 * j))

    per_file()
    ctx.sync()
    ctx.event(2)
    per_file()
    ctx.event(3)
    pf_ms = ctx.elapsed(2, 3)
    pf_gbs = float(cap[samp].sum()) / (pf_ms / 1e3) / 1e9
    if not np.all(d_st1.download(4 * len(samp), np.int32) == 0):
        raise SystemExit("c4: per-file launch failed")

    alg = int(slen.sum()) + total + n * (32 + 16)
    cpu = None
    if not args.no_cpu:
        import oracle
        so = soff[:-1].astype(np.uint64)
        out_h = np.empty(total, np.uint8)
        cs = samp[:64]
        src_s = np.concatenate([src_h[soff[k]:soff[k] + slen[k]] for k in cs])
        so_s = np.concatenate([[0], np.cumsum(slen[cs])[:-1]]).astype(np.uint64)
        do_s = np.concatenate([[0], np.cumsum(cap[cs])[:-1]]).astype(np.uint64)
        s_bytes = int(cap[cs].sum())
        cpu = cpu_leg(lambda: oracle.inflate_batch(src_h, so, slen, doff[:-1], cap, total,
                                                   host_threads(), out=out_h),
                      total, f"the whole batch ({n} streams, {total >> 20} MiB out)",
                      lambda: oracle.inflate_batch(src_s, so_s, slen[cs], do_s, cap[cs], s_bytes, 1,
                                                   out=out_h),
                      s_bytes, f"{len(cs)} of the streams ({s_bytes >> 20} MiB out)",
                      f"system zlib {oracle.zlib_version()} inflate, one stream per task")
    return {
        "workload": f"batched uneven deflate, {n} zlib streams log-uniform 4 KiB-4 MiB, "
                    "block mix fixed/dynamic/stored 20/75/5, one persistent graph-replayed launch",
        "value": round(total / (ms / 1e3) / 1e9, 2), "unit": "GB/s", "ms_per_step": round(ms, 4),
        "dtype": "u8", "bit_exact": True, "streams": n,
        "decompressed_bytes": total, "compressed_bytes": int(slen.sum()),
        "blocks_fixed_dynamic_stored": [int(x) for x in mix], "crc32": crc,
        "groups_per_cu": int(best.groups_per_cu), "threads_per_group": int(best.threads_per_group),
        "autotune": tuned,
        "wall_s": round(wall, 3),
        "per_file_launch_gbs": round(pf_gbs, 3),
        "per_file_sample": f"{len(samp)} streams, one inflate launch each",
        "roofline": roofline(kname, ms, alg, f"{kname}:c4:{ctx.info()[1] * best.groups_per_cu}"),
        "cpu_baseline": cpu,
    }


# ------------------------------------------------------------- served (c5)

def bench_served(ctx, args):
    """Served mix (40% RLE v1 u64, 30% RLE v2 u32, 30% Deflate; sizes log-uniform
    64 KiB..256 MiB) through the continuous-batching server, device-resident,
    200 GB admission budget. Arrivals draw from a pool of distinct requests,
    each arrival with its own output buffer; Poisson arrivals at 10..90% of
    the closed-loop rate measured in the same run."""
    import ctypes
    rng = np.random.default_rng(0xC0DA6 + 5)
    n_pool, n_arr = args.pool, args.requests
    sizes = np.exp(rng.uniform(np.log(64 << 10), np.log(256 << 20), n_pool))
    kinds = rng.choice(3, size=n_pool, p=[0.4, 0.3, 0.3])
    log(f"[bench] c5: generating a pool of {n_pool} requests")
    pool = []
    for i in range(n_pool):
        nch = max(1, int(round(sizes[i] / CH)))
        seed = 0xC0DA6 + 500 + i
        if kinds[i] == 0:
            raw, cont, blk, cl, n = ca.synth_rlev1(nch * CH // 8, 8, CH, 1 << 16, seed=seed)
            comp, off, ln, w, codec = cont[8:], blk[:n].copy(), cl[:n].copy(), 8, ca.RLEV1
        elif kinds[i] == 1:
            raw, comp, blk, cl, n = ca.synth_rlev2(nch * CH // 4, 4, CH, seed=seed)
            off, ln, w, codec = blk[:n].copy(), cl.copy(), 4, ca.RLEV2
        else:
            raw = ca.synth_deflate_payload(nch * CH, CH, seed=seed)
            comp, blk, _, n = ca.deflate_compress_chunks(raw, CH, level=6, threads=host_threads())
            off, ln, w, codec = blk[1:].copy(), np.zeros(1, np.uint64), 1, ca.DEFLATE
        d = [ctx.upload(np.ascontiguousarray(comp)), ctx.upload(off), ctx.upload(ln)]
        pool.append(dict(d=d, codec=int(codec), w=w, n=n, comp_bytes=len(comp), raw=raw,
                         bytes=len(raw)))
    # the arrival sequence: pool entries drawn uniformly, each with its own output
    pick = np.random.default_rng(0xC0DA6 + 6).integers(0, n_pool, n_arr)
    out_off = np.concatenate([[0], np.cumsum([pool[p]["n"] * CH for p in pick])]).astype(np.int64)
    log(f"[bench] c5: {n_arr} arrivals, {out_off[-1] >> 20} MiB of output buffers")
    d_arena = ctx.alloc(int(out_off[-1]))
    reqs = []
    for a, p in enumerate(pick):
        e = pool[p]
        q = ca.Request()
        q.codec, q.flags, q.elem_bytes, q.chunk_bytes, q.n_chunks = (e["codec"], ca.REQ_DEVICE_IO,
                                                                    e["w"], CH, e["n"])
        q.comp, q.comp_bytes = e["d"][0].ptr.value, e["comp_bytes"]
        q.blk_off, q.col_len = e["d"][1].ptr.value, e["d"][2].ptr.value
        q.out, q.user_tag = d_arena.offset(int(out_off[a])), a
        reqs.append(q)
    arr_bytes = np.array([pool[p]["bytes"] for p in pick], np.float64)
    total_bytes = float(arr_bytes.sum())

    def run(inflight, arrivals=None):
        cfg = ca.ServerConfig(0, 200 * 10**9, inflight)
        h = ctypes.c_void_p()
        ca.check(ca.lib().codag_server_create(ctypes.byref(cfg), ctypes.byref(h)), "server_create")
        buf = (ca.Completion * 64)()
        done = []
        try:
            t0 = time.perf_counter()
            rid = ctypes.c_uint64()
            for k, q in enumerate(reqs):
                if arrivals is not None:
                    dt = t0 + arrivals[k] - time.perf_counter()
                    if dt > 0:
                        time.sleep(dt)
                ca.check(ca.lib().codag_server_submit(h, ctypes.byref(q), ctypes.byref(rid)),
                         "server_submit")
            while len(done) < len(reqs):
                got = ctypes.c_uint32()
                st = ca.lib().codag_server_poll(h, buf, 64, ctypes.byref(got), 120000)
                ca.check(st, "server_poll")
                done += [ca.Completion.from_buffer_copy(buf[j]) for j in range(got.value)]
            wall = time.perf_counter() - t0
        finally:
            ca.lib().codag_server_destroy(h)
        if any(c.status != 0 or c.failed_chunks for c in done):
            raise SystemExit("c5: served request failed")
        lat = np.array([c.latency_seconds for c in done]) * 1e3
        return wall, lat

    def check():
        for a, p in enumerate(pick):
            e = pool[p]
            out = _slice(d_arena, int(out_off[a]), e["bytes"])
            if not np.array_equal(out, e["raw"]):
                raise SystemExit(f"c5: served output mismatch (arrival {a})")

    run(8)  # warm-up (code objects, the server's memory pool)
    check()
    log("[bench] c5: bit-exact check of every arrival passed")
    wall1, _ = run(1)
    wall8, lat8 = run(8)
    peak = total_bytes / wall8
    levels = []
    for load in (0.1, 0.3, 0.5, 0.7, 0.9):
        rate = load * peak / (total_bytes / n_arr)  # requests per second
        arr = np.cumsum(np.random.default_rng(int(load * 100)).exponential(1.0 / rate, n_arr))
        arr -= arr[0]
        wall, lat = run(8, arr)
        levels.append({"load": load, "offered_gbs": round(load * peak / 1e9, 2),
                       "achieved_gbs": round(total_bytes / wall / 1e9, 2),
                       "p50_ms": round(float(np.percentile(lat, 50)), 3),
                       "p99_ms": round(float(np.percentile(lat, 99)), 3)})
        log(f"[bench] c5: load {load}: {levels[-1]}")
    check()
    return {
        "workload": f"served mixed-codec decode, {n_arr} Poisson arrivals per load level drawn from "
                    f"a pool of {n_pool} requests (40% RLE v1 u64, 30% RLE v2 u32, 30% deflate), "
                    "log-uniform 64 KiB-256 MiB, device-resident, 200 GB HBM budget, 8 slots",
        "value": round(peak / 1e9, 2), "unit": "GB/s", "ms_per_step": round(wall8 * 1e3, 3),
        "dtype": "mixed", "bit_exact": True, "decompressed_bytes_per_level": int(total_bytes),
        "closed_loop_gbs": {"slots_8": round(peak / 1e9, 2),
                            "slots_1_sequential": round(total_bytes / wall1 / 1e9, 2)},
        "closed_loop_p50_ms": round(float(np.percentile(lat8, 50)), 3),
        "closed_loop_p99_ms": round(float(np.percentile(lat8, 99)), 3),
        "poisson": levels, "roofline": None, "cpu_baseline": None,
    }


def _slice(buf, off, nbytes):
    import ctypes
    out = np.empty(nbytes, np.uint8)
    if nbytes:
        ca.check(ca.lib().codag_copy_d2h(buf.ctx.h, out.ctypes.data_as(ctypes.c_void_p),
                                         ctypes.c_void_p(buf.offset(off)), nbytes), "copy_d2h")
    return out


# -------------------------------------------------------------------- main

def headline(rec, args, extra_configs, device):
    cfg = {k: v for k, v in rec.items()
           if k not in ("value", "unit", "ms_per_step", "dtype", "roofline", "cpu_baseline",
                        "bit_exact")}
    cfg["device"] = device
    if extra_configs:
        cfg["configs"] = extra_configs
    return {
        "metric": METRIC, "value": rec["value"], "unit": rec["unit"], "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": rec["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": rec["dtype"],
        "data": "synthetic", "bit_exact": True, "config": cfg,
        "roofline": rec["roofline"], "cpu_baseline": rec["cpu_baseline"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="all",
                    choices=["all", "c1", "c2", "c3", "c4", "c5", "brle", "lzss"])
    ap.add_argument("--streams", type=int, default=10000, help="c4: number of zlib streams")
    ap.add_argument("--pool", type=int, default=64, help="c5: distinct requests in the pool")
    ap.add_argument("--requests", type=int, default=500, help="c5: arrivals per load level")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-autotune", action="store_true",
                    help="c1/c2: keep the default launch shape instead of tuning per column")
    ap.add_argument("--c4-shape", default=None, metavar="THREADS,PER_CU",
                    help="c4: batch launch shape (64 = one wave per stream, 128 = a decoding "
                         "and a writing wave), workgroups per CU; default: the library's")
    ap.add_argument("--c3-waves", type=int, default=0, choices=[0, 1, 2],
                    help="c3: waves per chunk (2 = a decoding and a writing wave; "
                         "0 = the library's choice by chunk count)")
    ap.add_argument("--copy-ref", type=int, default=0, metavar="MIB",
                    help="first run the plain copy kernel over MIB MiB (traffic calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}). The decode shards "
              "by chunk across GPUs without communication, so no multi-GPU figure is measured; "
              "run it with --gpus 1.", file=sys.stderr)
        sys.exit(2)
    ctx = ca.Context(0)
    name, cus = ctx.info()
    device = f"{name} ({cus} CUs)"
    log(f"[bench] device {device}; host: {host_threads()} cores usable, {cpu_model()}")
    if args.copy_ref:
        # one launch of the plain copy kernel moving a known byte count: under
        # rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE it scales the decode kernels'
        # counters to bytes
        nb = args.copy_ref << 20
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        a.fill(1)
        ctx.copy(b, a, nb)
        ctx.sync()
        a.free()
        b.free()
        log(f"[bench] reference copy of {nb} bytes done")
    run = {"c1": lambda: bench_rle(ctx, args, "c1"), "c2": lambda: bench_rle(ctx, args, "c2"),
           "c3": lambda: bench_deflate(ctx, args), "c4": lambda: bench_batch(ctx, args),
           "c5": lambda: bench_served(ctx, args),
           "brle": lambda: bench_bytecodec(ctx, args, "brle"),
           "lzss": lambda: bench_bytecodec(ctx, args, "lzss")}
    if args.config != "all":
        print(json.dumps(headline(run[args.config](), args, None, device)), flush=True)
        return
    t0 = time.perf_counter()
    head = run["c4"]()
    log(f"[bench] c4 done ({time.perf_counter() - t0:.0f} s)")
    others = {}
    for c in ("c1", "c2", "c3", "c5", "brle", "lzss"):
        others[c] = run[c]()
        log(f"[bench] {c} done ({time.perf_counter() - t0:.0f} s)")
    print(json.dumps(headline(head, args, others, device)), flush=True)


if __name__ == "__main__":
    main()
