"""Continuous-batching server: every request's result is bit-identical whether
it runs alone or mixed with others, in host-I/O and device-resident modes."""
import ctypes

import numpy as np
import pytest

import codag_amd as ca
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu
CH = 131072


class Server:
    def __init__(self, budget, inflight):
        cfg = ca.ServerConfig(0, budget, inflight)
        self.h = ctypes.c_void_p()
        ca.check(ca.lib().codag_server_create(ctypes.byref(cfg), ctypes.byref(self.h)),
                 "server_create")

    def submit(self, req):
        rid = ctypes.c_uint64()
        st = ca.lib().codag_server_submit(self.h, ctypes.byref(req), ctypes.byref(rid))
        return st, rid.value

    def wait_all(self, n, timeout_ms=60000):
        got = []
        buf = (ca.Completion * 64)()
        while len(got) < n:
            k = ctypes.c_uint32()
            st = ca.lib().codag_server_poll(self.h, buf, 64, ctypes.byref(k), timeout_ms)
            assert st in (ca.OK, ca.ERR_TIMEOUT)
            assert st == ca.OK, "server poll timed out"
            got += [ca.Completion.from_buffer_copy(buf[i]) for i in range(k.value)]
        return got

    def close(self):
        ca.lib().codag_server_destroy(self.h)


def make_requests(seed=1):
    """(request struct, keep-alive arrays, expected output) per request."""
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(12):
        kind = i % 3
        nch = int(rng.integers(1, 6))
        if kind == 0:
            raw, cont, blk, cl, n = ca.synth_rlev1(nch * CH // 8, 8, CH, 1 << 16, seed=seed + i)
            comp, off, ln, w, codec = cont[8:], blk[:n].copy(), cl[:n].copy(), 8, ca.RLEV1
        elif kind == 1:
            raw, comp, blk, cl, n = ca.synth_rlev2(nch * CH // 4, 4, CH, seed=seed + i)
            off, ln, w, codec = blk[:n].copy(), cl.copy(), 4, ca.RLEV2
        else:
            raw = ca.synth_deflate_payload(nch * CH, CH, seed=seed + i)
            comp, blk, _, n = ca.deflate_compress_chunks(raw, CH, level=6)
            off, ln, w, codec = blk[1:].copy(), np.zeros(1, np.uint64), 1, ca.DEFLATE
        out = np.zeros(n * CH, np.uint8)
        reqs.append(dict(codec=codec, w=w, n=n, comp=np.ascontiguousarray(comp), off=off, ln=ln,
                         out=out, raw=raw))
    return reqs


def to_struct(r, tag):
    q = ca.Request()
    q.codec, q.flags, q.elem_bytes, q.chunk_bytes, q.n_chunks = r["codec"], 0, r["w"], CH, r["n"]
    q.comp, q.comp_bytes = r["comp"].ctypes.data, len(r["comp"])
    q.blk_off, q.col_len = r["off"].ctypes.data, r["ln"].ctypes.data
    q.out, q.user_tag = r["out"].ctypes.data, tag
    return q


def test_alone_equals_mixed(gpu_required):
    say("server alone vs mixed")
    reqs = make_requests()
    srv = Server(1 << 30, 4)
    try:
        alone = []
        for i, r in enumerate(reqs):
            st, _ = srv.submit(to_struct(r, i))
            assert st == ca.OK
            (c,) = srv.wait_all(1)
            assert c.status == ca.OK and c.failed_chunks == 0
            alone.append(r["out"].copy())
            r["out"][:] = 0
        structs = [to_struct(r, i) for i, r in enumerate(reqs)]
        for q in structs:
            assert srv.submit(q)[0] == ca.OK
        done = srv.wait_all(len(reqs))
        assert sorted(c.user_tag for c in done) == list(range(len(reqs)))
        for c in done:
            assert c.status == ca.OK and c.latency_seconds >= c.queue_seconds >= 0
        for i, r in enumerate(reqs):
            assert r["out"].tobytes() == alone[i].tobytes()
            assert r["out"][: len(r["raw"])].tobytes() == r["raw"].tobytes()
    finally:
        srv.close()


def test_device_resident_requests(gpu_required):
    say("server device-resident")
    ctx = ca.Context(0)
    reqs = make_requests(seed=7)
    srv = Server(1 << 30, 3)
    keep = []
    try:
        structs = []
        for i, r in enumerate(reqs):
            d_comp, d_off, d_ln = ctx.upload(r["comp"]), ctx.upload(r["off"]), ctx.upload(r["ln"])
            d_out = ctx.alloc(r["n"] * CH)
            keep += [d_comp, d_off, d_ln, d_out]
            q = to_struct(r, i)
            q.flags = ca.REQ_DEVICE_IO
            q.comp, q.blk_off, q.col_len, q.out = (d_comp.ptr.value, d_off.ptr.value,
                                                   d_ln.ptr.value, d_out.ptr.value)
            structs.append((q, d_out))
        for q, _ in structs:
            assert srv.submit(q)[0] == ca.OK
        done = srv.wait_all(len(reqs))
        assert all(c.status == ca.OK for c in done)
        for (q, d_out), r in zip(structs, reqs):
            assert d_out.download(len(r["raw"])).tobytes() == r["raw"].tobytes()
    finally:
        srv.close()
        ctx.close()


def test_admission_budget(gpu_required):
    say("server budget")
    reqs = make_requests(seed=3)[:3]
    srv = Server(4 * CH, 2)  # smaller than any request's footprint
    try:
        st, _ = srv.submit(to_struct(reqs[0], 0))
        assert st == ca.ERR_INVALID
        bad = to_struct(reqs[1], 1)
        bad.comp_bytes = 3  # metadata points outside the stream
        srv2 = Server(1 << 30, 1)
        try:
            assert srv2.submit(bad)[0] == ca.ERR_FORMAT
        finally:
            srv2.close()
    finally:
        srv.close()


def _deflate_request(n_chunks, seed):
    raw = ca.synth_deflate_payload(n_chunks * CH, CH, seed=seed)
    comp, blk, _, n = ca.deflate_compress_chunks(raw, CH, level=6, threads=8)
    return dict(codec=ca.DEFLATE, w=1, n=n, comp=np.ascontiguousarray(comp), off=blk[1:].copy(),
                ln=np.zeros(1, np.uint64), out=np.zeros(n * CH, np.uint8), raw=raw)


def _footprint(r):
    meta = (r["n"] + 1) * 8 if r["codec"] == ca.DEFLATE else r["n"] * 16
    return len(r["comp"]) + meta + r["n"] * CH + r["n"] * 4


def test_blocked_head_is_passed_within_the_aging_bound(gpu_required):
    say("server: 40 MiB requests, a budget-blocked head, small requests passing it")
    big1, big2 = _deflate_request(320, 51), _deflate_request(320, 52)  # 40 MiB each
    smalls = [_deflate_request(1, 60 + i) for i in range(12)]
    fb = max(_footprint(big1), _footprint(big2))
    budget = fb + 8 * (1 << 20)  # one big request plus every small one, never both bigs
    assert _footprint(big1) + _footprint(big2) > budget
    assert fb + sum(_footprint(r) for r in smalls) <= budget
    srv = Server(budget, 6)
    try:
        import time
        t_sub = {}
        reqs = [big1, big2] + smalls
        structs = [to_struct(r, i) for i, r in enumerate(reqs)]
        for i, q in enumerate(structs):
            t_sub[i] = time.perf_counter()
            assert srv.submit(q)[0] == ca.OK
        done = {c.user_tag: c for c in srv.wait_all(len(reqs))}
        assert all(c.status == ca.OK and c.failed_chunks == 0 for c in done.values())
        for r in reqs:
            assert r["out"][: len(r["raw"])].tobytes() == r["raw"].tobytes()
        # admission time of request i, on the submitter's clock (the server's
        # clock differs by a constant, which cancels in the comparisons)
        adm = {i: t_sub[i] + done[i].queue_seconds for i in done}
        # big2 cannot start before big1 has finished (the budget holds one)
        assert adm[1] >= t_sub[0] + done[0].latency_seconds - 2e-3
        # the first kMaxPass (8) small requests may pass the blocked big2;
        # the 9th to 12th are never admitted before it
        for i in range(2 + 8, 2 + 12):
            assert adm[i] >= adm[1] - 2e-3, (i, adm[i] - adm[1])
    finally:
        srv.close()
