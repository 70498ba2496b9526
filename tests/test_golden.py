"""Known-answer vectors (tests/golden/kats.json) against the oracle (CPU) and
against the GPU decoders through the C ABI (marked gpu)."""
import json
import os

import numpy as np
import pytest

import codag_amd as ca
import oracle
from tests.gpu_helpers import inflate, rle_decode, say, single_chunk

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "kats.json")) as f:
    KATS = json.load(f)


def _ids(vs):
    return [v["name"] for v in vs]


@pytest.mark.parametrize("v", KATS["rlev1"], ids=_ids(KATS["rlev1"]))
def test_oracle_rlev1_kat(v):
    w = v["width"]
    rc, out = oracle.rlev1_decode_chunk(bytes.fromhex(v["hex"]), w, len(v["values"]) * w)
    assert rc == 0 and out.tolist() == v["values"]


@pytest.mark.parametrize("v", KATS["rlev2"], ids=_ids(KATS["rlev2"]))
def test_oracle_rlev2_kat(v):
    w = v["width"]
    rc, out = oracle.rlev2_decode_chunk(bytes.fromhex(v["hex"]), w, len(v["values"]) * w)
    assert rc == 0 and out.tolist() == v["values"]


@pytest.mark.parametrize("v", KATS["zlib"], ids=_ids(KATS["zlib"]))
def test_oracle_zlib_kat(v):
    rc, out = oracle.zlib_inflate(bytes.fromhex(v["hex"]), 4096)
    assert rc == 0 and out.tobytes().hex() == v["out_hex"]


@pytest.fixture(scope="module")
def ctx(gpu_required):
    c = ca.Context(0)
    yield c
    c.close()


_DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


@pytest.mark.gpu
@pytest.mark.parametrize("codec", ["rlev1", "rlev2"])
def test_gpu_rle_kats(ctx, codec):
    say(f"golden {codec}")
    cid = ca.RLEV1 if codec == "rlev1" else ca.RLEV2
    for v in KATS[codec]:
        w = v["width"]
        chunk = len(v["values"]) * w
        comp, blk, cl = single_chunk(bytes.fromhex(v["hex"]), pad=16)
        st, out = rle_decode(ctx, cid, comp, blk, cl, w, chunk)
        assert st.tolist() == [0], v["name"]
        assert out.view(_DT[w]).tolist() == v["values"], v["name"]


@pytest.mark.gpu
def test_gpu_zlib_kats(ctx):
    say("golden zlib")
    streams = [bytes.fromhex(v["hex"]) for v in KATS["zlib"]]
    pad = [s + b"\0" * (-len(s) % 8) for s in streams]
    off = np.cumsum([0] + [len(s) for s in pad]).astype(np.uint64)
    comp = np.frombuffer(b"".join(pad), np.uint8)
    st, ln, out = inflate(ctx, comp, off, 4096)
    for i, v in enumerate(KATS["zlib"]):
        n = len(v["out_hex"]) // 2
        assert st[i] == 0 and int(ln[i]) == n, v["name"]
        assert out[i * 4096: i * 4096 + n].tobytes().hex() == v["out_hex"], v["name"]
