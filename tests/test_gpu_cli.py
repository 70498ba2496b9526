"""The drop-in CLIs and the reference-shaped host wrappers on the GPU.

* rlev1 -f <col> -t <w>: exit 0, the reference's stdout lines (a header, then
  "NUM WARP: 1" and "compressed\\tmeta\\tratio\\tseconds" for each of the five
  NUM_CHUNK variants, src/rlev1/rlev1.cu:168-193, rlev1.h:507,576-579), the
  ./input_data side files in the reference's sizes;
* rlev2 in both argv forms: two groups of five "<seconds> secs" lines after a
  "131072 \\t48\\t<encoded>\\t<meta>\\t" prefix (src/rlev2/rlev2.cu:85-110);
* zlib with 4 and 3 arguments: "\\t<seconds>", the output file (4 arguments);
* deflate_compressor -> zlib round trip;
* codag_rlev1_decode_host / codag_rlev2_decode_host on reference-layout side
  files, against the oracle.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import codag_amd as ca
import oracle
from tests.gpu_helpers import say

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "codag_amd", "bin")
CH = 131072


@pytest.fixture(scope="module")
def ctx(gpu_required):
    c = ca.Context(0)
    yield c
    c.close()


def run(tool, args, cwd):
    say(f"  {tool} {' '.join(map(str, args))}")
    return subprocess.run([os.path.join(BIN, tool)] + [str(a) for a in args], cwd=cwd,
                          capture_output=True, text=True, timeout=120)


def v2_round_trippable(raw, width):
    """The column with 64-bit values cut to 32 bits: the reference's RLE v2
    encoder writes PATCHED_BASE patches up to 56 bits wide, which its decoder
    (and this library, bit-exact to it) masks to 32 bits (decoder_new.h:1105),
    so only such columns round-trip."""
    if width != 8:
        return raw
    return (raw.view(np.uint64) & np.uint64(0xFFFFFFFF)).view(np.uint8)


def rle_column(width, n_values, seed):
    raw, *_ = ca.synth_rlev1(n_values, width, CH, 1 << min(16, 8 * width - 1), seed=seed)
    return raw


@pytest.mark.parametrize("width", [1, 4, 8])
def test_rlev1_cli(tmp_path, width):
    say(f"rlev1 CLI width {width}")
    raw = rle_column(width, 3 * CH // width + 1234 // width, seed=40 + width)
    col = tmp_path / "col.bin"
    col.write_bytes(raw.tobytes())
    p = run("rlev1", ["-f", col, "-t", width], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = p.stdout.splitlines()
    assert lines[0].startswith("CHUNK_SIZE\tCOLUMN WIDTH")
    body = lines[1:]
    assert len(body) == 10 and "File Differ" not in p.stdout
    cont, blk, cl, n = ca.rlev1_encode(raw, width, CH)
    blk_f = (tmp_path / "input_data" / "rle_v1_blk_offset.bin").read_bytes()
    col_f = (tmp_path / "input_data" / "rle_v1_col_len.bin").read_bytes()
    assert blk_f == blk.tobytes() and col_f == cl.tobytes()
    assert len(blk_f) == 8 * (n * 32 + 1) and len(col_f) == 8 * n * 32 * 32  # rlev1.h:178-186
    for k in range(5):
        assert body[2 * k] == "NUM WARP: 1"
        f = body[2 * k + 1].split("\t")
        assert len(f) == 4 and int(f[0]) == len(cont) and int(f[1]) == len(blk_f) + len(col_f)
        assert abs(float(f[2]) - len(cont) / len(raw)) < 1e-4 and float(f[3]) > 0


def test_rlev1_cli_unsupported_width(tmp_path):
    col = tmp_path / "c.bin"
    col.write_bytes(b"\0" * 64)
    p = run("rlev1", ["-f", col, "-t", 3], tmp_path)
    assert p.returncode != 0 and "Not supported input bytes" in p.stdout


@pytest.mark.parametrize("form", ["positional", "flags"])
@pytest.mark.parametrize("width", [4, 8])
def test_rlev2_cli(tmp_path, form, width):
    say(f"rlev2 CLI {form} width {width}")
    raw, *_ = ca.synth_rlev2(2 * CH // width + 77, width, CH, seed=50 + width)
    raw = v2_round_trippable(raw, width)
    col = tmp_path / "col.bin"
    col.write_bytes(raw.tobytes())
    args = [col, width] if form == "positional" else ["-f", col, "-t", width]
    p = run("rlev2", args, tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    lines = p.stdout.splitlines()
    assert lines[0].startswith("CHUNK_SIZE\t BUFFER SIZE")
    body = lines[1:]
    assert len(body) == 10 and "fail" not in p.stdout
    comp, blk, cl, n = ca.rlev2_encode(raw, width, CH)
    meta = 8 * n * 32 + 8 * (n + 1)  # encoder_trans.h:505
    prefix = re.compile(rf"^131072 \t48\t{len(comp)}\t{meta}\t([0-9.e-]+) secs$")
    for k, line in enumerate(body):
        if k in (0, 5):
            assert prefix.match(line), line
        else:
            assert re.match(r"^[0-9.e-]+ secs$", line), line


def _container(tmp_path, raw, name="data"):
    src = tmp_path / f"{name}.dat"
    src.write_bytes(raw.tobytes())
    p = subprocess.run([os.path.join(BIN, "deflate_compressor"), str(src), str(tmp_path)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return tmp_path / f"{name}_comp.bin", tmp_path / f"{name}_blk_offset.bin"


def test_zlib_cli_four_and_three_arguments(tmp_path):
    say("zlib CLI")
    raw = ca.synth_deflate_payload(5 * CH + 4321, CH, seed=61)
    comp, blk = _container(tmp_path, raw)
    out = tmp_path / "out.bin"
    p = run("zlib", [comp, out, blk, CH], tmp_path)
    assert p.returncode == 0, p.stdout + p.stderr
    assert re.match(r"^\t[0-9.e-]+$", p.stdout.splitlines()[0])
    assert out.read_bytes() == raw.tobytes()
    before = set(os.listdir(tmp_path))
    p = run("zlib", [comp, blk, CH], tmp_path)
    assert p.returncode == 0 and re.match(r"^\t[0-9.e-]+$", p.stdout.splitlines()[0])
    assert set(os.listdir(tmp_path)) == before  # no output file in the 3-argument form


def test_zlib_cli_corrupt_container_fails(tmp_path):
    raw = ca.synth_deflate_payload(2 * CH, CH, seed=62)
    comp, blk = _container(tmp_path, raw)
    b = bytearray(comp.read_bytes())
    b[40] ^= 0xFF
    comp.write_bytes(bytes(b))
    p = run("zlib", [comp, tmp_path / "o.bin", blk, CH], tmp_path)
    assert p.returncode != 0 and "Fatal Error" in p.stderr


def test_rlev1_host_wrapper_reference_side_files(ctx):
    say("codag_rlev1_decode_host on reference-layout side files")
    for width, seed in ((8, 70), (2, 71)):
        raw = rle_column(width, 5 * CH // width + 99, seed)
        cont, blk, cl, n = ca.rle_v1.compress_gpu_orig(raw, width, CH)
        assert len(blk) == n * 32 + 1 and len(cl) == n * 32 * 32  # the files' entry counts
        out, secs = ca.rle_v1.decompress_gpu_orig(ctx, cont, blk, cl, width, CH)
        st, ref = oracle.rlev1_decode_chunks(cont[8:], blk[:n], cl[:n], width, CH)
        assert np.all(st == 0) and secs > 0
        assert out.tobytes() == ref[: len(raw)].tobytes() == raw.tobytes()
    # a damaged stream: the wrapper reports the decode error, as the oracle does
    bad = cont.copy()
    bad[8 + int(blk[1]): 8 + int(blk[1]) + 16] = 0x80  # chunk 1: a varint of 15+ bytes
    st, _ = oracle.rlev1_decode_chunks(bad[8:], blk[:n], cl[:n], width, CH)
    assert np.any(st != 0)
    with pytest.raises(ca.CodagError) as e:
        ca.rle_v1.decompress_gpu_orig(ctx, bad, blk, cl, width, CH)
    assert e.value.status == ca.ERR_DECODE


def test_rlev2_host_wrapper_reference_side_files(ctx):
    say("codag_rlev2_decode_host on reference-layout buffers")
    for width, seed in ((4, 72), (8, 73)):
        raw, *_ = ca.synth_rlev2(3 * CH // width + 55, width, CH, seed=seed)
        raw = v2_round_trippable(raw, width)
        comp, blk, cl, n = ca.rlev2.compress_gpu_orig(raw, width, CH)
        assert int(blk[n]) == len(raw)  # encoder_trans.h:545
        out, secs = ca.rlev2.decompress_gpu_orig(ctx, comp, n, blk, cl, width, CH)
        st, ref = oracle.rlev2_decode_chunks(comp, blk[:n], cl, width, CH)
        assert np.all(st == 0) and secs > 0
        assert out.tobytes() == ref[: len(raw)].tobytes() == raw.tobytes()
