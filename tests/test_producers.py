"""CPU tests of the producer side (encoders, container writers, synthetic
generators). Each container is decoded by the oracle and must reproduce the
raw column exactly -- a closed-form round trip through independent code."""
import zlib

import numpy as np
import pytest

import codag_amd as ca
import oracle

CH = 131072


def _column(width, n, seed):
    rng = np.random.default_rng(seed)
    parts = []
    total = 0
    while total < n:
        k = int(rng.integers(1, 300))
        kind = rng.integers(0, 4)
        if kind == 0:
            v = np.full(k, rng.integers(0, 1 << min(8 * width, 62)))
        elif kind == 1:
            v = rng.integers(0, 1 << 30, 1) + np.arange(k) * int(rng.integers(-5, 6))
        elif kind == 2:
            v = rng.integers(0, 1 << min(8 * width, 62), k)
        else:
            v = rng.integers(0, 200, k)
        parts.append(v.astype(np.int64).view(np.uint64))
        total += k
    col = np.concatenate(parts)[:n]
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[width]
    return col.astype(dt)


@pytest.mark.parametrize("width", [1, 2, 4, 8])
def test_rlev1_encode_round_trip(width):
    col = _column(width, 3 * CH // width + 1234, width)
    cont, blk, cl, n = ca.rlev1_encode(col, width)
    assert n % 32 == 0 and len(blk) == n * 32 + 1 and len(cl) == n * 32 * 32
    assert int(cont[:8].view(np.uint64)[0]) == col.nbytes
    assert np.all(blk[: n] % 4 == 0) and blk[n] == len(cont) - 8
    st, out = oracle.rlev1_decode_chunks(cont[8:], blk[:n], cl[:n], width, CH, threads=4)
    assert np.all(st == 0)
    assert out[: col.nbytes].tobytes() == col.tobytes()
    assert not out[col.nbytes:].any()  # zero padding (rlev1.h:127)


@pytest.mark.parametrize("width", [1, 2, 4, 8])
def test_rlev2_encode_round_trip(width):
    col = _column(width, 2 * CH // width + 777, 10 + width)
    if width == 8:
        # values below 2^32: the reference encoder writes PATCHED_BASE
        # patches up to 56 bits wide, which its decoder masks to 32 bits
        # (decoder_new.h:1105), see test_rlev2_wide_patch_is_masked_as_the_reference_decodes
        col = col & np.uint64(0xFFFFFFFF)
    comp, blk, cl, n = ca.rlev2_encode(col, width)
    assert n % 32 == 0 and blk[n] == col.nbytes and np.all(blk[:n] % 128 == 0)
    st, out = oracle.rlev2_decode_chunks(comp, blk[:n], cl, width, CH, threads=4)
    assert np.all(st == 0)
    assert out[: col.nbytes].tobytes() == col.tobytes()


@pytest.mark.parametrize("width", [8, 4])
def test_synth_rlev1_matches_oracle(width):
    raw, cont, blk, cl, n = ca.synth_rlev1(40000, width, CH, 1 << 16, seed=3)
    assert len(raw) == 40000 * width
    st, out = oracle.rlev1_decode_chunks(cont[8:], blk[:n], cl[:n], width, CH, threads=4)
    assert np.all(st == 0)
    assert out[: len(raw)].tobytes() == raw.tobytes()


@pytest.mark.parametrize("width", [8, 4, 2, 1])
def test_synth_rlev2_matches_oracle(width):
    nvals = 70000
    raw, comp, blk, cl, n = ca.synth_rlev2(nvals, width, CH, seed=4)
    st, out = oracle.rlev2_decode_chunks(comp, blk[:n], cl, width, CH, threads=4)
    assert np.all(st == 0)
    assert out[: len(raw)].tobytes() == raw.tobytes()


def test_synth_rlev2_mix_has_all_block_types():
    raw, comp, blk, cl, n = ca.synth_rlev2(CH // 8, 8, CH, seed=5)
    s = comp[int(blk[0]): int(blk[0]) + int(cl[0])].tobytes()
    kinds = set()
    # walk headers with the oracle-independent size rules of the format
    p = 0
    dw = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24,
          26, 28, 30, 32, 40, 48, 56, 64]
    while p < len(s):
        h = s[p]
        k = h >> 6
        kinds.add(k)
        if k == 0:
            p += 2 + ((h >> 3) & 7)
            continue
        ln = ((((h & 1) << 8) | s[p + 1]) + 1) & 0xFF
        w = dw[(h >> 1) & 31]
        if k == 1:
            p += 2 + (ln * w + 7) // 8
        elif k == 3:
            q = p + 2
            for _ in range(2):
                while s[q] >= 0x80:
                    q += 1
                q += 1
            p = q + (((ln - 2) & 0xFF) * w + 7) // 8 if (h >> 1) & 31 else q
        else:
            bw = (s[p + 2] >> 5) + 1
            pw = dw[s[p + 2] & 31]
            pgw = (s[p + 3] >> 5) + 1
            pll = s[p + 3] & 31
            cfb = next(x for x in dw if x >= pw + pgw)
            p += 4 + bw + (ln * w + 7) // 8 + (pll * cfb + 7) // 8
    assert kinds == {0, 1, 2, 3}


def test_deflate_container_layout_and_oracle():
    raw = ca.synth_deflate_payload(5 * CH + 1000, CH, seed=7)
    comp, blk, cl, n = ca.deflate_compress_chunks(raw, CH, level=6, threads=4)
    assert n == 6 and blk[0] == len(raw) and blk[1] == 0 and len(blk) == n + 2
    assert np.all(cl % 8 == 0) and np.array_equal(np.diff(blk[1:]), cl)
    st, out = oracle.inflate_chunks(comp, blk[1:], CH, threads=4)
    assert np.all(st == 0)
    assert out[: len(raw)].tobytes() == raw.tobytes()
    # Python's zlib agrees on every chunk
    for i in range(n):
        s = comp[int(blk[1 + i]): int(blk[2 + i])].tobytes()
        assert zlib.decompressobj().decompress(s) == out[i * CH:(i + 1) * CH].tobytes()


def test_deflate_payload_is_half_text_half_int32():
    raw = ca.synth_deflate_payload(4 * CH, CH, seed=8)
    text = raw[:CH]
    assert len(set(text.tolist())) <= 64
    ints = raw[CH:2 * CH].view(np.int32)
    assert ints.min() >= 1 and np.mean(ints == 1) > 0.1  # Zipf: 1 is the mode


def test_deflate_mixed_blocks():
    raw = ca.synth_deflate_payload(600000, CH, seed=9)
    s, counts = ca.deflate_compress_mixed(raw, 20, 5, 16384, seed=2)
    assert sum(counts) > 20 and counts[1] > counts[0] > 0
    assert zlib.decompress(s.tobytes()) == raw.tobytes()


def test_rlev2_wide_patch_is_masked_as_the_reference_decodes():
    # A uint64 block of small values and one outlier: the reference encoder
    # (encoder.h:196-266) patches it with a 56-bit patch (header byte 3 =
    # 0x1e: one base byte, patch width code 30); the reference decoder keeps
    # only the low 32 bits of a patch (uint32 patch mask, decoder_new.h:1105),
    # and so does the oracle. Byte identity with the reference therefore
    # means this column does not round-trip, in the reference either.
    small = [121, 160, 96, 122, 150, 173, 117, 101, 197, 193, 36, 178, 92, 5, 16, 180, 77, 27,
             180, 72, 11, 17, 45, 145, 132, 21, 153, 178, 101, 74, 192, 163, 167, 75, 162, 104,
             79, 137, 176, 6, 110, 5]
    big = 237880511751946841
    col = np.array(small + [big], np.uint64)
    comp, blk, cl, n = ca.rlev2_encode(col, 8)
    s = comp[: int(cl[0])].tobytes()
    assert s[:4] == bytes.fromhex("8e2a1ea1")  # PATCHED_BASE, 8-bit values, 56-bit patch
    st, dec = oracle.rlev2_decode_chunk(s, 8, CH // 8)
    assert st == 0 and dec[:42].tolist() == small
    red = big - min(small)
    assert int(dec[42]) == min(small) + ((red & 0xFF) | (((red >> 8) & 0xFFFFFFFF) << 8))
