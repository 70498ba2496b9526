"""Small pure-Python encoders used by the CPU tests to pin the oracle.

These are written from the stream formats (ORC RLE v1 / v2 as the reference
decoders read them) independently of the oracle's decoder code: a decoder that
drops a term, swaps a shift direction or mis-indexes a width table fails the
round trips below. Sizes are kept small (pure-Python loops).
"""
from __future__ import annotations

import numpy as np


def uleb(v: int) -> bytes:
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def zigzag(v: int) -> int:
    return (v << 1) ^ (v >> 63) if v < 0 else v << 1


def rlev1_encode(values, width: int) -> bytes:
    """Greedy ORC v1 encoder: runs of >=3 equal-delta values (|delta|<=127,
    delta computed in the column's unsigned arithmetic as a signed int8),
    literal groups of <=128 values otherwise. Emits header -128 too, which the
    reference encoder never does (encoder.h caps literals at 127)."""
    mask = (1 << (8 * width)) - 1
    vals = [int(v) & mask for v in values]
    out = bytearray()
    i, n = 0, len(vals)
    lits: list[int] = []

    def flush():
        while lits:
            g = lits[:128]
            del lits[:128]
            out.append((256 - len(g)) & 0xFF)
            for v in g:
                out.extend(uleb(v))

    while i < n:
        if i + 2 < n:
            d = (vals[i + 1] - vals[i]) & mask
            sd = d - (1 << (8 * width)) if d >> (8 * width - 1) else d
            if -128 <= sd <= 127 and ((vals[i + 2] - vals[i + 1]) & mask) == d:
                j = i + 2
                while j + 1 < n and j + 1 - i < 130 and ((vals[j + 1] - vals[j]) & mask) == d:
                    j += 1
                flush()
                ln = j - i + 1
                out.append(ln - 3)
                out.append(sd & 0xFF)
                out.extend(uleb(vals[i]))
                i = j + 1
                continue
        lits.append(vals[i])
        i += 1
    flush()
    return bytes(out)


def pack_msb(values, bits: int) -> bytes:
    """MSB-first bit packing, zero-padded to a byte."""
    acc, nacc, out = 0, 0, bytearray()
    for v in values:
        acc = (acc << bits) | (int(v) & ((1 << bits) - 1))
        nacc += bits
        while nacc >= 8:
            nacc -= 8
            out.append((acc >> nacc) & 0xFF)
        acc &= (1 << nacc) - 1
    if nacc:
        out.append((acc << (8 - nacc)) & 0xFF)
    return bytes(out)


DECODE_WIDTHS = [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22,
                 23, 24, 26, 28, 30, 32, 40, 48, 56, 64]


def v2_direct(values, code: int) -> bytes:
    n = len(values)
    assert 1 <= n <= 512
    h = 0x40 | (code << 1) | (((n - 1) >> 8) & 1)
    return bytes([h, (n - 1) & 0xFF]) + pack_msb(values, DECODE_WIDTHS[code])


def v2_short_repeat(value: int, count: int, nbytes: int) -> bytes:
    assert 3 <= count <= 10 and 1 <= nbytes <= 8
    h = ((nbytes - 1) << 3) | (count - 3)
    return bytes([h]) + int(value).to_bytes(nbytes, "big")


def v2_delta(base: int, base_delta: int, deltas, code: int | None) -> bytes:
    """code None -> fixed delta (encoded width 0); len = 2 + len(deltas)."""
    n = 2 + len(deltas)
    enc = 0 if code is None else code
    h = 0xC0 | (enc << 1) | (((n - 1) >> 8) & 1)
    body = uleb(base) + uleb(zigzag(base_delta) & ((1 << 64) - 1))
    if code is not None:
        body += pack_msb(deltas, DECODE_WIDTHS[code])
    return bytes([h, (n - 1) & 0xFF]) + body


def v2_patched_base(reduced, base: int, code: int, bw: int, patches, pw_code: int, pgw: int):
    """reduced: low bits of each value (before base); patches: list of
    (gap, patch_value) with gap relative to the previous patch."""
    n = len(reduced)
    h = 0x80 | (code << 1) | (((n - 1) >> 8) & 1)
    third = ((bw - 1) << 5) | pw_code
    fourth = ((pgw - 1) << 5) | len(patches)
    pw = DECODE_WIDTHS[pw_code]
    cfb = closest_fixed(pw + pgw)
    entries = [(g << pw) | p for g, p in patches]
    return (bytes([h, (n - 1) & 0xFF, third, fourth]) + int(base).to_bytes(bw, "big")
            + pack_msb(reduced, DECODE_WIDTHS[code]) + pack_msb(entries, cfb))


def closest_fixed(bits: int) -> int:
    for w in [1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23,
              24, 26, 28, 30, 32, 40, 48, 56, 64]:
        if bits <= w:
            return w
    raise ValueError(bits)


def as_np(values, width: int) -> np.ndarray:
    dt = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[width]
    return np.array([int(v) & ((1 << (8 * width)) - 1) for v in values], dtype=dt)
