"""Byte identity of codag_rlev1_encode / codag_rlev2_encode with the
reference encoders.

RLE v1: the reference RLE v1 encoder
(include/rlev1/encoder.h:1352-1695, compression_warp_orig), pinned by vectors
derived by hand from that state machine -- one per encoder choice:

* a run needs three values with one signed 8-bit step, and runs hold at most
  129 values: the 130th value closes the run with header 126 and starts the
  next sequence (encoder.h:1509-1523);
* literal groups hold at most 127 values (encoder.h:1447, 1573); the header
  is a placeholder overwritten with -count when the group closes;
* the step is computed in the column type before widening (encoder.h:1416,
  1459, 1491): a decreasing uint32 column is all literals, the same uint64
  or uint8 column is a run; uint64 steps wrap, uint8 steps do not.

Each chunk is encoded whole; a vector of n values uses chunk_bytes = n * w so
that chunk 0 holds exactly the vector. Every vector also round-trips through
the oracle decoder, which restates include/rlev1/decoder.h:671-791.

RLE v2: the reference's block_encode_orig (include/rlev2/encoder_trans.h:
184-357) over include/rlev2/encoder.h:83-591, values held as uint64
(utils.h:4-8). Vectors derived by hand from that code, one per choice:
SHORT_REPEAT for 3..10 equal values; DIRECT for a block whose minimum is 0
and maximum at most 128 (is_safe_subtract, encoder.h:64-67), with widths
rounded to the aligned widths; DELTA for monotone blocks, its deltas the
unsigned 64-bit differences, a varying width 1 written as 2; PATCHED_BASE
when the 100th and 90th percentile widths differ by more than one. The
SHORT_REPEAT and DELTA vectors equal the ORC specification's examples. Each
also decodes through the oracle, which restates decoder_new.h:909-1187.
"""
import numpy as np
import pytest

import codag_amd as ca
import oracle

DT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def chunk0(values, w):
    col = np.array(values, dtype=DT[w])
    cont, blk, cl, n = ca.rlev1_encode(col, w, len(values) * w)
    s = cont[8 + int(blk[0]): 8 + int(blk[0]) + int(cl[0])].tobytes()
    st, dec = oracle.rlev1_decode_chunk(s, w, len(values) * w)
    assert st == 0 and dec.tolist() == col.tolist()
    return s


def test_run_of_four():
    assert chunk0([5, 6, 7, 8], 8) == bytes.fromhex("010105")


def test_runs_hold_at_most_129_values():
    # 129 values (header 126), then the 130th starts a 71-value run (header 68)
    assert chunk0([9] * 200, 8) == bytes.fromhex("7e0009" "440009")


def test_literal_groups_hold_at_most_127_values():
    vals = [0 if k % 2 == 0 else 1000 for k in range(130)]
    lit = lambda v: b"\x00" if v == 0 else bytes.fromhex("e807")
    exp = b"\x81" + b"".join(lit(v) for v in vals[:127]) + b"\xfd" + \
        b"".join(lit(v) for v in vals[127:])
    assert chunk0(vals, 8) == exp


def test_literals_then_run():
    # the literal header placeholder (0x03) is overwritten with -2
    assert chunk0([1, 100, 200, 201, 202, 203], 8) == bytes.fromhex("fe0164" "0101c801")


@pytest.mark.parametrize("w,exp", [(8, "01ff0a"), (1, "01ff0a"), (2, "01ff0a"),
                                   (4, "fc0a090807")])
def test_decreasing_column_depends_on_the_type(w, exp):
    # uint32: 9 - 10 wraps to 2^32 - 1 before widening, so no step fits
    assert chunk0([10, 9, 8, 7], w) == bytes.fromhex(exp)


def test_uint64_steps_wrap_uint8_steps_do_not():
    top = (1 << 64) - 1
    assert chunk0([top, 0, 1], 8) == bytes.fromhex("0001") + bytes([0xFF] * 9 + [0x01])
    # uint8 255 -> 0 is a step of -255 in int arithmetic: literals
    assert chunk0([255, 0, 1], 1) == bytes.fromhex("fd" "ff01" "00" "01")


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_reference_encoder_round_trips_synthetic_columns(w):
    raw, *_ = ca.synth_rlev1(5000, w, 4096, 1 << min(16, 8 * w - 1), seed=7 + w)
    noise = np.random.default_rng(w).integers(0, 300, len(raw) // w).astype(DT[w])
    for col in (raw.view(DT[w]), noise):
        cont, blk, cl, n = ca.rlev1_encode(col, w, 4096)
        st, out = oracle.rlev1_decode_chunks(cont[8:], blk[:n], cl[:n], w, 4096)
        assert np.all(st == 0)
        assert out[: col.nbytes].tobytes() == col.tobytes()
        assert not out[col.nbytes:].any()  # the zero padding of the last chunks
        # side files: padded offsets, unpadded lengths (encoder.h:780-783)
        assert all(int(blk[i + 1]) - int(blk[i]) == (int(cl[i]) + 3) // 4 * 4 for i in range(n))


def v2chunk0(values, w=8):
    col = np.array(values, dtype=DT[w])
    comp, blk, cl, n = ca.rlev2_encode(col, w, 131072)
    s = comp[int(blk[0]): int(blk[0]) + int(cl[0])].tobytes()
    st, dec = oracle.rlev2_decode_chunk(s, w, 131072 // w)
    assert st == 0 and dec[: len(values)].tolist() == col.tolist()
    return s


@pytest.mark.parametrize("values,exp", [
    ([10000] * 5, "0a2710"),                                       # SHORT_REPEAT (ORC example)
    ([0, 5, 3, 100], "4e03" "00050364"),                           # min 0, max <= 128: DIRECT 8-bit
    ([1, 5, 3, 100], "84030321" "01" "1130" "f0"),                 # PATCHED_BASE: 3-bit, one patch
    ([2, 3, 5, 7, 11, 13, 17, 19, 23, 29], "c6090202" "22424246"),  # DELTA (ORC example)
    ([1, 2, 2, 3], "c2030102" "10"),                               # varying DELTA, width 1 -> 2
    ([10, 8, 6, 5], "fe030a03" + "ff" * 7 + "fe" + "ff" * 8),       # unsigned 64-bit deltas
    ([7] * 200, "c07f0700" "c0470700"),                            # fixed DELTA runs of 128 + 72
])
def test_v2_reference_choices(values, exp):
    assert v2chunk0(values).hex() == exp


def test_v2_direct_quirk_then_delta():
    # 0..127 is one DIRECT block (minimum 0, maximum 127 <= 128), the rest DELTA
    s = v2chunk0(list(range(300)))
    assert s[:2] == bytes.fromhex("4e7f") and s[2:130] == bytes(range(128))
    assert s[130:] == bytes.fromhex("c07f800102" "c02b800202")


@pytest.mark.parametrize("w", [1, 2, 4, 8])
def test_v2_reference_encoder_round_trips(w):
    raw, *_ = ca.synth_rlev2(3 * 8192 // w + 11, w, 8192, seed=30 + w)
    rng = np.random.default_rng(w)
    cols = [raw.view(DT[w]),
            rng.integers(0, 1 << min(8 * w, 20), 5000).astype(DT[w]),
            np.repeat(rng.integers(0, 1 << min(8 * w - 1, 62), 700),
                      rng.integers(1, 15, 700)).astype(DT[w]),
            (np.cumsum(rng.integers(0, 5, 6000)) % (1 << min(8 * w - 1, 62))).astype(DT[w])]
    for col in cols:
        comp, blk, cl, n = ca.rlev2_encode(col, w, 8192)
        assert n % 32 == 0 and int(blk[n]) == col.nbytes  # encoder_trans.h:490, 545
        st, out = oracle.rlev2_decode_chunks(comp, blk[:n], cl, w, 8192)
        assert np.all(st == 0)
        assert out[: col.nbytes].tobytes() == col.tobytes()
        assert all(int(blk[i + 1]) % 128 == 0 for i in range(n - 1))  # 128-byte aligned streams
