"""Pins of the oracle that do not come from the oracle's own formulas:
brute force over every bit width with an independent bit-string packer,
the reference's RLE walk rules stated as cases, ALP/ALPRD against numpy
IEEE arithmetic, FSST symbol-table import corner cases, encoder round trips
(independent C++ code) and the generator's PRNG against a Python one."""
import struct

import numpy as np
import pytest

import sirius_amd._abi as A
import sirius_amd.encode as E
from oracle import oracle as O

RNG = np.random.default_rng(20240601)
UT = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def bitstring_pack(values, w, pad32=True):
    """Independent packer: build the LSB-first stream as a '0'/'1' string."""
    bits = "".join(format(int(v), f"0{w}b")[::-1] if w else "" for v in values)
    n = len(values)
    total = ((n + 31) // 32) * 32 * w if pad32 else len(bits)
    bits = bits.ljust(max(total, len(bits)) + 64, "0")
    nbytes = ((n + 31) // 32) * 32 * w // 8 if pad32 else (len(bits) + 7) // 8
    return bytes(int(bits[8 * i:8 * i + 8][::-1], 2) for i in range(nbytes))


def bp_segment(ts, mode, hdr_vals, packed=b""):
    t = {1: "B", 2: "H", 4: "I", 8: "Q"}[ts]
    data = b"".join(struct.pack("<" + t, v & ((1 << (8 * ts)) - 1)) for v in hdr_vals) + packed
    md_end = 8 + len(data) + 4
    return struct.pack("<Q", md_end) + data + struct.pack("<I", (mode << 24) | 8)


@pytest.mark.parametrize("ts", [1, 2, 4, 8])
def test_bitpacking_for_every_width(ts):
    bits = 8 * ts
    for w in range(0, bits + 1):
        n = int(RNG.integers(1, 200))
        vals = [int(x) for x in RNG.integers(0, 2 ** w if w else 1, n, dtype=np.uint64)] if w < 64 else \
            [int(x) for x in RNG.integers(0, 2 ** 63, n, dtype=np.uint64) * 2 + RNG.integers(0, 2, n, dtype=np.uint64)]
        frame = int(RNG.integers(0, 2 ** min(bits, 63)))
        seg = bp_segment(ts, 5, [frame, w], bitstring_pack(vals, w))
        got = O.decode_segment(A.BITPACKING, seg, n, ts).view(UT[ts])
        want = np.array([(frame + v) % (1 << bits) for v in vals], dtype=UT[ts])
        np.testing.assert_array_equal(got, want, err_msg=f"ts={ts} w={w}")


@pytest.mark.parametrize("ts", [1, 2, 4, 8])
def test_bitpacking_delta_for_wraps(ts):
    bits = 8 * ts
    w = min(bits, 7)
    n = 2048
    vals = [int(x) for x in RNG.integers(0, 2 ** w, n)]
    frame, doff = (1 << (bits - 1)) + 3, (1 << bits) - 5
    seg = bp_segment(ts, 4, [frame, w, doff], bitstring_pack(vals, w))
    got = O.decode_segment(A.BITPACKING, seg, n, ts).view(UT[ts])
    run, want = doff, []
    for v in vals:
        run = (run + frame + v) % (1 << bits)
        want.append(run)
    np.testing.assert_array_equal(got, np.array(want, dtype=UT[ts]))


def test_bitpacking_width_stored_as_t_truncates_to_32_bits():
    """The kernel reads the width as (uint32_t)T (gpu_decode_bitpacking.cu:163);
    an int64 width field 2^32 + 5 is width 5."""
    vals = [3, 1, 4, 1, 5]
    seg = bp_segment(8, 5, [10, (1 << 32) + 5], bitstring_pack(vals, 5))
    got = O.decode_segment(A.BITPACKING, seg, 5, 8).view(np.uint64)
    np.testing.assert_array_equal(got, np.array([10 + v for v in vals], np.uint64))


def rle_seg(values, counts, ts=4):
    t = {1: "B", 2: "H", 4: "I", 8: "Q"}[ts]
    vb = b"".join(struct.pack("<" + t, v) for v in values)
    off = (8 + len(vb) + 7) & ~7
    return struct.pack("<Q", off) + vb + b"\0" * (off - 8 - len(vb)) + b"".join(struct.pack("<H", c) for c in counts)


def test_rle_zero_count_in_earlier_tile_is_malformed():
    """A zero count in a 2048-entry tile with no exact match makes the
    segment malformed even if a later tile matches (gpu_decode_rle.cu:262-316)."""
    n = 3000
    counts = [1] * n
    counts[100] = 0
    rows = n - 1
    seg = rle_seg(list(range(n)), counts)
    got = O.decode_segment(A.RLE, seg, rows, 4).view(np.uint32)
    assert not got.any()


def test_rle_zero_count_in_matching_tile_is_skipped():
    """In the tile where the running count first equals row_count, a zero
    count is accepted (match wins) and its run covers no rows."""
    n = 3000
    counts = [1] * n
    counts[2100] = 0
    rows = n - 1
    vals = list(range(n))
    seg = rle_seg(vals, counts)
    got = O.decode_segment(A.RLE, seg, rows, 4).view(np.uint32)
    want = [v for v, c in zip(vals, counts) for _ in range(c)]
    np.testing.assert_array_equal(got, np.array(want, np.uint32))


def test_rle_overshoot_is_malformed_and_exact_match_stops_walk():
    seg = rle_seg([1, 2, 3], [4, 4, 4])
    assert not O.decode_segment(A.RLE, seg, 6, 4).any()  # 4, 8: passes 6 without matching
    got = O.decode_segment(A.RLE, seg, 8, 4).view(np.uint32)  # match at entry 1; entry 2 unused
    np.testing.assert_array_equal(got, np.array([1] * 4 + [2] * 4, np.uint32))


def test_rle_long_runs_and_all_widths():
    for ts in (1, 2, 4, 8):
        vals = [int(x) for x in RNG.integers(0, 2 ** (8 * ts - 1), 50, dtype=np.uint64)]
        counts = [int(x) for x in RNG.integers(1, 65536, 50)]
        seg = rle_seg(vals, counts, ts)
        rows = sum(counts)
        got = O.decode_segment(A.RLE, seg, rows, ts).view(UT[ts])
        want = np.repeat(np.array(vals, dtype=UT[ts]), counts)
        np.testing.assert_array_equal(got, want)


def alp_vec_bytes(exp, fac, forv, bw, mant, exc_vals, exc_pos, t):
    return (bytes([exp, fac]) + struct.pack("<HQ", len(exc_vals), forv) + bytes([bw]) +
            bitstring_pack(mant, bw) + b"".join(struct.pack("<" + t, x) for x in exc_vals) +
            b"".join(struct.pack("<H", p) for p in exc_pos))


def alp_seg(vectors):
    body = b""
    offs = []
    for vb in vectors:
        offs.append(4 + len(body))
        body += vb
    trailer = b"".join(struct.pack("<I", o) for o in reversed(offs))
    return struct.pack("<I", 4 + len(body) + len(trailer)) + body + trailer


@pytest.mark.parametrize("ts", [4, 8])
def test_alp_against_numpy_ieee(ts):
    ft, t = (np.float64, "d") if ts == 8 else (np.float32, "f")
    max_e = 18 if ts == 8 else 10
    for trial in range(40):
        e = int(RNG.integers(0, max_e + 1))
        f = int(RNG.integers(0, 19))
        bw = int(RNG.integers(0, min(40, 8 * ts) + 1))
        n = int(RNG.integers(1, 1025))
        forv = int(RNG.integers(0, 2 ** 63, dtype=np.uint64)) * int(RNG.integers(0, 3))
        mant = [int(x) for x in RNG.integers(0, 2 ** bw if bw else 1, n, dtype=np.uint64)]
        seg = alp_seg([alp_vec_bytes(e, f, forv & (2 ** 64 - 1), bw, mant, [], [], t)])
        got = O.decode_segment(A.ALP, seg, n, ts).view(ft)
        enc = (np.array(mant, np.uint64) + np.uint64(forv & (2 ** 64 - 1))).view(np.int64)
        frac = ft(float(f"1e-{e}"))
        want = enc.astype(ft) * ft(10 ** f) * frac
        np.testing.assert_array_equal(got.view(UT[ts]), want.view(UT[ts]), err_msg=f"e={e} f={f} bw={bw}")


def test_alp_exceptions_last_wins_and_out_of_range_ignored():
    mant = list(range(16))
    seg = alp_seg([alp_vec_bytes(0, 0, 0, 4, mant, [1.5, 2.5, 3.5, 9.0], [3, 3, 20, 5], "d")])
    got = O.decode_segment(A.ALP, seg, 16, 8).view(np.float64)
    want = np.arange(16, dtype=np.float64)
    want[3], want[5] = 2.5, 9.0
    np.testing.assert_array_equal(got, want)


def test_alprd_bit_assembly_and_full_right():
    n = 100
    dct = [0x3FF0, 0x4000, 0x0001]
    left = [int(x) for x in RNG.integers(0, 3, n)]
    right = [int(x) for x in RNG.integers(0, 2 ** 48, n, dtype=np.uint64)]
    hdr = struct.pack("<I", 0) + bytes([48, 2, 3]) + b"".join(struct.pack("<H", d) for d in dct)
    vec = struct.pack("<H", 1) + bitstring_pack(left, 2) + bitstring_pack(right, 48) + struct.pack("<HH", 0xBEEF, 7)
    body = hdr + vec
    seg = bytearray(body + struct.pack("<I", len(hdr)))
    seg[0:4] = struct.pack("<I", len(seg))
    got = O.decode_segment(A.ALPRD, bytes(seg), n, 8).view(np.uint64)
    want = np.array([(dct[l] << 48) | r for l, r in zip(left, right)], np.uint64)
    want[7] = (0xBEEF << 48) | right[7]
    np.testing.assert_array_equal(got, want)
    # right width 64: the left index and exceptions do not change the bits
    full = [int(x) for x in RNG.integers(0, 2 ** 63, n, dtype=np.uint64)]
    hdr = struct.pack("<I", 0) + bytes([64, 1, 1]) + struct.pack("<H", 0xFFFF)
    vec = struct.pack("<H", 1) + bitstring_pack([1] * n, 1) + bitstring_pack(full, 64) + struct.pack("<HH", 0x1234, 3)
    seg = bytearray(hdr + vec + struct.pack("<I", len(hdr)))
    seg[0:4] = struct.pack("<I", len(seg))
    got = O.decode_segment(A.ALPRD, bytes(seg), n, 8).view(np.uint64)
    np.testing.assert_array_equal(got, np.array(full, np.uint64))


def fsst_seg(symtab: bytes, rows_comp):
    lens = [len(r) for r in rows_comp]
    w = max(lens + [0]).bit_length()
    lp = bitstring_pack(lens, w) if w else b""
    sym_off = 16 + len(lp)
    total = sym_off + len(symtab) + sum(lens)
    seg = bytearray(total)
    seg[16:16 + len(lp)] = lp
    seg[sym_off:sym_off + len(symtab)] = symtab
    cum = 0
    for r in rows_comp:
        cum += len(r)
        seg[total - cum:total - cum + len(r)] = r
    seg[0:16] = struct.pack("<IIII", sum(lens), total, w, sym_off)
    return bytes(seg)


def fsst_decode(seg, n):
    lengths = np.zeros(n, np.uint32)
    chars = np.zeros(4096, np.uint8)
    lib = O.load()
    buf = np.frombuffer(seg, np.uint8)
    assert lib.oracle_fsst_decode(buf.ctypes.data, len(seg), n, lengths.ctypes.data, chars.ctypes.data, 4096) == 0
    out, o = [], 0
    for l in lengths:
        out.append(bytes(chars[o:o + l]))
        o += int(l)
    return out


def symtab(histo, body, zero_term=0, version=20190218):
    return struct.pack("<Q", version << 32) + bytes([zero_term]) + bytes(histo) + body


def test_fsst_import_group_order_and_corrupt_fill():
    # histo[k] counts symbols of length k+1; code order is lengths 2..8 then 1
    st = symtab([2, 1, 0, 0, 0, 0, 0, 1], b"xy" + b"abcdefgh" + b"P" + b"Q")
    # codes: 0 "xy", 1 "abcdefgh", 2 "P", 3 "Q", 4.. "corrupt\0"
    got = fsst_decode(fsst_seg(st, [bytes([0, 1]), bytes([2, 3, 255, 0x41]), bytes([4])]), 3)
    assert got == [b"xyabcdefgh", b"PQA", b"corrupt\0"]


def test_fsst_version_mismatch_keeps_only_literals():
    st = symtab([1, 0, 0, 0, 0, 0, 0, 0], b"Z", version=20190219)
    got = fsst_decode(fsst_seg(st, [bytes([0, 255, 0x42, 0, 255])]), 1)
    assert got == [b"B"]


def test_fsst_escaped_escape_byte_is_literal():
    st = symtab([1, 0, 0, 0, 0, 0, 0, 0], b"Z")
    got = fsst_decode(fsst_seg(st, [bytes([255, 255, 0, 255, 255, 255, 0])]), 1)
    assert got == [b"\xffZ\xff\x00"]  # ESC 255, code 0, ESC 255, ESC 0


def test_fsst_histogram_past_255_symbols_is_capped():
    st = symtab([200, 100, 0, 0, 0, 0, 0, 0], b"ab" * 100 + bytes(range(200)))
    got = fsst_decode(fsst_seg(st, [bytes([0, 99, 100, 254])]), 1)
    assert got == [b"ab" + b"ab" + bytes([0]) + bytes([154])]


# ---- encoder round trips (C++ encoders vs C oracle) --------------------------

SPECIAL = np.array([0.0, 1.0, -1.0, np.inf, -np.inf, np.nan, -0.0, 5e-324, 1.7976931348623157e308, 0.1, 1e-300,
                    123456789.125, -2.5e-5], np.float64)


@pytest.mark.parametrize("dtype", [np.int8, np.int16, np.int32, np.int64, np.uint8, np.uint16, np.uint32, np.uint64])
def test_bitpacking_encoder_roundtrip_extremes(dtype):
    info = np.iinfo(dtype)
    n = 50_000
    v = RNG.integers(info.min, info.max, n, dtype=dtype, endpoint=True)
    v[::7] = info.min
    v[::11] = info.max
    v[:5000] = np.arange(5000, dtype=dtype) * 3  # constant delta groups
    v[5000:8000] = 42  # constant groups
    enc = E.encode_bitpacking(v)
    np.testing.assert_array_equal(O.decode_encoded(enc).view(dtype), v)


@pytest.mark.parametrize("ts", [4, 8])
def test_alp_encoder_roundtrip_special_values(ts):
    ft = np.float64 if ts == 8 else np.float32
    n = 30_000
    v = np.round(RNG.random(n) * 1000, 2).astype(ft)
    with np.errstate(over="ignore"):  # the float64 maximum becomes inf in float32, on purpose
        v[RNG.integers(0, n, 500)] = np.resize(SPECIAL.astype(ft), 500)
    for enc in (E.encode_alp(v), E.encode_alprd(v)):
        np.testing.assert_array_equal(O.decode_encoded(enc).view(UT[ts]), v.view(UT[ts]))


def test_rle_encoder_splits_runs_longer_than_65535():
    v = np.repeat(np.array([7, 8, 7], np.int32), [70000, 3, 200000])
    enc = E.encode_rle(v, rg_rows=1 << 20)
    np.testing.assert_array_equal(O.decode_encoded(enc).view(np.int32), v)


def test_prng_matches_python_xoshiro():
    def splitmix(x):
        x = (x + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        return x, z ^ (z >> 31)

    def rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & (2 ** 64 - 1)

    x, s = 0xC0DE0001, []
    for _ in range(4):
        x, z = splitmix(x)
        s.append(z)
    out = []
    for _ in range(100):
        r = (rotl((s[1] * 5) & (2 ** 64 - 1), 7) * 9) & (2 ** 64 - 1)
        t = (s[1] << 17) & (2 ** 64 - 1)
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45)
        out.append(r)
    np.testing.assert_array_equal(E.gen_u64(0xC0DE0001, 100), np.array(out, np.uint64))
