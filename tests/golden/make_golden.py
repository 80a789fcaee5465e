"""Write tests/golden/golden.npz: known-answer vectors restated from the
reference's decode unit tests (ducndh/sirius test/cpp/scan/).

Each case holds segment bytes built by the byte layouts the reference's test
synthesizers use (bitpacking_synth.hpp, rle_synth.hpp, alp_synth.hpp) and
the expected output taken from that test's closed form (the REQUIRE lines),
never from the oracle.  Source file:line is stored with every case.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import struct
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "golden.npz"

BITPACKING, RLE, ALP, ALPRD, FSST, UNCOMPRESSED, CONSTANT = 6, 3, 10, 11, 7, 1, 2
BP_CONSTANT, BP_CONSTANT_DELTA, BP_DELTA_FOR, BP_FOR, BP_INVALID = 2, 3, 4, 5, 0

NP = {(1, False): np.uint8, (2, False): np.uint16, (4, True): np.int32, (8, True): np.int64,
      (2, True): np.int16, (4, False): np.uint32, (8, False): np.uint64, (1, True): np.int8}

cases: dict[str, dict] = {}


def add(name, src, codec, ts, segs, expected, canary=False, strings=None):
    """segs: list of (bytes, row_offset, row_count[, bytes_size])"""
    blob = bytearray()
    table = []
    for s in segs:
        b, ro, rc = s[0], s[1], s[2]
        bs = s[3] if len(s) > 3 else len(b)
        while len(blob) % 16:
            blob.append(0)
        table.append((len(blob), bs, ro, rc))
        blob += b
    total = max((ro + rc for _, _, ro, rc in table), default=0)
    cases[name] = dict(src=src, codec=codec, ts=ts, blob=np.frombuffer(bytes(blob), np.uint8),
                       table=np.array(table, np.uint64).reshape(-1, 4), total=total,
                       expected=expected, canary=canary, strings=strings)


def le(fmt, *v):
    return struct.pack("<" + fmt, *v)


def tcode(ts, signed=True):
    return {1: "b", 2: "h", 4: "i", 8: "q"}[ts] if signed else {1: "B", 2: "H", 4: "I", 8: "Q"}[ts]


# ---- LSB-first packing (bitpacking_synth.hpp:40-62, alp_synth.hpp:40-70) ----
def pack(values, width, pad32=False):
    if width == 0 or len(values) == 0:
        return b"" if pad32 else b"\0\0\0\0"
    n = len(values)
    nbits = n * width
    acc = 0
    for i, v in enumerate(values):
        acc |= (int(v) & ((1 << width) - 1)) << (i * width)
    if pad32:  # DuckDB GetRequiredSize: ceil32(n) * width / 8 bytes
        nbytes = ((n + 31) // 32) * 32 * width // 8
    else:  # pack_values: ceil(bits / 32) + 1 guard word
        nbytes = ((nbits + 31) // 32 + 1) * 4
    return acc.to_bytes(nbytes, "little")


# ---- BITPACKING blocks (bitpacking_synth.hpp:64-131) ----------------------
def bp_constant(ts, value, signed=True):
    b = bytearray(64)
    b[0:8] = le("Q", 32)
    b[8:8 + ts] = le(tcode(ts, signed), value)
    b[28:32] = le("I", (BP_CONSTANT << 24) | 8)
    return bytes(b)


def bp_constant_delta(ts, frame, delta):
    b = bytearray(64)
    b[0:8] = le("Q", 32)
    b[8:8 + ts] = le(tcode(ts), frame)
    b[8 + ts:8 + 2 * ts] = le(tcode(ts), delta)
    b[28:32] = le("I", (BP_CONSTANT_DELTA << 24) | 8)
    return bytes(b)


def bp_for(ts, frame, width, values, signed=True):
    p = pack(values, width)
    md_end = 8 + 2 * ts + len(p) + 4
    b = bytearray(max(md_end, 64))
    b[0:8] = le("Q", md_end)
    b[8:8 + ts] = le(tcode(ts, signed), frame)
    b[8 + ts:8 + 2 * ts] = le(tcode(ts, signed), width)
    b[8 + 2 * ts:8 + 2 * ts + len(p)] = p
    b[md_end - 4:md_end] = le("I", (BP_FOR << 24) | 8)
    return bytes(b)


def bp_delta_for(ts, frame, delta_offset, width, values, signed=True):
    p = pack(values, width)
    md_end = 8 + 3 * ts + len(p) + 4
    b = bytearray(max(md_end, 64))
    b[0:8] = le("Q", md_end)
    t = tcode(ts, signed)
    b[8:8 + ts] = le(t, frame)
    b[8 + ts:8 + 2 * ts] = le(t, width)
    b[8 + 2 * ts:8 + 3 * ts] = le(t, delta_offset)
    b[8 + 3 * ts:8 + 3 * ts + len(p)] = p
    b[md_end - 4:md_end] = le("I", (BP_DELTA_FOR << 24) | 8)
    return bytes(b)


def arr(ts, vals, signed=True):
    return np.array(vals, dtype=NP[(ts, signed)]).view(np.uint8)


SRC_BP = "test/cpp/scan/test_gpu_decode_bitpacking.cpp"


def bitpacking_cases():
    add("bp_constant_i32", SRC_BP + ":100-107", BITPACKING, 4, [(bp_constant(4, 42), 0, 100)], arr(4, [42] * 100))
    v = 5_000_000_000
    add("bp_constant_i64", SRC_BP + ":109-117", BITPACKING, 8, [(bp_constant(8, v), 0, 257)], arr(8, [v] * 257))
    add("bp_constant_u8", SRC_BP + ":119-128", BITPACKING, 1, [(bp_constant(1, 0x5A, False), 0, 200)],
        arr(1, [0x5A] * 200, False))
    add("bp_cdelta_i32", SRC_BP + ":134-141", BITPACKING, 4, [(bp_constant_delta(4, 10, 3), 0, 50)],
        arr(4, [10 + 3 * i for i in range(50)]))
    add("bp_cdelta_i64", SRC_BP + ":143-150", BITPACKING, 8, [(bp_constant_delta(8, -1000, 7), 0, 1024)],
        arr(8, [-1000 + 7 * i for i in range(1024)]))
    add("bp_for_w5", SRC_BP + ":156-165", BITPACKING, 4, [(bp_for(4, 100, 5, list(range(32))), 0, 32)],
        arr(4, [100 + i for i in range(32)]))
    d = [(i * 7) % 4096 for i in range(100)]
    add("bp_for_w12", SRC_BP + ":167-176", BITPACKING, 4, [(bp_for(4, -50, 12, d), 0, 100)],
        arr(4, [-50 + x for x in d]))
    d = [(i * 0x1234567) & ((1 << 33) - 1) for i in range(64)]
    add("bp_for_i64_w33", SRC_BP + ":178-191", BITPACKING, 8, [(bp_for(8, 1_000_000, 33, d), 0, 64)],
        arr(8, [1_000_000 + x for x in d]))
    d = [(i * 3) % 1024 for i in range(200)]
    add("bp_for_u16_w10", SRC_BP + ":193-202", BITPACKING, 2, [(bp_for(2, 500, 10, d, False), 0, 200)],
        arr(2, [500 + x for x in d], False))
    d = [i & 1 for i in range(64)]
    add("bp_for_w1", SRC_BP + ":204-215", BITPACKING, 4, [(bp_for(4, 50, 1, d), 0, 64)], arr(4, [50 + x for x in d]))
    d = [0x10000000 + i for i in range(33)]
    add("bp_for_w32", SRC_BP + ":217-229", BITPACKING, 4, [(bp_for(4, 0, 32, d), 0, 33)], arr(4, d))
    d = [(1 << 50) + i for i in range(40)]
    add("bp_for_i64_w64", SRC_BP + ":231-243", BITPACKING, 8, [(bp_for(8, 0, 64, d), 0, 40)], arr(8, d))
    d = [(i * 0x123456789) & ((1 << 63) - 1) for i in range(100)]
    add("bp_for_i64_w63", SRC_BP + ":245-261", BITPACKING, 8, [(bp_for(8, 0, 63, d), 0, 100)], arr(8, d))
    p = [1, 2, 3, 4, 5, 6, 7, 8, 1, 1, 1, 1, 1, 1, 1, 1]
    run, exp = 10, []
    for x in p:
        run += x
        exp.append(run)
    add("bp_dfor_w4", SRC_BP + ":267-277", BITPACKING, 4, [(bp_delta_for(4, 0, 10, 4, p), 0, 16)], arr(4, exp))
    p = [(i * 17 + 3) & 0xFF for i in range(1000)]
    run, exp = 1000, []
    for x in p:
        run += 1 + x
        exp.append(run)
    add("bp_dfor_i64_w8", SRC_BP + ":279-291", BITPACKING, 8, [(bp_delta_for(8, 1, 1000, 8, p), 0, 1000)], arr(8, exp))
    # multi-segment (:294-322)
    add("bp_multi_segment", SRC_BP + ":294-322", BITPACKING, 4,
        [(bp_constant(4, 7), 0, 50), (bp_for(4, 100, 4, [3] * 50), 50, 50)], arr(4, [7] * 50 + [103] * 50))
    # unaligned destination row offset (:324-382)
    a = bp_for(4, 0, 4, [1] * 50)
    add("bp_unaligned_constant", SRC_BP + ":341-360", BITPACKING, 4,
        [(a, 0, 50), (bp_constant(4, 77), 50, 50)], arr(4, [1] * 50 + [77] * 50))
    add("bp_unaligned_cdelta", SRC_BP + ":362-381", BITPACKING, 4,
        [(a, 0, 50), (bp_constant_delta(4, 1000, 3), 50, 50)], arr(4, [1] * 50 + [1000 + 3 * i for i in range(50)]))
    # multi-group segment, two CONSTANT groups (:384-432)
    md_end = 8 + 8 + 8
    b = bytearray(md_end)
    b[0:8] = le("Q", md_end)
    b[8:12] = le("i", 7)
    b[12:16] = le("i", 99)
    b[md_end - 4:md_end] = le("I", (BP_CONSTANT << 24) | 8)
    b[md_end - 8:md_end - 4] = le("I", (BP_CONSTANT << 24) | 12)
    add("bp_two_constant_groups", SRC_BP + ":384-432", BITPACKING, 4, [(bytes(b), 0, 2548)],
        arr(4, [7] * 2048 + [99] * 500))

    def groups_block(groups):
        """groups: list of (mode, header_ts_values, packed_bytes)"""
        data = bytearray()
        entries = []
        for mode, hdr, pk in groups:
            entries.append((mode << 24) | (8 + len(data)))
            for x in hdr:
                data += le("i", x)
            data += pk
        md_end = 8 + len(data) + 4 * len(groups)
        b = bytearray(md_end)
        b[0:8] = le("Q", md_end)
        b[8:8 + len(data)] = data
        for k, e in enumerate(entries):
            b[md_end - 4 * (k + 1):md_end - 4 * k] = le("I", e)
        return bytes(b)

    g1 = [(i * 3 + 1) & 0xFF for i in range(500)]
    blk = groups_block([(BP_CONSTANT, [7], b""), (BP_FOR, [100, 8], pack(g1, 8))])
    add("bp_group1_for", SRC_BP + ":434-481", BITPACKING, 4, [(blk, 0, 2548)], arr(4, [7] * 2048 + [100 + x for x in g1]))
    g1 = [(i * 5 + 7) & 0xFF for i in range(500)]
    blk = groups_block([(BP_CONSTANT, [0], b""), (BP_DELTA_FOR, [2, 8, 1000], pack(g1, 8))])
    run, exp = 1000, []
    for x in g1:
        run += 2 + x
        exp.append(run)
    add("bp_group1_delta_for", SRC_BP + ":483-538", BITPACKING, 4, [(blk, 0, 2548)], arr(4, [0] * 2048 + exp))
    d = [i & 0xFF for i in range(2048)]
    add("bp_full_group", SRC_BP + ":540-555", BITPACKING, 4, [(bp_for(4, 100, 8, d), 0, 2048)], arr(4, [100 + x for x in d]))
    g1 = [(i * 3 + 1) & 0xFF for i in range(2048)]
    g2 = [(i * 5 + 7) & 0xFF for i in range(500)]
    blk = groups_block([(BP_CONSTANT, [7], b""), (BP_FOR, [100, 8], pack(g1, 8)),
                        (BP_DELTA_FOR, [2, 8, 1000], pack(g2, 8))])
    run, e2 = 1000, []
    for x in g2:
        run += 2 + x
        e2.append(run)
    add("bp_three_groups_mixed", SRC_BP + ":557-630", BITPACKING, 4, [(blk, 0, 4596)],
        arr(4, [7] * 2048 + [100 + x for x in g1] + e2))
    p = [(i * 7 + 13) & 0x1F for i in range(1024)]
    run, exp = 100, []
    for x in p:
        run = (run + 1 + x) & 0xFFFF
        exp.append(run)
    add("bp_dfor_u16", SRC_BP + ":632-678", BITPACKING, 2, [(bp_delta_for(2, 1, 100, 8, p, False), 0, 1024)],
        arr(2, exp, False))
    # defensive zero-fill (:680-773)
    b = bytearray(64)
    b[0:8] = le("Q", 32)
    b[28:32] = le("I", (BP_INVALID << 24) | 8)
    add("bp_invalid_mode", SRC_BP + ":680-702", BITPACKING, 4, [(bytes(b), 0, 32)], arr(4, [0] * 32), canary=True)
    b = bytearray(64)
    b[0:8] = le("Q", 1 << 30)
    add("bp_corrupt_metadata_end", SRC_BP + ":704-726", BITPACKING, 4, [(bytes(b), 0, 16)], arr(4, [0] * 16), canary=True)
    b = bytearray(bp_for(4, 0, 5, [1] * 8))
    b[12:16] = le("i", 100)
    add("bp_width_too_large", SRC_BP + ":728-750", BITPACKING, 4, [(bytes(b), 0, 8)], arr(4, [0] * 8), canary=True)
    add("bp_packed_past_end", SRC_BP + ":752-773", BITPACKING, 4, [(bp_for(4, 0, 4, [1] * 8), 0, 2048)],
        arr(4, [0] * 2048), canary=True)


# ---- RLE (rle_synth.hpp:44-60) ----------------------------------------------
SRC_RLE = "test/cpp/scan/test_gpu_decode_rle.cpp"


def rle_block(ts, values, counts, signed=True):
    vb = b"".join(le(tcode(ts, signed), v) for v in values)
    off = (8 + len(vb) + 7) & ~7
    b = bytearray(off + 2 * len(counts))
    b[0:8] = le("Q", off)
    b[8:8 + len(vb)] = vb
    b[off:] = b"".join(le("H", c) for c in counts)
    return bytes(b)


def expand(values, counts):
    out = []
    for v, c in zip(values, counts):
        out += [v] * c
    return out


def rle_cases():
    add("rle_single_i32", SRC_RLE + ":99-106", RLE, 4, [(rle_block(4, [42], [100]), 0, 100)], arr(4, [42] * 100))
    v = 5_000_000_000
    add("rle_single_i64", SRC_RLE + ":108-115", RLE, 8, [(rle_block(8, [v], [50]), 0, 50)], arr(8, [v] * 50))
    add("rle_single_u8", SRC_RLE + ":117-123", RLE, 1, [(rle_block(1, [200], [33], False), 0, 33)],
        arr(1, [200] * 33, False))
    vals, cnts = [10, 20, 30, 40], [5, 3, 8, 4]
    add("rle_four_runs", SRC_RLE + ":128-136", RLE, 4, [(rle_block(4, vals, cnts), 0, 20)], arr(4, expand(vals, cnts)))
    vals, cnts = [-1000, 1000, -1000, 1000, -1000], [7] * 5
    add("rle_i16_alternating", SRC_RLE + ":138-146", RLE, 2, [(rle_block(2, vals, cnts), 0, 35)],
        arr(2, expand(vals, cnts)))
    vals, cnts = [i * 11 for i in range(12)], [256] * 12
    add("rle_cross_chunk", SRC_RLE + ":149-164", RLE, 4, [(rle_block(4, vals, cnts), 0, 3072)], arr(4, expand(vals, cnts)))
    vals, cnts = list(range(64)), [33] * 64
    add("rle_warp_straddle", SRC_RLE + ":166-181", RLE, 4, [(rle_block(4, vals, cnts), 0, 2112)],
        arr(4, expand(vals, cnts)))
    vals = list(range(1, 4001))
    add("rle_near_cap", SRC_RLE + ":183-194", RLE, 4, [(rle_block(4, vals, [1] * 4000), 0, 4000)], arr(4, vals))
    add("rle_multi_segment", SRC_RLE + ":196-225", RLE, 4,
        [(rle_block(4, [7], [50]), 0, 50), (rle_block(4, [100, 200], [30, 20]), 50, 50)],
        arr(4, [7] * 50 + [100] * 30 + [200] * 20))
    off = 8 + 8 + 4
    b = bytearray(off + 4)
    b[0:8] = le("Q", off)
    b[8:16] = le("ii", 77, 88)
    b[off:off + 4] = le("HH", 10, 20)
    add("rle_ondisk_padding", SRC_RLE + ":227-250", RLE, 4, [(bytes(b), 0, 30)], arr(4, [77] * 10 + [88] * 20))
    blk = rle_block(4, [1, 2, 3], [10, 20, 30]) + bytes(4096)
    add("rle_trailing_slack", SRC_RLE + ":252-267", RLE, 4, [(blk, 0, 60)], arr(4, expand([1, 2, 3], [10, 20, 30])))
    seg = rle_block(1, [11, 22, 33], [5, 10, 15], False)
    blk = seg + b"\xff" * 64
    add("rle_narrow_padding", SRC_RLE + ":269-318", RLE, 1, [(blk, 0, 30)],
        arr(1, expand([11, 22, 33], [5, 10, 15]), False))
    vals = list(range(1, 12001))
    add("rle_many_entries", SRC_RLE + ":320-332", RLE, 4, [(rle_block(4, vals, [1] * 12000), 0, 12000)], arr(4, vals))
    vals = list(range(1, 100001))
    add("rle_over_cap", SRC_RLE + ":334-355", RLE, 4, [(rle_block(4, vals, [1] * 100000), 0, 100000)],
        arr(4, [0] * 100000), canary=True)
    b = bytearray(64)
    b[0:8] = le("Q", 1 << 30)
    add("rle_offset_past_segment", SRC_RLE + ":357-378", RLE, 4, [(bytes(b), 0, 32)], arr(4, [0] * 32), canary=True)
    b = bytearray(64)
    b[0:8] = le("Q", 4)
    add("rle_offset_below_header", SRC_RLE + ":380-401", RLE, 4, [(bytes(b), 0, 16)], arr(4, [0] * 16), canary=True)
    add("rle_sum_underflow", SRC_RLE + ":403-422", RLE, 4, [(rle_block(4, [1, 2], [5, 5]), 0, 100)],
        arr(4, [0] * 100), canary=True)
    add("rle_sum_overflow", SRC_RLE + ":424-443", RLE, 4, [(rle_block(4, [1, 2], [50, 60]), 0, 100)],
        arr(4, [0] * 100), canary=True)
    add("rle_zero_count_underflow", SRC_RLE + ":445-466", RLE, 4, [(rle_block(4, [1, 99, 2], [10, 0, 5]), 0, 20)],
        arr(4, [0] * 20), canary=True)
    add("rle_too_small", SRC_RLE + ":468-488", RLE, 4, [(bytes(4), 0, 8)], arr(4, [0] * 8), canary=True)


# ---- ALP / ALPRD (alp_synth.hpp:109-245) ------------------------------------
SRC_ALP = "test/cpp/scan/test_gpu_decode_alp.cpp"
FRAC64 = [float(f"1e-{e}") for e in range(21)]
FRAC32 = [np.float32(float(f"1e-{e}")) for e in range(11)]


def alp_block(ts, vectors):
    """vectors: ('c', exp, fac, for, bw, mantissas, exc_vals, exc_pos) or ('u', raw_values)"""
    t = "d" if ts == 8 else "f"
    seg = bytearray(4)
    offs = []
    for v in vectors:
        offs.append(len(seg))
        if v[0] == "c":
            _, exp, fac, forv, bw, man, ev, ep = v
            seg += bytes([exp, fac]) + le("H", len(ev)) + le("Q", forv) + bytes([bw])
            seg += pack(man, bw, pad32=True)
            seg += b"".join(le(t, x) for x in ev)
            seg += b"".join(le("H", p) for p in ep)
        else:
            seg += b"\xff" + b"".join(le(t, x) for x in v[1])
    for o in reversed(offs):
        seg += le("I", o)
    seg[0:4] = le("I", len(seg))
    return bytes(seg)


def alp_f64(forv, m, fac, exp):
    enc = np.int64(np.uint64((m + forv) & (2 ** 64 - 1)).view(np.int64))
    return np.float64(enc) * np.float64(10 ** fac) * np.float64(FRAC64[exp])


def alp_f32(forv, m, fac, exp):
    enc = np.int64(np.uint64((m + forv) & (2 ** 64 - 1)).view(np.int64))
    return np.float32(enc) * np.float32(10 ** fac) * FRAC32[exp]


def alprd_block(ts, rbw, lbw, dct, vectors):
    """vectors: ('c', left_idx, right, exc_left, exc_pos) or ('u', raw)"""
    t = "d" if ts == 8 else "f"
    seg = bytearray(7) + b"".join(le("H", d) for d in dct)
    seg[4], seg[5], seg[6] = rbw, lbw, len(dct)
    offs = []
    for v in vectors:
        offs.append(len(seg))
        if v[0] == "c":
            _, li, rp, el, ep = v
            seg += le("H", len(ep)) + pack(li, lbw, pad32=True) + pack(rp, rbw, pad32=True)
            seg += b"".join(le("H", x) for x in el) + b"".join(le("H", p) for p in ep)
        else:
            seg += le("H", 0xFFFF) + b"".join(le(t, x) for x in v[1])
    for o in reversed(offs):
        seg += le("I", o)
    seg[0:4] = le("I", len(seg))
    return bytes(seg)


def alp_cases():
    m = list(range(100))
    blk = alp_block(8, [("c", 2, 0, 0, 7, m, [], [])])
    add("alp_f64_exp2", SRC_ALP + ":143-170", ALP, 8, [(blk, 0, 100)],
        np.array([alp_f64(0, x, 0, 2) for x in m], np.float64).view(np.uint8))
    m = [i * 7 for i in range(64)]
    blk = alp_block(4, [("c", 1, 0, 100, 10, m, [3.1415926, 2.7182818], [5, 17])])
    exp = [alp_f32(100, x, 0, 1) for x in m]
    exp[5], exp[17] = np.float32(3.1415926), np.float32(2.7182818)
    add("alp_f32_exceptions", SRC_ALP + ":172-211", ALP, 4, [(blk, 0, 64)], np.array(exp, np.float32).view(np.uint8))
    raw = [i * 3.14159 + 0.5 for i in range(50)]
    add("alp_f64_uncompressed", SRC_ALP + ":213-232", ALP, 8, [(alp_block(8, [("u", raw)]), 0, 50)],
        np.array(raw, np.float64).view(np.uint8))
    m0, m1 = list(range(1024)), [i + 5000 for i in range(300)]
    blk = alp_block(8, [("c", 0, 0, 0, 11, m0, [], []), ("c", 0, 0, 0, 13, m1, [], [])])
    add("alp_two_vectors", SRC_ALP + ":234-277", ALP, 8, [(blk, 0, 1324)], np.array(m0 + m1, np.float64).view(np.uint8))
    dct = [0xABCD, 0x1234]
    li = [i & 1 for i in range(80)]
    rp = [(i * 0x1234567) & ((1 << 48) - 1) for i in range(80)]
    blk = alprd_block(8, 48, 1, dct, [("c", li, rp, [], [])])
    add("alprd_canonical", SRC_ALP + ":279-312", ALPRD, 8, [(blk, 0, 80)],
        np.array([(dct[a] << 48) | b for a, b in zip(li, rp)], np.uint64).view(np.uint8))
    dct = [0x1111, 0x2222, 0x3333, 0x4444]
    rp = [i * 9 for i in range(64)]
    blk = alprd_block(8, 48, 2, dct, [("c", [0] * 64, rp, [0xDEAD], [7])])
    add("alprd_exception", SRC_ALP + ":314-347", ALPRD, 8, [(blk, 0, 64)],
        np.array([((0xDEAD if i == 7 else dct[0]) << 48) | rp[i] for i in range(64)], np.uint64).view(np.uint8))
    raw = [np.sqrt(np.float32(i + 1)) for i in range(30)]
    blk = alprd_block(4, 24, 0, [0], [("u", raw)])
    add("alprd_f32_uncompressed", SRC_ALP + ":349-370", ALPRD, 4, [(blk, 0, 30)], np.array(raw, np.float32).view(np.uint8))
    # defensive zero-fills (:406-680)
    b = bytearray(64)
    b[0:4] = le("I", 1 << 30)
    add("alp_metadata_end_past", SRC_ALP + ":406-427", ALP, 8, [(bytes(b), 0, 32)], arr(8, [0] * 32), canary=True)
    good = alp_block(8, [("c", 0, 0, 0, 8, list(range(32)), [], [])])
    b = bytearray(good)
    md = struct.unpack_from("<I", b, 0)[0]
    b[md - 4:md] = le("I", 2)  # vector pointer below the 4-byte header
    add("alp_pointer_below_header", SRC_ALP + ":429-453", ALP, 8, [(bytes(b), 0, 32)], arr(8, [0] * 32), canary=True)
    b = bytearray(good)
    b[4 + 12] = 65
    add("alp_bw_too_large", SRC_ALP + ":455-488", ALP, 8, [(bytes(b), 0, 32)], arr(8, [0] * 32), canary=True)
    exc = alp_block(8, [("c", 0, 0, 0, 8, list(range(32)), [1.5, 2.5], [1, 2])])
    b = bytearray(exc)
    b[4 + 2:4 + 4] = le("H", 5000)  # exception count runs past metadata_end
    add("alp_exceptions_past_end", SRC_ALP + ":490-523", ALP, 8, [(bytes(b), 0, 32)], arr(8, [0] * 32), canary=True)
    rd = alprd_block(8, 48, 1, [1, 2], [("c", [0] * 16, list(range(16)), [], [])])
    b = bytearray(rd)
    b[6] = 9
    add("alprd_dict_too_large", SRC_ALP + ":525-551", ALPRD, 8, [(bytes(b), 0, 16)], arr(8, [0] * 16), canary=True)
    b = bytearray(good)
    b[4 + 1] = 19
    add("alp_factor_too_large", SRC_ALP + ":553-582", ALP, 8, [(bytes(b), 0, 32)], arr(8, [0] * 32), canary=True)
    b = bytearray(good)
    b[4] = 19
    add("alp_exponent_too_large", SRC_ALP + ":584-613", ALP, 8, [(bytes(b), 0, 32)], arr(8, [0] * 32), canary=True)
    g32 = alp_block(4, [("c", 0, 0, 0, 8, list(range(32)), [], [])])
    b = bytearray(g32)
    b[4] = 11  # float MAX_EXPONENT is 10
    add("alp_f32_exponent_too_large", SRC_ALP + ":584-613", ALP, 4, [(bytes(b), 0, 32)], arr(4, [0] * 32), canary=True)
    b = bytearray(good)
    md = struct.unpack_from("<I", b, 0)[0]
    b[md - 4:md] = le("I", md + 1000)
    add("alp_pointer_past_end", SRC_ALP + ":615-650", ALP, 8, [(bytes(b), 0, 32)], arr(8, [0] * 32), canary=True)
    b = bytearray(rd)
    b[4] = 65
    add("alprd_rbw_too_large", SRC_ALP + ":652-678", ALPRD, 8, [(bytes(b), 0, 16)], arr(8, [0] * 16), canary=True)


# ---- FSST (fsst.cu:17-44, detail/fsst.cuh:50-96) ------------------------------
SRC_STR = "test/cpp/scan/test_gpu_decode_strings.cpp"


def fsst_symtab(symbols, zero_terminated=False):
    """duckdb_fsst_export layout; symbols listed in code order (lengths 2..8, then 1)."""
    histo = [0] * 8
    for s in symbols:
        histo[len(s) - 1] += 1
    if zero_terminated:  # the terminator "\0" is counted but not serialised
        histo[0] += 1
    version = (20190218 << 32) | (len(symbols) << 8) | 1
    body = b"".join(symbols)
    return le("Q", version) + bytes([1 if zero_terminated else 0]) + bytes(histo) + body


def fsst_segment(symtab, comp_rows, width=None, dict_end_override=None):
    """comp_rows: list of compressed byte strings (row order)."""
    lens = [len(r) for r in comp_rows]
    w = max(lens, default=0).bit_length() if width is None else width
    lp = pack(lens, w, pad32=True) if w else b""
    sym_off = 16 + len(lp)
    dict_beg = sym_off + len(symtab)
    total = dict_beg + sum(lens)
    seg = bytearray(total)
    seg[16:16 + len(lp)] = lp
    seg[sym_off:sym_off + len(symtab)] = symtab
    cum = 0
    for r in comp_rows:
        cum += len(r)
        seg[total - cum:total - cum + len(r)] = r
    seg[0:16] = le("IIII", sum(lens), total if dict_end_override is None else dict_end_override, w, sym_off)
    return bytes(seg)


def strings_expected(rows):
    offs = [0]
    for r in rows:
        offs.append(offs[-1] + len(r))
    return np.array(offs, np.int32), np.frombuffer(b"".join(rows), np.uint8).copy()


def fsst_cases():
    # symbols by code: 0 "ab", 1 "cde", 2 "hello wo", 3 "x", 4 " "
    syms = [b"ab", b"cde", b"hello wo", b"x", b" "]
    st = fsst_symtab(syms)
    rows = [b"abcde", b"hello world", b"", b"x xab", b"zq\xe2\x82\xac", b"cdecdecde"]
    comp = [bytes([0, 1]),
            bytes([2, 255, ord("r"), 255, ord("l"), 255, ord("d")]),
            b"",
            bytes([3, 4, 3, 0]),
            bytes([255, ord("z"), 255, ord("q"), 255, 0xE2, 255, 0x82, 255, 0xAC]),
            bytes([1, 1, 1])]
    seg = fsst_segment(st, comp)
    offs, chars = strings_expected(rows)
    add("fsst_handmade", "fsst.cu:17-44 layout; round trip as test_gpu_decode_strings.cpp:422-460", FSST, 0,
        [(seg, 0, len(rows))], chars, strings=offs)
    # trailing escape is dropped (fsst.cu:273-281)
    seg = fsst_segment(st, [bytes([0, 255])])
    offs, chars = strings_expected([b"ab"])
    add("fsst_trailing_escape", "fsst.cu:273-281", FSST, 0, [(seg, 0, 1)], chars, strings=offs)
    # codes with no symbol decode to "corrupt\0" (detail/fsst.cuh:88-92)
    seg = fsst_segment(st, [bytes([7])])
    offs, chars = strings_expected([b"corrupt\0"])
    add("fsst_unused_code", "detail/fsst.cuh:88-92", FSST, 0, [(seg, 0, 1)], chars, strings=offs)
    # zero-terminated table: code 0 is the 1-byte "\0", stored symbols follow
    st0 = fsst_symtab([b"ab"], zero_terminated=True)
    seg = fsst_segment(st0, [bytes([1, 0, 1])])
    offs, chars = strings_expected([b"ab\0ab"])
    add("fsst_zero_terminated", "detail/fsst.cuh:67-77", FSST, 0, [(seg, 0, 1)], chars, strings=offs)
    # defensive (test_gpu_decode_strings.cpp:351-376): empty strings
    good = fsst_segment(st, [bytes([0]), bytes([1])])
    b = bytearray(good)
    b[4:8] = le("I", len(good) + 100)
    offs, chars = strings_expected([b"", b""])
    add("fsst_dict_end_past", SRC_STR + ":351-363", FSST, 0, [(bytes(b), 0, 2)], chars, strings=offs)
    b = bytearray(good)
    b[8:12] = le("I", 33)
    add("fsst_width_too_large", SRC_STR + ":365-374", FSST, 0, [(bytes(b), 0, 2)], chars, strings=offs)
    # multi-segment column: offsets continue across segments
    s1 = fsst_segment(st, [bytes([0, 0]), bytes([3])])
    s2 = fsst_segment(st, [bytes([1]), b"", bytes([2])])
    offs, chars = strings_expected([b"abab", b"x", b"cde", b"", b"hello wo"])
    add("fsst_two_segments", "gpu_decode_strings.cu:257-309 (one column-wide scan)", FSST, 0,
        [(s1, 0, 2), (s2, 2, 3)], chars, strings=offs)


def main():
    bitpacking_cases()
    rle_cases()
    alp_cases()
    fsst_cases()
    out = {}
    for name, c in cases.items():
        p = f"{name}/"
        out[p + "src"] = np.array(c["src"])
        out[p + "meta"] = np.array([c["codec"], c["ts"], c["total"], int(c["canary"])], np.int64)
        out[p + "blob"] = c["blob"]
        out[p + "table"] = c["table"]
        out[p + "expected"] = np.asarray(c["expected"], np.uint8)
        if c["strings"] is not None:
            out[p + "offsets"] = c["strings"]
    np.savez_compressed(OUT, **out)
    print(f"wrote {len(cases)} cases to {OUT}")


if __name__ == "__main__":
    main()
