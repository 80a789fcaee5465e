"""The oracle's DICT_FSST decoder (src/cuda/scan/strings/dict_fsst.cu):
known-answer segments assembled byte by byte for modes 0, 1 and 2 (header,
aligned regions, bit-packed lengths and indices, an FSST symbol table with
an escape and a trailing escape), the special cases the reference's rules
fix (NULL entry 0, index past the dictionary, invalid headers, inline NULLs
and the mask they need), and round trips of encoded columns."""
import ctypes as C
import struct

import numpy as np
import pytest

import sirius_amd._abi as A
import sirius_amd.encode as E
from oracle import oracle as O

FSST_VERSION = 20190218


def symtab_ab_c():
    """Exported FSST table with "ab" (a 2-byte symbol: code 0) and "c" (a
    1-byte symbol, whose length class comes last: code 1)."""
    version = (FSST_VERSION << 32) | (2 << 8)
    histo = [1, 1, 0, 0, 0, 0, 0, 0]  # [len 1, len 2, ..., len 8]
    return struct.pack("<Q", version) + bytes([0]) + bytes(histo) + b"ab" + b"c"


def mode0_segment():
    # dict "", "ab", "xyz"; rows select 1, 2, 0 (NULL), 1, 2
    hdr = struct.pack("<IIBBBBI", 5, 3, 0, 2, 2, 0, 0)
    dict_ = b"abxyz" + bytes(3)          # 16..20, padded to 24
    slens = bytes([0x38]) + bytes(7)     # [0, 2, 3] x 2 bits at 24, padded to 32
    didx = bytes([0x49, 0x02]) + bytes(6)  # [1, 2, 0, 1, 2] x 2 bits at 32, padded to 40
    return hdr + dict_ + slens + didx


def mode1_segment():
    # dict "", "abc" = [0, 1], "cz" = [1, ESC, 'z'], "ab" + trailing escape = [0, ESC]
    st = symtab_ab_c()
    assert len(st) == 20
    hdr = struct.pack("<IIBBBBI", 7, 4, 1, 2, 2, 0, len(st))
    dict_ = bytes([0, 1, 1, 255, ord("z"), 0, 255]) + bytes(1)  # 16..22, padded to 24
    tab = st + bytes(4)                                           # 24..43, padded to 48
    slens = bytes([0xB8]) + bytes(7)                              # [0, 2, 3, 2] at 48
    didx = bytes([0x39, 0x02]) + bytes(6)                         # [1, 2, 3, 0, 2] at 56
    return hdr + dict_ + tab + slens + didx


def mode2_segment():
    # entries "", "abc" = [0, 1], "c" = [1]; row i is entry i + 1
    st = symtab_ab_c()
    hdr = struct.pack("<IIBBBBI", 3, 3, 2, 2, 2, 0, len(st))
    dict_ = bytes([0, 1, 1]) + bytes(5)  # 16..18, padded to 24
    tab = st + bytes(4)                  # 24..43, padded to 48
    slens = bytes([0x18]) + bytes(7)     # [0, 2, 1] at 48; no index region
    return hdr + dict_ + tab + slens


def rows_of(lengths, chars):
    out, o = [], 0
    for n in lengths:
        out.append(bytes(chars[o:o + int(n)]))
        o += int(n)
    return out


def test_mode0_known_answer():
    seg = mode0_segment()
    assert len(seg) == 40
    lengths, chars, null = O.dict_fsst_segment(seg, 5)
    assert rows_of(lengths, chars) == [b"ab", b"xyz", b"", b"ab", b"xyz"]
    assert null  # entry 0 is empty and the dictionary has more than one entry


def test_mode1_known_answer_with_escapes():
    seg = mode1_segment()
    assert len(seg) == 64
    lengths, chars, null = O.dict_fsst_segment(seg, 5)
    # a trailing escape still counts one decoded byte (written as 0 here)
    assert rows_of(lengths, chars) == [b"abc", b"cz", b"ab\x00", b"", b"cz"]
    assert null


def test_mode2_known_answer():
    seg = mode2_segment()
    lengths, chars, null = O.dict_fsst_segment(seg, 2)
    assert rows_of(lengths, chars) == [b"abc", b"c"]
    assert not null  # mode 2 never has inline NULLs
    # a row past the dictionary (row 2 would be entry 3 of 3) is empty
    lengths, chars, _ = O.dict_fsst_segment(seg, 3)
    assert list(lengths) == [3, 1, 0]


def test_index_past_the_dictionary_is_empty():
    seg = bytearray(mode0_segment())
    seg[32] = 0x4B  # row 0 selects 3 (binary 11): dict_count is 3
    lengths, chars, _ = O.dict_fsst_segment(bytes(seg), 5)
    assert rows_of(lengths, chars) == [b"", b"xyz", b"", b"ab", b"xyz"]


@pytest.mark.parametrize("patch", ["mode3", "short", "didx_outside", "wide"])
def test_invalid_segments_decode_empty(patch):
    seg = bytearray(mode0_segment())
    rows = 5
    if patch == "mode3":
        seg[8] = 3
    elif patch == "short":
        seg = seg[:15]
    elif patch == "didx_outside":
        seg = seg[:31]  # the index region would start at 32
    elif patch == "wide":
        seg[10] = 33  # index width above 32
    lengths, chars, null = O.dict_fsst_segment(bytes(seg), rows)
    assert list(lengths) == [0] * rows and not null


def test_small_dictionary_in_fsst_mode_uses_raw_lengths():
    # mode 1 with only the NULL entry: no symbol table is read, rows are empty
    hdr = struct.pack("<IIBBBBI", 0, 1, 1, 1, 1, 0, 0)
    seg = hdr + bytes([0]) * 8 + bytes(8)  # slens [0] at 16, indices [0, 0] at 24
    lengths, chars, null = O.dict_fsst_segment(seg, 2)
    assert list(lengths) == [0, 0] and not null


def _column(segs_and_rows, mask=True):
    blob = b"".join(s for s, _ in segs_and_rows)
    buf = np.frombuffer(blob, np.uint8).copy()
    segs, off, row = [], 0, 0
    for s, n in segs_and_rows:
        segs.append(A.Segment(buf.ctypes.data + off, len(s), row, n, 0))
        off += len(s)
        row += n
    runs, keep = A.make_runs([(A.DICT_FSST, segs)])
    offs = np.zeros(row + 1, np.int32)
    chars = np.zeros(4096, np.uint8)
    m = np.zeros((row + 7) // 8, np.uint8)
    col = A.StringColumn()
    col.total_rows, col.data, col.n_data = row, runs, 1
    col.d_offsets, col.d_chars, col.chars_capacity = offs.ctypes.data, chars.ctypes.data, chars.size
    col.d_null_mask = m.ctypes.data if mask else None
    rc = O.load().oracle_decode_strings_column_mt(C.byref(col), 1)
    return rc, col, offs, chars, np.unpackbits(m, bitorder="little")[:row].astype(bool), (buf, keep)


def test_inline_nulls_mark_index_zero_rows_of_every_non_fsst_only_segment():
    bad = bytearray(mode0_segment())
    bad[8] = 3  # invalid: all its rows are NULL once the column has inline NULLs
    rc, col, offs, chars, valid, _ = _column([(mode0_segment(), 5), (mode2_segment(), 2), (bytes(bad), 3),
                                              (mode1_segment(), 5)])
    assert rc == A.SDR_OK
    assert list(valid) == [1, 1, 0, 1, 1] + [1, 1] + [0, 0, 0] + [1, 1, 1, 0, 1]
    assert col.null_count == 5
    assert rows_of(np.diff(offs), chars) == [b"ab", b"xyz", b"", b"ab", b"xyz", b"abc", b"c", b"", b"", b"",
                                             b"abc", b"cz", b"ab\x00", b"", b"cz"]


def test_inline_nulls_need_a_mask():
    rc, *_ = _column([(mode0_segment(), 5)], mask=False)
    assert rc == A.SDR_E_INVALID_ARG
    rc, col, *_ = _column([(mode2_segment(), 2)], mask=False)  # no inline NULLs: no mask needed
    assert rc == A.SDR_OK and col.chars_used == 4


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_encoded_round_trip(mode):
    chars, offs = E.gen_zipf_strings(11 + mode, 30000)
    n = offs.size - 1
    valid = None if mode == 2 else (np.arange(n) % 17 != 3).astype(np.uint8)
    enc = E.encode_dict_fsst(chars, offs, mode, valid=valid, seg_rows=4096)
    assert len(enc.segs) >= 7
    got_offs, got_chars, got_valid, nulls = O.decode_strings_masked(enc, threads=4)
    want_len = np.diff(offs.astype(np.int64))
    if valid is not None:
        want_len[valid == 0] = 0
        assert nulls == int((valid == 0).sum()) and np.array_equal(got_valid, valid.astype(bool))
    else:
        assert nulls == 0 and got_valid.all()
    assert np.array_equal(np.diff(got_offs.astype(np.int64)), want_len)
    keep = np.repeat(valid.astype(bool) if valid is not None else np.ones(n, bool), np.diff(offs.astype(np.int64)))
    assert np.array_equal(got_chars, chars[:int(offs[-1])][keep])
