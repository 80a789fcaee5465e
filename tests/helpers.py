"""Shared test helpers: golden fixtures, device upload, column decode calls."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import sirius_amd._abi as A

GOLDEN = Path(__file__).resolve().parent / "golden" / "golden.npz"


def golden_cases():
    z = np.load(GOLDEN, allow_pickle=False)
    names = sorted({k.split("/")[0] for k in z.files})
    out = []
    for n in names:
        codec, ts, total, canary = (int(x) for x in z[f"{n}/meta"])
        out.append(dict(name=n, src=str(z[f"{n}/src"]), codec=codec, ts=ts, total=total, canary=bool(canary),
                        blob=z[f"{n}/blob"], table=z[f"{n}/table"], expected=z[f"{n}/expected"],
                        offsets=z[f"{n}/offsets"] if f"{n}/offsets" in z.files else None))
    return out


def segs_from_table(table, base_ptr):
    return [A.Segment(base_ptr + int(o), int(bs), int(ro), int(rc), 0) for o, bs, ro, rc in table]


def fixed_column(runs, type_size, total_rows, d_out, validity=None, d_mask=None):
    """runs: [(codec, [Segment])] -> (Column, keepalive)"""
    r, keep = A.make_runs(runs)
    col = A.Column()
    col.type_size = type_size
    col.total_rows = total_rows
    col.data = r
    col.n_data = len(runs)
    col.d_out = d_out
    if validity:
        v, k2 = A.make_runs(validity)
        keep += k2
        col.validity = v
        col.n_validity = len(validity)
        col.has_nulls = 1
        col.d_null_mask = d_mask
    return col, keep


def string_column(runs, total_rows, d_offsets, d_chars, cap, validity=None, d_mask=None):
    r, keep = A.make_runs(runs)
    col = A.StringColumn()
    col.total_rows = total_rows
    col.data = r
    col.n_data = len(runs)
    col.d_offsets = d_offsets
    col.d_chars = d_chars
    col.chars_capacity = cap
    if validity:
        v, k2 = A.make_runs(validity)
        keep += k2
        col.validity = v
        col.n_validity = len(validity)
        col.has_nulls = 1
        col.d_null_mask = d_mask
    return col, keep


DUCKDB = Path(__file__).resolve().parent / "golden" / "duckdb_segments.npz"


def duckdb_cases():
    """Segments written by DuckDB itself (tests/golden/make_duckdb_fixtures.py)."""
    import json
    z = np.load(DUCKDB, allow_pickle=False)
    names = sorted({k.split("/")[0] for k in z.files})
    out = []
    for n in names:
        kind, ts, total, nseg, nval = (int(x) for x in z[f"{n}/meta"])
        get = lambda k: z[f"{n}/{k}"] if f"{n}/{k}" in z.files and z[f"{n}/{k}"].size else None  # noqa: E731
        out.append(dict(name=n, src=json.loads(str(z[f"{n}/src"])), string=bool(kind), ts=ts, total=total,
                        blob=z[f"{n}/blob"], table=z[f"{n}/table"], stats=z[f"{n}/stats"],
                        vblob=get("vblob"), vtable=get("vtable"), expected=get("expected"),
                        offsets=get("offsets"), chars=get("chars")))
    return out


def duckdb_segments(table, base_ptr):
    """Fixture table rows -> (codec, Segment) runs, one run per codec in row order."""
    runs = []
    for o, bs, ro, rc, codec, msl in table:
        seg = A.Segment(base_ptr + int(o), int(bs), int(ro), int(rc), int(msl))
        if runs and runs[-1][0] == int(codec):
            runs[-1][1].append(seg)
        else:
            runs.append((int(codec), [seg]))
    return runs


def validity_bits(case):
    """Per-row validity (1 = valid) from the fixture's UNCOMPRESSED validity segments, else all ones."""
    valid = np.ones(case["total"], np.uint8)
    if case["vtable"] is not None:
        for o, bs, ro, rc, codec, _ in case["vtable"]:
            bits = np.unpackbits(case["vblob"][int(o):int(o) + int(bs)], bitorder="little")[:int(rc)]
            valid[int(ro):int(ro) + int(rc)] = bits
    return valid


# ---------------------------------------------------------------------------
# DICTIONARY / UNCOMPRESSED string segments in DuckDB's layout (test inputs;
# the layout is the one the DuckDB-written fixtures decode with)
# ---------------------------------------------------------------------------
def _pack_lsb(values, width):
    """LSB-first bit stream of `values`, padded to whole groups of 32 values."""
    n = len(values)
    total_bits = ((n + 31) // 32) * 32 * width
    buf = np.zeros((total_bits + 7) // 8 + 8, np.uint8)
    acc = 0
    for i, v in enumerate(values):
        acc |= int(v) << (i * width)
    raw = acc.to_bytes(len(buf), "little")
    return raw[:(total_bits + 7) // 8]


def uncompressed_string_segment(strings, negate=()):
    """UNCOMPRESSED: {dict_size, dict_end}, int32 cumulative ends (rows in `negate` stored negated),
    chars back to front ending at dict_end."""
    n = len(strings)
    ends = np.cumsum([len(s) for s in strings]).astype(np.int64) if n else np.zeros(0, np.int64)
    total = int(ends[-1]) if n else 0
    dict_end = 8 + 4 * n + total
    buf = bytearray(dict_end)
    buf[0:4] = total.to_bytes(4, "little")
    buf[4:8] = dict_end.to_bytes(4, "little")
    prev = 0
    for i, s in enumerate(strings):
        e = int(ends[i])
        v = -e if i in negate else e
        buf[8 + 4 * i:12 + 4 * i] = int(v).to_bytes(4, "little", signed=True)
        buf[dict_end - e:dict_end - prev] = s
        prev = e
    return bytes(buf)


def dictionary_string_segment(strings):
    """DICTIONARY: 20-byte header, bit-packed selection (0 = NULL/None), u32 cumulative entry ends,
    entries back to front ending at dict_end."""
    entries, index, sel = [], {}, []
    for s in strings:
        if s is None:
            sel.append(0)
            continue
        if s not in index:
            entries.append(s)
            index[s] = len(entries)
        sel.append(index[s])
    width = max(sel, default=0).bit_length()
    sel_bytes = _pack_lsb(sel, width)
    ibo = 20 + len(sel_bytes)
    ibo += (-ibo) % 4
    idx = np.concatenate([[0], np.cumsum([len(e) for e in entries])]).astype(np.int64)
    ibc = len(idx)
    dict_end = ibo + 4 * ibc + int(idx[-1])
    buf = bytearray(dict_end)
    for k, v in enumerate((int(idx[-1]), dict_end, ibo, ibc, width)):
        buf[4 * k:4 * k + 4] = int(v).to_bytes(4, "little")
    buf[20:20 + len(sel_bytes)] = sel_bytes
    for k, v in enumerate(idx):
        buf[ibo + 4 * k:ibo + 4 * k + 4] = int(v).to_bytes(4, "little")
    for k, e in enumerate(entries, start=1):
        buf[dict_end - int(idx[k]):dict_end - int(idx[k - 1])] = e
    return bytes(buf)


class StringBlob:
    """Segments laid out 16-byte aligned in one blob, with their table rows
    (offset, bytes_size, row_offset, row_count, codec, max_string_length)."""

    def __init__(self):
        self.parts, self.table, self.off = [], [], 0

    def add(self, codec, seg: bytes, row_offset: int, rows: int, bytes_size=None, max_len=0):
        self.table.append((self.off, len(seg) if bytes_size is None else bytes_size, row_offset, rows, codec,
                           max_len))
        self.parts.append(seg + bytes((-len(seg)) % 16))
        self.off += len(self.parts[-1])

    @property
    def blob(self):
        return np.frombuffer(b"".join(self.parts), np.uint8).copy() if self.parts else np.zeros(16, np.uint8)


def oracle_string_runs(blob, table, total_rows):
    """Decode (codec, Segment) runs held in host memory with the oracle -> (offsets, chars)."""
    import ctypes as C

    from oracle import oracle as O
    cap = int(sum(int(t[1]) for t in table)) * 4 + total_rows * 64 + 1024
    offs = np.zeros(total_rows + 1, np.int32)
    chars = np.zeros(cap, np.uint8)
    col, keep = string_column(duckdb_segments(table, blob.ctypes.data), total_rows, offs.ctypes.data,
                              chars.ctypes.data, cap)
    rc = O.load().oracle_decode_strings_column(C.byref(col))
    if rc != 0:
        raise RuntimeError(f"oracle rc {rc}")
    return offs, chars[:col.chars_used]
