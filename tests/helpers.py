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
