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
