"""Write tests/golden/duckdb_segments.npz: column segments written by DuckDB's
own encoders, cut from the two database files the reference's integration
tests attach (test/cpp/integration/data/duckdb/integration.duckdb, TPC-H,
DuckDB v1.4.3; tpcds.duckdb, TPC-DS SF 0.01, DuckDB v1.5.4), with expected
values where an independent source exists.

Run in a checkout that has the reference tree at REF (this script is the
only reader of those files; the tests use the npz):

    python tests/golden/make_duckdb_fixtures.py [REF]

Segments are read with sirius_amd.duckdb_file and trimmed to the bytes their
headers address (BITPACKING: up to metadata_end; FSST, DICTIONARY and
UNCOMPRESSED strings: up to dict_end; RLE: values and counts; UNCOMPRESSED
validity: the row bits), so every decode also proves it reads nothing past
the header's bound.

Expected values:
  * TPC-H columns: the same rows of the reference's TPC-H Parquet files
    (test/cpp/integration/data/parquet/*.parquet), read with pyarrow:
    DECIMAL as the unscaled int64, DATE as int32 days, strings as Arrow
    offsets + chars.  The DuckDB file and the Parquet files hold the same
    rows in the same order (tests/test_duckdb_segments.py checks every
    fixture column against them through the oracle).
  * TPC-DS columns: no independent values; the tests check each segment
    against its own statistics (min/max, or the maximum string length) and
    that NULL rows are empty, and the GPU against the oracle.

npz layout, per fixture NAME:
  NAME/meta     int64 [kind (0 fixed, 1 string), type_size, total_rows, n_segments, n_validity]
  NAME/src      str   file, table, column, type, row range, DuckDB version
  NAME/blob     uint8 segments, each 16-byte aligned
  NAME/table    int64 [n, 6] offset, bytes_size, row_offset, row_count, codec, max_string_length
  NAME/stats    int64 [n, 3] has_min_max, min, max (fixed-width); has_max_len, 0, max_string_length (strings)
  NAME/vblob, NAME/vtable   validity segments (UNCOMPRESSED), same layout
  NAME/expected uint8 decoded values (TPC-H fixed)    NAME/offsets int32, NAME/chars uint8 (TPC-H strings)
"""
from __future__ import annotations

import json
import struct
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from sirius_amd.duckdb_file import DuckDBFile  # noqa: E402

OUT = Path(__file__).resolve().parent / "duckdb_segments.npz"

# (file, table, column, first row, end row) -- row ranges cover whole row groups
TPCH = [
    ("lineitem", "l_orderkey", 0, 122880),
    ("lineitem", "l_discount", 0, 122880),
    ("lineitem", "l_shipdate", 0, 122880),
    ("orders", "o_totalprice", 122880, 150000),
    ("customer", "c_acctbal", 0, 15000),
    ("customer", "c_nationkey", 0, 15000),
    ("customer", "c_name", 0, 15000),
    ("customer", "c_phone", 0, 15000),
    ("part", "p_comment", 0, 20000),
    ("supplier", "s_name", 0, 1000),
    ("lineitem", "l_returnflag", 0, 122880),
    ("lineitem", "l_shipmode", 0, 122880),
    ("orders", "o_clerk", 122880, 150000),
    ("part", "p_type", 0, 20000),
    ("nation", "n_name", 0, 25),
    ("nation", "n_comment", 0, 25),
    ("supplier", "s_address", 0, 1000),
    ("supplier", "s_comment", 0, 1000),
]
TPCDS = [
    ("customer_demographics", "cd_dep_count"),
    ("household_demographics", "hd_vehicle_count"),
    ("date_dim", "d_quarter_seq"),
    ("store_sales", "ss_ext_discount_amt"),
    ("store_sales", "ss_coupon_amt"),
    ("store_sales", "ss_ext_sales_price"),
    ("store_sales", "ss_net_profit"),
    ("time_dim", "t_second"),
    ("catalog_page", "cp_department"),
    ("customer", "c_first_name"),
    ("customer", "c_login"),
    ("customer", "c_email_address"),
    ("customer_address", "ca_zip"),
]


def used_bytes(seg, b: bytes, ts: int, kind: str) -> int:
    """Bytes of the segment its header addresses (kind: fixed, string or validity)."""
    if kind == "string" and seg.codec in (1, 4):  # UNCOMPRESSED / DICTIONARY strings: u32 dict_end at byte 4
        return min(len(b), struct.unpack_from("<I", b, 4)[0])
    if seg.codec == 6:  # BITPACKING: u64 metadata_end
        return min(len(b), struct.unpack_from("<Q", b, 0)[0])
    if seg.codec == 7:  # FSST: u32 dict_end at byte 4
        return min(len(b), struct.unpack_from("<I", b, 4)[0])
    if seg.codec == 3:  # RLE: u64 counts offset, values T[ec], counts u16[ec]
        co = struct.unpack_from("<Q", b, 0)[0]
        return min(len(b), co + 2 * ((co - 8) // ts))
    if seg.codec == 1 and kind == "validity":  # UNCOMPRESSED validity: one bit per row in u64 words
        return min(len(b), (seg.tuple_count + 63) // 64 * 8)
    return len(b)


def pack(f, segs, r0, ts, kind):
    parts, table, stats, off = [], [], [], 0
    for s in segs:
        b = f.segment_bytes(s)
        b = b[:used_bytes(s, b, ts, kind)]
        table.append((off, len(b), s.row_start - r0, s.tuple_count, s.codec, s.stats.get("max_string_length", 0)))
        st = s.stats
        if "min" in st and "max" in st:
            stats.append((1, st["min"], st["max"]))
        elif "max_string_length" in st:
            stats.append((1, 0, st["max_string_length"]))
        else:
            stats.append((0, 0, 0))
        parts.append(b + bytes((-len(b)) % 16))
        off += len(parts[-1])
    blob = np.frombuffer(b"".join(parts), np.uint8).copy() if parts else np.zeros(0, np.uint8)
    return blob, np.array(table, np.int64).reshape(-1, 6), np.array(stats, np.int64).reshape(-1, 3)


def in_range(segs, r0, r1):
    out = [s for s in segs if r0 <= s.row_start < r1]
    if any(s.row_start + s.tuple_count > r1 for s in out):
        raise SystemExit("row range splits a segment")
    return out


def parquet_values(ref: Path, table: str, column: str, r0: int, r1: int, col):
    import pyarrow as pa
    import pyarrow.parquet as pq
    a = pq.read_table(ref / "test/cpp/integration/data/parquet" / f"{table}.parquet",
                      columns=[column]).column(0).combine_chunks().slice(r0, r1 - r0)
    if col.type.is_string:
        o = np.frombuffer(a.buffers()[1], np.int32)[a.offset:a.offset + len(a) + 1]
        chars = np.frombuffer(a.buffers()[2], np.uint8)[o[0]:o[-1]].copy()
        return (o - o[0]).astype(np.int32), chars
    ts = col.type.type_size
    if pa.types.is_decimal(a.type):
        w = np.frombuffer(a.buffers()[1], np.int64)[2 * a.offset:2 * (a.offset + len(a))].reshape(-1, 2)[:, 0]
        v = w.astype({4: np.int32, 8: np.int64}[ts])
    elif pa.types.is_date32(a.type):
        v = np.asarray(a.cast(pa.int32()))
    else:
        v = np.asarray(a).astype({1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}[ts])
    return np.ascontiguousarray(v).view(np.uint8).copy(), None


def main():
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    z = {}
    db = ref / "test/cpp/integration/data/duckdb"
    with DuckDBFile(str(db / "integration.duckdb")) as f:
        for table, column, r0, r1 in TPCH:
            col = next(c for c in f.tables[table].columns if c.name == column)
            segs = in_range(col.segments, r0, r1)
            ts = col.type.type_size
            name = f"tpch_{column}"
            blob, tab, st = pack(f, segs, r0, ts, "string" if col.type.is_string else "fixed")
            z[f"{name}/blob"], z[f"{name}/table"], z[f"{name}/stats"] = blob, tab, st
            z[f"{name}/meta"] = np.array([int(col.type.is_string), ts, r1 - r0, len(segs), 0], np.int64)
            z[f"{name}/src"] = np.array(json.dumps(dict(
                file="integration.duckdb", duckdb=f.library_version, table=table, column=column,
                type=col.type.name, rows=[r0, r1], codecs=sorted({s.codec_name for s in segs}),
                expected=f"parquet/{table}.parquet")))
            v, chars = parquet_values(ref, table, column, r0, r1, col)
            if chars is None:
                z[f"{name}/expected"] = v
            else:
                z[f"{name}/offsets"], z[f"{name}/chars"] = v, chars
            print(name, col.type.name, len(segs), "segments", len(blob), "bytes")
    with DuckDBFile(str(db / "tpcds.duckdb")) as f:
        for table, column in TPCDS:
            t = f.tables[table]
            col = next(c for c in t.columns if c.name == column)
            ts = col.type.type_size
            name = f"tpcds_{column}"
            blob, tab, st = pack(f, col.segments, 0, ts, "string" if col.type.is_string else "fixed")
            if any(s.codec not in (1, 2) for s in col.validity):  # only UNCOMPRESSED bytes / CONSTANT stats
                raise ValueError(f"{table}.{column}: validity codecs {sorted({s.codec_name for s in col.validity})}")
            valid = [s for s in col.validity if s.codec == 1]
            vblob, vtab, _ = pack(f, valid, 0, ts, "validity")
            z[f"{name}/blob"], z[f"{name}/table"], z[f"{name}/stats"] = blob, tab, st
            z[f"{name}/vblob"], z[f"{name}/vtable"] = vblob, vtab
            z[f"{name}/meta"] = np.array([int(col.type.is_string), ts, t.rows, len(col.segments), len(valid)],
                                         np.int64)
            z[f"{name}/src"] = np.array(json.dumps(dict(
                file="tpcds.duckdb", duckdb=f.library_version, table=table, column=column, type=col.type.name,
                rows=[0, t.rows], codecs=sorted({s.codec_name for s in col.segments}),
                validity=sorted({s.codec_name for s in col.validity}), expected="segment statistics")))
            print(name, col.type.name, [s.codec_name for s in col.segments], len(blob), "bytes",
                  "validity", len(vblob))
    np.savez_compressed(OUT, **z)
    print(OUT, OUT.stat().st_size, "bytes")


if __name__ == "__main__":
    main()
