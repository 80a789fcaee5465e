"""Segments written by DuckDB's own encoders (TPC-H from DuckDB v1.4.3, TPC-DS
from v1.5.4), cut from the reference's vendored database files
(test/cpp/integration/data/duckdb/*.duckdb) by the file reader.

Pins the oracle on real-encoder bytes: TPC-H columns must decode to the same
rows of the reference's TPC-H Parquet files; TPC-DS columns (RLE,
BITPACKING, NULLs) must meet each segment's own min/max statistics.  Also
covers the reader (sirius_amd.duckdb_file) against the database files when
the reference tree is present.
"""
import os
from pathlib import Path

import numpy as np
import pytest

import sirius_amd._abi as A
from oracle import oracle as O
from sirius_amd.duckdb_file import DuckDBFile, FormatError
from tests.helpers import duckdb_cases, validity_bits

CASES = duckdb_cases()
FIXED = [c for c in CASES if not c["string"]]
STRINGS = [c for c in CASES if c["string"]]
DTYPE = {1: np.int8, 2: np.int16, 4: np.int32, 8: np.int64}
REF_DB = Path(os.environ.get("SDR_REFERENCE", "/root/reference")) / "test/cpp/integration/data/duckdb"


def oracle_fixed(case):
    ts = case["ts"]
    out = np.zeros(case["total"] * ts, np.uint8)
    for o, bs, ro, rc, codec, _ in case["table"]:
        seg = case["blob"][int(o):int(o) + int(bs)]
        out[int(ro) * ts:(int(ro) + int(rc)) * ts] = O.decode_segment(int(codec), seg, int(rc), ts)
    return out


def oracle_strings(case):
    from sirius_amd.encode import SEG_DTYPE, Encoded
    t = case["table"]
    segs = np.zeros(len(t), SEG_DTYPE)
    for i, (o, bs, ro, rc, codec, msl) in enumerate(t):
        segs[i] = (o, bs, ro, rc, msl)
    return O.decode_strings(Encoded(int(t[0, 4]), 0, case["total"], case["blob"], segs))


def test_fixture_inventory():
    codecs = {tuple(c["src"]["codecs"]) for c in CASES}
    for codec in ("BitPacking", "RLE", "FSST", "Dictionary", "Uncompressed"):
        assert (codec,) in codecs
    assert any(c["vtable"] is not None for c in CASES)


@pytest.mark.parametrize("case", [c for c in FIXED if c["expected"] is not None], ids=lambda c: c["name"])
def test_tpch_fixed_matches_parquet(case):
    np.testing.assert_array_equal(oracle_fixed(case), case["expected"])


@pytest.mark.parametrize("case", [c for c in STRINGS if c["chars"] is not None], ids=lambda c: c["name"])
def test_tpch_strings_match_parquet(case):
    offs, chars = oracle_strings(case)
    np.testing.assert_array_equal(offs, case["offsets"])
    np.testing.assert_array_equal(chars, case["chars"])


@pytest.mark.parametrize("case", [c for c in FIXED if c["expected"] is None], ids=lambda c: c["name"])
def test_tpcds_segments_meet_their_statistics(case):
    v = oracle_fixed(case).view(DTYPE[case["ts"]])
    valid = validity_bits(case).astype(bool)
    assert case["vtable"] is None or not valid.all()  # these columns hold NULLs
    checked = 0
    for (o, bs, ro, rc, codec, _), (has, lo, hi) in zip(case["table"], case["stats"]):
        rows = slice(int(ro), int(ro) + int(rc))
        live = v[rows][valid[rows]]
        if has and len(live):
            assert live.min() == lo and live.max() == hi
            checked += 1
    assert checked == len(case["table"])


@pytest.mark.parametrize("case", [c for c in STRINGS if c["chars"] is None], ids=lambda c: c["name"])
def test_tpcds_strings_meet_their_statistics(case):
    offs, chars = oracle_strings(case)
    lens = np.diff(offs)
    valid = validity_bits(case).astype(bool)
    assert (lens[~valid] == 0).all()  # NULL rows are stored as empty strings
    for (o, bs, ro, rc, codec, _), (has, _, max_len) in zip(case["table"], case["stats"]):
        seg_lens = lens[int(ro):int(ro) + int(rc)]
        assert has and seg_lens.max() == max_len


def test_rle_fixture_has_many_runs():
    case = next(c for c in FIXED if c["name"] == "tpcds_ss_coupon_amt")
    v = oracle_fixed(case).view(np.int32)
    assert np.count_nonzero(np.diff(v)) > 1000


@pytest.mark.skipif(not REF_DB.exists(), reason="reference database files not present")
def test_reader_walks_the_tpch_file():
    with DuckDBFile(str(REF_DB / "integration.duckdb")) as f:
        assert f.library_version == "v1.4.3" and f.storage_version == 64
        rows = {n: t.rows for n, t in f.tables.items()}
        assert rows == {"customer": 15000, "lineitem": 600572, "nation": 25, "orders": 150000, "part": 20000,
                        "partsupp": 80000, "region": 5, "supplier": 1000}
        li = f.tables["lineitem"]
        assert [c.name for c in li.columns][:3] == ["l_orderkey", "l_partkey", "l_suppkey"]
        assert li.row_groups[0] == (0, 122880) and sum(n for _, n in li.row_groups) == 600572
        for c in li.columns:
            covered = sum(s.tuple_count for s in c.segments)
            assert covered == 600572, c.name
            assert all(s.bytes_size > 0 for s in c.segments if s.block_id >= 0)


@pytest.mark.skipif(not REF_DB.exists(), reason="reference database files not present")
def test_reader_reproduces_fixture_bytes():
    files = {n: DuckDBFile(str(REF_DB / n)) for n in ("integration.duckdb", "tpcds.duckdb")}
    try:
        for case in CASES:
            f = files[case["src"]["file"]]
            col = next(c for c in f.tables[case["src"]["table"]].columns if c.name == case["src"]["column"])
            r0 = case["src"]["rows"][0]
            segs = [s for s in col.segments if r0 <= s.row_start < case["src"]["rows"][1]]
            assert len(segs) == len(case["table"])
            for s, (o, bs, ro, rc, codec, _) in zip(segs, case["table"]):
                assert (s.row_start - r0, s.tuple_count, s.codec) == (ro, rc, codec)
                assert f.segment_bytes(s)[:int(bs)] == case["blob"][int(o):int(o) + int(bs)].tobytes()
    finally:
        for f in files.values():
            f.close()


def test_reader_rejects_other_files(tmp_path):
    p = tmp_path / "x.duckdb"
    p.write_bytes(b"\0" * 20000)
    with pytest.raises(FormatError):
        DuckDBFile(str(p))
    q = tmp_path / "y.duckdb"
    head = bytearray(12288)
    head[8:12] = b"DUCK"
    import struct
    struct.pack_into("<qqqqqqq", head, 4096 + 8, 1, 0, -1, 5, 262144, 2048, 1)  # 5 blocks, none present
    q.write_bytes(bytes(head))
    with pytest.raises(FormatError):
        DuckDBFile(str(q))


# ---------------------------------------------------------------------------
# a database file written around the fixture segments (tests/duckdb_writer.py)
# ---------------------------------------------------------------------------
def fixture_tables(cases):
    """One table per fixture column, its segments and validity as the fixture holds them."""
    tables = []
    for c in cases:
        src = c["src"]
        ty = {"BIGINT": 14, "INTEGER": 13, "DATE": 15, "VARCHAR": 25}.get(src["type"].split("(")[0], 21)
        width, scale = 0, 0
        if ty == 21:
            width, scale = (int(x) for x in src["type"][8:-1].split(","))
        segs = []
        for (o, bs, ro, rc, codec, msl), (has, lo, hi) in zip(c["table"], c["stats"]):
            s = {"row_start": int(ro), "rows": int(rc), "codec": int(codec),
                 "bytes": c["blob"][int(o):int(o) + int(bs)].tobytes()}
            if ty == 25:
                s["max_string_length"] = int(hi)
            elif has:
                s["min"], s["max"] = int(lo), int(hi)
            segs.append(s)
        val = []
        if c["vtable"] is not None:
            for o, bs, ro, rc, codec, _ in c["vtable"]:
                val.append({"row_start": int(ro), "rows": int(rc), "codec": 1, "has_null": True,
                            "bytes": c["vblob"][int(o):int(o) + int(bs)].tobytes()})
        tables.append({"name": c["name"], "rows": c["total"],
                       "columns": [{"name": src["column"], "type_id": ty, "width": width, "scale": scale,
                                    "segments": segs, "validity": val}]})
    # a CONSTANT column: its value lives in the statistics only
    tables.append({"name": "constants", "rows": 5000,
                   "columns": [{"name": "k", "type_id": 14, "segments": [
                       {"row_start": 0, "rows": 3000, "codec": 2, "min": -77, "max": -77},
                       {"row_start": 3000, "rows": 2000, "codec": 2, "min": 123456789012, "max": 123456789012}]}]})
    return tables


def test_reader_on_a_written_file(tmp_path):
    from tests.duckdb_writer import write_database
    path = write_database(tmp_path / "fixtures.duckdb", fixture_tables(CASES))
    with DuckDBFile(str(path)) as f:
        assert f.library_version == "v1.4.3"
        assert set(f.tables) == {c["name"] for c in CASES} | {"constants"}
        for c in CASES:
            t = f.tables[c["name"]]
            assert t.rows == c["total"] and len(t.columns) == 1
            col = t.columns[0]
            assert col.name == c["src"]["column"]
            assert col.type.is_string == c["string"]
            if not c["string"]:
                assert col.type.type_size == c["ts"]
            assert len(col.segments) == len(c["table"])
            for s, (o, bs, ro, rc, codec, _) in zip(col.segments, c["table"]):
                assert (s.row_start, s.tuple_count, s.codec) == (ro, rc, codec)
                assert f.segment_bytes(s)[:int(bs)] == c["blob"][int(o):int(o) + int(bs)].tobytes()
            assert len([v for v in col.validity if v.codec == 1]) == (0 if c["vtable"] is None else len(c["vtable"]))
        k = f.tables["constants"].columns[0]
        assert [(s.codec, s.stats["min"]) for s in k.segments] == [(2, -77), (2, 123456789012)]


def test_reader_rejects_a_cyclic_metadata_chain(tmp_path):
    from tests.duckdb_writer import write_database
    path = write_database(tmp_path / "c.duckdb", fixture_tables([c for c in CASES if c["name"] == "tpch_s_name"]))
    raw = bytearray(path.read_bytes())
    # point the catalog piece's next pointer at itself and make its list count huge
    import struct
    head = struct.unpack_from("<q", raw, 4096 + 16)[0]
    piece = head >> 56
    base = 3 * 4096 + 8 + piece * ((262136 // 64) & ~7)
    struct.pack_into("<Q", raw, base, head)
    raw[base + 8 + 2:base + 8 + 7] = b"\xff\xff\xff\xff\x0f"  # catalog list count: a huge varint
    bad = tmp_path / "bad.duckdb"
    bad.write_bytes(bytes(raw))
    with pytest.raises(FormatError):
        DuckDBFile(str(bad))
