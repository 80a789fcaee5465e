"""Staging segment bytes from a file (sdr_stage_plan / sdr_stage_file): the
layout on the host, and on the GPU the staged image decoded bit-exactly
(DuckDB-written segments from tests/golden/duckdb_segments.npz placed in a
scratch file the way a database file holds them)."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
from sirius_amd import lib as L
from sirius_amd import stage as S
from tests.helpers import duckdb_cases, duckdb_segments, fixed_column, string_column


def test_plan_sorts_coalesces_and_aligns():
    # (file_offset, bytes): two ranges 8 bytes apart coalesce, a far one starts a new piece,
    # an overlapping one shares bytes, an empty one stays at 0
    ranges = [(10_000, 100), (0, 50), (58, 30), (40, 20), (5000, 0)]
    total, offs = S.plan(ranges, max_gap=8)
    assert offs[1] == 0                # first piece starts at 0
    assert offs[3] == 40               # overlaps the first range: same piece
    assert offs[2] == 58               # gap of 8 from 50: coalesced
    assert offs[0] % 16 == 0 and offs[0] >= 88  # new piece, 16-byte aligned
    assert offs[4] == 0
    assert total == offs[0] + 112      # 100 bytes, padded to 16
    # a gap of 8 bytes: one piece at max_gap 8, two 16-byte aligned pieces at 0
    assert S.plan([(0, 50), (58, 30)], max_gap=8) == (96, [0, 58])
    assert S.plan([(0, 50), (58, 30)], max_gap=0) == (96, [0, 64])


def test_plan_of_nothing():
    assert S.plan([])[0] == 0


def _scratch_file(tmp_path, cases):
    """Write each fixture's segments into one file with gaps like a database file; return the
    path and, per case, the [(file_offset, bytes)] of its segments."""
    rng = np.random.default_rng(9)
    parts, pos, where = [], 0, []
    for c in cases:
        offs = []
        for o, bs, ro, rc, codec, _ in c["table"]:
            gap = int(rng.choice([0, 8, 4096]))
            parts.append(bytes(gap)); pos += gap
            parts.append(c["blob"][int(o):int(o) + int(bs)].tobytes())
            offs.append((pos, int(bs))); pos += int(bs)
        where.append(offs)
    p = tmp_path / "segments.bin"
    p.write_bytes(b"".join(parts))
    return p, where


@pytest.mark.gpu
def test_staged_segments_decode(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    dev = torch.device("cuda:0")
    cases = [c for c in duckdb_cases() if c["name"] in
             ("tpch_l_orderkey", "tpch_l_shipdate", "tpch_c_name", "tpch_p_type", "tpch_s_comment",
              "tpcds_ss_coupon_amt")]
    path, where = _scratch_file(tmp_path, cases)
    ranges = [r for w in where for r in w]
    buf, offs = S.stage_file(path, ranges, dev)
    torch.cuda.synchronize()
    k = 0
    for c, w in zip(cases, where):
        table = c["table"].copy()
        for i in range(len(table)):
            table[i, 0] = offs[k]
            k += 1
        n = c["total"]
        print(f"[gpu] staged {c['name']} rows={n}", flush=True)
        if c["string"]:
            cap = len(c["chars"]) + 64
            o = torch.full((n + 1,), -1, dtype=torch.int32, device=dev)
            ch = torch.zeros(cap, dtype=torch.uint8, device=dev)
            col, keep = string_column(duckdb_segments(table, buf.data_ptr()), n, o.data_ptr(), ch.data_ptr(), cap)
            L.check(L.load().sdr_decode_strings_column(C.byref(col), torch.cuda.current_stream().cuda_stream))
            np.testing.assert_array_equal(o.cpu().numpy(), c["offsets"])
            np.testing.assert_array_equal(ch.cpu().numpy()[:col.chars_used], c["chars"])
        else:
            out = torch.zeros(n * c["ts"], dtype=torch.uint8, device=dev)
            col, keep = fixed_column(duckdb_segments(table, buf.data_ptr()), c["ts"], n, out.data_ptr())
            arr = (A.Column * 1)(col)
            L.check(L.load().sdr_decode_table(arr, 1, torch.cuda.current_stream().cuda_stream))
            if c["expected"] is not None:
                np.testing.assert_array_equal(out.cpu().numpy(), c["expected"])
            else:
                from tests.test_duckdb_segments import oracle_fixed
                np.testing.assert_array_equal(out.cpu().numpy(), oracle_fixed(c))


@pytest.mark.gpu
def test_stage_file_errors(tmp_path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    lib = L.load()
    p = tmp_path / "small.bin"
    p.write_bytes(bytes(range(256)) * 4)
    d = torch.zeros(4096, dtype=torch.uint8, device="cuda:0")
    sh = torch.cuda.current_stream().cuda_stream
    ext = (A.Extent * 2)(A.Extent(0, 100, 0), A.Extent(1000, 100, 0))  # second range ends past the file
    assert lib.sdr_stage_file(str(p).encode(), ext, 2, 8, C.c_void_p(d.data_ptr()), 4096, sh) == A.SDR_E_BOUNDS
    ext = (A.Extent * 1)(A.Extent(0, 1000, 0))
    assert lib.sdr_stage_file(str(p).encode(), ext, 1, 8, C.c_void_p(d.data_ptr()), 500, sh) == A.SDR_E_CAPACITY
    assert lib.sdr_stage_file(str(tmp_path / "none").encode(), ext, 1, 8, C.c_void_p(d.data_ptr()), 4096,
                              sh) == A.SDR_E_INVALID_ARG
    assert lib.sdr_stage_file(str(p).encode(), ext, 1, 8, C.c_void_p(d.data_ptr()), 4096, sh) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d[:1000].cpu().numpy(), np.frombuffer(p.read_bytes()[:1000], np.uint8))


def test_scan_column_refuses_roaring_validity(tmp_path):
    """A validity segment in a codec the decoder does not read (ROARING, 13)
    is refused before any device work instead of being read as all-valid."""
    from sirius_amd.duckdb_file import DuckDBFile, FormatError
    from sirius_amd.stage import scan_column
    from tests.duckdb_writer import write_database

    tables = [{"name": "t", "rows": 4096, "columns": [{
        "name": "k", "type_id": 14,
        "segments": [{"row_start": 0, "rows": 4096, "codec": 2, "min": 5, "max": 5}],
        "validity": [{"row_start": 0, "rows": 4096, "codec": 13, "has_null": True, "bytes": bytes(64)}]}]}]
    path = write_database(tmp_path / "roaring.duckdb", tables)
    with DuckDBFile(str(path)) as f:
        assert [s.codec for s in f.tables["t"].columns[0].validity] == [13]
        with pytest.raises(FormatError, match="validity segment at row 0"):
            scan_column(f, "t", "k", "cpu")
