"""scan_column end to end: a database file written around the DuckDB-written
fixture segments (tests/duckdb_writer.py) is read by the file reader, staged
with sdr_stage_file and decoded on the GPU, bit-exact against the rows of the
reference's TPC-H Parquet files or the oracle; CONSTANT segments decode from
their statistics."""
import numpy as np
import pytest

from sirius_amd.duckdb_file import DuckDBFile
from sirius_amd.stage import scan_column
from tests.duckdb_writer import write_database
from tests.helpers import duckdb_cases, validity_bits
from tests.test_duckdb_segments import fixture_tables, oracle_fixed, oracle_strings

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
CASES = duckdb_cases()


@pytest.fixture(scope="module")
def dbfile(tmp_path_factory):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    path = write_database(tmp_path_factory.mktemp("db") / "fixtures.duckdb", fixture_tables(CASES))
    with DuckDBFile(str(path)) as f:
        yield f


@pytest.mark.parametrize("case", CASES, ids=lambda c: c["name"])
def test_scan_column(dbfile, case):
    print(f"[gpu] scan {case['name']}", flush=True)
    out = scan_column(dbfile, case["name"], case["src"]["column"], torch.device("cuda:0"))
    torch.cuda.synchronize()
    if case["string"]:
        want_o, want_c = (case["offsets"], case["chars"]) if case["chars"] is not None else oracle_strings(case)
        np.testing.assert_array_equal(out["offsets"].cpu().numpy(), want_o)
        np.testing.assert_array_equal(out["chars"].cpu().numpy(), want_c)
    else:
        want = case["expected"] if case["expected"] is not None else oracle_fixed(case)
        np.testing.assert_array_equal(out["values"].cpu().numpy(), want)
    if case["vtable"] is not None:
        valid = validity_bits(case)
        assert out["null_count"] == int((valid == 0).sum()) > 0
        got = np.unpackbits(out["mask"].cpu().numpy(), bitorder="little")[:case["total"]]
        np.testing.assert_array_equal(got, valid)
    else:
        assert out["mask"] is None and out["null_count"] == 0


def test_scan_constant_column(dbfile):
    out = scan_column(dbfile, "constants", "k", torch.device("cuda:0"))
    torch.cuda.synchronize()
    v = out["values"].cpu().numpy().view(np.int64)
    np.testing.assert_array_equal(v, np.r_[np.full(3000, -77), np.full(2000, 123456789012)])


def test_rowids_per_range():
    import ctypes as C

    import sirius_amd._abi as A
    from sirius_amd import lib as L
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    ranges = [(0, 0, 122880), (245760, 122880, 1000), (10**12, 123880, 0), (-5, 123880, 300000)]
    rr = (A.RowRange * len(ranges))(*[A.RowRange(a, b, c, 0) for a, b, c in ranges])
    total = 123880 + 300000
    out = torch.full((total + 8,), -1, dtype=torch.int64, device="cuda:0")
    L.check(L.load().sdr_rowids(rr, len(ranges), C.c_void_p(out.data_ptr()), torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    want = np.r_[np.arange(122880), 245760 + np.arange(1000), -5 + np.arange(300000), np.full(8, -1)]
    np.testing.assert_array_equal(out.cpu().numpy(), want)


def test_scan_table_with_rowids(dbfile):
    from sirius_amd.stage import scan_table
    out = scan_table(dbfile, "constants", device=torch.device("cuda:0"), rowid=True)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(out["rowid"].cpu().numpy(), np.arange(5000))
    assert set(out) == {"k", "rowid"}


def test_scan_on_a_side_stream(dbfile):
    """A scan given its own stream allocates, copies and decodes there: the
    results match the same scans on the current stream, with a CONSTANT
    column (host-built tail), a VARCHAR column and a column with NULLs."""
    from sirius_amd.stage import scan_table
    dev = torch.device("cuda:0")
    names = ["constants"]
    names += [next(c["name"] for c in CASES if c["string"])]
    names += [next(c["name"] for c in CASES if c["vtable"] is not None)]
    side = torch.cuda.Stream(device=dev)
    for name in names:
        print(f"[gpu] side-stream scan {name}", flush=True)
        want = scan_table(dbfile, name, device=dev, rowid=True)
        torch.cuda.synchronize()
        got = scan_table(dbfile, name, device=dev, stream=side, rowid=True)
        side.synchronize()
        for col, w in want.items():
            for key in ("values", "offsets", "chars", "mask"):
                if isinstance(w, dict) and w.get(key) is not None:
                    np.testing.assert_array_equal(got[col][key].cpu().numpy(), w[key].cpu().numpy(), err_msg=f"{name}.{col}.{key}")
            if not isinstance(w, dict):
                np.testing.assert_array_equal(got[col].cpu().numpy(), w.cpu().numpy())


def test_arrow_export_of_scanned_columns(dbfile):
    """Scanned columns leave as Arrow device arrays on ROCm memory without a
    copy (buffer addresses are the decode's tensors); host copies of those
    buffers, imported by pyarrow, hold the fixture's values and nulls."""
    pa = pytest.importorskip("pyarrow")
    import sirius_amd.arrow as X
    dev = torch.device("cuda:0")
    picks = [next(c for c in CASES if c["string"] and c["vtable"] is not None),
             next(c for c in CASES if not c["string"] and c["vtable"] is not None),
             next(c for c in CASES if c["string"] and c["vtable"] is None)]
    for case in picks:
        name, column = case["name"], case["src"]["column"]
        print(f"[gpu] arrow export {name}", flush=True)
        out = scan_column(dbfile, name, column, dev)
        ctype = next(c for c in dbfile.tables[name].columns if c.name == column).type
        exp = X.export_scan(out, ctype, name=column)
        a = exp.array
        assert a.device_type == X.ARROW_DEVICE_ROCM and a.device_id == 0 and a.sync_event
        main = out["offsets"] if ctype.is_string else out["values"]
        assert a.array.buffers[1] == main.data_ptr()
        assert (a.array.buffers[0] or 0) == (out["mask"].data_ptr() if out["mask"] is not None else 0)
        assert a.array.null_count == out["null_count"] and exp.schema.name == column.encode()
        torch.cuda.synchronize()
        # the same layout from host copies, through pyarrow
        host = {k: out[k].cpu().numpy() for k in ("values", "offsets", "chars", "mask") if out.get(k) is not None}
        t, vb, prec, scale = X.arrow_type(ctype)
        n = a.array.length
        hexp = X.export(t, n, (host["offsets"] if ctype.is_string else host["values"]).ctypes.data,
                        validity=host["mask"].ctypes.data if "mask" in host else 0,
                        chars=host["chars"].ctypes.data if ctype.is_string else 0, null_count=out["null_count"],
                        value_bytes=vb, precision=prec, scale=scale, device_type=X.ARROW_DEVICE_CPU, keep=host)
        arr = pa.Array._import_from_c_device(hexp.array_ptr, hexp.schema_ptr)
        valid = validity_bits(case) if case["vtable"] is not None else np.ones(n, dtype=np.uint8)
        assert arr.null_count == int((valid == 0).sum())
        if ctype.is_string:
            want_o, want_c = (case["offsets"], case["chars"]) if case["chars"] is not None else oracle_strings(case)
            want = [bytes(want_c[want_o[i]:want_o[i + 1]]).decode() if valid[i] else None for i in range(n)]
            assert arr.to_pylist() == want
        else:
            raw = case["expected"] if case["expected"] is not None else oracle_fixed(case)
            storage = pa.Array.from_buffers(pa.uint8(), n * ctype.type_size, [None, arr.buffers()[1]])
            np.testing.assert_array_equal(np.asarray(storage), np.asarray(raw).view(np.uint8)[:n * ctype.type_size])
            assert [v is not None for v in arr.to_pylist()] == [bool(x) for x in valid]
            del storage
        del arr
        exp.release()
    assert X.live_exports() == 0
