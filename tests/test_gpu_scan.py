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
