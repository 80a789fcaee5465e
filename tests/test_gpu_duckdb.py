"""GPU decode of segments written by DuckDB's own encoders (TPC-H, DuckDB
v1.4.3; TPC-DS, v1.5.4), through the C ABI, bit-exact against the rows of
the reference's TPC-H Parquet files or, for TPC-DS, against the oracle.
Validity comes from the files' own UNCOMPRESSED validity segments."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
from sirius_amd import lib as L
from tests.helpers import duckdb_cases, duckdb_segments, fixed_column, string_column, validity_bits
from tests.test_duckdb_segments import oracle_fixed, oracle_strings

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
CASES = duckdb_cases()
FIXED = [c for c in CASES if not c["string"]]
STRINGS = [c for c in CASES if c["string"]]
MODES = [A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT]


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def upload(blob):
    t = torch.zeros(len(blob) + 256, dtype=torch.uint8, device=_dev())
    if len(blob):
        t[:len(blob)] = torch.from_numpy(np.ascontiguousarray(blob)).to(_dev())
    return t, t.data_ptr()


@pytest.mark.parametrize("case", FIXED, ids=lambda c: c["name"])
@pytest.mark.parametrize("mode", MODES, ids=["batched", "plain"])
def test_duckdb_fixed_columns(case, mode):
    n, ts = case["total"], case["ts"]
    d, base = upload(case["blob"])
    out = torch.full((n * ts,), 0xCC, dtype=torch.uint8, device=_dev())
    validity, mask, keep_v = None, None, None
    if case["vtable"] is not None:
        dv, vbase = upload(case["vblob"])
        keep_v = dv
        validity = duckdb_segments(case["vtable"], vbase)
        mask = torch.zeros((n + 7) // 8, dtype=torch.uint8, device=_dev())
    col, keep = fixed_column(duckdb_segments(case["table"], base), ts, n, out.data_ptr(), validity,
                             mask.data_ptr() if mask is not None else None)
    arr = (A.Column * 1)(col)
    rc = L.load().sdr_decode_table_mode(arr, 1, torch.cuda.current_stream().cuda_stream, mode)
    torch.cuda.synchronize()
    L.check(rc)
    want = case["expected"] if case["expected"] is not None else oracle_fixed(case)
    np.testing.assert_array_equal(out.cpu().numpy(), want)
    if mask is not None:
        valid = validity_bits(case)
        assert arr[0].null_count == int((valid == 0).sum()) > 0
        # bits past the last row are whatever the segment's last byte holds (copied as a byte, as the
        # reference's gpu_native_decode.cu:418-425 does); only rows [0, n) are defined
        got = np.unpackbits(mask.cpu().numpy(), bitorder="little")[:n]
        np.testing.assert_array_equal(got, valid)
    del keep_v


@pytest.mark.parametrize("case", STRINGS, ids=lambda c: c["name"])
@pytest.mark.parametrize("mode", MODES + [A.LAUNCH_PERSISTENT], ids=["batched", "plain", "persistent"])
def test_duckdb_string_columns(case, mode):
    n = case["total"]
    if case["chars"] is not None:
        want_offs, want_chars = case["offsets"], case["chars"]
    else:
        want_offs, want_chars = oracle_strings(case)
    d, base = upload(case["blob"])
    cap = len(want_chars) + 64
    offs = torch.full((n + 1,), -7, dtype=torch.int32, device=_dev())
    chars = torch.full((cap,), 0xCC, dtype=torch.uint8, device=_dev())
    validity, mask, keep_v = None, None, None
    if case["vtable"] is not None:
        keep_v, vbase = upload(case["vblob"])
        validity = duckdb_segments(case["vtable"], vbase)
        mask = torch.zeros((n + 7) // 8, dtype=torch.uint8, device=_dev())
    col, keep = string_column(duckdb_segments(case["table"], base), n, offs.data_ptr(), chars.data_ptr(), cap,
                              validity, mask.data_ptr() if mask is not None else None)
    rc = L.load().sdr_decode_strings_column_mode(C.byref(col), torch.cuda.current_stream().cuda_stream, mode)
    torch.cuda.synchronize()
    L.check(rc)
    np.testing.assert_array_equal(offs.cpu().numpy(), want_offs)
    assert col.chars_used == len(want_chars)
    np.testing.assert_array_equal(chars.cpu().numpy()[:col.chars_used], want_chars)
    if mask is not None:
        valid = validity_bits(case)
        assert col.null_count == int((valid == 0).sum())
        np.testing.assert_array_equal(np.unpackbits(mask.cpu().numpy(), bitorder="little")[:n], valid)
