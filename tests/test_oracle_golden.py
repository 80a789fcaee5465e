"""The oracle against the known-answer vectors restated from the
reference's unit tests (tests/golden/make_golden.py)."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
from oracle import oracle as O
from tests.helpers import fixed_column, golden_cases, segs_from_table, string_column

CASES = golden_cases()


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_oracle_matches_golden(case):
    blob = np.ascontiguousarray(case["blob"])
    segs = segs_from_table(case["table"], blob.ctypes.data)
    if case["codec"] == A.FSST:
        n = case["total"]
        offs = np.full(n + 1, -1, np.int32)
        chars = np.full(max(1, len(case["expected"]) + 64), 0xCC, np.uint8)
        col, keep = string_column([(case["codec"], segs)], n, offs.ctypes.data, chars.ctypes.data, len(chars))
        assert O.load().oracle_decode_strings_column(C.byref(col)) == 0
        np.testing.assert_array_equal(offs, case["offsets"])
        assert col.chars_used == len(case["expected"])
        np.testing.assert_array_equal(chars[: col.chars_used], case["expected"])
        return
    ts, n = case["ts"], case["total"]
    out = np.full(n * ts, 0xCC if case["canary"] else 0x5A, np.uint8)
    col, keep = fixed_column([(case["codec"], segs)], ts, n, out.ctypes.data)
    assert O.load().oracle_decode_table(C.byref(col), 1) == 0
    np.testing.assert_array_equal(out, case["expected"])


def test_golden_covers_every_codec():
    codecs = {c["codec"] for c in CASES}
    assert {A.BITPACKING, A.RLE, A.ALP, A.ALPRD, A.FSST} <= codecs
    assert sum(c["canary"] for c in CASES) >= 15
