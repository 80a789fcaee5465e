"""sdr_decode_strings_column_noalloc, the allocation-free string decode:
through the C ABI its offsets and chars equal sdr_decode_strings_column's,
bit for bit, on two calls in a row (the second with a smaller column, so any
state the first left in the shared scratch array would show), and a column
whose descriptors exceed the scratch is refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
import sirius_amd.encode as E
from sirius_amd import lib as L
from tests.helpers import string_column

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def _column(n, seed, gap):
    """An FSST column of n rows whose first `gap` rows no segment covers
    (they decode to empty strings)."""
    chars, offs = E.gen_zipf_strings(seed, n - gap)
    enc = E.encode_fsst(chars, offs)
    blob = torch.from_numpy(enc.blob).to(_dev())
    segs = enc.segments(blob.data_ptr())
    for s in segs:
        s.row_offset += gap
    return blob, segs, len(chars)


def _decode(entry, segs, n, cap):
    offs = torch.full((n + 1,), -7, dtype=torch.int32, device=_dev())
    chars = torch.full((cap + 32,), 0xCC, dtype=torch.uint8, device=_dev())
    col, keep = string_column([(A.FSST, segs)], n, offs.data_ptr(), chars.data_ptr(), cap)
    rc = entry(C.byref(col), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return rc, col.chars_used, offs.cpu().numpy(), chars.cpu().numpy()


def test_noalloc_matches_allocating_entry_twice():
    lib = L.load()
    cases = [(300_000, 0xA11C, 1000), (40_000, 0xB22D, 0)]  # the second call is smaller
    for n, seed, gap in cases:
        print(f"[no_alloc] rows={n} seed={seed:#x} gap={gap}", flush=True)
        blob, segs, n_chars = _column(n, seed, gap)
        cap = n_chars + 64
        rc0, used0, offs0, chars0 = _decode(lib.sdr_decode_strings_column, segs, n, cap)
        rc1, used1, offs1, chars1 = _decode(lib.sdr_decode_strings_column_noalloc, segs, n, cap)
        L.check(rc0)
        L.check(rc1)
        assert used0 == used1 == n_chars
        np.testing.assert_array_equal(offs1, offs0)
        np.testing.assert_array_equal(chars1, chars0)
        del blob


def test_noalloc_refuses_columns_past_the_scratch():
    lib = L.load()
    _dev()
    # descriptors of 8000 one-row segments: about 8000 * 2.3 KiB > 16 MiB
    # (each points at 64 real zero bytes, so a decode could only read those)
    n = 8000
    zeros = torch.zeros(64, dtype=torch.uint8, device=_dev())
    segs = [A.Segment(zeros.data_ptr(), 64, i, 1, 8) for i in range(n)]
    offs = torch.full((n + 1,), -7, dtype=torch.int32, device=_dev())
    col, keep = string_column([(A.FSST, segs)], n, offs.data_ptr(), 0, 0)
    rc = lib.sdr_decode_strings_column_noalloc(C.byref(col), torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == A.SDR_E_CAPACITY
    assert b"scratch" in lib.sdr_last_error()
    assert (offs.cpu().numpy() == -7).all()  # nothing was written
