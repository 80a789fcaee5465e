"""DICT_FSST on the GPU (sdr_decode_strings_column in every launch mode) is
bit-exact to the oracle: the byte-assembled known-answer segments of modes
0, 1 and 2 beside an invalid one, encoded columns with inline NULLs
(offsets, chars, mask and null count), a column mixing DICT_FSST with FSST
and an uncovered gap under validity runs, the mask that inline NULLs need,
and the entries that refuse DICT_FSST (persistent batches, the
no-allocation entry)."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
import sirius_amd.encode as E
from oracle import oracle as O
from sirius_amd import lib as L
from tests.helpers import string_column
from tests.test_oracle_dict_fsst import mode0_segment, mode1_segment, mode2_segment

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

MODES = [A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT, A.LAUNCH_PERSISTENT]
MODE_IDS = ["batched", "plain", "persistent"]


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def upload(blob, skew=0):
    t = torch.zeros(len(blob) + skew + 256, dtype=torch.uint8, device=_dev())
    if len(blob):
        t[skew:skew + len(blob)] = torch.from_numpy(np.ascontiguousarray(blob)).to(_dev())
    return t, t.data_ptr() + skew


def gpu_decode(runs, n, cap, mode=A.LAUNCH_BATCHED, validity=None, mask=True):
    """-> rc, col, offsets, chars, valid (bool[n]) with the mask preset all valid."""
    offs = torch.full((n + 1,), -7, dtype=torch.int32, device=_dev())
    chars = torch.full((cap + 32,), 0xCC, dtype=torch.uint8, device=_dev())
    m = torch.full(((n + 7) // 8 + 8,), 0xFF, dtype=torch.uint8, device=_dev())
    col, keep = string_column(runs, n, offs.data_ptr(), chars.data_ptr(), cap, validity,
                              m.data_ptr() if mask else None)
    col.d_null_mask = m.data_ptr() if mask else None
    rc = L.load().sdr_decode_strings_column_mode(C.byref(col), torch.cuda.current_stream().cuda_stream, mode)
    torch.cuda.synchronize()
    valid = np.unpackbits(m.cpu().numpy(), bitorder="little")[:n].astype(bool)
    return rc, col, offs.cpu().numpy(), chars.cpu().numpy()[:cap], valid


def oracle_decode(runs_host, n, cap, validity_host=None):
    """The oracle's column decode over host copies of the same segments."""
    offs = np.zeros(n + 1, np.int32)
    chars = np.zeros(cap, np.uint8)
    m = np.full((n + 7) // 8 + 8, 0xFF, np.uint8)
    col, keep = string_column(runs_host, n, offs.ctypes.data, chars.ctypes.data, cap, validity_host, m.ctypes.data)
    col.d_null_mask = m.ctypes.data
    rc = O.load().oracle_decode_strings_column_mt(C.byref(col), 4)
    valid = np.unpackbits(m, bitorder="little")[:n].astype(bool)
    return rc, col, offs, chars, valid


def segments_at(blobs_rows, base, start_row=0):
    out, off, row = [], 0, start_row
    for b, n in blobs_rows:
        out.append(A.Segment(base + off, len(b), row, n, 0))
        off += len(b)
        row += n
    return out


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
def test_known_answer_segments(mode):
    bad = bytearray(mode0_segment())
    bad[8] = 3  # invalid header: its rows are empty, and NULL since the column has inline NULLs
    parts = [(mode0_segment(), 5), (mode2_segment(), 2), (bytes(bad), 3), (mode1_segment(), 5)]
    blob = b"".join(b for b, _ in parts)
    d, base = upload(np.frombuffer(blob, np.uint8), skew=3)
    rc, col, offs, chars, valid = gpu_decode([(A.DICT_FSST, segments_at(parts, base))], 15, 256, mode)
    L.check(rc)
    want = [b"ab", b"xyz", b"", b"ab", b"xyz", b"abc", b"c", b"", b"", b"", b"abc", b"cz", b"ab\x00", b"", b"cz"]
    got = [bytes(chars[offs[i]:offs[i + 1]]) for i in range(15)]
    assert got == want
    assert list(valid) == [1, 1, 0, 1, 1] + [1, 1] + [0, 0, 0] + [1, 1, 1, 0, 1]
    assert col.null_count == 5 and col.chars_used == sum(map(len, want))


@pytest.mark.parametrize("mode", MODES, ids=MODE_IDS)
@pytest.mark.parametrize("dmode", [0, 1, 2])
def test_encoded_column_matches_oracle(dmode, mode):
    chars, offs = E.gen_zipf_strings(0xD1C7 + dmode, 60_000)
    n = offs.size - 1
    valid_in = None if dmode == 2 else (np.arange(n) % 17 != 3).astype(np.uint8)
    enc = E.encode_dict_fsst(chars, offs, dmode, valid=valid_in, seg_rows=4096)
    o_offs, o_chars, o_valid, o_nulls = O.decode_strings_masked(enc, threads=8)
    d, base = upload(enc.blob, skew=5)
    rc, col, g_offs, g_chars, g_valid = gpu_decode([(A.DICT_FSST, enc.segments(base))], n, len(chars) + 64, mode)
    L.check(rc)
    np.testing.assert_array_equal(g_offs, o_offs)
    assert col.chars_used == len(o_chars)
    np.testing.assert_array_equal(g_chars[: col.chars_used], o_chars)
    np.testing.assert_array_equal(g_valid, o_valid)
    assert col.null_count == o_nulls


def test_mixed_with_fsst_gap_and_validity():
    c1, o1 = E.gen_zipf_strings(21, 9_000)
    c2, o2 = E.gen_zipf_strings(22, 7_000)
    v1 = (np.arange(9_000) % 11 != 0).astype(np.uint8)
    e1 = E.encode_dict_fsst(c1, o1, 1, valid=v1, seg_rows=3000)
    e2 = E.encode_fsst(c2, o2)
    gap = 2_000
    n = 9_000 + gap + 7_000
    rng = np.random.default_rng(4)
    vbits = np.packbits(rng.random(n) > 0.2, bitorder="little")
    cap = len(c1) + len(c2) + 64
    d1, b1 = upload(e1.blob)
    d2, b2 = upload(e2.blob)
    dv, bv = upload(vbits)
    s2 = e2.segments(b2)
    for s in s2:
        s.row_offset += 9_000 + gap
    rc, col, g_offs, g_chars, g_valid = gpu_decode([(A.DICT_FSST, e1.segments(b1)), (A.FSST, s2)], n, cap,
                                                   validity=[(A.UNCOMPRESSED, [A.Segment(bv, len(vbits), 0, n, 0)])])
    L.check(rc)
    h2 = e2.segments(e2.blob.ctypes.data)
    for s in h2:
        s.row_offset += 9_000 + gap
    orc, ocol, o_offs, o_chars, o_valid = oracle_decode(
        [(A.DICT_FSST, e1.segments(e1.blob.ctypes.data)), (A.FSST, h2)], n, cap,
        [(A.UNCOMPRESSED, [A.Segment(vbits.ctypes.data, len(vbits), 0, n, 0)])])
    assert orc == A.SDR_OK
    np.testing.assert_array_equal(g_offs, o_offs)
    assert col.chars_used == ocol.chars_used
    np.testing.assert_array_equal(g_chars[: col.chars_used], o_chars[: ocol.chars_used])
    np.testing.assert_array_equal(g_valid, o_valid)
    assert col.null_count == ocol.null_count
    # every inline NULL of the DICT_FSST run is NULL in the mask
    assert not g_valid[:9_000][v1 == 0].any()


def test_inline_nulls_without_a_mask_are_refused():
    d, base = upload(np.frombuffer(mode0_segment(), np.uint8))
    rc, col, offs, chars, _ = gpu_decode([(A.DICT_FSST, segments_at([(mode0_segment(), 5)], base))], 5, 64,
                                         mask=False)
    assert rc == A.SDR_E_INVALID_ARG and "d_null_mask" in L.load().sdr_last_error().decode()
    assert (offs == -7).all()  # nothing written
    d, base = upload(np.frombuffer(mode2_segment(), np.uint8))  # no inline NULLs: no mask needed
    rc, col, offs, chars, _ = gpu_decode([(A.DICT_FSST, segments_at([(mode2_segment(), 2)], base))], 2, 64,
                                         mask=False)
    L.check(rc)
    assert bytes(chars[: col.chars_used]) == b"abcc"


def test_batch_and_noalloc_refuse_dict_fsst():
    lib = L.load()
    d, base = upload(np.frombuffer(mode2_segment(), np.uint8))
    offs = torch.zeros(3, dtype=torch.int32, device=_dev())
    chars = torch.zeros(64, dtype=torch.uint8, device=_dev())
    col, keep = string_column([(A.DICT_FSST, segments_at([(mode2_segment(), 2)], base))], 2, offs.data_ptr(),
                              chars.data_ptr(), 64)
    b = C.c_void_p()
    L.check(lib.sdr_batch_create(C.byref(b)))
    try:
        assert lib.sdr_batch_add_strings(b, C.byref(col)) == A.SDR_E_UNSUPPORTED_CODEC
    finally:
        lib.sdr_batch_destroy(b)
    rc = lib.sdr_decode_strings_column_noalloc(C.byref(col), torch.cuda.current_stream().cuda_stream)
    assert rc == A.SDR_E_UNSUPPORTED_CODEC
