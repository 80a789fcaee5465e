"""GPU parity of the FSST string path through the C ABI (offsets and chars
bit-exact against the oracle and against the encoder's input strings)."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd._abi as A
import sirius_amd.encode as E
from oracle import oracle as O
from sirius_amd import lib as L
from tests.helpers import golden_cases, segs_from_table, string_column

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def upload(blob, skew=0):
    t = torch.zeros(len(blob) + skew + 256, dtype=torch.uint8, device=_dev())
    if len(blob):
        t[skew:skew + len(blob)] = torch.from_numpy(np.ascontiguousarray(blob)).to(_dev())
    return t, t.data_ptr() + skew


def decode_strings_gpu(runs, n, cap, mode=A.LAUNCH_BATCHED, chars_skew=0, validity=None, d_mask=None):
    offs = torch.full((n + 1,), -7, dtype=torch.int32, device=_dev())
    chars = torch.full((cap + chars_skew + 32,), 0xCC, dtype=torch.uint8, device=_dev())
    col, keep = string_column(runs, n, offs.data_ptr(), chars.data_ptr() + chars_skew, cap, validity, d_mask)
    rc = L.load().sdr_decode_strings_column_mode(C.byref(col), torch.cuda.current_stream().cuda_stream, mode)
    torch.cuda.synchronize()
    return rc, col, offs.cpu().numpy(), chars.cpu().numpy()[chars_skew:chars_skew + cap]


STR = [c for c in golden_cases() if c["codec"] == A.FSST]


@pytest.mark.parametrize("case", STR, ids=[c["name"] for c in STR])
@pytest.mark.parametrize("mode", [A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT], ids=["batched", "plain"])
def test_golden_strings_on_gpu(case, mode):
    d, base = upload(case["blob"])
    n = case["total"]
    cap = len(case["expected"]) + 16
    rc, col, offs, chars = decode_strings_gpu([(A.FSST, segs_from_table(case["table"], base))], n, cap, mode)
    L.check(rc)
    np.testing.assert_array_equal(offs, case["offsets"])
    assert col.chars_used == len(case["expected"])
    np.testing.assert_array_equal(chars[: col.chars_used], case["expected"])


def _zipf(n, seed):
    chars, offs = E.gen_zipf_strings(seed, n)
    return chars, offs, E.encode_fsst(chars, offs)


@pytest.mark.parametrize("mode", [A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT], ids=["batched", "plain"])
@pytest.mark.parametrize("skew,chars_skew", [(0, 0), (5, 3), (11, 1)])
def test_zipf_column_matches_oracle(mode, skew, chars_skew):
    chars, offs, enc = _zipf(250_000, 0xC0DE0004)
    o_offs, o_chars = O.decode_strings(enc, threads=8)
    np.testing.assert_array_equal(o_chars, chars)  # oracle round trip
    d, base = upload(enc.blob, skew)
    rc, col, g_offs, g_chars = decode_strings_gpu([(A.FSST, enc.segments(base))], enc.n_rows, len(chars) + 64,
                                                  mode, chars_skew)
    L.check(rc)
    np.testing.assert_array_equal(g_offs, o_offs)
    assert col.chars_used == len(o_chars)
    np.testing.assert_array_equal(g_chars[: col.chars_used], o_chars)


def test_binary_strings_with_escaped_0xff():
    """Rows holding byte 255 (escaped literal 255 followed by codes)."""
    rng = np.random.default_rng(5)
    n = 20_000
    lens = rng.integers(0, 40, n)
    offs = np.zeros(n + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    chars = rng.choice(np.array([0xFF, 0x41, 0x42, 0x43, 0xFF, 0x00, 0x20], np.uint8), int(offs[-1]))
    enc = E.encode_fsst(chars, offs)
    o_offs, o_chars = O.decode_strings(enc)
    np.testing.assert_array_equal(o_chars, chars)
    d, base = upload(enc.blob)
    rc, col, g_offs, g_chars = decode_strings_gpu([(A.FSST, enc.segments(base))], n, len(chars) + 16)
    L.check(rc)
    np.testing.assert_array_equal(g_offs, o_offs)
    np.testing.assert_array_equal(g_chars[: col.chars_used], chars)


def test_gaps_multi_run_and_nulls():
    """Two runs, an uncovered row range between them (empty strings) and a
    validity mask with its null count."""
    c1, o1, e1 = _zipf(5_000, 1)
    c2, o2, e2 = _zipf(7_000, 2)
    d1, b1 = upload(e1.blob)
    d2, b2 = upload(e2.blob)
    s1 = e1.segments(b1)
    s2 = e2.segments(b2)
    gap = 3_000
    for s in s2:
        s.row_offset += 5_000 + gap
    n = 5_000 + gap + 7_000
    valid = np.random.default_rng(9).random(n) > 0.1
    mb = np.packbits(valid, bitorder="little")
    dm, mbase = upload(mb)
    mask = torch.zeros(len(mb), dtype=torch.uint8, device=_dev())
    vseg = [A.Segment(mbase, len(mb), 0, n, 0)]
    cap = len(c1) + len(c2) + 64
    rc, col, g_offs, g_chars = decode_strings_gpu([(A.FSST, s1), (A.FSST, s2)], n, cap,
                                                  validity=[(A.UNCOMPRESSED, vseg)], d_mask=mask.data_ptr())
    L.check(rc)
    want_lens = np.concatenate([np.diff(o1), np.zeros(gap, np.uint64), np.diff(o2)])
    want_offs = np.zeros(n + 1, np.int64)
    want_offs[1:] = np.cumsum(want_lens)
    np.testing.assert_array_equal(g_offs, want_offs.astype(np.int32))
    np.testing.assert_array_equal(g_chars[: col.chars_used], np.concatenate([c1, c2]))
    assert col.null_count == int((~valid).sum())
    np.testing.assert_array_equal(mask.cpu().numpy(), mb)


def test_capacity_overflow_reports_total():
    chars, offs, enc = _zipf(10_000, 3)
    d, base = upload(enc.blob)
    rc, col, g_offs, _ = decode_strings_gpu([(A.FSST, enc.segments(base))], enc.n_rows, 1000)
    assert rc == A.SDR_E_OVERFLOW
    assert col.chars_used == len(chars)
    np.testing.assert_array_equal(g_offs.astype(np.uint64), offs)


@pytest.mark.parametrize("codec", [A.RLE, A.BITPACKING, A.ALP])
def test_unsupported_string_codec(codec):
    chars, offs, enc = _zipf(100, 4)
    d, base = upload(enc.blob)
    rc, col, _, _ = decode_strings_gpu([(codec, enc.segments(base))], enc.n_rows, 10_000)
    assert rc == A.SDR_E_UNSUPPORTED_CODEC


def test_config4_full_size():
    n = 16 << 20
    chars, offs, enc = _zipf(n, 0xC0DE0004)
    d, base = upload(enc.blob)
    rc, col, g_offs, g_chars = decode_strings_gpu([(A.FSST, enc.segments(base))], n, len(chars) + 64)
    L.check(rc)
    assert col.chars_used == len(chars)
    assert np.array_equal(g_offs.astype(np.uint64), offs)
    assert np.array_equal(g_chars[: col.chars_used], chars)


def _shape(name, seed):
    """String columns that steer chunks onto each decode path: short rows
    (one staged walk), long incompressible rows (compressed bytes larger than
    the stage), long repetitive rows (output larger than one window), one
    row longer than the output window, mostly-empty and all-empty rows."""
    rng = np.random.default_rng(seed)
    if name == "long_random":
        lens = rng.integers(60, 300, 6_000)
    elif name == "repetitive":
        lens = rng.integers(40, 120, 12_000)
    elif name == "huge_row":
        lens = rng.integers(0, 30, 5_000)
        lens[1234] = 70_000
    elif name == "mostly_empty":
        lens = np.where(rng.random(40_000) < 0.7, 0, rng.integers(1, 40, 40_000))
    elif name == "all_empty":
        lens = np.zeros(9_000, np.int64)
    else:
        raise ValueError(name)
    offs = np.zeros(len(lens) + 1, np.uint64)
    offs[1:] = np.cumsum(lens)
    total = int(offs[-1])
    if name == "long_random":
        chars = rng.integers(0, 256, total, dtype=np.uint8)
    else:
        words = [b"abcdefgh", b"ijklmnop", b"qrstuvwx", b"yz012345", b"\xff\xfe\x00\x01"]
        pick = rng.integers(0, len(words), total // 4 + 2)
        chars = np.frombuffer(b"".join(words[i] for i in pick), np.uint8)[:total].copy()
    return chars, offs


SHAPES = ["long_random", "repetitive", "huge_row", "mostly_empty", "all_empty"]


@pytest.mark.parametrize("mode", [A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT, A.LAUNCH_PERSISTENT],
                         ids=["batched", "plain", "persistent"])
@pytest.mark.parametrize("shape", SHAPES)
def test_string_shapes_match_oracle(shape, mode):
    chars, offs = _shape(shape, 0x5A17)
    enc = E.encode_fsst(chars, offs)
    o_offs, o_chars = O.decode_strings(enc)
    np.testing.assert_array_equal(o_chars, chars)
    d, base = upload(enc.blob, 3)
    rc, col, g_offs, g_chars = decode_strings_gpu([(A.FSST, enc.segments(base))], enc.n_rows, len(chars) + 64,
                                                  mode, 5)
    L.check(rc)
    np.testing.assert_array_equal(g_offs, o_offs)
    assert col.chars_used == len(o_chars)
    np.testing.assert_array_equal(g_chars[: col.chars_used], o_chars)


@pytest.mark.parametrize("mode", [A.LAUNCH_BATCHED, A.LAUNCH_PERSISTENT], ids=["batched", "persistent"])
def test_lookback_give_up_reaches_the_caller(mode):
    """With the look-back's wait bound forced to zero from the host, every
    chunk after the first gives up at once: the kernels exit normally, the
    call reports SDR_E_HIP, and with the bound restored the same column
    decodes bit-exactly."""
    lib = L.load()
    chars, offs, enc = _zipf(20_000, 0xC0DE0044)
    d, base = upload(enc.blob)
    runs = [(A.FSST, enc.segments(base))]
    old = lib.sdr_tune_get(A.TUNE_FSST_SPIN_LIMIT)
    try:
        L.check(lib.sdr_tune_set(A.TUNE_FSST_SPIN_LIMIT, 0))
        rc, col, _, _ = decode_strings_gpu(runs, enc.n_rows, len(chars) + 64, mode)
        assert rc == A.SDR_E_HIP
        assert b"look-back did not complete" in lib.sdr_last_error()
    finally:
        L.check(lib.sdr_tune_set(A.TUNE_FSST_SPIN_LIMIT, old))
    rc, col, g_offs, g_chars = decode_strings_gpu(runs, enc.n_rows, len(chars) + 64, mode)
    L.check(rc)
    np.testing.assert_array_equal(g_offs, offs.astype(np.int32))
    np.testing.assert_array_equal(g_chars[: col.chars_used], chars)
