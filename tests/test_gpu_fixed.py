"""GPU parity of the fixed-width path (BITPACKING, RLE, ALP, ALPRD,
UNCOMPRESSED, CONSTANT) through the C ABI, bit-exact against the oracle."""
import ctypes as C
import functools

import numpy as np
import pytest

import sirius_amd._abi as A
import sirius_amd.encode as E
from oracle import oracle as O
from sirius_amd import lib as L
from tests.helpers import fixed_column, golden_cases, segs_from_table

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def upload(blob: np.ndarray, skew: int = 0):
    """Device copy of `blob` starting `skew` bytes into a padded buffer."""
    dev = _dev()
    t = torch.zeros(len(blob) + skew + 256, dtype=torch.uint8, device=dev)
    if len(blob):
        t[skew:skew + len(blob)] = torch.from_numpy(np.ascontiguousarray(blob)).to(dev)
    return t, t.data_ptr() + skew


def stream():
    return torch.cuda.current_stream().cuda_stream


def run_table(cols, mode=A.LAUNCH_BATCHED):
    arr = (A.Column * len(cols))(*cols)
    rc = L.load().sdr_decode_table_mode(arr, len(cols), stream(), mode)
    torch.cuda.synchronize()
    L.check(rc)
    return arr


def decode_gpu(enc, mode=A.LAUNCH_BATCHED, skew=0, out_skew=0, fill=0xCC):
    d_blob, base = upload(enc.blob, skew)
    n, ts = enc.n_rows, enc.type_size
    out = torch.full((n * ts + out_skew + 16,), fill, dtype=torch.uint8, device=_dev())
    col, keep = fixed_column([(enc.codec, enc.segments(base))], ts, n, out.data_ptr() + out_skew)
    run_table([col], mode)
    return out[out_skew:out_skew + n * ts].cpu().numpy()


# ---------------------------------------------------------------------------
# golden vectors
# ---------------------------------------------------------------------------
FIXED = [c for c in golden_cases() if c["codec"] != A.FSST]


@pytest.mark.parametrize("case", FIXED, ids=[c["name"] for c in FIXED])
@pytest.mark.parametrize("mode", [A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT], ids=["batched", "plain"])
def test_golden_on_gpu(case, mode):
    d_blob, base = upload(case["blob"])
    ts, n = case["ts"], case["total"]
    out = torch.full((n * ts,), 0xCC, dtype=torch.uint8, device=_dev())
    col, keep = fixed_column([(case["codec"], segs_from_table(case["table"], base))], ts, n, out.data_ptr())
    run_table([col], mode)
    np.testing.assert_array_equal(out.cpu().numpy(), case["expected"])


# ---------------------------------------------------------------------------
# encoded columns vs oracle
# ---------------------------------------------------------------------------
def _columns():
    rng = np.random.default_rng(1234)
    n = 300_000
    cols = {
        "bp_i32_uniform17": E.encode_bitpacking(E.gen_bp_int32(11, n)),
        "bp_i32_qty": E.encode_bitpacking(E.gen_uniform_int32(12, n, 1, 50)),
        "bp_i32_dates": E.encode_bitpacking(E.gen_dates_sorted(13, n)),
        "bp_i64_wide": E.encode_bitpacking(rng.integers(-(1 << 62), 1 << 62, n, dtype=np.int64)),
        "bp_i64_w40": E.encode_bitpacking(rng.integers(0, 1 << 40, n, dtype=np.int64)),
        "bp_u8": E.encode_bitpacking(rng.integers(0, 256, n, dtype=np.uint8)),
        "bp_i16_sorted": E.encode_bitpacking(np.sort(rng.integers(-30000, 30000, n, dtype=np.int16))),
        "bp_i32_forced_dfor": E.encode_bitpacking(np.cumsum(rng.integers(-5, 9, n)).astype(np.int32), force_mode=4),
        "bp_i64_cdelta": E.encode_bitpacking((np.arange(n, dtype=np.int64) * 7 - 12345)),
        "rle_i64": E.encode_rle(E.gen_runs_int64(21, n)),
        "rle_i32_long": E.encode_rle(E.gen_runs_int64(22, n, mean_run=700.0, value_mod=1000).astype(np.int32)),
        "rle_u8_short": E.encode_rle(E.gen_runs_int64(23, n, mean_run=1.5, value_mod=3).astype(np.uint8)),
        "rle_i16": E.encode_rle(E.gen_runs_int64(24, n, mean_run=40.0, value_mod=60000).astype(np.int16)),
        "alp_f64": E.encode_alp(E.gen_alp_doubles(31, n)),
        "alp_f32": E.encode_alp((rng.integers(0, 100000, n) / 100).astype(np.float32)),
        "alprd_f64": E.encode_alprd(E.gen_unit_doubles(32, n)),
        "alprd_f32": E.encode_alprd(rng.random(n, dtype=np.float32)),
        "alp_f64_mixed": E.encode_alp(np.where(rng.random(n) < 0.3, rng.random(n), np.round(rng.random(n) * 1e4, 3))),
        "unc_i64": E.encode_uncompressed(rng.integers(0, 1 << 60, n, dtype=np.int64)),
    }
    return cols


COLS = None


def cols():
    global COLS
    if COLS is None:
        COLS = _columns()
    return COLS


NAMES = ["bp_i32_uniform17", "bp_i32_qty", "bp_i32_dates", "bp_i64_wide", "bp_i64_w40", "bp_u8", "bp_i16_sorted",
         "bp_i32_forced_dfor", "bp_i64_cdelta", "rle_i64", "rle_i32_long", "rle_u8_short", "rle_i16", "alp_f64",
         "alp_f32", "alprd_f64", "alprd_f32", "alp_f64_mixed", "unc_i64"]


@pytest.mark.parametrize("name", NAMES)
@pytest.mark.parametrize("mode", [A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT], ids=["batched", "plain"])
def test_encoded_column_matches_oracle(name, mode):
    enc = cols()[name]
    want = O.decode_encoded(enc, threads=8)
    got = decode_gpu(enc, mode)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("name", ["bp_i32_uniform17", "bp_i64_wide", "rle_i64", "alp_f64", "alprd_f64", "bp_u8"])
@pytest.mark.parametrize("skew", [1, 3, 6, 13])
def test_unaligned_segment_bytes(name, skew):
    """Segments at every byte alignment decode identically."""
    enc = cols()[name]
    np.testing.assert_array_equal(decode_gpu(enc, skew=skew), O.decode_encoded(enc, threads=8))


@pytest.mark.parametrize("name", ["bp_i32_uniform17", "rle_i64", "alp_f32", "bp_u8", "bp_i16_sorted"])
def test_unaligned_output(name):
    """A destination that is not 16-byte aligned takes the element-store path."""
    enc = cols()[name]
    ts = enc.type_size
    np.testing.assert_array_equal(decode_gpu(enc, out_skew=ts), O.decode_encoded(enc, threads=8))


def test_mixed_table_multi_column():
    """Several columns and codecs in one call (gpu_decode_table's shape)."""
    n = 200_000
    a = E.encode_bitpacking(E.gen_bp_int32(41, n))
    b = E.encode_rle(E.gen_runs_int64(42, n))
    c = E.encode_alp(E.gen_alp_doubles(43, n))
    keep, cols_, outs = [], [], []
    for enc in (a, b, c):
        d, base = upload(enc.blob)
        out = torch.full((n * enc.type_size,), 0xCC, dtype=torch.uint8, device=_dev())
        col, k = fixed_column([(enc.codec, enc.segments(base))], enc.type_size, n, out.data_ptr())
        keep += [d, k]
        cols_.append(col)
        outs.append(out)
    run_table(cols_)
    for enc, out in zip((a, b, c), outs):
        np.testing.assert_array_equal(out.cpu().numpy(), O.decode_encoded(enc, threads=8))


def test_mixed_codec_runs_in_one_column():
    """One column whose segments alternate codecs (several runs)."""
    n = 122880 * 4
    v = E.gen_runs_int64(51, n, mean_run=30.0, value_mod=1 << 20).astype(np.int32)
    parts = []
    for i in range(4):
        sl = v[i * 122880:(i + 1) * 122880]
        enc = E.encode_rle(sl) if i % 2 else E.encode_bitpacking(sl)
        parts.append(enc)
    keep, runs = [], []
    for i, enc in enumerate(parts):
        d, base = upload(enc.blob)
        keep.append(d)
        segs = enc.segments(base)
        for s in segs:
            s.row_offset += i * 122880
        runs.append((enc.codec, segs))
    out = torch.full((n * 4,), 0xCC, dtype=torch.uint8, device=_dev())
    col, k = fixed_column(runs, 4, n, out.data_ptr())
    run_table([col])
    np.testing.assert_array_equal(out.cpu().numpy().view(np.int32), v)


def test_validity_and_null_count():
    n = 10_000
    enc = E.encode_bitpacking(E.gen_bp_int32(61, n))
    rng = np.random.default_rng(7)
    valid = rng.random(n) > 0.2
    maskbytes = np.packbits(valid, bitorder="little")
    d_blob, base = upload(enc.blob)
    d_mask_src, mbase = upload(maskbytes)
    out = torch.zeros(n * 4, dtype=torch.uint8, device=_dev())
    mask = torch.zeros((n + 7) // 8, dtype=torch.uint8, device=_dev())
    vseg = [A.Segment(mbase, len(maskbytes), 0, n, 0)]
    col, keep = fixed_column([(enc.codec, enc.segments(base))], 4, n, out.data_ptr(),
                             validity=[(A.UNCOMPRESSED, vseg)], d_mask=mask.data_ptr())
    arr = run_table([col])
    assert arr[0].null_count == int((~valid).sum())
    np.testing.assert_array_equal(mask.cpu().numpy(), maskbytes)


def test_per_codec_entry_points():
    n = 100_000
    enc = E.encode_bitpacking(E.gen_bp_int32(71, n))
    d_blob, base = upload(enc.blob)
    out = torch.zeros(n * 4, dtype=torch.uint8, device=_dev())
    runs, keep = A.make_runs([(A.BITPACKING, enc.segments(base))])
    rc = L.load().sdr_decode_bitpacking_data(runs, out.data_ptr(), 4, stream())
    torch.cuda.synchronize()
    L.check(rc)
    np.testing.assert_array_equal(out.cpu().numpy(), O.decode_encoded(enc))
    rc = L.load().sdr_decode_rle_data(runs, out.data_ptr(), 4, stream())
    assert rc == A.SDR_E_INVALID_ARG  # run codec does not match the entry point


def test_uncompressed_and_constant():
    n = 5000
    rng = np.random.default_rng(3)
    for ts in (1, 2, 4, 8, 16):
        raw = rng.integers(0, 256, n * ts, dtype=np.uint8)
        const = rng.integers(0, 256, ts, dtype=np.uint8)
        d1, b1 = upload(raw)
        d2, b2 = upload(const)
        out = torch.zeros(2 * n * ts, dtype=torch.uint8, device=_dev())
        col, keep = fixed_column([(A.UNCOMPRESSED, [A.Segment(b1, n * ts, 0, n, 0)]),
                                  (A.CONSTANT, [A.Segment(b2, ts, n, n, 0)])], ts, 2 * n, out.data_ptr())
        run_table([col])
        got = out.cpu().numpy()
        np.testing.assert_array_equal(got[: n * ts], raw)
        np.testing.assert_array_equal(got[n * ts:], np.tile(const, n))


@pytest.mark.parametrize("mode", [A.LAUNCH_BATCHED, A.LAUNCH_PERSISTENT], ids=["batched", "persistent"])
@pytest.mark.parametrize("skew,out_skew", [(0, 0), (0, 5), (3, 3), (8, 1), (13, 0), (8, 0), (4, 0), (5, 9), (2, 0)])
def test_uncompressed_and_constant_any_alignment(mode, skew, out_skew):
    """Copies and fills at phases of 16 between source and destination that
    take each word width (16, 8, 4 bytes after a byte head, or bytes
    throughout), spans of several units with odd tails; bytes outside the
    output keep their fill."""
    rng = np.random.default_rng(11 + skew + 7 * out_skew)
    for ts in (1, 2, 4, 8, 16):
        n = 2101 if ts >= 8 else 9001
        raw = rng.integers(0, 256, n * ts, dtype=np.uint8)
        const = rng.integers(0, 256, ts, dtype=np.uint8)
        d1, b1 = upload(raw, skew)
        d2, b2 = upload(const, skew + 1)
        out = torch.full((2 * n * ts + out_skew + 64,), 0xCC, dtype=torch.uint8, device=_dev())
        col, keep = fixed_column([(A.UNCOMPRESSED, [A.Segment(b1, n * ts, 0, n, 0)]),
                                  (A.CONSTANT, [A.Segment(b2, ts, n, n, 0)])], ts, 2 * n, out.data_ptr() + out_skew)
        run_table([col], mode)
        got = out.cpu().numpy()
        np.testing.assert_array_equal(got[out_skew:out_skew + n * ts], raw, err_msg=f"copy ts={ts}")
        np.testing.assert_array_equal(got[out_skew + n * ts:out_skew + 2 * n * ts], np.tile(const, n),
                                      err_msg=f"constant ts={ts}")
        assert (got[:out_skew] == 0xCC).all() and (got[out_skew + 2 * n * ts:] == 0xCC).all()


def test_errors_are_returned_not_thrown():
    n = 100
    enc = E.encode_bitpacking(E.gen_bp_int32(81, n))
    d_blob, base = upload(enc.blob)
    out = torch.zeros(n * 4, dtype=torch.uint8, device=_dev())
    segs = enc.segments(base)
    col, keep = fixed_column([(A.FSST, segs)], 4, n, out.data_ptr())
    arr = (A.Column * 1)(col)
    assert L.load().sdr_decode_table(arr, 1, stream()) == A.SDR_E_UNSUPPORTED_CODEC
    col, keep = fixed_column([(A.ALP, segs)], 2, n, out.data_ptr())
    arr = (A.Column * 1)(col)
    assert L.load().sdr_decode_table(arr, 1, stream()) == A.SDR_E_BAD_WIDTH
    col, keep = fixed_column([(A.BITPACKING, segs)], 4, n - 1, out.data_ptr())
    arr = (A.Column * 1)(col)
    assert L.load().sdr_decode_table(arr, 1, stream()) == A.SDR_E_BOUNDS
    assert b"row_count" in L.load().sdr_last_error()


# ---------------------------------------------------------------------------
# configs at full size: bit-exact against the generated values (the encoders
# and the oracle round-trip these exactly; see the CPU tests)
# ---------------------------------------------------------------------------
def test_config1_plain_full_size():
    v = E.gen_bp_int32(0xC0DE0001, 8 << 20)
    enc = E.encode_bitpacking(v)
    assert len(enc.segs) == 69
    got = decode_gpu(enc, A.LAUNCH_PER_SEGMENT)
    np.testing.assert_array_equal(got.view(np.int32), v)


def test_config2_columns_full_size():
    n = 64 << 20
    for gen, f in ((lambda: E.gen_runs_int64(0xC0DE0002, n), E.encode_rle),
                   (lambda: E.gen_uniform_int32(0xC0DE0012, n, 1, 50), E.encode_bitpacking),
                   (lambda: E.gen_dates_sorted(0xC0DE0022, n), E.encode_bitpacking)):
        v = gen()
        enc = f(v)
        got = decode_gpu(enc)
        assert np.array_equal(got, v.view(np.uint8))


@pytest.mark.parametrize("mode", [A.LAUNCH_BATCHED, A.LAUNCH_PERSISTENT, A.LAUNCH_PER_SEGMENT],
                         ids=["batched", "persistent", "per_segment"])
def test_config3_alp_full_size(mode):
    """32 Mi ALP doubles and 8 Mi ALP-RD doubles, the benchmark's seeds:
    every bit pattern equals the generated value's (-0.0, NaN payloads and
    exceptions included)."""
    for v, enc in _config3_columns():
        got = decode_gpu(enc, mode)
        assert np.array_equal(got.view(np.uint64), v.view(np.uint64))


@functools.lru_cache(maxsize=1)
def _config3_columns():
    a = E.gen_alp_doubles(0xC0DE0003, 32 << 20)
    r = E.gen_unit_doubles(0xC0DE0013, 8 << 20)
    return ((a, E.encode_alp(a)), (r, E.encode_alprd(r)))


@pytest.mark.parametrize("dtype,bits", [(np.float64, b) for b in (31, 32, 45, 48, 52)] +
                         [(np.float32, b) for b in (31, 32)],  # a float column tops out at 32-bit codes
                         ids=lambda x: {np.float64: "f64", np.float32: "f32"}.get(x, str(x)))
def test_alp_wide_codes_staged_in_chunks(dtype, bits):
    """ALP vectors whose codes do not fit one 4 KiB stage (widths above 31 bits) decode in
    row chunks."""
    rng = np.random.default_rng(bits)
    n = 100_000
    k = rng.integers(0, 2 ** min(bits, 53), n, dtype=np.int64)
    vals = (k.astype(np.float64) / 100.0).astype(dtype)
    enc = E.encode_alp(vals)
    np.testing.assert_array_equal(O.decode_encoded(enc).view(dtype), vals)  # oracle round trip
    for mode in (A.LAUNCH_BATCHED, A.LAUNCH_PERSISTENT):
        got = decode_gpu(enc, mode)
        np.testing.assert_array_equal(got, O.decode_encoded(enc))


@pytest.mark.parametrize("n_segs", [1, 100, 128, 129, 256, 257, 300])
@pytest.mark.parametrize("codec", ["bp", "rle", "alp", "unc"])
def test_descriptors_in_kernel_arguments_and_on_the_device(codec, n_segs):
    """A group of up to 256 segments travels in the kernel arguments (128
    per launch), a larger one through the uploaded descriptor array; both
    decode the same column bit-exact, batched and one launch per segment."""
    rows = n_segs * 2048
    if codec == "bp":
        enc = E.encode_bitpacking(E.gen_bp_int32(0xD0 + n_segs, rows), rg_rows=2048)
    elif codec == "rle":
        enc = E.encode_rle(E.gen_runs_int64(0xD1 + n_segs, rows, 16.0, 1 << 40), rg_rows=2048)
    elif codec == "alp":
        enc = E.encode_alp(E.gen_alp_doubles(0xD2 + n_segs, rows), rg_rows=2048)
    else:
        enc = E.encode_uncompressed(E.gen_u64(0xD3 + n_segs, rows).view(np.int64), rg_rows=2048)
    assert len(enc.segs) == n_segs
    want = O.decode_encoded(enc, threads=8)
    for mode in (A.LAUNCH_BATCHED, A.LAUNCH_PER_SEGMENT):
        np.testing.assert_array_equal(decode_gpu(enc, mode), want, err_msg=f"mode {mode}")
