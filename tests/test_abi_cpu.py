"""The C-ABI libraries load and export every symbol the headers declare;
argument validation answers without touching a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import sirius_amd._abi as A
from sirius_amd import lib as L

ROOT = Path(__file__).resolve().parent.parent


def declared(header: Path):
    text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(sdr_[a-z0-9_]+)\s*\(", text)))


def test_decode_header_symbols_exported():
    names = sorted(set().union(*(declared(h) for h in (ROOT / "include" / "sdr").glob("*.h"))))
    assert "sdr_decode_table" in names and "sdr_decode_strings_column" in names and "sdr_arrow_export" in names
    lib = L.load()
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/sdr/ but not exported"
    assert sorted(L.DECLARED) == names


def test_encoder_header_symbols_exported():
    import sirius_amd.encode as E

    lib = E._load()
    for n in declared(ROOT / "sirius_amd" / "csrc" / "encode" / "sdr_encode.h"):
        assert hasattr(lib, n), n


def test_codec_names_roundtrip():
    lib = L.load()
    for cid, name in A.CODEC_NAMES.items():
        assert lib.sdr_codec_from_name(name.encode()) == cid
        assert lib.sdr_codec_name(cid).decode() == name
    assert lib.sdr_codec_from_name(b"Chimp") == A.SDR_E_UNSUPPORTED_CODEC
    assert lib.sdr_abi_version() == 1
    assert lib.sdr_build_target() == b"gfx950"


def _col(codec, ts, total, rows, row_offset=0, bytes_size=64):
    seg = A.Segment(0x1000, bytes_size, row_offset, rows, 0)
    runs, keep = A.make_runs([(codec, [seg])])
    col = A.Column()
    col.type_size, col.total_rows, col.data, col.n_data, col.d_out = ts, total, runs, 1, 0x2000
    return col, keep


@pytest.mark.parametrize("codec,ts,total,rows,off,code", [
    (A.FSST, 4, 10, 10, 0, A.SDR_E_UNSUPPORTED_CODEC),
    (A.DICTIONARY, 4, 10, 10, 0, A.SDR_E_UNSUPPORTED_CODEC),
    (A.ALP, 2, 10, 10, 0, A.SDR_E_BAD_WIDTH),
    (A.BITPACKING, 16, 10, 10, 0, A.SDR_E_BAD_WIDTH),
    (A.RLE, 3, 10, 10, 0, A.SDR_E_BAD_WIDTH),
    (A.BITPACKING, 4, 10, 8, 5, A.SDR_E_BOUNDS),
])
def test_validation_errors_without_device_work(codec, ts, total, rows, off, code):
    col, keep = _col(codec, ts, total, rows, off)
    arr = (A.Column * 1)(col)
    assert L.load().sdr_decode_table(arr, 1, None) == code
    assert L.load().sdr_last_error()


def test_mismatched_total_rows_and_short_uncompressed():
    a, k1 = _col(A.BITPACKING, 4, 10, 10)
    b, k2 = _col(A.BITPACKING, 4, 11, 11)
    arr = (A.Column * 2)(a, b)
    assert L.load().sdr_decode_table(arr, 2, None) == A.SDR_E_BOUNDS
    c, k3 = _col(A.UNCOMPRESSED, 8, 10, 10, 0, bytes_size=79)
    arr = (A.Column * 1)(c)
    assert L.load().sdr_decode_table(arr, 1, None) == A.SDR_E_BOUNDS
    d, k4 = _col(A.CONSTANT, 8, 10, 10, 0, bytes_size=7)
    arr = (A.Column * 1)(d)
    assert L.load().sdr_decode_table(arr, 1, None) == A.SDR_E_BOUNDS


def test_empty_table_is_ok():
    assert L.load().sdr_decode_table(None, 0, None) == A.SDR_OK


def test_strings_upper_bound():
    segs = [A.Segment(0x1000, 100, 0, 10, 44), A.Segment(0x2000, 100, 10, 5, 20)]
    runs, keep = A.make_runs([(A.FSST, segs)])
    assert L.load().sdr_strings_upper_bound(runs, 1) == 10 * 44 + 5 * 20
    segs[1].max_string_length = 0
    runs, keep = A.make_runs([(A.FSST, segs)])
    assert L.load().sdr_strings_upper_bound(runs, 1) == 2 ** 64 - 1


def test_tuning_knobs_set_get_load(tmp_path):
    lib = L.load()
    defaults = {k: lib.sdr_tune_get(k) for k in A.TUNE_NAMES}
    assert defaults[A.TUNE_RLE_SLICE_ROWS] == 16384 and defaults[A.TUNE_BP_CHUNK_ROWS] == 2048
    try:
        assert lib.sdr_tune_set(A.TUNE_RLE_SLICE_ROWS, 8192) == A.SDR_OK
        assert lib.sdr_tune_get(A.TUNE_RLE_SLICE_ROWS) == 8192
        # out of range or not a multiple of 64, unknown key, bad chunk size
        assert lib.sdr_tune_set(A.TUNE_RLE_SLICE_ROWS, 1000) == A.SDR_E_INVALID_ARG
        assert lib.sdr_tune_set(A.TUNE_RLE_SLICE_ROWS, 8200) == A.SDR_E_INVALID_ARG
        assert lib.sdr_tune_set(99, 1) == A.SDR_E_INVALID_ARG
        assert lib.sdr_tune_set(A.TUNE_BP_CHUNK_ROWS, 768) == A.SDR_E_INVALID_ARG
        assert lib.sdr_tune_get(A.TUNE_RLE_SLICE_ROWS) == 8192 and lib.sdr_tune_get(99) == 0
        t = tmp_path / "tune.txt"
        t.write_text("# table\nrle_slice_rows 32768\n\nbp_groups_per_wave 2  # two groups\n")
        assert lib.sdr_tune_load(str(t).encode()) == A.SDR_OK
        assert lib.sdr_tune_get(A.TUNE_RLE_SLICE_ROWS) == 32768
        assert lib.sdr_tune_get(A.TUNE_BP_GROUPS_PER_WAVE) == 2
        t.write_text("rle_slice_rows 4096\nno_such_knob 3\nbp_chunk_rows 512\n")
        assert lib.sdr_tune_load(str(t).encode()) == A.SDR_E_INVALID_ARG
        assert b"bad tuning line" in lib.sdr_last_error()
        assert lib.sdr_tune_get(A.TUNE_RLE_SLICE_ROWS) == 4096  # lines before the bad one applied
        assert lib.sdr_tune_get(A.TUNE_BP_CHUNK_ROWS) == 2048
        assert lib.sdr_tune_load(str(tmp_path / "missing.txt").encode()) == A.SDR_E_NOT_FOUND
        # per-width values: a plain line sets every width, "name@ts" one of them
        t.write_text("rle_slice_rows 8192\nrle_slice_rows@8 24576\nbatch_units_rle@2 3\n")
        assert lib.sdr_tune_load(str(t).encode()) == A.SDR_OK
        assert [lib.sdr_tune_get_width(A.TUNE_RLE_SLICE_ROWS, w) for w in A.TUNE_WIDTHS] == [8192, 8192, 8192, 24576]
        assert [lib.sdr_tune_get_width(A.TUNE_BATCH_UNITS_RLE, w) for w in A.TUNE_WIDTHS] == [2, 3, 2, 2]
        assert lib.sdr_tune_get(A.TUNE_RLE_SLICE_ROWS) == 8192  # the 4-byte value
        assert lib.sdr_tune_set_width(A.TUNE_BP_CHUNK_ROWS, 8, 1024) == A.SDR_OK
        assert lib.sdr_tune_get_width(A.TUNE_BP_CHUNK_ROWS, 8) == 1024 and lib.sdr_tune_get(A.TUNE_BP_CHUNK_ROWS) == 2048
        assert lib.sdr_tune_set_width(A.TUNE_BP_CHUNK_ROWS, 3, 1024) == A.SDR_E_BAD_WIDTH
        assert lib.sdr_tune_set_width(A.TUNE_BP_CHUNK_ROWS, 8, 700) == A.SDR_E_INVALID_ARG
        assert lib.sdr_tune_get_width(A.TUNE_BP_CHUNK_ROWS, 3) == 0
        for bad in ("rle_slice_rows@3 8192\n", "rle_slice_rows@ 8192\n", "rle_slice_rows@8x 8192\n"):
            t.write_text(bad)
            assert lib.sdr_tune_load(str(t).encode()) == A.SDR_E_INVALID_ARG, bad
    finally:
        for k, v in defaults.items():
            assert lib.sdr_tune_set(k, v) == A.SDR_OK
