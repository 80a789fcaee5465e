"""ctypes mirror of include/sdr/decode.h (structs, enums, status codes)."""
from __future__ import annotations

import ctypes as C

SDR_OK = 0
SDR_E_INVALID_ARG = -1
SDR_E_UNSUPPORTED_CODEC = -2
SDR_E_BAD_WIDTH = -3
SDR_E_BOUNDS = -4
SDR_E_OVERFLOW = -5
SDR_E_ALIGNMENT = -6
SDR_E_HIP = -7
SDR_E_CAPACITY = -8
SDR_E_NOT_FOUND = -9
SDR_E_STATE = -10

STATUS_NAMES = {
    0: "SDR_OK", -1: "SDR_E_INVALID_ARG", -2: "SDR_E_UNSUPPORTED_CODEC", -3: "SDR_E_BAD_WIDTH",
    -4: "SDR_E_BOUNDS", -5: "SDR_E_OVERFLOW", -6: "SDR_E_ALIGNMENT", -7: "SDR_E_HIP",
    -8: "SDR_E_CAPACITY", -9: "SDR_E_NOT_FOUND", -10: "SDR_E_STATE",
}

UNCOMPRESSED = 1
CONSTANT = 2
RLE = 3
DICTIONARY = 4
BITPACKING = 6
FSST = 7
ALP = 10
ALPRD = 11
DICT_FSST = 15

CODEC_NAMES = {
    UNCOMPRESSED: "Uncompressed", CONSTANT: "Constant", RLE: "RLE", DICTIONARY: "Dictionary",
    BITPACKING: "BitPacking", FSST: "FSST", ALP: "ALP", ALPRD: "ALPRD", DICT_FSST: "DICT_FSST",
}

LAUNCH_BATCHED = 0
LAUNCH_PER_SEGMENT = 1
LAUNCH_PERSISTENT = 2

BATCH_USE_GRAPH = 1
BATCH_GRID = 2
BATCH_FIXED_KERNELS = 3  # 0: per-codec grids (default), 1: persistent group kernels
BATCH_UNITS = {BITPACKING: 16, RLE: 17, ALP: 18, ALPRD: 19, UNCOMPRESSED: 20, CONSTANT: 21}

# sdr_tune_key
TUNE_RLE_SLICE_ROWS = 1
TUNE_BP_GROUPS_PER_WAVE = 2
TUNE_BP_CHUNK_ROWS = 3
TUNE_ALP_VECTORS_PER_WAVE = 4
TUNE_FSST_SPIN_LIMIT = 5
TUNE_BATCH_UNITS_BITPACKING = 6
TUNE_BATCH_UNITS_RLE = 7
TUNE_BATCH_UNITS_ALP = 8
TUNE_BATCH_UNITS_ALPRD = 9
TUNE_BATCH_UNITS_COPY = 10
TUNE_NAMES = {TUNE_RLE_SLICE_ROWS: "rle_slice_rows", TUNE_BP_GROUPS_PER_WAVE: "bp_groups_per_wave",
              TUNE_BP_CHUNK_ROWS: "bp_chunk_rows", TUNE_ALP_VECTORS_PER_WAVE: "alp_vectors_per_wave",
              TUNE_FSST_SPIN_LIMIT: "fsst_spin_limit", TUNE_BATCH_UNITS_BITPACKING: "batch_units_bitpacking",
              TUNE_BATCH_UNITS_RLE: "batch_units_rle", TUNE_BATCH_UNITS_ALP: "batch_units_alp",
              TUNE_BATCH_UNITS_ALPRD: "batch_units_alprd", TUNE_BATCH_UNITS_COPY: "batch_units_copy"}
TUNE_WIDTHS = (1, 2, 4, 8)


class RowRange(C.Structure):
    """sdr_row_range: table rows [first_row, first_row + count) written at out_offset."""
    _fields_ = [
        ("first_row", C.c_int64),
        ("out_offset", C.c_uint64),
        ("count", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class Extent(C.Structure):
    """sdr_extent: a byte range of a file and its position in the staged device image."""
    _fields_ = [
        ("file_offset", C.c_uint64),
        ("bytes", C.c_uint64),
        ("device_offset", C.c_uint64),
    ]


class Segment(C.Structure):
    _fields_ = [
        ("d_bytes", C.c_void_p),
        ("bytes_size", C.c_uint32),
        ("row_offset", C.c_uint32),
        ("row_count", C.c_uint32),
        ("max_string_length", C.c_uint32),
    ]


class Run(C.Structure):
    _fields_ = [
        ("codec", C.c_uint32),
        ("n_segments", C.c_uint32),
        ("segments", C.POINTER(Segment)),
    ]


class Column(C.Structure):
    _fields_ = [
        ("type_size", C.c_uint32),
        ("total_rows", C.c_uint32),
        ("data", C.POINTER(Run)),
        ("n_data", C.c_uint32),
        ("has_nulls", C.c_uint32),
        ("validity", C.POINTER(Run)),
        ("n_validity", C.c_uint32),
        ("reserved", C.c_uint32),
        ("d_out", C.c_void_p),
        ("d_null_mask", C.c_void_p),
        ("null_count", C.c_int64),
    ]


class StringColumn(C.Structure):
    _fields_ = [
        ("total_rows", C.c_uint32),
        ("has_nulls", C.c_uint32),
        ("data", C.POINTER(Run)),
        ("n_data", C.c_uint32),
        ("n_validity", C.c_uint32),
        ("validity", C.POINTER(Run)),
        ("d_offsets", C.c_void_p),
        ("d_chars", C.c_void_p),
        ("chars_capacity", C.c_uint64),
        ("d_null_mask", C.c_void_p),
        ("chars_used", C.c_uint64),
        ("null_count", C.c_int64),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("budget_bytes", C.c_uint64),
        ("max_inflight", C.c_uint32),
        ("mode", C.c_uint32),
        ("max_batch_requests", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("launches", C.c_uint64),
        ("requests", C.c_uint64),
        ("queue_entries", C.c_uint64),
        ("host_submit_ms", C.c_double),
        ("host_collect_ms", C.c_double),
        ("host_alloc_ms", C.c_double),
        ("host_launch_ms", C.c_double),
        ("host_release_ms", C.c_double),
        ("host_stage_ms", C.c_double),
        ("gpu_busy_ms", C.c_double),
        ("gpu_union_ms", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Request(C.Structure):
    _fields_ = [
        ("tag", C.c_uint64),
        ("is_string", C.c_uint32),
        ("type_size", C.c_uint32),
        ("total_rows", C.c_uint32),
        ("n_data", C.c_uint32),
        ("data", C.POINTER(Run)),
        ("chars_capacity", C.c_uint64),
    ]


class Completion(C.Structure):
    _fields_ = [
        ("tag", C.c_uint64),
        ("status", C.c_int32),
        ("reserved", C.c_uint32),
        ("d_out", C.c_void_p),
        ("d_chars", C.c_void_p),
        ("chars_used", C.c_uint64),
        ("out_bytes", C.c_uint64),
        ("submit_ms", C.c_double),
        ("start_ms", C.c_double),
        ("done_ms", C.c_double),
    ]


def make_runs(runs):
    """[(codec, [Segment...]), ...] -> (Run array, keepalive list)."""
    keep = []
    arr = (Run * max(1, len(runs)))()
    for i, (codec, segs) in enumerate(runs):
        sa = (Segment * max(1, len(segs)))(*segs)
        keep.append(sa)
        arr[i].codec = codec
        arr[i].n_segments = len(segs)
        arr[i].segments = C.cast(sa, C.POINTER(Segment))
    keep.append(arr)
    return arr, keep
