"""DuckDB-format encoders and seeded generators (ctypes over libsdr_encode.so).

An encoded column is one byte blob (segments 16-byte aligned, as the
reference's staging buffer lays them out) plus a segment table.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import _abi

LIB_PATH = Path(__file__).resolve().parent / "libsdr_encode.so"
ROW_GROUP_ROWS = 122880
BLOCK_SIZE = 262136

SEG_DTYPE = np.dtype([("offset", "<u8"), ("bytes_size", "<u4"), ("row_offset", "<u4"),
                      ("row_count", "<u4"), ("max_string_length", "<u4")])

_lib = None


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"encoder library not built: {LIB_PATH} (run python -m sirius_amd.build)")
    lib = C.CDLL(str(LIB_PATH))
    vp, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int
    lib.sdr_encode_bitpacking.argtypes = [vp, u64, u32, i32, u32, u32, i32]
    lib.sdr_encode_rle.argtypes = [vp, u64, u32, u32, u32]
    lib.sdr_encode_alp.argtypes = [vp, u64, u32, u32, u32]
    lib.sdr_encode_alprd.argtypes = [vp, u64, u32, u32, u32]
    lib.sdr_encode_uncompressed.argtypes = [vp, u64, u32, u32, u32]
    lib.sdr_encode_fsst.argtypes = [vp, vp, u64, u32, u32, u64]
    lib.sdr_encode_dict_fsst.argtypes = [vp, vp, vp, u64, u32, u32, u32, u64]
    for f in ("sdr_encode_bitpacking", "sdr_encode_rle", "sdr_encode_alp", "sdr_encode_alprd",
              "sdr_encode_uncompressed", "sdr_encode_fsst", "sdr_encode_dict_fsst"):
        getattr(lib, f).restype = vp
    lib.sdr_encoded_bytes.argtypes = [vp, C.POINTER(u64)]
    lib.sdr_encoded_bytes.restype = vp
    lib.sdr_encoded_segments.argtypes = [vp, C.POINTER(u32)]
    lib.sdr_encoded_segments.restype = vp
    lib.sdr_encoded_free.argtypes = [vp]
    lib.sdr_gen_u64.argtypes = [u64, u64, vp]
    lib.sdr_gen_bp_int32.argtypes = [u64, u64, vp]
    lib.sdr_gen_runs_int64.argtypes = [u64, u64, C.c_double, u64, vp]
    lib.sdr_gen_uniform_int32.argtypes = [u64, u64, C.c_int32, u32, vp]
    lib.sdr_gen_dates_sorted.argtypes = [u64, u64, vp]
    lib.sdr_gen_alp_doubles.argtypes = [u64, u64, vp]
    lib.sdr_gen_unit_doubles.argtypes = [u64, u64, vp]
    lib.sdr_gen_zipf_strings.argtypes = [u64, u64, u32, u32, C.c_double, vp, vp]
    lib.sdr_gen_zipf_strings.restype = u64
    _lib = lib
    return lib


@dataclass
class Encoded:
    """One encoded column: blob + segment table (offsets index the blob)."""
    codec: int
    type_size: int  # bytes per decoded value (0 for strings)
    n_rows: int
    blob: np.ndarray  # uint8
    segs: np.ndarray  # SEG_DTYPE

    @property
    def compressed_bytes(self) -> int:
        return int(self.segs["bytes_size"].sum())

    def segments(self, base_ptr: int) -> list:
        out = []
        for s in self.segs:
            out.append(_abi.Segment(base_ptr + int(s["offset"]), int(s["bytes_size"]), int(s["row_offset"]),
                                    int(s["row_count"]), int(s["max_string_length"])))
        return out


def _take(h, codec, ts, n) -> Encoded:
    lib = _load()
    if not h:
        raise ValueError("encoder rejected its arguments")
    try:
        size = C.c_uint64()
        p = lib.sdr_encoded_bytes(h, C.byref(size))
        blob = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), shape=(size.value,)).copy() \
            if size.value else np.zeros(0, np.uint8)
        ns = C.c_uint32()
        sp = lib.sdr_encoded_segments(h, C.byref(ns))
        raw = C.string_at(sp, ns.value * SEG_DTYPE.itemsize) if ns.value else b""
        segs = np.frombuffer(raw, dtype=SEG_DTYPE).copy()
    finally:
        lib.sdr_encoded_free(h)
    return Encoded(codec, ts, n, blob, segs)


def _c(a: np.ndarray):
    a = np.ascontiguousarray(a)
    return a, a.ctypes.data


def encode_bitpacking(values: np.ndarray, force_mode: int = 0, rg_rows: int = ROW_GROUP_ROWS,
                      block: int = BLOCK_SIZE) -> Encoded:
    a, p = _c(values)
    signed = np.issubdtype(a.dtype, np.signedinteger)
    h = _load().sdr_encode_bitpacking(p, a.size, a.dtype.itemsize, int(signed), rg_rows, block, force_mode)
    return _take(h, _abi.BITPACKING, a.dtype.itemsize, a.size)


def encode_rle(values: np.ndarray, rg_rows: int = ROW_GROUP_ROWS, block: int = BLOCK_SIZE) -> Encoded:
    a, p = _c(values)
    h = _load().sdr_encode_rle(p, a.size, a.dtype.itemsize, rg_rows, block)
    return _take(h, _abi.RLE, a.dtype.itemsize, a.size)


def encode_alp(values: np.ndarray, rg_rows: int = ROW_GROUP_ROWS, block: int = BLOCK_SIZE) -> Encoded:
    a, p = _c(values)
    h = _load().sdr_encode_alp(p, a.size, a.dtype.itemsize, rg_rows, block)
    return _take(h, _abi.ALP, a.dtype.itemsize, a.size)


def encode_alprd(values: np.ndarray, rg_rows: int = ROW_GROUP_ROWS, block: int = BLOCK_SIZE) -> Encoded:
    a, p = _c(values)
    h = _load().sdr_encode_alprd(p, a.size, a.dtype.itemsize, rg_rows, block)
    return _take(h, _abi.ALPRD, a.dtype.itemsize, a.size)


def encode_uncompressed(values: np.ndarray, rg_rows: int = ROW_GROUP_ROWS, block: int = BLOCK_SIZE) -> Encoded:
    a, p = _c(values)
    h = _load().sdr_encode_uncompressed(p, a.size, a.dtype.itemsize, rg_rows, block)
    return _take(h, _abi.UNCOMPRESSED, a.dtype.itemsize, a.size)


def encode_fsst(chars: np.ndarray, offsets: np.ndarray, seed: int = 0x5EED, rg_rows: int = ROW_GROUP_ROWS,
                block: int = BLOCK_SIZE) -> Encoded:
    c, pc = _c(chars.astype(np.uint8, copy=False))
    o, po = _c(offsets.astype(np.uint64, copy=False))
    n = o.size - 1
    h = _load().sdr_encode_fsst(pc, po, n, rg_rows, block, seed)
    return _take(h, _abi.FSST, 0, n)


def encode_dict_fsst(chars: np.ndarray, offsets: np.ndarray, mode: int, valid: np.ndarray | None = None,
                     seg_rows: int = 4096, block: int = BLOCK_SIZE, seed: int = 0x5EED) -> Encoded:
    """DICT_FSST segments (mode 0 raw dictionary, 1 FSST-compressed
    dictionary, 2 one FSST entry per row) of up to `seg_rows` rows each;
    rows whose `valid` byte is 0 select entry 0, the NULL slot."""
    c, pc = _c(chars.astype(np.uint8, copy=False))
    o, po = _c(offsets.astype(np.uint64, copy=False))
    n = o.size - 1
    v, pv = (None, None) if valid is None else _c(np.asarray(valid, dtype=np.uint8))
    h = _load().sdr_encode_dict_fsst(pc, po, pv, n, mode, seg_rows, block, seed)
    return _take(h, _abi.DICT_FSST, 0, n)


# ---- generators -------------------------------------------------------------

def gen_u64(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.uint64)
    _load().sdr_gen_u64(seed, n, out.ctypes.data)
    return out


def gen_bp_int32(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    _load().sdr_gen_bp_int32(seed, n, out.ctypes.data)
    return out


def gen_runs_int64(seed: int, n: int, mean_run: float = 16.0, value_mod: int = 1 << 40) -> np.ndarray:
    out = np.empty(n, np.int64)
    _load().sdr_gen_runs_int64(seed, n, mean_run, value_mod, out.ctypes.data)
    return out


def gen_uniform_int32(seed: int, n: int, lo: int, rng: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    _load().sdr_gen_uniform_int32(seed, n, lo, rng, out.ctypes.data)
    return out


def gen_dates_sorted(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.int32)
    _load().sdr_gen_dates_sorted(seed, n, out.ctypes.data)
    return out


def gen_alp_doubles(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    _load().sdr_gen_alp_doubles(seed, n, out.ctypes.data)
    return out


def gen_unit_doubles(seed: int, n: int) -> np.ndarray:
    out = np.empty(n, np.float64)
    _load().sdr_gen_unit_doubles(seed, n, out.ctypes.data)
    return out


def gen_zipf_strings(seed: int, n: int, min_len: int = 10, max_len: int = 44, zipf_s: float = 1.1):
    """-> (chars uint8[total], offsets uint64[n+1])"""
    lib = _load()
    offs = np.empty(n + 1, np.uint64)
    total = lib.sdr_gen_zipf_strings(seed, n, min_len, max_len, zipf_s, offs.ctypes.data, None)
    chars = np.empty(max(1, total), np.uint8)
    lib.sdr_gen_zipf_strings(seed, n, min_len, max_len, zipf_s, offs.ctypes.data, chars.ctypes.data)
    return chars[:total], offs
