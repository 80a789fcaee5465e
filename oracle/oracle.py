"""ctypes wrapper of the CPU oracle (libsdr_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker.  The product library never
loads it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import sirius_amd._abi as A

LIB_PATH = Path(__file__).resolve().parent / "libsdr_oracle.so"
_lib = None


def load():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle not built: {LIB_PATH} (run make -C oracle)")
        lib = C.CDLL(str(LIB_PATH))
        vp, u32 = C.c_void_p, C.c_uint32
        lib.oracle_decode_segment.argtypes = [u32, vp, u32, u32, u32, vp]
        lib.oracle_decode_table.argtypes = [C.POINTER(A.Column), u32]
        lib.oracle_decode_table_mt.argtypes = [C.POINTER(A.Column), u32, C.c_int]
        lib.oracle_decode_strings_column.argtypes = [C.POINTER(A.StringColumn)]
        lib.oracle_decode_strings_column_mt.argtypes = [C.POINTER(A.StringColumn), C.c_int]
        lib.oracle_fsst_lengths.argtypes = [vp, u32, u32, vp]
        lib.oracle_fsst_decode.argtypes = [vp, u32, u32, vp, vp, C.c_uint64]
        lib.oracle_dict_fsst_decode.argtypes = [vp, u32, u32, vp, vp, C.c_uint64, vp]
        _lib = lib
    return _lib


def decode_segment(codec: int, seg: bytes | np.ndarray, rows: int, type_size: int,
                   bytes_size: int | None = None) -> np.ndarray:
    """Decode one segment held in host memory -> uint8 array of rows*type_size."""
    buf = np.frombuffer(bytes(seg), np.uint8) if not isinstance(seg, np.ndarray) else np.ascontiguousarray(seg)
    out = np.zeros(max(1, rows * type_size), np.uint8)
    n = len(buf) if bytes_size is None else bytes_size
    rc = load().oracle_decode_segment(codec, buf.ctypes.data if len(buf) else None, n, rows, type_size,
                                      out.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle rc {A.STATUS_NAMES.get(rc, rc)}")
    return out[: rows * type_size]


def decode_encoded(enc, threads: int = 1) -> np.ndarray:
    """Decode an Encoded column (sirius_amd.encode) on the CPU -> uint8 bytes."""
    out = np.zeros(max(1, enc.n_rows * enc.type_size), np.uint8)
    segs = enc.segments(enc.blob.ctypes.data)
    runs, keep = A.make_runs([(enc.codec, segs)])
    col = A.Column()
    col.type_size = enc.type_size
    col.total_rows = enc.n_rows
    col.data = runs
    col.n_data = 1
    col.d_out = out.ctypes.data
    rc = load().oracle_decode_table_mt(C.byref(col), 1, threads)
    if rc != 0:
        raise RuntimeError(f"oracle rc {A.STATUS_NAMES.get(rc, rc)}")
    return out[: enc.n_rows * enc.type_size]


def decode_strings(enc, threads: int = 1):
    """FSST Encoded column -> (offsets int32[n+1], chars uint8)."""
    n = enc.n_rows
    segs = enc.segments(enc.blob.ctypes.data)
    runs, keep = A.make_runs([(enc.codec, segs)])
    offs = np.zeros(n + 1, np.int32)
    cap = int(sum(int(s["row_count"]) * int(s["max_string_length"]) for s in enc.segs)) or 1
    chars = np.zeros(cap, np.uint8)
    col = A.StringColumn()
    col.total_rows = n
    col.data = runs
    col.n_data = 1
    col.d_offsets = offs.ctypes.data
    col.d_chars = chars.ctypes.data
    col.chars_capacity = cap
    rc = load().oracle_decode_strings_column_mt(C.byref(col), threads)
    if rc != 0:
        raise RuntimeError(f"oracle rc {A.STATUS_NAMES.get(rc, rc)}")
    return offs, chars[: col.chars_used]


def decode_strings_masked(enc, threads: int = 1):
    """String Encoded column, decoded with a null mask (no validity runs, so
    only DICT_FSST inline NULLs clear bits) -> (offsets int32[n+1], chars
    uint8, valid bool[n], null_count)."""
    n = enc.n_rows
    segs = enc.segments(enc.blob.ctypes.data)
    runs, keep = A.make_runs([(enc.codec, segs)])
    offs = np.zeros(n + 1, np.int32)
    cap = int(sum(int(s["row_count"]) * int(s["max_string_length"]) for s in enc.segs)) or 1
    chars = np.zeros(cap, np.uint8)
    mask = np.full((n + 7) // 8, 0xFF, np.uint8)  # left all valid when no NULL is found
    col = A.StringColumn()
    col.total_rows = n
    col.data = runs
    col.n_data = 1
    col.d_offsets = offs.ctypes.data
    col.d_chars = chars.ctypes.data
    col.chars_capacity = cap
    col.d_null_mask = mask.ctypes.data
    rc = load().oracle_decode_strings_column_mt(C.byref(col), threads)
    if rc != 0:
        raise RuntimeError(f"oracle rc {A.STATUS_NAMES.get(rc, rc)}")
    valid = np.unpackbits(mask, bitorder="little")[:n].astype(bool)
    return offs, chars[: col.chars_used], valid, int(col.null_count)


def dict_fsst_segment(seg: bytes | np.ndarray, rows: int, cap: int = 1 << 20):
    """One DICT_FSST segment -> (lengths uint32[rows], chars uint8, inline_null)."""
    b = np.frombuffer(bytes(seg), np.uint8) if not isinstance(seg, np.ndarray) else np.ascontiguousarray(seg)
    lengths = np.zeros(max(1, rows), np.uint32)
    chars = np.zeros(cap, np.uint8)
    flag = np.zeros(1, np.uint8)
    rc = load().oracle_dict_fsst_decode(b.ctypes.data, b.size, rows, lengths.ctypes.data, chars.ctypes.data, cap,
                                        flag.ctypes.data)
    if rc != 0:
        raise RuntimeError(f"oracle rc {A.STATUS_NAMES.get(rc, rc)}")
    return lengths[:rows], chars[: int(lengths[:rows].sum())], bool(flag[0])
