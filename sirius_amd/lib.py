"""Loader for the HIP decode library (libsdr.so, built in-tree).

The product path has no fallback: if the library is missing or cannot be
loaded, every entry raises.  Build it with ``python -m sirius_amd.build``.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

from . import _abi

LIB_PATH = Path(__file__).resolve().parent / "libsdr.so"

# every symbol declared in include/sdr/decode.h (checked by the tests)
DECLARED = [
    "sdr_last_error", "sdr_codec_from_name", "sdr_codec_name", "sdr_abi_version", "sdr_build_target",
    "sdr_decode_table", "sdr_decode_table_mode",
    "sdr_decode_bitpacking_data", "sdr_decode_rle_data", "sdr_decode_alp_data", "sdr_decode_alprd_data",
    "sdr_strings_upper_bound", "sdr_decode_strings_column", "sdr_decode_strings_column_mode",
    "sdr_decode_strings_column_noalloc",
    "sdr_plan_create", "sdr_plan_launch", "sdr_plan_launch_codec", "sdr_plan_num_launches",
    "sdr_plan_null_counts", "sdr_plan_destroy",
    "sdr_batch_create", "sdr_batch_set_option", "sdr_batch_add_table", "sdr_batch_add_strings",
    "sdr_batch_num_entries", "sdr_batch_launch", "sdr_batch_finish", "sdr_batch_clear", "sdr_batch_destroy",
    "sdr_tune_set", "sdr_tune_get", "sdr_tune_load", "sdr_tune_set_width", "sdr_tune_get_width",
    "sdr_server_create", "sdr_server_stage", "sdr_server_unstage", "sdr_server_submit", "sdr_server_progress",
    "sdr_server_poll", "sdr_server_release", "sdr_server_pending", "sdr_server_held_bytes", "sdr_server_now_ms",
    "sdr_server_destroy", "sdr_server_get_stats", "sdr_server_stage_wait",
    "sdr_stage_plan", "sdr_stage_file", "sdr_rowids",
    "sdr_arrow_export",
]

_lib = None


class SdrError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_abi.STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"HIP decode library not built: {LIB_PATH} (run python -m sirius_amd.build)")
    # One HIP runtime per process: torch ships its own, and when torch loads
    # after this library the process ends up with two, the second of which
    # finds no device.  Loading torch first makes the library bind to it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(str(LIB_PATH))
    vp, u32, i32 = C.c_void_p, C.c_uint32, C.c_int
    P = C.POINTER
    lib.sdr_last_error.restype = C.c_char_p
    lib.sdr_codec_from_name.argtypes = [C.c_char_p]
    lib.sdr_codec_from_name.restype = i32
    lib.sdr_codec_name.argtypes = [u32]
    lib.sdr_codec_name.restype = C.c_char_p
    lib.sdr_abi_version.restype = i32
    lib.sdr_build_target.restype = C.c_char_p
    lib.sdr_decode_table.argtypes = [P(_abi.Column), u32, vp]
    lib.sdr_decode_table.restype = i32
    lib.sdr_decode_table_mode.argtypes = [P(_abi.Column), u32, vp, u32]
    lib.sdr_decode_table_mode.restype = i32
    for name in ("sdr_decode_bitpacking_data", "sdr_decode_rle_data", "sdr_decode_alp_data",
                 "sdr_decode_alprd_data"):
        f = getattr(lib, name)
        f.argtypes = [P(_abi.Run), vp, u32, vp]
        f.restype = i32
    lib.sdr_strings_upper_bound.argtypes = [P(_abi.Run), u32]
    lib.sdr_strings_upper_bound.restype = C.c_uint64
    lib.sdr_decode_strings_column.argtypes = [P(_abi.StringColumn), vp]
    lib.sdr_decode_strings_column.restype = i32
    lib.sdr_decode_strings_column_mode.argtypes = [P(_abi.StringColumn), vp, u32]
    lib.sdr_decode_strings_column_mode.restype = i32
    lib.sdr_decode_strings_column_noalloc.argtypes = [P(_abi.StringColumn), vp]
    lib.sdr_decode_strings_column_noalloc.restype = i32
    lib.sdr_plan_create.argtypes = [P(_abi.Column), u32, u32, vp, P(vp)]
    lib.sdr_plan_create.restype = i32
    lib.sdr_plan_launch.argtypes = [vp, vp]
    lib.sdr_plan_launch.restype = i32
    lib.sdr_plan_launch_codec.argtypes = [vp, u32, vp]
    lib.sdr_plan_launch_codec.restype = i32
    lib.sdr_plan_num_launches.argtypes = [vp]
    lib.sdr_plan_num_launches.restype = i32
    lib.sdr_plan_null_counts.argtypes = [vp, P(C.c_int64), u32, vp]
    lib.sdr_plan_null_counts.restype = i32
    lib.sdr_plan_destroy.argtypes = [vp]
    lib.sdr_plan_destroy.restype = i32
    lib.sdr_server_create.argtypes = [P(_abi.ServerConfig), P(vp)]
    lib.sdr_server_create.restype = i32
    lib.sdr_server_stage.argtypes = [vp, vp, C.c_uint64, P(vp)]
    lib.sdr_server_stage.restype = i32
    lib.sdr_server_unstage.argtypes = [vp, vp]
    lib.sdr_server_unstage.restype = i32
    lib.sdr_server_submit.argtypes = [vp, P(_abi.Request)]
    lib.sdr_server_submit.restype = i32
    lib.sdr_server_progress.argtypes = [vp, P(u32)]
    lib.sdr_server_progress.restype = i32
    lib.sdr_server_poll.argtypes = [vp, P(_abi.Completion), u32, P(u32)]
    lib.sdr_server_poll.restype = i32
    lib.sdr_server_release.argtypes = [vp, C.c_uint64, vp]
    lib.sdr_server_release.restype = i32
    lib.sdr_server_pending.argtypes = [vp]
    lib.sdr_server_pending.restype = i32
    lib.sdr_server_held_bytes.argtypes = [vp]
    lib.sdr_server_held_bytes.restype = C.c_uint64
    lib.sdr_server_now_ms.argtypes = [vp]
    lib.sdr_server_now_ms.restype = C.c_double
    lib.sdr_server_stage_wait.argtypes = [vp, vp]
    lib.sdr_server_stage_wait.restype = i32
    lib.sdr_server_get_stats.argtypes = [vp, P(_abi.ServerStats)]
    lib.sdr_server_get_stats.restype = i32
    lib.sdr_server_destroy.argtypes = [vp]
    lib.sdr_server_destroy.restype = i32
    lib.sdr_tune_set.argtypes = [u32, C.c_uint64]
    lib.sdr_tune_set.restype = i32
    lib.sdr_tune_get.argtypes = [u32]
    lib.sdr_tune_get.restype = C.c_uint64
    lib.sdr_tune_load.argtypes = [C.c_char_p]
    lib.sdr_tune_load.restype = i32
    lib.sdr_tune_set_width.argtypes = [u32, u32, C.c_uint64]
    lib.sdr_tune_set_width.restype = i32
    lib.sdr_tune_get_width.argtypes = [u32, u32]
    lib.sdr_tune_get_width.restype = C.c_uint64
    lib.sdr_arrow_export.argtypes = [vp, vp, vp]
    lib.sdr_arrow_export.restype = i32
    lib.sdr_stage_plan.argtypes = [P(_abi.Extent), u32, C.c_uint64]
    lib.sdr_stage_plan.restype = C.c_uint64
    lib.sdr_stage_file.argtypes = [C.c_char_p, P(_abi.Extent), u32, C.c_uint64, vp, C.c_uint64, vp]
    lib.sdr_stage_file.restype = i32
    lib.sdr_rowids.argtypes = [P(_abi.RowRange), u32, vp, vp]
    lib.sdr_rowids.restype = i32
    lib.sdr_batch_create.argtypes = [P(vp)]
    lib.sdr_batch_set_option.argtypes = [vp, u32, C.c_int64]
    lib.sdr_batch_add_table.argtypes = [vp, P(_abi.Column), u32]
    lib.sdr_batch_add_strings.argtypes = [vp, P(_abi.StringColumn)]
    lib.sdr_batch_num_entries.argtypes = [vp]
    lib.sdr_batch_launch.argtypes = [vp, vp]
    lib.sdr_batch_finish.argtypes = [vp, vp]
    lib.sdr_batch_clear.argtypes = [vp]
    lib.sdr_batch_destroy.argtypes = [vp]
    for f in ("sdr_batch_create", "sdr_batch_set_option", "sdr_batch_add_table", "sdr_batch_add_strings",
              "sdr_batch_num_entries", "sdr_batch_launch", "sdr_batch_finish", "sdr_batch_clear",
              "sdr_batch_destroy"):
        getattr(lib, f).restype = i32
    _lib = lib
    return lib


def check(code: int) -> None:
    if code != _abi.SDR_OK:
        raise SdrError(code, load().sdr_last_error().decode())
