"""Decoded columns as Arrow C device arrays (sdr_arrow_export).

The reference hands decoded buffers downstream as cuDF columns
(``gpu_native_decode.cu:546-608``, ``gpu_decode_strings.cu:100-402``,
``duckdb_native_gpu_ingestible.cpp:345-388``).  Here the hand-off is the
Arrow C Device Data Interface: the decode's own device buffers (validity
bitmap, values or int32 offsets, chars) described by an ArrowSchema and an
ArrowDeviceArray on ROCm memory, with nothing copied.  The buffers stay alive
until the consumer calls the array's release callback.
"""
from __future__ import annotations

import ctypes as C
import itertools
import threading

from . import lib as L
from .duckdb_file import ColumnType

ARROW_DEVICE_CPU = 1
ARROW_DEVICE_ROCM = 10

(INT8, INT16, INT32, INT64, UINT8, UINT16, UINT32, UINT64, FLOAT32, FLOAT64, DATE32, TIMESTAMP_US, TIME_US,
 DECIMAL, UTF8, TIMESTAMP_S, TIMESTAMP_MS, TIMESTAMP_NS, TIMESTAMP_US_UTC) = range(1, 20)


class ArrowSchema(C.Structure):
    _fields_ = [("format", C.c_char_p), ("name", C.c_char_p), ("metadata", C.c_char_p), ("flags", C.c_int64),
                ("n_children", C.c_int64), ("children", C.c_void_p), ("dictionary", C.c_void_p),
                ("release", C.c_void_p), ("private_data", C.c_void_p)]


class ArrowArray(C.Structure):
    _fields_ = [("length", C.c_int64), ("null_count", C.c_int64), ("offset", C.c_int64), ("n_buffers", C.c_int64),
                ("n_children", C.c_int64), ("buffers", C.POINTER(C.c_void_p)), ("children", C.c_void_p),
                ("dictionary", C.c_void_p), ("release", C.c_void_p), ("private_data", C.c_void_p)]


class ArrowDeviceArray(C.Structure):
    _fields_ = [("array", ArrowArray), ("device_id", C.c_int64), ("device_type", C.c_int32),
                ("sync_event", C.c_void_p), ("reserved", C.c_int64 * 3)]


RELEASE_FN = C.CFUNCTYPE(None, C.c_void_p)


class ArrowColumnDesc(C.Structure):
    """sdr_arrow_column (include/sdr/arrow.h)."""
    _fields_ = [("name", C.c_char_p), ("type", C.c_uint32), ("value_bytes", C.c_uint32), ("precision", C.c_uint32),
                ("scale", C.c_uint32), ("length", C.c_int64), ("null_count", C.c_int64), ("validity", C.c_void_p),
                ("values", C.c_void_p), ("chars", C.c_void_p), ("device_type", C.c_int32), ("reserved", C.c_int32),
                ("device_id", C.c_int64), ("sync_event", C.c_void_p), ("release", RELEASE_FN),
                ("owner", C.c_void_p)]


# Objects that keep exported buffers alive, by owner id, until the consumer's release.
_owners: dict = {}
_owners_mu = threading.Lock()
_ids = itertools.count(1)


@RELEASE_FN
def _on_release(owner):
    with _owners_mu:
        _owners.pop(owner, None)


def live_exports() -> int:
    """Exported arrays not yet released by their consumer."""
    with _owners_mu:
        return len(_owners)


# DuckDB logical type id -> Arrow type (BOOLEAN is one byte per value in DuckDB: exported as uint8)
_BY_TYPE_ID = {10: UINT8, 11: INT8, 12: INT16, 13: INT32, 14: INT64, 15: DATE32, 16: TIME_US, 17: TIMESTAMP_S,
               18: TIMESTAMP_MS, 19: TIMESTAMP_US, 20: TIMESTAMP_NS, 22: FLOAT32, 23: FLOAT64, 25: UTF8,
               28: UINT8, 29: UINT16, 30: UINT32, 31: UINT64, 32: TIMESTAMP_US_UTC}


def arrow_type(t: ColumnType):
    """(sdr_arrow_type, value_bytes, precision, scale) of a DuckDB column type."""
    if t.type_id == 21:
        return DECIMAL, t.type_size, t.width, t.scale
    if t.type_id not in _BY_TYPE_ID:
        raise ValueError(f"no Arrow export for {t.name}")
    return _BY_TYPE_ID[t.type_id], 0, 0, 0


class Exported:
    """One exported column: `schema` and `array` (ArrowSchema, ArrowDeviceArray), whose
    addresses a consumer imports (moving them, per the C Data Interface).  If no consumer
    took them, release() frees them."""

    def __init__(self, schema: ArrowSchema, array: ArrowDeviceArray):
        self.schema, self.array = schema, array

    @property
    def schema_ptr(self) -> int:
        return C.addressof(self.schema)

    @property
    def array_ptr(self) -> int:
        return C.addressof(self.array)

    def release(self):
        for s in (self.array.array, self.schema):
            if s.release:
                C.CFUNCTYPE(None, C.c_void_p)(s.release)(C.addressof(s))


def export(arrow_t: int, length: int, values: int, *, validity: int = 0, chars: int = 0, null_count: int = -1,
           name: str | None = None, value_bytes: int = 0, precision: int = 0, scale: int = 0,
           device_type: int = ARROW_DEVICE_ROCM, device_id: int = 0, sync_event: int = 0, keep=None) -> Exported:
    """sdr_arrow_export over raw buffer addresses; `keep` stays referenced until the
    consumer releases the array."""
    with _owners_mu:
        owner = next(_ids)
        _owners[owner] = keep
    d = ArrowColumnDesc(name.encode() if name is not None else None, arrow_t, value_bytes, precision, scale,
                        length, null_count, validity or None, values or None, chars or None, device_type, 0,
                        device_id, sync_event or None, _on_release, owner)
    schema, array = ArrowSchema(), ArrowDeviceArray()
    rc = L.load().sdr_arrow_export(C.byref(d), C.byref(schema), C.byref(array))
    if rc:
        with _owners_mu:
            _owners.pop(owner, None)
        L.check(rc)
    return Exported(schema, array)


def export_scan(out: dict, col_type: ColumnType, name: str | None = None, stream=None) -> Exported:
    """Export one scan_column output (device tensors) without a copy.  An event recorded
    on `stream` (the decode's) after the decode is the array's sync_event."""
    import torch

    t, vb, prec, scale = arrow_type(col_type)
    dev = (out["offsets"] if t == UTF8 else out["values"]).device
    st = stream or torch.cuda.current_stream(dev)
    ev = torch.cuda.Event()
    ev.record(st)
    handle = C.c_void_p(ev.cuda_event)  # sync_event points at the hipEvent_t
    mask = out.get("mask")
    keep = (out, ev, handle)
    if t == UTF8:
        n = int(out["offsets"].numel()) - 1
        return export(t, n, out["offsets"].data_ptr(), validity=mask.data_ptr() if mask is not None else 0,
                      chars=out["chars"].data_ptr() if out["chars"].numel() else out["offsets"].data_ptr(),
                      null_count=int(out["null_count"]), name=name, device_id=dev.index or 0,
                      sync_event=C.addressof(handle), keep=keep)
    n = int(out["values"].numel()) // col_type.type_size
    return export(t, n, out["values"].data_ptr(), validity=mask.data_ptr() if mask is not None else 0,
                  null_count=int(out["null_count"]), name=name, value_bytes=vb, precision=prec, scale=scale,
                  device_id=dev.index or 0, sync_event=C.addressof(handle), keep=keep)
