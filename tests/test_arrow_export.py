"""sdr_arrow_export on the CPU: host buffers exported with device type CPU
are imported by pyarrow through the C Device Data Interface, so the schema
formats, buffer order, validity bitmaps, null counts and the release
protocol are checked against an independent Arrow implementation."""
import ctypes as C

import numpy as np
import pytest

import sirius_amd.arrow as X
from sirius_amd import lib as L
from sirius_amd.duckdb_file import ColumnType

pa = pytest.importorskip("pyarrow")


def _bits(valid):
    return np.packbits(np.asarray(valid, dtype=np.uint8), bitorder="little")


def _import(exp):
    arr = pa.Array._import_from_c_device(exp.array_ptr, exp.schema_ptr)
    assert not exp.array.array.release and not exp.schema.release  # moved into pyarrow
    return arr


def test_fixed_width_types_round_trip():
    rng = np.random.default_rng(5)
    n = 1000
    cases = [
        (X.INT8, np.int8, pa.int8()), (X.INT16, np.int16, pa.int16()), (X.INT32, np.int32, pa.int32()),
        (X.INT64, np.int64, pa.int64()), (X.UINT8, np.uint8, pa.uint8()), (X.UINT16, np.uint16, pa.uint16()),
        (X.UINT32, np.uint32, pa.uint32()), (X.UINT64, np.uint64, pa.uint64()),
        (X.FLOAT32, np.float32, pa.float32()), (X.FLOAT64, np.float64, pa.float64()),
        (X.DATE32, np.int32, pa.date32()), (X.TIMESTAMP_US, np.int64, pa.timestamp("us")),
        (X.TIMESTAMP_S, np.int64, pa.timestamp("s")), (X.TIMESTAMP_MS, np.int64, pa.timestamp("ms")),
        (X.TIMESTAMP_NS, np.int64, pa.timestamp("ns")), (X.TIMESTAMP_US_UTC, np.int64, pa.timestamp("us", "UTC")),
        (X.TIME_US, np.int64, pa.time64("us")),
    ]
    for t, dt, want_type in cases:
        vals = (rng.integers(0, 100, n) if np.dtype(dt).kind in "iu" else rng.random(n)).astype(dt)
        valid = rng.random(n) < 0.8
        bm = _bits(valid)
        exp = X.export(t, n, vals.ctypes.data, validity=bm.ctypes.data, null_count=int((~valid).sum()),
                       name="c", device_type=X.ARROW_DEVICE_CPU, keep=(vals, bm))
        arr = _import(exp)
        assert arr.type == want_type, t
        assert arr.null_count == int((~valid).sum())
        want = pa.Array.from_buffers(want_type, n, [pa.py_buffer(bm), pa.py_buffer(vals.copy())],
                                     int((~valid).sum()))
        assert arr.equals(want), t
        storage = pa.Array.from_buffers(pa.from_numpy_dtype(dt), n, arr.buffers()[:2], arr.null_count)
        got = storage.to_pylist()
        assert got == [v if ok else None for v, ok in zip(vals.tolist(), valid)], t
        del arr, storage  # both hold the imported buffers
    assert X.live_exports() == 0  # every release reached the owner


def test_utf8_with_nulls_round_trip():
    strings = ["alpha", "", None, "é中", "x" * 300, None]
    data = b"".join((s or "").encode() for s in strings)
    offs = np.zeros(len(strings) + 1, dtype=np.int32)
    offs[1:] = np.cumsum([len((s or "").encode()) for s in strings])
    chars = np.frombuffer(data, dtype=np.uint8).copy()
    valid = np.array([s is not None for s in strings])
    bm = _bits(valid)
    exp = X.export(X.UTF8, len(strings), offs.ctypes.data, validity=bm.ctypes.data, chars=chars.ctypes.data,
                   null_count=2, name="s", device_type=X.ARROW_DEVICE_CPU, keep=(offs, chars, bm))
    arr = _import(exp)
    assert arr.type == pa.utf8() and arr.to_pylist() == strings
    del arr
    assert X.live_exports() == 0


def test_decimal_widths():
    n = 64
    for vb, prec, scale in ((4, 9, 2), (8, 18, 3), (16, 30, 4)):
        raw = np.arange(n, dtype=np.int64) * 12345 - 99999
        buf = raw.astype(np.int32).view(np.uint8) if vb == 4 else raw.view(np.uint8) if vb == 8 else \
            np.stack([raw, np.where(raw < 0, -1, 0)], axis=1).astype(np.int64).reshape(-1).view(np.uint8)
        buf = np.ascontiguousarray(buf)
        exp = X.export(X.DECIMAL, n, buf.ctypes.data, null_count=0, value_bytes=vb, precision=prec, scale=scale,
                       device_type=X.ARROW_DEVICE_CPU, keep=buf)
        arr = _import(exp)
        assert arr.type.precision == prec and arr.type.scale == scale and arr.type.byte_width == vb
        assert [int(v.as_py().scaleb(scale)) for v in arr] == raw.tolist()
        del arr


def test_no_validity_and_unknown_null_count():
    vals = np.arange(10, dtype=np.int32)
    exp = X.export(X.INT32, 10, vals.ctypes.data, device_type=X.ARROW_DEVICE_CPU, keep=vals)
    assert exp.array.array.null_count == 0 and exp.array.array.buffers[0] is None
    arr = _import(exp)
    assert arr.to_pylist() == list(range(10))


def test_device_fields_and_release_without_consumer():
    vals = np.arange(4, dtype=np.int64)
    ev = C.c_void_p(0x1234)
    exp = X.export(X.INT64, 4, vals.ctypes.data, device_type=X.ARROW_DEVICE_ROCM, device_id=3,
                   sync_event=C.addressof(ev), name="k", keep=vals)
    a = exp.array
    assert a.device_type == X.ARROW_DEVICE_ROCM and a.device_id == 3 and a.sync_event == C.addressof(ev)
    assert exp.schema.format == b"l" and exp.schema.name == b"k" and exp.schema.flags == 2
    assert a.array.n_buffers == 2 and a.array.buffers[1] == vals.ctypes.data
    assert X.live_exports() == 1
    exp.release()
    assert not a.array.release and not exp.schema.release and X.live_exports() == 0
    exp.release()  # a second release is a no-op


@pytest.mark.parametrize("kw,code", [
    (dict(arrow_t=99), "SDR_E_INVALID_ARG"),
    (dict(arrow_t=X.DECIMAL, value_bytes=2, precision=4), "SDR_E_BAD_WIDTH"),
    (dict(arrow_t=X.DECIMAL, value_bytes=8, precision=5, scale=6), "SDR_E_INVALID_ARG"),
    (dict(arrow_t=X.INT32, null_count=3), "SDR_E_INVALID_ARG"),      # nulls without a bitmap
    (dict(arrow_t=X.UTF8), "SDR_E_INVALID_ARG"),                      # no chars
    (dict(arrow_t=X.INT32, length=-1), "SDR_E_INVALID_ARG"),
    (dict(arrow_t=X.INT32, device_type=2), "SDR_E_INVALID_ARG"),
])
def test_errors(kw, code):
    vals = np.zeros(8, dtype=np.int64)
    args = dict(arrow_t=X.INT32, length=8, device_type=X.ARROW_DEVICE_CPU)
    args.update(kw)
    t, n = args.pop("arrow_t"), args.pop("length")
    with pytest.raises(L.SdrError, match=code):
        X.export(t, n, vals.ctypes.data, keep=vals, **args)
    assert X.live_exports() == 0


def test_duckdb_type_map():
    assert X.arrow_type(ColumnType(13)) == (X.INT32, 0, 0, 0)
    assert X.arrow_type(ColumnType(25))[0] == X.UTF8
    assert X.arrow_type(ColumnType(21, 15, 2)) == (X.DECIMAL, 8, 15, 2)
    assert X.arrow_type(ColumnType(21, 4, 1))[1] == 2  # exported only if Arrow had a 16-bit decimal: it fails
    with pytest.raises(ValueError):
        X.arrow_type(ColumnType(99))
