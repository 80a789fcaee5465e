"""Staging of DuckDB column segments from a database file into device
memory, then decode: the file reader, sdr_stage_file and the decode entry
points in one call.

The reference's scan path reads a split's segments into pinned memory
with coalesced reads and copies each segment to the device
(``src/op/scan/duckdb_native_decoder.cpp:675-855``); segments that are
not in the file (CONSTANT values, from statistics) are copied from host
memory (``:159-191``).  Here one host-to-device copy moves the packed
image (``sdr_stage_file``); CONSTANT values and all-NULL validity, which
have no bytes in the file, ride in a small host-built tail.
"""
from __future__ import annotations

import ctypes as C
import struct


from . import _abi as A
from . import lib as L
from .duckdb_file import BLOCK_HEADER, Column, DuckDBFile, FormatError

# Ranges closer than this share a read (the reference passes the block
# header size: segments split only by a block's checksum coalesce).
DEFAULT_MAX_GAP = BLOCK_HEADER

_PACK = {1: "<b", 2: "<h", 4: "<i", 8: "<q"}
_UPACK = {1: "<B", 2: "<H", 4: "<I", 8: "<Q"}


def plan(ranges, max_gap: int = DEFAULT_MAX_GAP):
    """[(file_offset, bytes)] -> (image bytes, [device_offset]) without I/O (sdr_stage_plan)."""
    ext = (A.Extent * max(1, len(ranges)))(*[A.Extent(int(o), int(n), 0) for o, n in ranges])
    total = L.load().sdr_stage_plan(ext, len(ranges), max_gap)
    return int(total), [int(ext[i].device_offset) for i in range(len(ranges))]


def stage_file(path: str, ranges, device, stream=None, max_gap: int = DEFAULT_MAX_GAP, extra: int = 0):
    """Stage byte ranges of `path` into a new device buffer (torch uint8 tensor with `extra`
    spare bytes after the image).  Returns (buffer, [device_offset])."""
    import torch

    lib = L.load()
    ext = (A.Extent * max(1, len(ranges)))(*[A.Extent(int(o), int(n), 0) for o, n in ranges])
    total = int(lib.sdr_stage_plan(ext, len(ranges), max_gap))
    buf = torch.empty(total + extra + 256, dtype=torch.uint8, device=device)
    sh = (stream or torch.cuda.current_stream(device)).cuda_stream
    L.check(lib.sdr_stage_file(str(path).encode(), ext, len(ranges), max_gap, C.c_void_p(buf.data_ptr()),
                               total, sh))
    return buf, [int(ext[i].device_offset) for i in range(len(ranges))]


def _constant_bytes(col: Column, seg) -> bytes:
    ts = col.type.type_size
    v = seg.stats.get("min")
    if v is None or seg.stats.get("max") != v:
        raise ValueError(f"{col.name}: CONSTANT segment at row {seg.row_start} without a single-valued statistic")
    return struct.pack(_PACK[ts] if col.type.signed else _UPACK[ts], v)


def scan_column(f: DuckDBFile, table: str, column: str, device, stream=None, max_gap: int = DEFAULT_MAX_GAP):
    """Decode one column of a table: stage its segments, then sdr_decode_table (fixed width)
    or sdr_decode_strings_column (VARCHAR).  Returns a dict with the device tensors:
    fixed width: values (uint8, rows * type_size), mask (or None), null_count;
    VARCHAR: offsets (int32, rows + 1), chars (uint8), mask (or None), null_count.
    With `stream`, every allocation, copy and launch runs on that stream."""
    import torch

    if stream is not None and stream != torch.cuda.current_stream(device):
        with torch.cuda.stream(stream):
            return scan_column(f, table, column, device, None, max_gap)

    t = f.tables[table]
    col = next(c for c in t.columns if c.name == column)
    n = t.rows
    ts = 0 if col.type.is_string else col.type.type_size
    # Validity decodes from UNCOMPRESSED bytes or a CONSTANT statistic only.
    # Any other codec (ROARING, which the reference decodes on the host with
    # DuckDB's own scan state, duckdb_native_decoder.cpp:194-243) is refused
    # rather than read as all-valid.
    for s in col.validity:
        if s.codec not in (A.UNCOMPRESSED, A.CONSTANT):
            raise FormatError(f"{column}: validity segment at row {s.row_start} uses {s.codec_name} "
                              f"(codec {s.codec}), which this decoder does not read")
    # bytes from the file, and host-built tails: CONSTANT values, all-NULL validity
    file_segs = [s for s in col.segments if s.codec != 2] + [s for s in col.validity if s.codec == 1]
    tail = bytearray()
    tail_at = {}
    for s in col.segments:
        if s.codec == 2:
            if col.type.is_string:
                raise ValueError(f"{column}: CONSTANT string segments are not supported")
            tail_at[id(s)] = len(tail)
            tail += _constant_bytes(col, s).ljust(16, b"\0")
    null_validity = [s for s in col.validity if s.codec == 2 and s.stats.get("has_null") and
                     not s.stats.get("has_no_null")]
    for s in null_validity:
        tail_at[id(s)] = len(tail)
        tail += bytes((s.tuple_count + 7) // 8 + 16 - ((s.tuple_count + 7) // 8) % 16)
    buf, offs = stage_file(f.path, [(f.payload_offset(s.block_id) + s.offset, s.bytes_size) for s in file_segs],
                           device, stream, max_gap, extra=len(tail))
    image = int(buf.numel()) - len(tail) - 256
    if tail:
        buf[image:image + len(tail)] = torch.frombuffer(bytearray(tail), dtype=torch.uint8).to(device)
    base = buf.data_ptr()
    where = {id(s): base + o for s, o in zip(file_segs, offs)}
    where.update({k: base + image + v for k, v in tail_at.items()})

    runs = []
    for s in col.segments:
        size = s.bytes_size if s.codec != 2 else 16
        seg = A.Segment(where[id(s)], size, s.row_start, s.tuple_count, s.stats.get("max_string_length", 0))
        if runs and runs[-1][0] == s.codec:
            runs[-1][1].append(seg)
        else:
            runs.append((s.codec, [seg]))
    vsegs = [A.Segment(where[id(s)], s.bytes_size if s.codec == 1 else (s.tuple_count + 7) // 8, s.row_start,
                       s.tuple_count, 0) for s in col.validity if id(s) in where]
    validity = [(A.UNCOMPRESSED, vsegs)] if vsegs else None
    mask = torch.zeros((n + 7) // 8, dtype=torch.uint8, device=device) if validity else None
    sh = (stream or torch.cuda.current_stream(device)).cuda_stream
    lib = L.load()
    rr, keep = A.make_runs(runs)
    vr, vkeep = A.make_runs(validity) if validity else (None, None)
    out = {"mask": mask, "buffer": buf}
    if col.type.is_string:
        cap = lib.sdr_strings_upper_bound(rr, len(runs))
        if cap == 2**64 - 1:
            cap = sum(int(s.bytes_size) for s in col.segments) * 8 + 1024
        offsets = torch.empty(n + 1, dtype=torch.int32, device=device)
        chars = torch.empty(max(1, int(cap)), dtype=torch.uint8, device=device)
        sc = A.StringColumn()
        sc.total_rows, sc.data, sc.n_data = n, rr, len(runs)
        sc.d_offsets, sc.d_chars, sc.chars_capacity = offsets.data_ptr(), chars.data_ptr(), int(cap)
        if validity:
            sc.has_nulls, sc.validity, sc.n_validity, sc.d_null_mask = 1, vr, len(validity), mask.data_ptr()
        rc = lib.sdr_decode_strings_column(C.byref(sc), sh)
        if rc == A.SDR_E_OVERFLOW and sc.chars_used <= 0x7FFFFFFF:  # the estimate was short: exact size
            chars = torch.empty(max(1, int(sc.chars_used)), dtype=torch.uint8, device=device)
            sc.d_chars, sc.chars_capacity = chars.data_ptr(), int(sc.chars_used)
            rc = lib.sdr_decode_strings_column(C.byref(sc), sh)
        L.check(rc)
        out.update(offsets=offsets, chars=chars[:sc.chars_used], null_count=sc.null_count)
    else:
        values = torch.empty(max(1, n * ts), dtype=torch.uint8, device=device)
        fc = A.Column()
        fc.type_size, fc.total_rows, fc.data, fc.n_data, fc.d_out = ts, n, rr, len(runs), values.data_ptr()
        if validity:
            fc.has_nulls, fc.validity, fc.n_validity, fc.d_null_mask = 1, vr, len(validity), mask.data_ptr()
        arr = (A.Column * 1)(fc)
        L.check(lib.sdr_decode_table(arr, 1, sh))
        out.update(values=values[:n * ts], null_count=arr[0].null_count)
    del keep, vkeep
    return out


def scan_table(f: DuckDBFile, table: str, columns=None, device="cuda:0", stream=None, rowid: bool = False,
               max_gap: int = DEFAULT_MAX_GAP):
    """Decode columns of a table (all by default): name -> scan_column's output.  With
    `rowid`, "rowid" holds each row's int64 row id, one range per row group (sdr_rowids).
    The outputs are flat device buffers (torch tensors, exportable without a copy through
    torch.utils.dlpack)."""
    import torch

    if stream is not None and stream != torch.cuda.current_stream(device):
        with torch.cuda.stream(stream):
            return scan_table(f, table, columns, device, None, rowid, max_gap)

    t = f.tables[table]
    names = list(columns) if columns is not None else [c.name for c in t.columns]
    out = {n: scan_column(f, table, n, device, stream, max_gap) for n in names}
    if rowid:
        groups = [(start, count) for start, count in t.row_groups if count]
        rr = (A.RowRange * max(1, len(groups)))()
        pos = 0
        for i, (start, count) in enumerate(groups):
            rr[i] = A.RowRange(start, pos, count, 0)
            pos += count
        ids = torch.empty(max(1, pos), dtype=torch.int64, device=device)
        sh = (stream or torch.cuda.current_stream(device)).cuda_stream
        L.check(L.load().sdr_rowids(rr, len(groups), C.c_void_p(ids.data_ptr()), sh))
        out["rowid"] = ids[:pos]
    return out
