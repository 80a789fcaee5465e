"""A writer of small DuckDB database files, for tests: the subset of the
format that sirius_amd.duckdb_file reads (main and database headers,
metadata pieces, the catalog, table statistics, row groups and column data
pointers), holding given segment bytes in data blocks.

The layout follows what the reader documents; segment bytes come from the
DuckDB-written fixtures, so a file written here carries real-encoder
segments behind synthetic metadata.
"""
from __future__ import annotations

import struct

BLOCK = 262144
PAYLOAD = BLOCK - 8
PIECE = (PAYLOAD // 64) & ~7
TERM = b"\xff\xff"


def uvar(v: int) -> bytes:
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if v:
            out.append(b | 0x80)
        else:
            out.append(b)
            return bytes(out)


def svar(v: int) -> bytes:
    out = bytearray()
    while True:
        b = v & 0x7F
        v >>= 7
        if (v == 0 and not b & 0x40) or (v == -1 and b & 0x40):
            out.append(b)
            return bytes(out)
        out.append(b | 0x80)


def fld(fid: int) -> bytes:
    return struct.pack("<H", fid)


def string(s: bytes | str) -> bytes:
    s = s.encode() if isinstance(s, str) else s
    return uvar(len(s)) + s


class Meta:
    """Metadata pieces of one metadata block (block 0): a byte stream with
    next-piece pointers; positions are (piece, offset) pointers."""

    def __init__(self, block: int = 0):
        self.block = block
        self.pieces = [bytearray(8)]

    def pointer(self):
        piece = len(self.pieces) - 1
        return self.block | (piece << 56), len(self.pieces[-1])

    def write(self, data: bytes):
        data = memoryview(data)
        while len(data):
            cur = self.pieces[-1]
            room = PIECE - len(cur)
            if room == 0:
                nxt = self.block | (len(self.pieces) << 56)
                cur[0:8] = struct.pack("<Q", nxt)
                self.pieces.append(bytearray(8))
                continue
            cur += data[:room]
            data = data[room:]

    def new_piece(self):
        """Continue in a fresh piece (a database header points at a piece's first byte)."""
        if len(self.pieces[-1]) > 8:
            self.pieces[-1][0:8] = struct.pack("<Q", self.block | (len(self.pieces) << 56))
            self.pieces.append(bytearray(8))
        return self.block | ((len(self.pieces) - 1) << 56)

    def payload(self) -> bytes:
        if len(self.pieces) > 64:
            raise ValueError("metadata does not fit one block")
        self.pieces[-1][0:8] = struct.pack("<Q", 0xFFFFFFFFFFFFFFFF)
        out = bytearray(PAYLOAD)
        for i, p in enumerate(self.pieces):
            out[i * PIECE:i * PIECE + len(p)] = p
        return bytes(out)


def _type(type_id: int, width: int = 0, scale: int = 0) -> bytes:
    b = fld(100) + uvar(type_id)
    if type_id == 21:
        b += fld(101) + b"\x01" + fld(100) + uvar(2) + fld(200) + uvar(width) + fld(201) + uvar(scale) + TERM
    return b + TERM


def _stats(col, seg) -> bytes:
    b = fld(100) + bytes([int(seg.get("has_null", False))]) + fld(101) + bytes([int(seg.get("has_no_null", True))])
    if col["type_id"] == 25:
        if "max_string_length" in seg:
            b += fld(103) + fld(203) + b"\x01" + fld(204) + uvar(seg["max_string_length"]) + TERM
    elif "min" in seg:
        v = lambda x: fld(100) + b"\x01" + fld(101) + svar(x) + TERM  # noqa: E731
        b += fld(103) + fld(200) + v(seg["min"]) + fld(201) + v(seg["max"]) + TERM
    return b + TERM


def _data_pointer(col, seg) -> bytes:
    b = fld(100) + uvar(seg["row_start"]) + fld(101) + uvar(seg["rows"])
    b += fld(102) + fld(100) + svar(seg.get("block", -1)) + fld(101) + uvar(seg.get("offset", 0)) + TERM
    b += fld(103) + uvar(seg["codec"]) + fld(104) + _stats(col, seg)
    return b + TERM


def write_database(path, tables, library_version: bytes = b"v1.4.3"):
    """tables: [{"name", "rows", "columns": [{"name", "type_id", "width", "scale",
    "segments": [{"row_start", "rows", "codec", "bytes" (absent for CONSTANT), "min"/"max" or
    "max_string_length", "has_null"}], "validity": [same, codec 1 with bytes or 2]}]}].
    Segment bytes go to data blocks from block 1 on, 8-byte aligned; returns the path."""
    data_blocks = [bytearray()]

    def place(b: bytes):
        cur = data_blocks[-1]
        off = (len(cur) + 7) & ~7
        if off + len(b) > PAYLOAD:
            data_blocks.append(bytearray())
            cur, off = data_blocks[-1], 0
        cur.extend(bytes(off - len(cur)))
        cur.extend(b)
        return len(data_blocks), off  # block ids from 1

    for t in tables:
        for c in t["columns"]:
            for s in c["segments"] + c.get("validity", []):
                if s.get("bytes") is not None:
                    s["block"], s["offset"] = place(s["bytes"])

    meta = Meta(0)
    # column data first, so the row groups can point at it
    col_ptrs = {}
    for ti, t in enumerate(tables):
        for ci, c in enumerate(t["columns"]):
            col_ptrs[ti, ci] = meta.pointer()
            b = fld(100) + uvar(len(c["segments"])) + b"".join(_data_pointer(c, s) for s in c["segments"])
            vcol = {"type_id": 10}
            vals = c.get("validity") or [{"row_start": 0, "rows": t["rows"], "codec": 2, "has_null": False}]
            b += fld(101) + fld(100) + uvar(len(vals)) + b"".join(_data_pointer(vcol, s) for s in vals) + TERM
            meta.write(b + TERM)
    table_ptrs = []
    for ti, t in enumerate(tables):
        table_ptrs.append(meta.pointer())
        b = fld(100) + uvar(len(t["columns"])) + b"\x00" * len(t["columns"]) + TERM  # no column statistics
        b += struct.pack("<Q", 1)  # one row group
        ptrs = b"".join(fld(100) + uvar(p) + fld(101) + uvar(o) + TERM
                        for p, o in (col_ptrs[ti, ci] for ci in range(len(t["columns"]))))
        b += fld(100) + uvar(0) + fld(101) + uvar(t["rows"]) + fld(102) + uvar(len(t["columns"])) + ptrs + TERM
        meta.write(b)
    root = meta.new_piece()
    entries = bytearray()
    for t, (p, o) in zip(tables, table_ptrs):
        cols = b"".join(fld(100) + string(c["name"]) + fld(101) + _type(c["type_id"], c.get("width", 0),
                                                                      c.get("scale", 0)) + TERM
                        for c in t["columns"])
        info = (fld(100) + uvar(1) + fld(102) + string("main") + fld(200) + string(t["name"]) +
                fld(201) + fld(100) + uvar(len(t["columns"])) + cols + TERM + TERM)
        entries += fld(99) + uvar(1) + fld(100) + b"\x01" + info
        entries += fld(101) + fld(100) + uvar(p) + fld(101) + uvar(o) + TERM + fld(102) + uvar(t["rows"]) + TERM
    meta.write(fld(100) + uvar(len(tables)) + bytes(entries) + TERM)

    n_blocks = 1 + len(data_blocks)
    head = bytearray(3 * 4096)
    head[8:12] = b"DUCK"
    head[12:20] = struct.pack("<Q", 64)
    head[52:52 + len(library_version)] = library_version
    struct.pack_into("<qQqqqqq", head, 4096 + 8, 1, root, -1, n_blocks, BLOCK, 2048, 1)
    struct.pack_into("<qqqqqqq", head, 8192 + 8, 0, -1, -1, 0, BLOCK, 2048, 1)
    with open(path, "wb") as fh:
        fh.write(head)
        fh.write(bytes(8) + meta.payload())
        for blk in data_blocks:
            fh.write(bytes(8) + bytes(blk) + bytes(PAYLOAD - len(blk)))
    return path
