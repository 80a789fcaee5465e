"""Read-only reader of DuckDB database files: tables, columns and the raw bytes
of every column segment, without DuckDB.

The reference reaches the same facts through DuckDB's own catalog and
segment trees (``src/op/scan/duckdb_native_metadata.cpp:519,660``;
``docs/super-sirius/scan.md:301-326``) and turns a block id into a file
offset with ``duckdb_block_payload_offset``
(``src/include/op/scan/duckdb_block_layout.hpp:25-63``).  This module parses
the file directly, so the decoders can be fed segments written by DuckDB's
real encoders.

File layout (single-file block manager):

* ``[0, 4096)`` main header: checksum u64, magic ``DUCK``, storage version
  u64, flags, library version strings;
* ``[4096, 8192)`` and ``[8192, 12288)`` two database headers (checksum,
  then iteration, meta_block, free_list, block_count, block_alloc_size,
  vector_size, serialization_compatibility); the one with the higher
  iteration is current;
* blocks from byte 12288, ``block_alloc_size`` bytes each: an 8-byte
  checksum, then the payload.

Metadata (catalog, table data, row groups, column data) lives in metadata
blocks: a block's payload is cut into 64 pieces of ``floor8(payload/64)``
bytes, each starting with the u64 pointer of the next piece (block id in
bits 0-55, piece index in bits 56-63; all ones ends the chain).  A
metadata pointer adds a byte offset inside its piece.

Serialized objects use DuckDB's binary format: a u16 field id, then the
value; integers are LEB128 varints (signed ones sign-extended), strings and
blobs are a varint length plus bytes, lists a varint count plus elements,
nullable objects a presence byte, and every object ends with field 0xFFFF.

Walk: catalog entries (field 100 of the root) -> for a table entry, the
CreateTableInfo (name, columns and types) and the table pointer -> table
statistics (one per column, typed by the column) and an optional sample ->
u64 row-group count -> row groups (row_start, tuple_count, one column-data
pointer per column) -> column data: data pointers (row_start, tuple_count,
block pointer, compression, statistics) and the validity child.

A segment's byte size is not stored; like the reference's walker, it is
bounded by the next segment start in the same block (over all segments of
the file), or by the block's end.  Decoders bound their reads by the
segment headers, so the slack is harmless.
"""
from __future__ import annotations

import mmap
import struct
from dataclasses import dataclass, field

MAIN_HEADER = 4096
BLOCK_START = 3 * MAIN_HEADER
BLOCK_HEADER = 8  # checksum
METADATA_PIECES = 64
END_OF_CHAIN = 0xFFFFFFFFFFFFFFFF
TERMINATOR = 0xFFFF

# DuckDB's compression enum; the values match sdr_codec.
CODEC_NAMES = {1: "Uncompressed", 2: "Constant", 3: "RLE", 4: "Dictionary", 5: "PFOR", 6: "BitPacking",
               7: "FSST", 8: "Chimp", 9: "Patas", 10: "ALP", 11: "ALPRD", 12: "ZSTD", 13: "Roaring",
               14: "Empty", 15: "DICT_FSST"}

# LogicalTypeId -> (name, physical bytes, signed); DECIMAL's width comes from its precision.
TYPES = {
    10: ("BOOLEAN", 1, False), 11: ("TINYINT", 1, True), 12: ("SMALLINT", 2, True), 13: ("INTEGER", 4, True),
    14: ("BIGINT", 8, True), 15: ("DATE", 4, True), 16: ("TIME", 8, True), 17: ("TIMESTAMP_S", 8, True),
    18: ("TIMESTAMP_MS", 8, True), 19: ("TIMESTAMP", 8, True), 20: ("TIMESTAMP_NS", 8, True),
    21: ("DECIMAL", 0, True), 22: ("FLOAT", 4, None), 23: ("DOUBLE", 8, None), 25: ("VARCHAR", 0, None),
    28: ("UTINYINT", 1, False), 29: ("USMALLINT", 2, False), 30: ("UINTEGER", 4, False),
    31: ("UBIGINT", 8, False), 32: ("TIMESTAMP_TZ", 8, True),
}


class FormatError(ValueError):
    """The file is not a DuckDB database this reader understands."""


@dataclass
class ColumnType:
    type_id: int
    width: int = 0  # DECIMAL precision
    scale: int = 0

    @property
    def name(self) -> str:
        n = TYPES.get(self.type_id, (f"type{self.type_id}",))[0]
        return f"DECIMAL({self.width},{self.scale})" if self.type_id == 21 else n

    @property
    def is_string(self) -> bool:
        return self.type_id == 25

    @property
    def type_size(self) -> int:
        """Bytes per stored value (0 for VARCHAR)."""
        if self.type_id == 21:
            return 2 if self.width <= 4 else 4 if self.width <= 9 else 8 if self.width <= 18 else 16
        if self.type_id not in TYPES:
            raise FormatError(f"unsupported column type id {self.type_id}")
        return TYPES[self.type_id][1]

    @property
    def signed(self):
        return True if self.type_id == 21 else TYPES.get(self.type_id, (0, 0, None))[2]


@dataclass
class Segment:
    """One column segment: rows [row_start, row_start + tuple_count) of its row group's column."""
    row_start: int
    tuple_count: int
    codec: int
    block_id: int
    offset: int
    stats: dict
    bytes_size: int = 0

    @property
    def codec_name(self) -> str:
        return CODEC_NAMES.get(self.codec, str(self.codec))


@dataclass
class Column:
    name: str
    type: ColumnType
    segments: list = field(default_factory=list)  # data segments, row order
    validity: list = field(default_factory=list)  # validity segments, row order


@dataclass
class Table:
    name: str
    rows: int
    columns: list
    row_groups: list  # (row_start, tuple_count)


class _Reader:
    """A byte stream over a chain of metadata pieces."""

    def __init__(self, f: "DuckDBFile", pointer: int, offset: int):
        self.f = f
        # no object spans more than all metadata pieces could hold: a cyclic
        # chain or a corrupt length ends in FormatError, not a loop
        self.budget = f.block_count * METADATA_PIECES * f.piece_size
        self._jump(pointer, offset)

    def _jump(self, pointer: int, offset: int):
        block, piece = pointer & ((1 << 56) - 1), pointer >> 56
        if piece >= METADATA_PIECES or block >= self.f.block_count:
            raise FormatError(f"metadata pointer out of range (block {block}, piece {piece})")
        self.base = self.f.payload_offset(block) + piece * self.f.piece_size
        self.next = struct.unpack_from("<Q", self.f.buf, self.base)[0]
        self.pos = max(offset, 8)

    def read(self, n: int) -> bytes:
        self.budget -= n
        if self.budget < 0:
            raise FormatError("metadata object longer than the file's metadata")
        out = bytearray()
        while n:
            if self.pos >= self.f.piece_size:
                if self.next == END_OF_CHAIN:
                    raise FormatError("metadata chain ended inside an object")
                self._jump(self.next, 0)
            k = min(n, self.f.piece_size - self.pos)
            out += self.f.buf[self.base + self.pos:self.base + self.pos + k]
            self.pos += k
            n -= k
        return bytes(out)

    def u8(self) -> int:
        return self.read(1)[0]

    def u16(self) -> int:
        return struct.unpack("<H", self.read(2))[0]

    def u64(self) -> int:
        return struct.unpack("<Q", self.read(8))[0]

    def uvar(self) -> int:
        r = s = 0
        while True:
            x = self.u8()
            r |= (x & 0x7F) << s
            s += 7
            if not x & 0x80:
                return r

    def svar(self) -> int:
        r = s = 0
        while True:
            x = self.u8()
            r |= (x & 0x7F) << s
            s += 7
            if not x & 0x80:
                return r - (1 << s) if x & 0x40 else r

    def blob(self) -> bytes:
        return self.read(self.uvar())


# ---- typed object parsers ---------------------------------------------------
# A schema maps field id -> parser(reader, ctx).  ctx carries the column type
# for statistics, whose value encoding depends on it.

def _obj(r: _Reader, kind: str, ctx=None) -> dict:
    out = {}
    schema = _SCHEMA[kind]
    while True:
        fid = r.u16()
        if fid == TERMINATOR:
            return out
        p = schema.get(fid)
        if p is None:
            raise FormatError(f"{kind}: unexpected field {fid}")
        out[fid] = p(r, ctx)


def _u(r, c): return r.uvar()
def _s(r, c): return r.svar()
def _b(r, c): return r.u8()
def _str(r, c): return r.blob()
def _dbl(r, c): return struct.unpack("<d", r.read(8))[0]
def _o(kind): return lambda r, c: _obj(r, kind, c)
def _n(kind): return lambda r, c: _obj(r, kind, c) if r.u8() else None
def _l(p): return lambda r, c: [p(r, c) for _ in range(r.uvar())]


def _stat_value(r: _Reader, ctx: ColumnType):
    t = ctx.type_id
    if t == 23:
        return _dbl(r, ctx)
    if t == 22:
        return struct.unpack("<f", r.read(4))[0]
    if t == 10:
        return r.u8()
    if ctx.type_size == 16:
        h = _obj(r, "hugeint")
        return (h.get(100, 0) << 64) | h.get(101, 0)
    return r.svar() if ctx.signed else r.uvar()


def _type_stats(r: _Reader, ctx: ColumnType):
    return _obj(r, "strstats" if ctx.is_string else "numstats", ctx)


_SCHEMA = {
    "root": {100: _l(_o("entry"))},
    "entry": {99: _u, 100: _n("createinfo"), 101: _o("mbp"), 102: _u, 103: _l(_o("mbp")), 104: _l(_o("indexinfo"))},
    "createinfo": {100: _u, 101: _str, 102: _str, 103: _b, 104: _b, 105: _u, 106: _str, 107: _o("value"),
                   108: _l(_o("kv")), 200: _str, 201: _o("columnlist")},
    "kv": {0: _str, 1: _str},
    "columnlist": {100: _l(_o("columndef"))},
    "columndef": {100: _str, 101: _o("type"), 103: _u, 104: _u, 105: _o("value"), 106: _l(_o("kv"))},
    "type": {100: _u, 101: _n("typeinfo")},
    "typeinfo": {100: _u, 101: _str, 200: _u, 201: _u},
    "value": {100: _o("type"), 101: _b},
    "mbp": {100: _u, 101: _u},
    "indexinfo": {100: _str, 101: _o("mbp")},
    "colstats": {100: _o("basestats"), 101: _n("distinct")},
    "basestats": {100: _b, 101: _b, 102: _u, 103: _type_stats},
    "numstats": {200: _o("numvalue"), 201: _o("numvalue")},
    "numvalue": {100: _b, 101: _stat_value},
    "hugeint": {100: _s, 101: _u},
    "strstats": {200: _str, 201: _str, 202: _b, 203: _b, 204: _u},
    "distinct": {100: _u, 101: _u, 102: _n("hll")},
    "hll": {100: _u, 101: _str},
    "sample": {100: _n("basesample"), 101: _u, 102: _b, 200: _u},
    "basesample": {100: _u, 101: _dbl, 102: _u, 103: _u, 104: _u, 105: _l(_o("weight"))},
    "weight": {0: _dbl, 1: _u},
    "rowgroup": {100: _u, 101: _u, 102: _l(_o("mbp")), 103: _l(_o("mbp"))},
    "columndata": {100: _l(_o("datapointer")), 101: _o("columndata")},
    "datapointer": {100: _u, 101: _u, 102: _o("blockpointer"), 103: _u, 104: _o("basestats"), 105: _n("segstate")},
    "blockpointer": {100: _s, 101: _u},
    "segstate": {100: _u, 101: _l(_s)},
}
_VALIDITY = ColumnType(10)  # validity statistics carry no type-specific part


def _stats_dict(bs: dict, ctx: ColumnType) -> dict:
    out = {"has_null": bool(bs.get(100, 0)), "has_no_null": bool(bs.get(101, 0))}
    ts = bs.get(103)
    if ts is not None and not ctx.is_string:
        mn, mx = ts.get(200, {}), ts.get(201, {})
        if mn.get(100):
            out["min"] = mn.get(101)
        if mx.get(100):
            out["max"] = mx.get(101)
    elif ts is not None:
        if ts.get(203):
            out["max_string_length"] = ts.get(204, 0)
    return out


class DuckDBFile:
    """A DuckDB database file opened read-only (memory-mapped)."""

    def __init__(self, path: str):
        self.path = path
        with open(path, "rb") as fh:
            self.buf = mmap.mmap(fh.fileno(), 0, access=mmap.ACCESS_READ)
        if len(self.buf) < BLOCK_START or self.buf[8:12] != b"DUCK":
            raise FormatError(f"{path}: not a DuckDB database file")
        self.storage_version = struct.unpack_from("<Q", self.buf, 12)[0]
        self.library_version = bytes(self.buf[52:84]).rstrip(b"\0").decode(errors="replace")
        heads = [struct.unpack_from("<qqqqqqq", self.buf, o + 8) for o in (MAIN_HEADER, 2 * MAIN_HEADER)]
        it, meta, _free, nblocks, alloc, vsize, _compat = max(heads, key=lambda h: h[0])
        if alloc <= BLOCK_HEADER or alloc % 4096 or nblocks < 0:
            raise FormatError(f"{path}: implausible block size {alloc}")
        self.block_alloc = alloc
        self.block_payload = alloc - BLOCK_HEADER
        self.block_count = nblocks
        self.vector_size = vsize
        self.piece_size = (self.block_payload // METADATA_PIECES) & ~7
        if BLOCK_START + nblocks * alloc > len(self.buf):
            raise FormatError(f"{path}: {nblocks} blocks do not fit in {len(self.buf)} bytes")
        self.tables: dict[str, Table] = {}
        if meta >= 0:
            try:
                self._walk(meta)
            except RecursionError:  # objects nested without end
                raise FormatError(f"{path}: metadata nests too deeply") from None
        self._bound_sizes()

    # byte offset of a block's payload (the reference's duckdb_block_payload_offset)
    def payload_offset(self, block_id: int) -> int:
        if block_id < 0 or block_id >= self.block_count:
            raise FormatError(f"block id {block_id} out of range")
        return BLOCK_START + block_id * self.block_alloc + BLOCK_HEADER

    def _walk(self, meta: int):
        root = _obj(_Reader(self, meta, 0), "root")
        for e in root.get(100, []):
            info = e.get(100)
            if e.get(99) != 1 or info is None:  # only table entries carry data
                continue
            cols = []
            for cd in info[201][100]:
                ty = cd[101]
                ti = ty.get(101) or {}
                cols.append(Column(cd[100].decode(), ColumnType(ty[100], ti.get(200, 0), ti.get(201, 0))))
            ptr = e[101]
            self.tables[info[200].decode()] = self._table(info[200].decode(), e.get(102, 0), cols, ptr)

    def _table(self, name: str, rows: int, cols: list, ptr: dict) -> Table:
        r = _Reader(self, ptr[100], ptr[101])
        # table statistics: field 100 is one (nullable) ColumnStatistics per column
        if r.u16() != 100 or r.uvar() != len(cols):
            raise FormatError(f"{name}: table statistics do not match the columns")
        for c in cols:
            if r.u8():
                _obj(r, "colstats", c.type if c.type.type_id in TYPES else _VALIDITY)
        fid = r.u16()
        if fid == 101:
            _n("sample")(r, None)
            fid = r.u16()
        if fid != TERMINATOR:
            raise FormatError(f"{name}: unexpected table-statistics field {fid}")
        row_groups = []
        for _ in range(r.u64()):
            rg = _obj(r, "rowgroup")
            start, count = rg.get(100, 0), rg.get(101, 0)
            row_groups.append((start, count))
            ptrs = rg.get(102, [])
            if len(ptrs) != len(cols):
                raise FormatError(f"{name}: row group holds {len(ptrs)} columns, table {len(cols)}")
            for c, p in zip(cols, ptrs):
                cd = _obj(_Reader(self, p[100], p[101]), "columndata", c.type)
                c.segments += self._segments(cd.get(100, []), c.type)
                c.validity += self._segments(cd.get(101, {}).get(100, []), _VALIDITY)
        for c in cols:
            c.segments.sort(key=lambda s: s.row_start)
            c.validity.sort(key=lambda s: s.row_start)
        return Table(name, rows, cols, row_groups)

    @staticmethod
    def _segments(dps: list, ty: ColumnType) -> list:
        out = []
        for d in dps:
            bp = d.get(102, {})
            out.append(Segment(d.get(100, 0), d.get(101, 0), d.get(103, 0), bp.get(100, -1), bp.get(101, 0),
                               _stats_dict(d.get(104, {}), ty)))
        return out

    def _bound_sizes(self):
        segs = [s for t in self.tables.values() for c in t.columns for s in c.segments + c.validity
                if s.block_id >= 0]
        by_block: dict[int, list] = {}
        for s in segs:
            by_block.setdefault(s.block_id, []).append(s)
        for blk, ss in by_block.items():
            offs = sorted({s.offset for s in ss})
            nxt = {o: (offs[i + 1] if i + 1 < len(offs) else self.block_payload) for i, o in enumerate(offs)}
            for s in ss:
                s.bytes_size = nxt[s.offset] - s.offset

    def segment_bytes(self, seg: Segment) -> bytes:
        """The segment's bytes (up to its size bound); b"" for segments with no block (CONSTANT)."""
        if seg.block_id < 0:
            return b""
        o = self.payload_offset(seg.block_id) + seg.offset
        return bytes(self.buf[o:o + seg.bytes_size])

    def close(self):
        self.buf.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
