"""Reader of qikkDB's on-disk database files: the .db catalogue (JSON) and one .data file per
column (Database.cpp:1258-1727 LoadDatabasesFromDisk, :1729-3000 LoadColumn).

Host code, for feeding the GPU path with a database's column blocks; nothing is written.

A numeric .data file is a sequence of blocks, each:
    uint32 block index, int32 group id, uint32 null-mask bytes, the null mask (LSB-first bits,
    bit set = NULL), uint64 entry count, bool compressed, statistics (min and max of the column
    type, float avg, sum of the column type), then block_size entries of the column type, of
    which the first `entry count` are data.
A POINT column has no statistics; its entries are (float latitude, float longitude) pairs.
String and polygon columns keep their entries in separate fragment files (.frag); they are not
read here (the reader raises for them).
"""
from __future__ import annotations

import json
import os
import struct
from dataclasses import dataclass, field

import numpy as np

INT, LONG, FLOAT, DOUBLE, POINT, POLYGON, STRING, INT8 = 8, 9, 10, 11, 12, 13, 14, 15
_NP = {INT: np.int32, LONG: np.int64, FLOAT: np.float32, DOUBLE: np.float64, INT8: np.int8}


@dataclass
class Block:
    index: int
    group_id: int
    data: np.ndarray          # entries (POINT: float32 array of shape (n, 2), latitude and longitude)
    nulls: np.ndarray         # bool per entry
    stats: tuple | None       # (min, max, avg, sum) as stored


@dataclass
class Column:
    name: str
    type: int
    nullable: bool
    blocks: list = field(default_factory=list)

    def values(self) -> np.ndarray:
        """Every block's entries, in block order."""
        parts = [b.data for b in self.blocks]
        if not parts:
            return np.empty(0, _NP.get(self.type, np.float32))
        return np.concatenate(parts)

    def null_flags(self) -> np.ndarray:
        parts = [b.nulls for b in self.blocks]
        return np.concatenate(parts) if parts else np.empty(0, bool)


@dataclass
class Table:
    name: str
    block_size: int
    columns: dict = field(default_factory=dict)


def _unpack_nulls(raw: bytes, n: int) -> np.ndarray:
    if not raw:
        return np.zeros(n, bool)
    bits = np.unpackbits(np.frombuffer(raw, np.uint8), bitorder="little")
    out = np.zeros(n, bool)
    m = min(n, bits.size)
    out[:m] = bits[:m].astype(bool)
    return out


def read_column(path: str, ctype: int, block_size: int) -> list:
    """The blocks of one .data file."""
    if ctype in (STRING, POLYGON):
        raise NotImplementedError("string and polygon columns are stored in fragment files, which this reader "
                                  "does not read")
    if ctype not in _NP and ctype != POINT:
        raise ValueError(f"unknown column type {ctype}")
    with open(path, "rb") as f:
        buf = f.read()
    blocks = []
    pos = 0
    while pos + 12 <= len(buf):
        index, group_id, nlen = struct.unpack_from("<IiI", buf, pos)
        pos += 12
        raw_nulls = buf[pos:pos + nlen]
        pos += nlen
        if pos + 9 > len(buf):
            raise ValueError(f"{path}: truncated block header")
        (count,) = struct.unpack_from("<Q", buf, pos)
        pos += 9  # entry count and the compressed flag
        if count > block_size:
            raise ValueError(f"{path}: block of {count} entries exceeds the block size {block_size}")
        stats = None
        if ctype == POINT:
            data = np.frombuffer(buf, np.float32, count * 2, pos).reshape(count, 2).copy()
            pos += block_size * 8
        else:
            dt = np.dtype(_NP[ctype])
            fmt = {INT: "<iifi", LONG: "<qqfq", FLOAT: "<ffff", DOUBLE: "<ddfd", INT8: "<bbfb"}[ctype]
            stats = struct.unpack_from(fmt, buf, pos)
            pos += struct.calcsize(fmt)
            data = np.frombuffer(buf, dt, count, pos).copy()
            pos += block_size * dt.itemsize
        if pos > len(buf):
            raise ValueError(f"{path}: truncated block data")
        blocks.append(Block(index, group_id, data, _unpack_nulls(raw_nulls, count), stats))
    return blocks


def load_database(db_path: str) -> dict:
    """{table name: Table} of a .db catalogue.  Column files are looked up by their file name
    next to the catalogue (the catalogue records the paths the server wrote them to)."""
    with open(db_path) as f:
        meta = json.load(f)
    here = os.path.dirname(os.path.abspath(db_path))
    tables = {}
    for t in meta["tables"]:
        bs = int(t.get("table_block_size") or meta["database_default_block_size"])
        table = Table(t["table_name"], bs)
        for c in t["columns"]:
            path = os.path.join(here, os.path.basename(c["file_path_data_file"]))
            col = Column(c["column_name"], int(c["column_type"]), bool(c["nullable"]))
            if col.type not in (STRING, POLYGON):
                col.blocks = read_column(path, col.type, bs)
            table.columns[col.name] = col
        tables[table.name] = table
    return tables


__all__ = ["Block", "Column", "Table", "load_database", "read_column"]
