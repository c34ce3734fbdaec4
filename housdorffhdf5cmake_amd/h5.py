"""Minimal HDF5 writer: contiguous datasets in the root group of a new file,
laid out as the HDF5 library writes them in its default (earliest) format --
superblock version 0, the root group as a symbol table (B-tree, local heap,
one symbol table node), version-1 object headers -- so libhdf5, h5py and the
library's own reader (hdm_h5_read) all read them.  numpy bool arrays are
stored as h5py stores them (an enumeration FALSE = 0, TRUE = 1 over int8).

Used to hand volumes to HDF5-based tools and to build benchmark inputs in
the layout of the reference's benchmark file (forBench/hdf5Manag.cu:196-220:
two flat boolean datasets of one file)."""
from __future__ import annotations

import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
LEAF_K, INTERNAL_K = 4, 16  # group B-tree K values (library defaults)
MAX_LINKS = 2 * LEAF_K      # one symbol table node


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


def _msg(mtype: int, body: bytes) -> bytes:
    body = _pad8(body)
    return struct.pack("<HHB3x", mtype, len(body), 0) + body


def _object_header(msgs) -> bytes:
    data = b"".join(msgs)
    return struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(data)) + data


def _int_type(size: int, signed: bool) -> bytes:
    return struct.pack("<B3sIHH", 0x10, bytes([0x08 if signed else 0, 0, 0]), size, 0, 8 * size)


def _datatype(dtype: np.dtype) -> bytes:
    if dtype == np.bool_:
        base = _int_type(1, True)
        names = _pad8(b"FALSE\0") + _pad8(b"TRUE\0")
        return struct.pack("<B3sI", 0x18, bytes([2, 0, 0]), 1) + base + names + b"\x00\x01"
    if dtype.kind in "iu" and dtype.itemsize in (1, 2, 4, 8):
        return _int_type(dtype.itemsize, dtype.kind == "i")
    if dtype.kind == "f" and dtype.itemsize in (4, 8):
        s = dtype.itemsize
        # IEEE little-endian: sign at the top bit, exponent and mantissa below
        exp_loc, exp_size, man_size, bias = (23, 8, 23, 127) if s == 4 else (52, 11, 52, 1023)
        bits = bytes([0x20, 8 * s - 1, 0])  # mantissa normalization implied, sign position
        return struct.pack("<B3sIHHBBBBI", 0x11, bits, s, 0, 8 * s, exp_loc, exp_size, 0,
                           man_size, bias)
    raise TypeError(f"unsupported dtype {dtype}")


def write(path: str, datasets: dict) -> None:
    """Writes {name: numpy array} (C order, at most 8 names, no '/') as
    contiguous datasets of a new file at path."""
    names = sorted(datasets)
    if not 1 <= len(names) <= MAX_LINKS:
        raise ValueError(f"1..{MAX_LINKS} datasets")
    arrays = {}
    for n in names:
        if not n or "/" in n:
            raise ValueError(f"bad dataset name {n!r}")
        a = np.ascontiguousarray(datasets[n])
        if a.dtype.byteorder == ">":
            a = a.astype(a.dtype.newbyteorder("<"))
        arrays[n] = a
    # local heap data: "" at 0, then the names
    heap = bytearray(8)
    name_off = {}
    for n in names:
        name_off[n] = len(heap)
        heap += _pad8(n.encode() + b"\0")
    heap_size = len(heap)

    SB = 96
    root_oh_size = 16 + 8 + 16
    o_root = SB
    o_heap = o_root + root_oh_size
    o_heap_data = o_heap + 32
    o_btree = o_heap_data + heap_size
    btree_size = 24 + 2 * INTERNAL_K * 8 + (2 * INTERNAL_K + 1) * 8
    o_snod = o_btree + btree_size
    snod_size = 8 + 2 * LEAF_K * 40
    pos = o_snod + snod_size
    ohs, o_ds = {}, {}
    for n in names:  # object headers, then the data
        a = arrays[n]
        space = struct.pack("<BBBx4x", 1, a.ndim, 0) + b"".join(
            struct.pack("<Q", d) for d in a.shape)
        fill = struct.pack("<BBBB", 2, 2, 0, 0)  # late allocation, no fill value
        ohs[n] = (space, _datatype(a.dtype), fill)
        size = len(_object_header([_msg(1, space), _msg(3, ohs[n][1]), _msg(5, fill),
                                   _msg(8, b"\0" * 18)]))
        o_ds[n] = pos
        pos += size
    o_data = {}
    for n in names:
        pos = (pos + 7) // 8 * 8
        o_data[n] = pos
        pos += arrays[n].nbytes
    eof = pos

    meta_end = min(o_data.values())
    out = bytearray(meta_end)  # everything before the first array

    def put(off, b):
        out[off:off + len(b)] = b

    put(0, b"\x89HDF\r\n\x1a\n" + bytes([0, 0, 0, 0, 0, 8, 8, 0]))
    put(16, struct.pack("<HHI", LEAF_K, INTERNAL_K, 0))
    put(24, struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF))
    put(56, struct.pack("<QQII", 0, o_root, 1, 0) + struct.pack("<QQ", o_btree, o_heap))
    put(o_root, _object_header([_msg(17, struct.pack("<QQ", o_btree, o_heap))]))
    # free list head 1: the library's "no free block" value
    put(o_heap, b"HEAP" + struct.pack("<B3xQQQ", 0, heap_size, 1, o_heap_data))
    put(o_heap_data, bytes(heap))
    last = name_off[names[-1]]
    put(o_btree, b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF) +
        struct.pack("<QQQ", 0, o_snod, last))
    ents = b"".join(struct.pack("<QQII16x", name_off[n], o_ds[n], 0, 0) for n in names)
    put(o_snod, b"SNOD" + struct.pack("<BBH", 1, 0, len(names)) + ents)
    for n in names:
        a = arrays[n]
        space, dt, fill = ohs[n]
        layout = struct.pack("<BBQQ", 3, 1, o_data[n], a.nbytes)
        put(o_ds[n], _object_header([_msg(1, space), _msg(3, dt), _msg(5, fill),
                                     _msg(8, layout)]))
    with open(path, "wb") as f:  # the arrays go straight from their buffers
        f.write(out)
        for n in names:
            f.seek(o_data[n])
            f.write(memoryview(arrays[n].reshape(-1).view(np.uint8)))
        f.truncate(eof)
