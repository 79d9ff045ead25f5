"""Writer of small HDF5 files in the layout OP-PIC's HDF5 mesh files use (superblock version 0, the root
group as a symbol table, version-1 object headers, contiguous datasets), from the published HDF5
file-format specification. Test infrastructure: it builds inputs for the drop-in's opp_decl_*_hdf5
path (the reader itself is checked against the reference's own HDF5 mesh files)."""

import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


def _msg(mtype: int, data: bytes) -> bytes:
    data = _pad8(data)
    return struct.pack("<HHB3x", mtype, len(data), 0) + data


def _object_header(msgs: list) -> bytes:
    body = b"".join(msgs)
    return struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(body)) + body


def _datatype(a: np.ndarray) -> bytes:
    if a.dtype.kind == "i":
        # fixed-point, version 1, little-endian, signed; offset 0, precision in bits
        return struct.pack("<B3sIHH", 0x10, bytes([0x08, 0, 0]), a.dtype.itemsize, 0, 8 * a.dtype.itemsize)
    if a.dtype == np.float64:
        # IEEE little-endian double: sign bit 63, exponent 52..62 (bias 1023), mantissa 0..51
        return struct.pack("<B3sIHHBBBBI", 0x11, bytes([0x20, 63, 0]), 8, 0, 64, 52, 11, 0, 52, 1023)
    if a.dtype == np.float32:
        return struct.pack("<B3sIHHBBBBI", 0x11, bytes([0x20, 31, 0]), 4, 0, 32, 23, 8, 0, 23, 127)
    raise ValueError(f"unsupported dtype {a.dtype}")


def write(path: str, datasets: dict):
    """datasets: name -> numpy array (int32, int64, float32 or float64; 1-D or 2-D)."""
    names = sorted(datasets)
    arrays = [np.ascontiguousarray(datasets[n]) for n in names]
    # local heap data: "" at 0, then every name, each padded to 8
    heap_data = b"\0" * 8
    name_off = []
    for n in names:
        name_off.append(len(heap_data))
        heap_data += _pad8(n.encode() + b"\0")
    out = bytearray()
    sb_size = 8 + 16 + 32 + 40  # signature + versions/sizes/K/flags + 4 addresses + root entry
    out += b"\0" * sb_size
    # dataset object headers and raw data
    oh_addr = []
    for a in arrays:
        shape = a.shape if a.ndim else (1,)
        space = struct.pack("<BBB5x", 1, len(shape), 0) + b"".join(struct.pack("<Q", d) for d in shape)
        # the object header, then the data right after it (the header's size does not depend on
        # the address it holds)
        start = len(out)
        size = len(_object_header([_msg(0x0001, space), _msg(0x0003, _datatype(a)),
                                   _msg(0x0008, struct.pack("<BBQQ", 3, 1, 0, a.nbytes))]))
        layout = struct.pack("<BBQQ", 3, 1, start + size, a.nbytes)
        hdr = _object_header([_msg(0x0001, space), _msg(0x0003, _datatype(a)), _msg(0x0008, layout)])
        out += hdr
        out += _pad8(a.tobytes())
        oh_addr.append(start)
    # local heap
    heap_addr = len(out)
    heap_seg = heap_addr + 32
    out += b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap_data), UNDEF, heap_seg) + heap_data
    # one symbol node with every entry (sorted by name)
    snod_addr = len(out)
    out += b"SNOD" + struct.pack("<BxH", 1, len(names))
    for off, oh in zip(name_off, oh_addr):
        out += struct.pack("<QQI4x16x", off, oh, 0)
    # one leaf B-tree node: key "" | child SNOD | key = last name
    btree_addr = len(out)
    out += b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
    out += struct.pack("<QQQ", 0, snod_addr, name_off[-1] if names else 0)
    # root group object header: the symbol table message
    root_addr = len(out)
    out += _object_header([_msg(0x0011, struct.pack("<QQ", btree_addr, heap_addr))])
    eof = len(out)
    sb = b"\x89HDF\r\n\x1a\n" + struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, 8, 8, 0, max(4, len(names)), 16, 0)
    sb += struct.pack("<QQQQ", 0, UNDEF, eof, UNDEF)
    sb += struct.pack("<QQI4x", 0, root_addr, 1) + struct.pack("<QQ", btree_addr, heap_addr)
    assert len(sb) == sb_size
    out[:sb_size] = sb
    with open(path, "wb") as f:
        f.write(bytes(out))
