"""A second, independent HDF5 reader in pure Python (numpy), written from the published HDF5
file-format specification for the subset OP-PIC's HDF5 mesh files use (superblock version 0 or 1,
symbol-table groups, version-1 object headers, contiguous or compact datasets). Test
infrastructure: make_h5_mesh_digests.py reads the reference's own HDF5 meshes with it, and the
library's C++ reader is checked against what it reads."""
import struct

import numpy as np

SIG = b"\x89HDF\r\n\x1a\n"

class H5:
    def __init__(self, data: bytes):
        self.b = data
        assert data[:8] == SIG, "not an HDF5 file"
        ver = data[8]
        assert ver in (0, 1), f"superblock version {ver}"
        self.so, self.sl = data[13], data[14]
        assert self.so == 8 and self.sl == 8
        p = 24 if ver == 0 else 28
        self.base, _, self.eof, _ = struct.unpack_from("<4Q", data, p)
        p += 32
        # root group symbol table entry
        _, ohdr, ctype = struct.unpack_from("<QQI", data, p)
        self.root = ohdr

    def u(self, fmt, off):
        return struct.unpack_from("<" + fmt, self.b, off)

    def messages(self, addr):
        b = self.b
        ver = b[addr]
        assert ver == 1, f"object header version {ver} at {addr}"
        nmsg, = self.u("H", addr + 2)
        size, = self.u("I", addr + 8)
        blocks = [(addr + 16, size)]
        out = []
        while blocks:
            start, length = blocks.pop(0)
            p = start
            while p < start + length and len(out) < nmsg:
                mtype, msize, flags = self.u("HHB", p)
                data = p + 8
                if mtype == 0x10:  # continuation
                    off, ln = self.u("QQ", data)
                    blocks.append((off, ln))
                else:
                    out.append((mtype, data, msize))
                p = data + msize
        return out

    def heap_name(self, heap, off):
        assert self.b[heap:heap + 4] == b"HEAP"
        dseg, = self.u("Q", heap + 24)
        s = dseg + off
        e = self.b.index(b"\0", s)
        return self.b[s:e].decode()

    def group(self, addr):
        """name -> object header address of a symbol-table group."""
        btree = heap = None
        for t, d, _ in self.messages(addr):
            if t == 0x11:
                btree, heap = self.u("QQ", d)
        assert btree is not None, "not a symbol-table group"
        out = {}
        def walk(node):
            assert self.b[node:node + 4] == b"TREE"
            ntype, level, used = self.b[node + 4], self.b[node + 5], self.u("H", node + 6)[0]
            p = node + 24
            for k in range(used):
                child, = self.u("Q", p + 8)
                p += 16
                if level > 0:
                    walk(child)
                else:
                    assert self.b[child:child + 4] == b"SNOD"
                    ns, = self.u("H", child + 6)
                    q = child + 8
                    for _ in range(ns):
                        name_off, oh = self.u("QQ", q)
                        out[self.heap_name(heap, name_off)] = oh
                        q += 40
        walk(btree)
        return out

    def dtype_of(self, d):
        cv = self.b[d]
        cls, ver = cv & 15, cv >> 4
        bits = self.b[d + 1]
        size, = self.u("I", d + 4)
        if cls == 0:
            signed = bool(bits & 8)
            order = ">" if bits & 1 else "<"
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}")
        if cls == 1:
            order = ">" if bits & 1 else "<"
            return np.dtype(f"{order}f{size}")
        if cls == 3:
            return np.dtype(f"S{size}")
        raise ValueError(f"datatype class {cls}")

    def dataspace(self, d):
        ver, ndim, flags = self.b[d], self.b[d + 1], self.b[d + 2]
        p = d + (8 if ver == 1 else 4)
        return tuple(self.u(f"{ndim}Q", p)) if ndim else ()

    def attrs(self, addr):
        out = {}
        for t, d, _ in self.messages(addr):
            if t != 0x0C:
                continue
            ver = self.b[d]
            assert ver == 1, f"attribute version {ver}"
            nsz, tsz, ssz = self.u("HHH", d + 2)
            pad = lambda x: (x + 7) & ~7
            p = d + 8
            name = self.b[p:p + nsz].split(b"\0")[0].decode()
            p += pad(nsz)
            dt = self.dtype_of(p)
            p += pad(tsz)
            shape = self.dataspace(p)
            p += pad(ssz)
            n = int(np.prod(shape)) if shape else 1
            out[name] = np.frombuffer(self.b, dt, n, p).copy()
        return out

    def dataset(self, addr):
        dt = shape = layout = None
        filters = False
        for t, d, sz in self.messages(addr):
            if t == 0x03: dt = self.dtype_of(d)
            elif t == 0x01: shape = self.dataspace(d)
            elif t == 0x0B: filters = True
            elif t == 0x08:
                ver = self.b[d]
                if ver == 3:
                    cls = self.b[d + 1]
                    if cls == 1:
                        a, n = self.u("QQ", d + 2); layout = ("contig", a, n)
                    elif cls == 0:
                        n, = self.u("H", d + 2); layout = ("compact", d + 4, n)
                    else:
                        layout = ("chunked", None, None)
                else:
                    ndim, cls = self.b[d + 1], self.b[d + 2]
                    if cls == 1:
                        a, = self.u("Q", d + 8); layout = ("contig", a, None)
                    else:
                        layout = (f"class{cls}", None, None)
        assert layout and layout[0] in ("contig", "compact") and not filters, f"layout {layout} filters {filters}"
        n = int(np.prod(shape)) if shape else 1
        if layout[1] == 0xFFFFFFFFFFFFFFFF:  # never written
            return np.zeros(shape, dt)
        return np.frombuffer(self.b, dt, n, layout[1]).reshape(shape).copy()
