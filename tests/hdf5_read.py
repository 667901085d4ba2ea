"""Minimal reader of the HDF5 files MPulse's HDFostream writes -- test infrastructure.

It follows the published HDF5 file-format specification for exactly the structures those
files use: a version-0 superblock, the root group as a symbol table (a version-1 group
B-tree whose leaves are symbol-table nodes, and a local heap of names), and datasets whose
object headers (version 1) carry a dataspace, an IEEE float64 or integer datatype, a
contiguous layout and, optionally, attributes.  Anything else raises.  It is checked against
the file the reference's own test program wrote (test/test.h5, test/testhdf.cpp).
"""
import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(Exception):
    pass


def _u(b, off, n):
    return int.from_bytes(b[off:off + n], "little")


def _datatype(b, off):
    """(numpy dtype, size) of a datatype message body at off."""
    cls_ver = b[off]
    cls, ver = cls_ver & 0x0F, cls_ver >> 4
    bits = b[off + 1] | (b[off + 2] << 8) | (b[off + 3] << 16)
    size = _u(b, off + 4, 4)
    if ver != 1:
        raise H5Error(f"datatype version {ver}")
    if bits & 1:
        raise H5Error("big-endian data")
    if cls == 1 and size == 8:
        return np.dtype("<f8")
    if cls == 1 and size == 4:
        return np.dtype("<f4")
    if cls == 0:
        signed = bool(bits & 0x08)
        return np.dtype(("<i" if signed else "<u") + str(size))
    raise H5Error(f"datatype class {cls} size {size}")


def _dataspace(b, off):
    ver, rank, flags = b[off], b[off + 1], b[off + 2]
    if ver != 1:
        raise H5Error(f"dataspace version {ver}")
    p = off + 8
    dims = [_u(b, p + 8 * i, 8) for i in range(rank)]
    return tuple(dims)


def _messages(b, addr):
    """Messages (type, flags, body offset, size) of a version-1 object header."""
    ver, nmsg, size = b[addr], _u(b, addr + 2, 2), _u(b, addr + 8, 4)
    if ver != 1:
        raise H5Error(f"object header version {ver}")
    out, p, end = [], addr + 16, addr + 16 + size
    while p < end and len(out) < nmsg:
        mtype, msize, mflags = _u(b, p, 2), _u(b, p + 2, 2), b[p + 4]
        if mtype == 0x10:  # continuation: not used by these files
            raise H5Error("object header continuation")
        out.append((mtype, mflags, p + 8, msize))
        p += 8 + msize
    return out


def _attribute(b, off):
    ver = b[off]
    if ver != 1:
        raise H5Error(f"attribute version {ver}")
    nlen, tlen, slen = _u(b, off + 2, 2), _u(b, off + 4, 2), _u(b, off + 6, 2)
    pad = lambda n: (n + 7) & ~7
    p = off + 8
    name = b[p:p + nlen].split(b"\0")[0].decode()
    p += pad(nlen)
    dt = _datatype(b, p)
    p += pad(tlen)
    dims = _dataspace(b, p)
    p += pad(slen)
    n = int(np.prod(dims)) if dims else 1
    return name, np.frombuffer(b, dt, n, p).reshape(dims)


class H5File:
    """Datasets of the root group by name: .names, .read(name) -> numpy array (C order over
    the dataspace dims, as the HDF5 library returns it), .attrs(name), .mtime(name)."""

    def __init__(self, data):
        b = bytes(data)
        self.b = b
        if b[:8] != b"\x89HDF\r\n\x1a\n":
            raise H5Error("no HDF5 signature")
        if b[8] != 0:
            raise H5Error(f"superblock version {b[8]}")
        if b[13] != 8 or b[14] != 8:
            raise H5Error("offsets / lengths not 8 bytes")
        self.leaf_k, self.internal_k = _u(b, 16, 2), _u(b, 18, 2)
        self.base, self.eof = _u(b, 24, 8), _u(b, 40, 8)
        if self.eof != len(b):
            raise H5Error(f"end-of-file address {self.eof} != file size {len(b)}")
        root = 56  # root group symbol-table entry
        self.root_header = _u(b, root + 8, 8)
        if _u(b, root + 16, 4) != 1:
            raise H5Error("root entry does not cache a symbol table")
        btree, heap = _u(b, root + 24, 8), _u(b, root + 32, 8)
        msgs = _messages(b, self.root_header)
        st = [m for m in msgs if m[0] == 0x11]
        if not st or (_u(b, st[0][2], 8), _u(b, st[0][2] + 8, 8)) != (btree, heap):
            raise H5Error("root symbol table message disagrees with the superblock cache")
        if b[heap:heap + 4] != b"HEAP":
            raise H5Error("no local heap")
        self.heap_data = _u(b, heap + 24, 8)
        self.entries = {}
        self._walk(btree)
        self.names = sorted(self.entries)

    def _name(self, off):
        p = self.heap_data + off
        return self.b[p:self.b.index(b"\0", p)].decode()

    def _walk(self, node):
        b = self.b
        if b[node:node + 4] != b"TREE":
            raise H5Error("no B-tree node")
        if b[node + 4] != 0:
            raise H5Error("not a group B-tree")
        level, used = b[node + 5], _u(b, node + 6, 2)
        p = node + 24
        for _ in range(used):
            child = _u(b, p + 8, 8)
            p += 16
            if level > 0:
                self._walk(child)
                continue
            if b[child:child + 4] != b"SNOD":
                raise H5Error("no symbol-table node")
            count = _u(b, child + 6, 2)
            for e in range(count):
                q = child + 8 + 40 * e
                self.entries[self._name(_u(b, q, 8))] = _u(b, q + 8, 8)

    def _header(self, name):
        return _messages(self.b, self.entries[name])

    def read(self, name):
        dt = dims = addr = None
        for mtype, _, off, _ in self._header(name):
            if mtype == 3:
                dt = _datatype(self.b, off)
            elif mtype == 1:
                dims = _dataspace(self.b, off)
            elif mtype == 8:
                ver, rank, cls = self.b[off], self.b[off + 1], self.b[off + 2]
                if ver not in (1, 2) or cls != 1:
                    raise H5Error(f"layout version {ver} class {cls}: only contiguous")
                addr = _u(self.b, off + 8, 8)
        if dt is None or dims is None or addr is None:
            raise H5Error(f"{name}: incomplete dataset header")
        n = int(np.prod(dims))
        return np.frombuffer(self.b, dt, n, self.base + addr).reshape(dims)

    def attrs(self, name):
        return dict(_attribute(self.b, off) for mtype, _, off, _ in self._header(name) if mtype == 0x0C)

    def mtime(self, name):
        for mtype, _, off, _ in self._header(name):
            if mtype == 0x12:
                return _u(self.b, off + 4, 4)
        return None


def read_file(path):
    with open(path, "rb") as f:
        return H5File(f.read())
