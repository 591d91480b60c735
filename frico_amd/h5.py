"""A small reader for the HDF5 files FRICO writes (antenna-analysis results and linear-system dumps).

FRICO writes HDF5 through HighFive/H5Easy (src/emw_ana_antenna.cpp:207-234, 260-268;
src/emw_solver.cpp:494-500), whose files use the format's original structures: a version-0
superblock, version-1 object headers, groups held as symbol tables (a version-1 B-tree of symbol
table nodes plus a local heap of names), contiguous or compact datasets of fixed-point, floating-point
or compound (std::complex -> {r, i}) elements.  This module reads exactly that subset, with numpy,
so that the files can be checked without h5py or the HDF5 library:

    f = H5File("biquad.h5")
    f.keys("/frico/antenna_analysis")        # names in a group
    f["/frico/antenna_analysis/0/gain_XY"]   # numpy array (scalars as 0-d arrays)

Unsupported structures (newer superblocks, link-message groups, chunked or compressed storage,
variable-length types) raise H5Error.  The layout follows the published HDF5 File Format
Specification, version 2.0 (superblock version 0, object header version 1, B-tree version 1,
symbol table node version 1, local heap version 0, dataspace versions 1-2, datatype versions
1-3, data layout versions 1-3).
"""
from __future__ import annotations

import struct

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"


class H5Error(Exception):
    pass


class _Obj:
    """An object header: its messages as (type, flags, raw bytes)."""

    def __init__(self, msgs):
        self.msgs = msgs

    def find(self, mtype):
        return [m for t, fl, m in self.msgs if t == mtype]


class H5File:
    def __init__(self, path_or_bytes):
        if isinstance(path_or_bytes, (bytes, bytearray)):
            self.b = bytes(path_or_bytes)
        else:
            with open(path_or_bytes, "rb") as fh:
                self.b = fh.read()
        b = self.b
        if b[:8] != SIGNATURE:
            raise H5Error("not an HDF5 file (signature)")
        if b[8] != 0:
            raise H5Error(f"superblock version {b[8]} not supported (only 0)")
        self.so, self.sl = b[13], b[14]
        if self.so != 8 or self.sl != 8:
            raise H5Error("only 8-byte offsets and lengths are supported")
        self.leaf_k, self.internal_k = struct.unpack_from("<HH", b, 16)
        base, _fs, self.eof, _drv = struct.unpack_from("<QQQQ", b, 24)
        if base != 0:
            raise H5Error("non-zero base address")
        # root group symbol table entry
        _name, self.root_addr, _ctype = struct.unpack_from("<QQI", b, 56)

    # ---- low-level structures ---------------------------------------------------------------
    def _u(self, fmt, off):
        return struct.unpack_from("<" + fmt, self.b, off)

    def _object(self, addr):
        b = self.b
        if b[addr] != 1:
            raise H5Error(f"object header version {b[addr]} at {addr} not supported (only 1)")
        nmsg, _ref, size = self._u("HII", addr + 2)
        msgs = []
        blocks = [(addr + 16, size)]
        while blocks and len(msgs) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(msgs) < nmsg:
                mtype, msize, flags = self._u("HHB", p)
                data = b[p + 8:p + 8 + msize]
                p += 8 + msize
                if mtype == 0x0010:  # continuation
                    caddr, clen = struct.unpack_from("<QQ", data, 0)
                    blocks.append((caddr, clen))
                msgs.append((mtype, flags, data))
        return _Obj(msgs)

    def _heap_name(self, heap_addr, off):
        b = self.b
        if b[heap_addr:heap_addr + 4] != b"HEAP":
            raise H5Error("local heap signature")
        _size, _free, data = self._u("QQQ", heap_addr + 8)
        end = b.index(b"\0", data + off)
        return b[data + off:end].decode()

    def _group_entries(self, btree, heap):
        """(name, object header address) of a symbol-table group, in B-tree order."""
        b = self.b
        out = []
        if b[btree:btree + 4] != b"TREE":
            raise H5Error("B-tree signature")
        ntype, level, used = b[btree + 4], b[btree + 5], self._u("H", btree + 6)[0]
        if ntype != 0:
            raise H5Error("not a group B-tree")
        p = btree + 24  # after signature, type, level, used, left and right siblings
        children = []
        for i in range(used):
            p += self.sl  # key
            children.append(self._u("Q", p)[0])
            p += self.so
        for c in children:
            if level > 0:
                out += self._group_entries(c, heap)
                continue
            if b[c:c + 4] != b"SNOD":
                raise H5Error("symbol table node signature")
            nsym = self._u("H", c + 6)[0]
            q = c + 8
            for _ in range(nsym):
                name_off, obj = self._u("QQ", q)
                out.append((self._heap_name(heap, name_off), obj))
                q += 40
        return out

    def _children(self, addr):
        o = self._object(addr)
        st = o.find(0x0011)
        if not st:
            if o.find(0x0002) or o.find(0x0006):
                raise H5Error("groups stored as links are not supported")
            raise H5Error("not a group")
        btree, heap = struct.unpack_from("<QQ", st[0], 0)
        return self._group_entries(btree, heap)

    def _resolve(self, path):
        addr = self.root_addr
        for part in [p for p in path.split("/") if p]:
            kids = dict(self._children(addr))
            if part not in kids:
                raise KeyError(path)
            addr = kids[part]
        return addr

    # ---- messages ----------------------------------------------------------------------------
    @staticmethod
    def _dataspace(m):
        ver, rank, flags = m[0], m[1], m[2]
        if ver == 1:
            p = 8
        elif ver == 2:
            p = 4
            if flags == 0 and m[3] == 0 and rank == 0:  # type 0: scalar
                return ()
        else:
            raise H5Error(f"dataspace version {ver}")
        return tuple(struct.unpack_from("<%dQ" % rank, m, p)) if rank else ()

    def _dtype(self, m, off=0):
        """numpy dtype of a datatype message starting at m[off]; returns (dtype, bytes used)."""
        cv = m[off]
        cls, ver = cv & 0x0F, cv >> 4
        bits = m[off + 1] | (m[off + 2] << 8) | (m[off + 3] << 16)
        size = struct.unpack_from("<I", m, off + 4)[0]
        p = off + 8
        if cls == 0:  # fixed point
            order = ">" if bits & 1 else "<"
            signed = bool(bits & 8)
            return np.dtype(f"{order}{'i' if signed else 'u'}{size}"), 8 + 4
        if cls == 1:  # floating point
            order = ">" if bits & 1 else "<"
            if size not in (4, 8):
                raise H5Error(f"float of {size} bytes")
            return np.dtype(f"{order}f{size}"), 8 + 12
        if cls == 6:  # compound
            nmemb = bits & 0xFFFF
            names, formats, offsets = [], [], []
            for _ in range(nmemb):
                end = m.index(b"\0", p)
                name = m[p:end].decode()
                if ver < 3:
                    p += ((end + 1 - p + 7) // 8) * 8  # name padded to a multiple of 8
                    moff = struct.unpack_from("<I", m, p)[0]
                    p += 4
                    if ver == 1:
                        p += 1 + 3 + 4 + 4 + 16  # dimensionality, reserved, permutation, reserved, sizes
                    else:
                        raise H5Error("compound datatype version 2")
                else:
                    p = end + 1
                    nb = 1 if size < 256 else 2 if size < 65536 else 3 if size < (1 << 24) else 4
                    moff = int.from_bytes(m[p:p + nb], "little")
                    p += nb
                dt, used = self._dtype(m, p)
                p += used
                names.append(name)
                formats.append(dt)
                offsets.append(moff)
            return np.dtype({"names": names, "formats": formats, "offsets": offsets, "itemsize": size}), p - off
        raise H5Error(f"datatype class {cls} not supported")

    def _layout(self, m):
        ver = m[0]
        if ver == 3:
            cls = m[1]
            if cls == 0:  # compact
                size = struct.unpack_from("<H", m, 2)[0]
                return ("compact", m[4:4 + size])
            if cls == 1:
                addr, size = struct.unpack_from("<QQ", m, 2)
                return ("contiguous", addr, size)
            raise H5Error("chunked or virtual storage not supported")
        if ver in (1, 2):
            rank, cls = m[1], m[2]
            if cls != 1:
                raise H5Error("only contiguous storage (layout versions 1-2)")
            addr = struct.unpack_from("<Q", m, 8)[0]
            return ("contiguous", addr, None)
        raise H5Error(f"layout version {ver}")

    # ---- public interface ----------------------------------------------------------------------
    def keys(self, path="/"):
        return [n for n, _ in self._children(self._resolve(path))]

    def is_group(self, path):
        return bool(self._object(self._resolve(path)).find(0x0011))

    def describe(self, path):
        """(message types of the object header, shape, dtype) of a dataset."""
        o = self._object(self._resolve(path))
        shape = self._dataspace(o.find(0x0001)[0])
        dt, _ = self._dtype(o.find(0x0003)[0])
        return [t for t, _, _ in o.msgs], shape, dt

    def __getitem__(self, path):
        o = self._object(self._resolve(path))
        if o.find(0x0011):
            raise H5Error(f"{path} is a group")
        shape = self._dataspace(o.find(0x0001)[0])
        dt, _ = self._dtype(o.find(0x0003)[0])
        lay = self._layout(o.find(0x0008)[0])
        count = int(np.prod(shape)) if shape else 1
        nbytes = count * dt.itemsize
        if lay[0] == "compact":
            raw = lay[1][:nbytes]
        else:
            addr = lay[1]
            if addr == UNDEF:
                raw = b"\0" * nbytes
            else:
                raw = self.b[addr:addr + nbytes]
        arr = np.frombuffer(raw, dtype=dt, count=count).reshape(shape)
        if dt.names and set(dt.names) == {"r", "i"}:
            arr = arr["r"].astype(np.float64) + 1j * arr["i"].astype(np.float64)
        return arr

    def walk(self, path="/"):
        """All dataset paths below a group, depth first in name order."""
        out = []
        for name in self.keys(path):
            full = path.rstrip("/") + "/" + name
            if self.is_group(full):
                out += self.walk(full)
            else:
                out.append(full)
        return out
