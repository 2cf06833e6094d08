"""A small reader and writer for the HDF5 files DIABLO exchanges.

DIABLO reads its wall-normal grid from grid.h5 (/grids/y, domain.f90:429-483)
and reads and writes flow files (/Timestep/{U,V,W,TH1..,P} with file and
group attributes, phdf5.f90:913-1459).  Neither h5py nor an HDF5 binding is
importable here, so this module implements the part of the HDF5 file format
(version 0 superblock) that those files use:

* reading: groups stored as symbol tables (v1 B-tree + local heap), object
  headers of version 1 with continuation blocks, datasets with contiguous or
  chunked (v1 B-tree, unfiltered) storage, IEEE float and integer data,
  attributes (messages of version 1-3) with numeric or fixed-string values;
* writing: the same structures with contiguous datasets; every group must hold
  at most 8 members (one symbol node).

Files are little-endian with 8-byte addresses and lengths.
"""
import mmap
import struct

import numpy as np

SIG = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(ValueError):
    pass


def _pad8(n):
    return (n + 7) & ~7


# ----------------------------------------------------------------- reading --
class _Reader:
    def __init__(self, data):
        self.d = data
        if data[:8] != SIG:
            raise H5Error("not an HDF5 file")
        ver = data[8]
        if ver != 0:
            raise H5Error(f"superblock version {ver} is not supported (only version 0)")
        so, sl = data[13], data[14]
        if so != 8 or sl != 8:
            raise H5Error("only 8-byte addresses and lengths are supported")
        self.base = self.u64(24)
        # root group symbol table entry at 56: name offset, object header address
        self.root = self.u64(64)

    def u8(self, o):
        return self.d[o]

    def u16(self, o):
        return struct.unpack_from("<H", self.d, o)[0]

    def u32(self, o):
        return struct.unpack_from("<I", self.d, o)[0]

    def u64(self, o):
        return struct.unpack_from("<Q", self.d, o)[0]

    def cstr(self, o):
        e = self.d.find(b"\0", o)
        return self.d[o:e].decode("ascii")

    # object header version 1 messages, following continuation blocks
    def messages(self, addr):
        d = self.d
        if d[addr] != 1:
            raise H5Error(f"object header version {d[addr]} at {addr} is not supported")
        nmsg = self.u16(addr + 2)
        size = self.u32(addr + 8)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            o = start
            while o + 8 <= start + length and len(out) < nmsg:
                mtype = self.u16(o)
                msize = self.u16(o + 2)
                body = o + 8
                if mtype == 0x10:  # continuation
                    blocks.append((self.u64(body), self.u64(body + 8)))
                out.append((mtype, body, msize))
                o = body + msize
        return out

    def dataspace(self, o):
        ver = self.d[o]
        rank = self.d[o + 1]
        if ver == 1:
            p = o + 8
        elif ver == 2:
            if self.d[o + 3] == 0:
                return ()
            p = o + 4
        else:
            raise H5Error(f"dataspace version {ver}")
        return tuple(self.u64(p + 8 * i) for i in range(rank))

    def datatype(self, o):
        cv = self.d[o]
        cls = cv & 0x0F
        bits = self.d[o + 1] | (self.d[o + 2] << 8) | (self.d[o + 3] << 16)
        size = self.u32(o + 4)
        big = bits & 1
        end = ">" if big else "<"
        if cls == 0:
            signed = (bits >> 3) & 1
            return np.dtype(f"{end}{'i' if signed else 'u'}{size}")
        if cls == 1:
            return np.dtype(f"{end}f{size}")
        if cls == 3:
            return np.dtype(f"S{size}")
        raise H5Error(f"datatype class {cls} is not supported")

    def attribute(self, o, size):
        ver = self.d[o]
        nlen = self.u16(o + 2)
        tlen = self.u16(o + 4)
        slen = self.u16(o + 6)
        if ver == 1:
            p = o + 8
            name = self.d[p:p + nlen - 1].decode("ascii")
            p += _pad8(nlen)
            dt = self.datatype(p)
            p += _pad8(tlen)
            shape = self.dataspace(p)
            p += _pad8(slen)
        elif ver in (2, 3):
            p = o + 8 + (1 if ver == 3 else 0)
            name = self.d[p:p + nlen - 1].decode("ascii")
            p += nlen
            dt = self.datatype(p)
            p += tlen
            shape = self.dataspace(p)
            p += slen
        else:
            raise H5Error(f"attribute version {ver}")
        n = int(np.prod(shape)) if shape else 1
        a = np.frombuffer(self.d, dtype=dt, count=n, offset=p).copy()
        if dt.kind == "S":
            vals = [v.split(b"\0")[0].decode("ascii", "replace") for v in a]
            return name, (vals[0] if not shape else vals)
        a = a.astype(dt.newbyteorder("="))
        return name, (a[0] if not shape else a.reshape(shape))

    def obj(self, addr):
        """('group', members, attrs) or ('dataset', array, attrs)."""
        attrs = {}
        shape = dt = layout = None
        symtab = None
        for mtype, o, size in self.messages(addr):
            if mtype == 0x01:
                shape = self.dataspace(o)
            elif mtype == 0x03:
                dt = self.datatype(o)
            elif mtype == 0x08:
                layout = o
            elif mtype == 0x0C:
                k, v = self.attribute(o, size)
                attrs[k] = v
            elif mtype == 0x11:
                symtab = (self.u64(o), self.u64(o + 8))
        if symtab is not None:
            return "group", self.group_members(*symtab), attrs
        if layout is None or dt is None or shape is None:
            raise H5Error(f"object at {addr} is neither a group nor a dataset")
        return "dataset", self.read_data(layout, shape, dt), attrs

    def heap_data(self, heap):
        if self.d[heap:heap + 4] != b"HEAP":
            raise H5Error("bad local heap")
        return self.u64(heap + 24)

    def group_members(self, btree, heap):
        hdata = self.heap_data(heap)
        out = {}

        def walk(node):
            if self.d[node:node + 4] != b"TREE":
                raise H5Error("bad group B-tree node")
            level = self.d[node + 5]
            used = self.u16(node + 6)
            p = node + 24
            for i in range(used):
                child = self.u64(p + 8 + 16 * i)
                if level > 0:
                    walk(child)
                else:
                    snod(child)

        def snod(a):
            if self.d[a:a + 4] != b"SNOD":
                raise H5Error("bad symbol table node")
            n = self.u16(a + 6)
            for i in range(n):
                e = a + 8 + 40 * i
                name = self.cstr(hdata + self.u64(e))
                out[name] = self.u64(e + 8)

        walk(btree)
        return out

    def read_data(self, lo, shape, dt):
        ver = self.d[lo]
        n = int(np.prod(shape)) if shape else 1
        if ver != 3:
            raise H5Error(f"data layout version {ver} is not supported")
        cls = self.d[lo + 1]
        if cls == 1:  # contiguous
            addr = self.u64(lo + 2)
            if addr == UNDEF:
                return np.zeros(shape, dt.newbyteorder("="))
            a = np.frombuffer(self.d, dtype=dt, count=n, offset=addr).copy()
            return a.astype(dt.newbyteorder("=")).reshape(shape)
        if cls == 0:  # compact
            size = self.u16(lo + 2)
            a = np.frombuffer(self.d, dtype=dt, count=n, offset=lo + 4).copy()
            return a.astype(dt.newbyteorder("=")).reshape(shape)
        if cls == 2:  # chunked
            rank1 = self.d[lo + 2]
            btree = self.u64(lo + 3)
            cdims = [self.u32(lo + 11 + 4 * i) for i in range(rank1)][:-1]
            out = np.zeros(shape, dt.newbyteorder("="))
            rank = len(shape)

            def walk(node):
                if self.d[node:node + 4] != b"TREE" or self.d[node + 4] != 1:
                    raise H5Error("bad chunk B-tree node")
                level = self.d[node + 5]
                used = self.u16(node + 6)
                ksize = 8 + 8 * (rank + 1)
                p = node + 24
                for i in range(used):
                    key = p + i * (ksize + 8)
                    child = self.u64(key + ksize)
                    if level > 0:
                        walk(child)
                        continue
                    csize, fmask = self.u32(key), self.u32(key + 4)
                    if fmask != 0:
                        raise H5Error("filtered chunks are not supported")
                    offs = [self.u64(key + 8 + 8 * d) for d in range(rank)]
                    cnt = int(np.prod(cdims))
                    c = np.frombuffer(self.d, dtype=dt, count=cnt, offset=child).reshape(cdims)
                    sl = tuple(slice(o, min(o + cd, s)) for o, cd, s in zip(offs, cdims, shape))
                    out[sl] = c[tuple(slice(0, x.stop - x.start) for x in sl)]

            if btree != UNDEF:
                walk(btree)
            return out
        raise H5Error(f"layout class {cls} is not supported")


def read(path):
    """The whole file as nested dicts: {'attrs': {...}, name: subtree or array}."""
    with open(path, "rb") as f:
        data = mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ)
    r = _Reader(data)

    def tree(addr):
        kind, val, attrs = r.obj(addr)
        if kind == "dataset":
            return val
        node = {"attrs": attrs}
        for name, a in val.items():
            node[name] = tree(a)
        return node

    try:
        return tree(r.root)
    finally:
        data.close()


def read_dataset(path, name):
    node = read(path)
    for part in name.strip("/").split("/"):
        node = node[part]
    return node


# ----------------------------------------------------------------- writing --
class _Writer:
    """Lays the file out in memory as a list of parts (dataset arrays are kept
    as references, not copied) and writes it in one pass."""

    def __init__(self):
        self.parts = [None]  # the superblock, filled in by finish()
        self.size = 96

    def alloc(self, data):
        pad = _pad8(self.size) - self.size
        if pad:
            self.parts.append(b"\0" * pad)
            self.size += pad
        a = self.size
        self.parts.append(data)
        self.size += len(memoryview(data).cast("B"))
        return a

    @staticmethod
    def msg(mtype, body):
        body = bytes(body) + b"\0" * (_pad8(len(body)) - len(body))
        return struct.pack("<HHB3x", mtype, len(body), 0) + body

    @staticmethod
    def dtype_msg(dt):
        dt = np.dtype(dt)
        if dt.kind == "f" and dt.itemsize == 8:
            return struct.pack("<BBBBI", 0x11, 0x20, 63, 0, 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
        if dt.kind in "iu":
            flags = 0x08 if dt.kind == "i" else 0
            return struct.pack("<BBBBI", 0x10, flags, 0, 0, dt.itemsize) + struct.pack("<HH", 0, 8 * dt.itemsize)
        if dt.kind == "S":
            return struct.pack("<BBBBI", 0x13, 0x00, 0, 0, dt.itemsize)  # ASCII, null-terminated
        raise H5Error(f"cannot write dtype {dt}")

    @staticmethod
    def space_msg(shape):
        b = struct.pack("<BBBB4x", 1, len(shape), 0, 0)
        for s in shape:
            b += struct.pack("<Q", s)
        return b

    @staticmethod
    def as_attr(value):
        if isinstance(value, str):
            raw = value.encode("ascii")
            return np.array(raw, dtype=f"S{max(1, len(raw))}")
        arr = np.asarray(value)
        if arr.dtype.kind == "f":
            return arr.astype("<f8")
        if arr.dtype.kind in "iu":
            return arr.astype("<i4")
        if arr.dtype.kind == "S":
            return arr
        raise H5Error(f"cannot write attribute of dtype {arr.dtype}")

    def attr_msg(self, name, value):
        arr = self.as_attr(value)
        nm = name.encode("ascii") + b"\0"
        dt = self.dtype_msg(arr.dtype)
        sp = self.space_msg(arr.shape)
        body = struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(sp))
        for part in (nm, dt, sp):
            body += part + b"\0" * (_pad8(len(part)) - len(part))
        body += arr.tobytes()
        return self.msg(0x0C, body)

    def ohdr(self, msgs):
        body = b"".join(msgs)
        head = struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(body))
        return self.alloc(head + body)

    def dataset(self, arr):
        arr = np.asarray(arr)
        if arr.dtype.kind == "f":
            arr = arr.astype("<f8", copy=False)
        elif arr.dtype.kind in "iu":
            arr = arr.astype(arr.dtype.newbyteorder("<"), copy=False)
        arr = np.ascontiguousarray(arr)
        data_addr = self.alloc(arr) if arr.nbytes else UNDEF
        msgs = [
            self.msg(0x01, self.space_msg(arr.shape)),
            self.msg(0x03, self.dtype_msg(arr.dtype)),
            # fill value, version 2: space allocated early, fill written if set, no fill value defined
            self.msg(0x05, struct.pack("<BBBB", 2, 1, 2, 0)),
            self.msg(0x08, struct.pack("<BBQQ", 3, 1, data_addr, arr.nbytes)),
        ]
        return self.ohdr(msgs)

    def group(self, node):
        """node: dict of name -> array or dict; the key 'attrs' holds the
        group's attributes.  Returns (object header, B-tree, heap) addresses."""
        subs = {}
        for k, v in node.items():
            if k == "attrs":
                continue
            subs[k] = self.group(v)[0] if isinstance(v, dict) else self.dataset(v)
        if len(subs) > 8:
            raise H5Error("at most 8 members per group")
        # local heap: the empty name at offset 0, then the member names
        names = b"\0" * 8
        offs = {}
        for k in subs:
            offs[k] = len(names)
            nm = k.encode("ascii") + b"\0"
            names += nm + b"\0" * (_pad8(len(nm)) - len(nm))
        names += b"\0" * 8
        data_addr = self.alloc(names)
        # free-list head 1 is the format's "no free block" value
        heap = self.alloc(b"HEAP" + struct.pack("<B3xQQQ", 0, len(names), 1, data_addr))
        # one symbol node (room for 2*4 entries) with the members in name order
        order = sorted(subs)
        tree = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1 if order else 0, UNDEF, UNDEF)
        if order:
            snod = b"SNOD" + struct.pack("<BBH", 1, 0, len(order))
            for k in order:
                snod += struct.pack("<QQII16x", offs[k], subs[k], 0, 0)
            snod += b"\0" * (8 + 40 * 8 - len(snod))
            snod_addr = self.alloc(snod)
            # leaf with one child, keyed by heap offsets (empty name, last name)
            tree += struct.pack("<QQQ", 0, snod_addr, offs[order[-1]])
        tree += b"\0" * (24 + 32 * 8 + 33 * 8 - len(tree))  # 2*16 children, 2*16+1 keys
        tree_addr = self.alloc(tree)
        msgs = [self.msg(0x11, struct.pack("<QQ", tree_addr, heap))]
        msgs += [self.attr_msg(k, v) for k, v in node.get("attrs", {}).items()]
        return self.ohdr(msgs), tree_addr, heap

    def write(self, path, root):
        addr, tree, heap = root
        sb = SIG + struct.pack("<BBBBBBBBHHI", 0, 0, 0, 0, 0, 8, 8, 0, 4, 16, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, self.size, UNDEF)
        sb += struct.pack("<QQII", 0, addr, 1, 0) + struct.pack("<QQ", tree, heap)
        assert len(sb) == 96
        self.parts[0] = sb
        with open(path, "wb") as f:
            for p in self.parts:
                f.write(p)


def write(path, tree):
    """Write nested dicts (the form read() returns) to an HDF5 file.  Datasets
    are stored contiguously; float arrays are written as IEEE float64 LE."""
    w = _Writer()
    w.write(path, w.group(tree))
