"""Minimal HDF5 reader for the scan metadata RecView takes from APS DXchange
files: the frame count of /exchange/data and the angles in /exchange/theta.

The reference carries its own small parser for this (chdf5.cpp, used by
ReadHDF5Theta, general.cpp:1668-1683, and LoadLogFile, gazoDoc.cpp:2229-2312):
walk from the superblock through the root group to "exchange", find a child
dataset, read its data object header (dimensions, element size, raw data
address) and convert float / double elements to float.  This module does the
same from the HDF5 file-format specification, for both group encodings the
HDF5 library writes:

  * superblock version 0/1, object header version 1, groups as symbol tables
    (version-1 B-tree of symbol-table nodes, names in a local heap);
  * superblock version 2/3, object header version 2 ("OHDR"), groups as link
    messages in the header (compact storage).

Datasets: dataspace (dimensions), datatype (IEEE float 4 or 8 bytes, or
unsigned/signed integers for dimension queries) and contiguous or compact
layout for reading values.  Dense link storage (fractal heaps), filters and
chunked reads are not needed for the metadata and raise H5Error.
"""
import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"


class H5Error(ValueError):
    pass


class _Dataset:
    def __init__(self):
        self.dims = None
        self.dtype = None
        self.layout = None          # ("contiguous", addr, size) / ("compact", bytes) / ("chunked",)


class H5File:
    """Read-only view of an HDF5 file (path or bytes)."""

    def __init__(self, source):
        if isinstance(source, (bytes, bytearray)):
            self._buf = bytes(source)
            self._f = None
        else:
            self._f = open(source, "rb")
            self._buf = None
        self._superblock()

    def close(self):
        if self._f:
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ------------------------------------------------------------ raw access
    def _read(self, addr, n):
        if self._buf is not None:
            if addr < 0 or addr + n > len(self._buf):
                raise H5Error(f"read past the end of the file at {addr}")
            return self._buf[addr:addr + n]
        self._f.seek(addr)
        b = self._f.read(n)
        if len(b) != n:
            raise H5Error(f"read past the end of the file at {addr}")
        return b

    def _uint(self, b, off, size):
        return int.from_bytes(b[off:off + size], "little")

    def _is_undef(self, v, size):
        return v == (1 << (8 * size)) - 1

    # ------------------------------------------------------------ superblock
    def _superblock(self):
        base = None
        for off in (0, 512, 1024, 2048, 4096, 8192, 16384, 32768, 65536):
            try:
                if self._read(off, 8) == SIGNATURE:
                    base = off
                    break
            except H5Error:
                break
        if base is None:
            raise H5Error("not an HDF5 file (no superblock signature)")
        hdr = self._read(base, 16)
        ver = hdr[8]
        if ver in (0, 1):
            self.so, self.sl = hdr[13], hdr[14]
            b = self._read(base, 24 + (4 if ver == 1 else 0) + 4 * self.so + 40 + 2 * self.so)
            p = 24 + (4 if ver == 1 else 0)
            self.base = self._uint(b, p, self.so)
            p += 4 * self.so                      # base, free space, EOF, driver info
            # root group symbol table entry: link name offset, object header address, ...
            self.root = self._uint(b, p + self.so, self.so)
        elif ver in (2, 3):
            self.so, self.sl = hdr[9], hdr[10]
            b = self._read(base, 12 + 4 * self.so)
            self.base = self._uint(b, 12, self.so)
            self.root = self._uint(b, 12 + 3 * self.so, self.so)
        else:
            raise H5Error(f"unsupported superblock version {ver}")
        if self.so not in (2, 4, 8) or self.sl not in (2, 4, 8):
            raise H5Error("unsupported offset / length sizes")

    def _addr(self, a):
        return self.base + a

    # --------------------------------------------------------- object header
    @staticmethod
    def _visit(seen, start, size):
        """Continuation blocks that repeat would make the walk loop forever."""
        if (start, size) in seen:
            raise H5Error("object header continuation blocks form a cycle")
        seen.add((start, size))

    def _messages(self, addr):
        """(type, data) of every message of the object header at addr."""
        a = self._addr(addr)
        head = self._read(a, 16)
        out = []
        if head[:4] == b"OHDR":
            ver, flags = head[4], head[5]
            if ver != 2:
                raise H5Error(f"unsupported object header version {ver}")
            p = 6 + (16 if flags & 0x20 else 0) + (4 if flags & 0x10 else 0)
            csize = 1 << (flags & 3)
            b = self._read(a, p + csize)
            size0 = self._uint(b, p, csize)
            # chunk 0: the messages follow the size field (its checksum comes after);
            # continuation chunks: "OCHK", messages, checksum
            blocks = [(a + p + csize, size0, False)]
            seen = set()
            track = 2 if flags & 0x04 else 0
            while blocks:
                start, size, cont = blocks.pop(0)
                self._visit(seen, start, size)
                data = self._read(start, size)
                if cont and data[:4] != b"OCHK":
                    raise H5Error("bad object header continuation block")
                q, end = (4, size - 4) if cont else (0, size)
                while q + 4 + track <= end:
                    mtype, msize = data[q], self._uint(data, q + 1, 2)
                    q += 4 + track
                    body = data[q:q + msize]
                    q += msize
                    if mtype == 0x10:
                        blocks.append((self._addr(self._uint(body, 0, self.so)), self._uint(body, self.so, self.sl),
                                       True))
                    elif mtype != 0:
                        out.append((mtype, body))
            return out
        ver = head[0]
        if ver != 1:
            raise H5Error(f"unsupported object header version {ver}")
        nmsg = self._uint(head, 2, 2)
        hsize = self._uint(head, 8, 4)
        blocks = [(a + 16, hsize)]
        seen = set()
        while blocks and len(out) < nmsg:
            start, size = blocks.pop(0)
            self._visit(seen, start, size)
            data = self._read(start, size)
            q = 0
            while q + 8 <= size:
                mtype, msize = self._uint(data, q, 2), self._uint(data, q + 2, 2)
                body = data[q + 8:q + 8 + msize]
                q += 8 + msize
                if mtype == 0x10:
                    blocks.append((self._addr(self._uint(body, 0, self.so)), self._uint(body, self.so, self.sl)))
                elif mtype != 0:
                    out.append((mtype, body))
        return out

    # ---------------------------------------------------------------- groups
    def _children(self, addr):
        """{name: object header address} of the group whose header is at addr."""
        msgs = self._messages(addr)
        kids = {}
        for mtype, body in msgs:
            if mtype == 0x11:                         # symbol table: B-tree + local heap
                btree = self._uint(body, 0, self.so)
                heap = self._uint(body, self.so, self.so)
                self._walk_btree(btree, self._heap_data(heap), kids)
            elif mtype == 0x06:                       # link message
                name, target = self._link(body)
                if target is not None:
                    kids[name] = target
            elif mtype == 0x02:                       # link info: dense storage if a heap is set
                p = 2 + (8 if body[1] & 1 else 0)
                fheap = self._uint(body, p, self.so)
                if not self._is_undef(fheap, self.so):
                    raise H5Error("dense link storage (fractal heap) is not supported")
        return kids

    def _heap_data(self, heap):
        h = self._read(self._addr(heap), 8 + 2 * self.sl + self.so)
        if h[:4] != b"HEAP":
            raise H5Error("bad local heap")
        size = self._uint(h, 8, self.sl)
        daddr = self._uint(h, 8 + 2 * self.sl, self.so)
        return self._read(self._addr(daddr), size)

    def _walk_btree(self, node, heap, kids, parent_level=None):
        h = self._read(self._addr(node), 8 + 2 * self.so)
        if h[:4] != b"TREE" or h[4] != 0:
            raise H5Error("bad group B-tree node")
        level, used = h[5], self._uint(h, 6, 2)
        if parent_level is not None and level >= parent_level:
            raise H5Error("group B-tree child is not below its parent (cycle)")
        body = self._read(self._addr(node) + 8 + 2 * self.so, (used + 1) * self.sl + used * self.so)
        p = self.sl                                   # key 0
        for _ in range(used):
            child = self._uint(body, p, self.so)
            p += self.so + self.sl                    # child, next key
            if level > 0:
                self._walk_btree(child, heap, kids, level)
            else:
                self._read_snod(child, heap, kids)

    def _read_snod(self, addr, heap, kids):
        h = self._read(self._addr(addr), 8)
        if h[:4] != b"SNOD":
            raise H5Error("bad symbol table node")
        n = self._uint(h, 6, 2)
        esize = 2 * self.so + 24
        ent = self._read(self._addr(addr) + 8, n * esize)
        for i in range(n):
            off = self._uint(ent, i * esize, self.so)
            obj = self._uint(ent, i * esize + self.so, self.so)
            name = heap[off:heap.index(b"\0", off)].decode("utf-8", "replace")
            kids[name] = obj

    def _link(self, body):
        ver, flags = body[0], body[1]
        if ver != 1:
            raise H5Error(f"unsupported link message version {ver}")
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = body[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        lsize = 1 << (flags & 3)
        n = self._uint(body, p, lsize)
        p += lsize
        name = body[p:p + n].decode("utf-8", "replace")
        p += n
        if ltype != 0:                                # soft / external links are not followed
            return name, None
        return name, self._uint(body, p, self.so)

    def _lookup(self, path):
        addr = self.root
        for part in [s for s in path.split("/") if s]:
            kids = self._children(addr)
            if part not in kids:
                raise KeyError(path)
            addr = kids[part]
        return addr

    # -------------------------------------------------------------- datasets
    def _dataset(self, path):
        ds = _Dataset()
        for mtype, body in self._messages(self._lookup(path)):
            if mtype == 0x01:
                ver, rank = body[0], body[1]
                p = 8 if ver == 1 else 4
                ds.dims = tuple(self._uint(body, p + i * self.sl, self.sl) for i in range(rank))
            elif mtype == 0x03:
                cls = body[0] & 0x0F
                size = self._uint(body, 4, 4)
                big = body[1] & 1
                if cls == 1 and size in (4, 8):
                    ds.dtype = np.dtype((">" if big else "<") + ("f4" if size == 4 else "f8"))
                elif cls == 0 and size in (1, 2, 4, 8):
                    signed = body[1] & 0x08
                    ds.dtype = np.dtype((">" if big else "<") + ("i" if signed else "u") + str(size))
            elif mtype == 0x08:
                ver = body[0]
                if ver in (3, 4):                     # version 4 changes only the chunked form
                    cls = body[1]
                    if cls == 0:
                        n = self._uint(body, 2, 2)
                        ds.layout = ("compact", body[4:4 + n])
                    elif cls == 1:
                        ds.layout = ("contiguous", self._uint(body, 2, self.so), self._uint(body, 2 + self.so, self.sl))
                    else:
                        ds.layout = ("chunked",)
                elif ver in (1, 2):
                    ndim, cls = body[1], body[2]
                    if cls == 1:
                        a = self._uint(body, 8, self.so)
                        ds.layout = ("contiguous", a, None)
                    elif cls == 0:
                        p = 8 + 4 * ndim
                        n = self._uint(body, p, 4)
                        ds.layout = ("compact", body[p + 4:p + 4 + n])
                    else:
                        ds.layout = ("chunked",)
                else:
                    raise H5Error(f"{path}: unsupported data layout version {ver}")
        if ds.dims is None:
            raise H5Error(f"{path}: no dataspace")
        return ds

    def shape(self, path):
        """Dimensions of a dataset."""
        return self._dataset(path).dims

    def exists(self, path):
        try:
            self._lookup(path)
            return True
        except KeyError:
            return False

    def read(self, path):
        """Values of a contiguous or compact numeric dataset."""
        ds = self._dataset(path)
        if ds.dtype is None:
            raise H5Error(f"{path}: unsupported datatype")
        n = int(np.prod(ds.dims)) if ds.dims else 1
        nbytes = n * ds.dtype.itemsize
        if ds.layout is None or ds.layout[0] == "chunked":
            raise H5Error(f"{path}: only contiguous or compact datasets can be read")
        if ds.layout[0] == "compact":
            raw = ds.layout[1][:nbytes]
        else:
            addr = ds.layout[1]
            if self._is_undef(addr, self.so):
                return np.zeros(ds.dims, ds.dtype)    # never written: fill value 0
            raw = self._read(self._addr(addr), nbytes)
        return np.frombuffer(raw, ds.dtype).reshape(ds.dims).copy()
