"""A small read-only HDF5 reader for CGNS mesh files.

The reference reads its meshes with HDF5.jl (src/processmesh.jl:14,
src/extractcoordinates.jl:1-13, src/extractelements.jl:1-19); h5py is not
available here, so this module parses the subset of the HDF5 file format that
CGNS/HDF5 files written by current libraries use:

* superblock versions 2 and 3,
* version 2 object headers ("OHDR") with continuation blocks ("OCHK"),
* groups whose links are stored compactly as link messages (no fractal heap),
* datasets with contiguous or compact layout, no filters,
* little- or big-endian fixed-point and IEEE floating-point element types.

Anything else raises NotImplementedError naming the structure.  The reader
only interprets bytes; it executes nothing from the file.
"""
from __future__ import annotations

import numpy as np

UNDEF = (1 << 64) - 1

# object header message types
_DATASPACE, _LINK_INFO, _DATATYPE, _FILL_OLD, _FILL, _LINK, _LAYOUT = 0x01, 0x02, 0x03, 0x04, 0x05, 0x06, 0x08
_FILTERS, _CONTINUATION, _SYMBOL_TABLE = 0x0B, 0x10, 0x11


class H5File:
    """Open an HDF5 file; `read(path)` returns a dataset as a numpy array and
    `children(path)` the link names of a group.  Paths are '/'-separated and
    may contain spaces (CGNS names its arrays ' data')."""

    def __init__(self, path: str):
        with open(path, "rb") as fh:
            self.d = fh.read()
        d = self.d
        if d[:8] != b"\x89HDF\r\n\x1a\n":
            raise ValueError(f"{path} is not an HDF5 file")
        ver = d[8]
        if ver not in (2, 3):
            raise NotImplementedError(f"HDF5 superblock version {ver}")
        self.O, self.L = d[9], d[10]
        base = self._u(12, self.O)
        if base != 0:
            raise NotImplementedError("HDF5 base address other than 0")
        self.root = self._u(12 + 3 * self.O, self.O)

    # ---- primitives
    def _u(self, off, n):
        return int.from_bytes(self.d[off:off + n], "little")

    def _messages(self, addr):
        d = self.d
        if d[addr:addr + 4] != b"OHDR":
            raise NotImplementedError(f"object header at {addr} is not version 2")
        flags = d[addr + 5]
        p = addr + 6
        if flags & 0x20:
            p += 16          # access, modification, change and birth times
        if flags & 0x10:
            p += 4           # attribute phase change values
        size_bytes = 1 << (flags & 3)
        chunk = self._u(p, size_bytes)
        p += size_bytes
        blocks = [(p, p + chunk)]
        track = bool(flags & 0x04)
        out = []
        while blocks:
            p, end = blocks.pop(0)
            while p + 4 <= end:
                mtype, msize = d[p], self._u(p + 1, 2)
                p += 4 + (2 if track else 0)
                body = d[p:p + msize]
                if mtype == _CONTINUATION:
                    ca, cl = self._u(p, self.O), self._u(p + self.O, self.L)
                    if d[ca:ca + 4] != b"OCHK":
                        raise ValueError(f"bad continuation block at {ca}")
                    blocks.append((ca + 4, ca + cl - 4))   # signature, checksum
                else:
                    out.append((mtype, body))
                p += msize
        return out

    def _links(self, msgs):
        links = {}
        for mtype, body in msgs:
            if mtype == _SYMBOL_TABLE:
                raise NotImplementedError("groups with symbol tables (old-style groups)")
            if mtype == _LINK_INFO:
                flags = body[1]
                p = 2 + (8 if flags & 1 else 0)
                heap = int.from_bytes(body[p:p + self.O], "little")
                if heap != UNDEF:
                    raise NotImplementedError("groups with dense link storage")
            if mtype == _LINK:
                flags = body[1]
                p = 2
                ltype = 0
                if flags & 0x08:
                    ltype = body[p]
                    p += 1
                if flags & 0x04:
                    p += 8       # creation order
                if flags & 0x10:
                    p += 1       # name character set
                lsize = 1 << (flags & 3)
                nlen = int.from_bytes(body[p:p + lsize], "little")
                p += lsize
                name = body[p:p + nlen].decode("utf-8")
                p += nlen
                if ltype == 0:   # hard link
                    links[name] = int.from_bytes(body[p:p + self.O], "little")
        return links

    def _lookup(self, path):
        addr = self.root
        for part in [s for s in path.split("/") if s != ""]:
            links = self._links(self._messages(addr))
            if part not in links:
                raise KeyError(f"{path}: no link named {part!r}")
            addr = links[part]
        return addr

    # ---- public
    def children(self, path: str = "/"):
        return list(self._links(self._messages(self._lookup(path))).keys())

    def read(self, path: str) -> np.ndarray:
        msgs = self._messages(self._lookup(path))
        dims = dtype = layout = None
        for mtype, body in msgs:
            if mtype == _DATASPACE:
                ver, nd, flags = body[0], body[1], body[2]
                p = 4 if ver == 2 else 8
                dims = tuple(int.from_bytes(body[p + 8 * i:p + 8 * i + 8], "little") for i in range(nd))
            elif mtype == _DATATYPE:
                cls = body[0] & 0x0F
                bits = body[1]
                size = int.from_bytes(body[4:8], "little")
                order = ">" if bits & 1 else "<"
                if cls == 0:
                    dtype = np.dtype(f"{order}{'i' if bits & 0x08 else 'u'}{size}")
                elif cls == 1:
                    if size not in (4, 8):
                        raise NotImplementedError(f"{size}-byte floating point")
                    dtype = np.dtype(f"{order}f{size}")
                else:
                    raise NotImplementedError(f"HDF5 datatype class {cls}")
            elif mtype == _FILTERS:
                raise NotImplementedError("filtered (compressed) datasets")
            elif mtype == _LAYOUT:
                ver, cls = body[0], body[1]
                if ver not in (3, 4):
                    raise NotImplementedError(f"data layout message version {ver}")
                if cls == 1:
                    a = int.from_bytes(body[2:2 + self.O], "little")
                    n = int.from_bytes(body[2 + self.O:2 + self.O + self.L], "little")
                    layout = (a, n)
                elif cls == 0:
                    n = int.from_bytes(body[2:4], "little")
                    layout = ("compact", body[4:4 + n])
                else:
                    raise NotImplementedError("chunked or virtual datasets")
        if dims is None or dtype is None or layout is None:
            raise ValueError(f"{path} is not a dataset")
        count = int(np.prod(dims)) if dims else 1
        if layout[0] == "compact":
            raw = layout[1]
        else:
            a, n = layout
            if a == UNDEF:   # never written: HDF5's default fill value is zero
                return np.zeros(dims, dtype=dtype.newbyteorder("="))
            raw = self.d[a:a + n]
        arr = np.frombuffer(raw, dtype=dtype, count=count).astype(dtype.newbyteorder("="))
        # dimensions in HDF5's row-major order (HDF5.jl, column-major, reports
        # them reversed)
        return arr.reshape(dims)
