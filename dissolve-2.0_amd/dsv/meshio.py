"""Mesh ingest: the reference's XDMF + HDF5 tetrahedral meshes (`domain/mesh_utils.py:155-267`),
read without h5py or DOLFINx.

* `read_xdmf(path)` parses the XDMF file the reference writes (`_write_xdmf`, `:155-194`): the
  Tetrahedron topology, the XYZ geometry and the cell Attribute, each a DataItem naming
  `file.h5:/dataset` (Format="HDF") or holding the numbers inline (Format="XML").
* `H5File(path)[name]` reads one dataset of an HDF5 file: superblock versions 0-3, object
  headers versions 1 and 2 (with continuation blocks), old-style groups (symbol-table B-trees
  and local heaps) and new-style compact groups (link messages), fixed-point and IEEE floating
  datatypes of either byte order, and compact, contiguous or unfiltered chunked (version-1
  B-tree) storage.  Filtered (compressed) datasets and dense link storage raise
  NotImplementedError.
* `load_mesh(xdmf_path, comm=None)` returns (dsv.core.Mesh, cell tags) like the reference's
  loader; the tags are the cell Attribute (region labels) or None.
"""
from __future__ import annotations

import os
import struct
import xml.etree.ElementTree as ET

import numpy as np

_SIG = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF


class H5File:
    """Read-only access to the datasets of an HDF5 file (see the module docstring)."""

    def __init__(self, path):
        with open(path, "rb") as f:
            self.buf = f.read()
        b = self.buf
        base = b.find(_SIG)
        if base < 0 or base % 512:
            raise ValueError(f"{path}: not an HDF5 file")
        ver = b[base + 8]
        if ver in (0, 1):
            self.so, self.sl = b[base + 13], b[base + 14]
            p = base + 24 + (4 if ver == 1 else 0)
            self.base = self._off(p)
            p += 4 * self.so  # base, free space, end of file, driver info addresses
            # root group symbol-table entry: name offset, object header address, ...
            self.root = self._off(p + self.so)
        elif ver in (2, 3):
            self.so, self.sl = b[base + 9], b[base + 10]
            p = base + 12
            self.base = self._off(p)
            self.root = self._off(p + 3 * self.so)
        else:
            raise NotImplementedError(f"HDF5 superblock version {ver}")

    # ---- primitives
    def _u(self, p, n):
        return int.from_bytes(self.buf[p:p + n], "little")

    def _off(self, p):
        return self._u(p, self.so)

    def _len(self, p):
        return self._u(p, self.sl)

    def _addr(self, a):
        return self.base + a

    # ---- object headers
    def _messages(self, addr):
        """[(type, data offset, size)] of the object header at addr, continuations included."""
        b, p = self.buf, self._addr(addr)
        out = []
        if b[p:p + 4] == b"OHDR":  # version 2
            flags = b[p + 5]
            q = p + 6 + (16 if flags & 0x20 else 0) + (4 if flags & 0x10 else 0)
            nsz = 1 << (flags & 3)
            size = self._u(q, nsz)
            q += nsz
            blocks = [(q, q + size)]
            while blocks:
                s, e = blocks.pop(0)
                while s + 4 <= e:
                    mt, ms, mf = b[s], self._u(s + 1, 2), b[s + 3]
                    s += 4 + (2 if flags & 0x04 else 0)
                    if mt == 0x10:
                        ca, cl = self._off(s), self._len(s + self.so)
                        cs = self._addr(ca)
                        if b[cs:cs + 4] != b"OCHK":
                            raise ValueError("bad object header continuation block")
                        blocks.append((cs + 4, cs + cl - 4))
                    elif mt:
                        out.append((mt, s, ms))
                    s += ms
            return out
        if b[p] != 1:
            raise NotImplementedError(f"object header version {b[p]}")
        nmsg = self._u(p + 2, 2)
        size = self._u(p + 8, 4)
        blocks = [(p + 16, p + 16 + size)]
        while blocks and len(out) < nmsg:
            s, e = blocks.pop(0)
            while s + 8 <= e:
                mt, ms = self._u(s, 2), self._u(s + 2, 2)
                s += 8
                if mt == 0x10:
                    blocks.append((self._addr(self._off(s)),
                                   self._addr(self._off(s)) + self._len(s + self.so)))
                elif mt:
                    out.append((mt, s, ms))
                s += ms
        return out

    # ---- groups
    def _children(self, addr):
        """{name: object header address} of the group at addr."""
        kids = {}
        for mt, p, _ in self._messages(addr):
            if mt == 0x11:  # symbol table: v1 B-tree + local heap
                self._btree_group(self._off(p), self._heap_data(self._off(p + self.so)), kids)
            elif mt == 0x06:  # link message (compact new-style group)
                flags = self.buf[p + 1]
                q = p + 2
                ltype = 0
                if flags & 0x08:
                    ltype = self.buf[q]
                    q += 1
                if flags & 0x04:
                    q += 8
                if flags & 0x10:
                    q += 1
                nl = 1 << (flags & 3)
                ln = self._u(q, nl)
                q += nl
                name = self.buf[q:q + ln].decode()
                q += ln
                if ltype == 0:
                    kids[name] = self._off(q)
            elif mt == 0x02:  # link info: dense storage when the fractal heap is defined
                if self._off(p + 2 + (8 if self.buf[p + 1] & 1 else 0)) != _UNDEF:
                    raise NotImplementedError("dense link storage")
        return kids

    def _heap_data(self, addr):
        p = self._addr(addr)
        if self.buf[p:p + 4] != b"HEAP":
            raise ValueError("bad local heap")
        return self._addr(self._off(p + 8 + 2 * self.sl))

    def _btree_group(self, addr, heap, kids):
        b, p = self.buf, self._addr(addr)
        if b[p:p + 4] != b"TREE" or b[p + 4] != 0:
            raise ValueError("bad group B-tree node")
        level, n = b[p + 5], self._u(p + 6, 2)
        q = p + 8 + 2 * self.so + self.sl  # siblings, then key 0
        for _ in range(n):
            child = self._off(q)
            q += self.so + self.sl
            if level > 0:
                self._btree_group(child, heap, kids)
                continue
            s = self._addr(child)
            if b[s:s + 4] != b"SNOD":
                raise ValueError("bad symbol table node")
            ent = s + 8
            esz = 2 * self.so + 24
            for k in range(self._u(s + 6, 2)):
                e = ent + k * esz
                name_at = heap + self._off(e)
                name = b[name_at:b.index(b"\0", name_at)].decode()
                kids[name] = self._off(e + self.so)

    def _lookup(self, path):
        addr = self.root
        for part in [x for x in path.split("/") if x]:
            kids = self._children(addr)
            if part not in kids:
                raise KeyError(path)
            addr = kids[part]
        return addr

    # ---- datasets
    def __getitem__(self, path):
        b = self.buf
        shape = dtype = layout = None
        for mt, p, _ in self._messages(self._lookup(path)):
            if mt == 0x01:  # dataspace
                ver, rank, flags = b[p], b[p + 1], b[p + 2]
                q = p + (8 if ver == 1 else 4)
                shape = tuple(self._len(q + k * self.sl) for k in range(rank))
            elif mt == 0x03:  # datatype
                cls, bits, size = b[p] & 0x0F, b[p + 1], self._u(p + 4, 4)
                order = ">" if bits & 1 else "<"
                if cls == 0:
                    dtype = np.dtype(f"{order}{'i' if bits & 0x08 else 'u'}{size}")
                elif cls == 1:
                    dtype = np.dtype(f"{order}f{size}")
                else:
                    raise NotImplementedError(f"HDF5 datatype class {cls}")
            elif mt == 0x08:  # data layout
                layout = (p, b[p], b[p + 1])
            elif mt == 0x0B:
                raise NotImplementedError("filtered (compressed) datasets")
        if shape is None or dtype is None or layout is None:
            raise ValueError(f"{path}: not a dataset")
        count = int(np.prod(shape)) if shape else 1
        p, ver, cls = layout
        if ver not in (3, 4):
            raise NotImplementedError(f"data layout message version {ver}")
        if cls == 0:  # compact
            n = self._u(p + 2, 2)
            raw = b[p + 4:p + 4 + n]
        elif cls == 1:  # contiguous
            a = self._off(p + 2)
            if a == _UNDEF:
                return np.zeros(shape, dtype=dtype.newbyteorder("="))
            raw = b[self._addr(a):self._addr(a) + count * dtype.itemsize]
        elif cls == 2 and ver == 3:  # chunked, version-1 B-tree index
            rank = b[p + 2] - 1
            tree = self._off(p + 3)
            chunk = tuple(self._u(p + 3 + self.so + 4 * k, 4) for k in range(rank))
            out = np.zeros(shape, dtype=dtype)
            if tree != _UNDEF:
                self._btree_chunks(tree, rank, chunk, dtype, out)
            return out.astype(dtype.newbyteorder("="))
        else:
            raise NotImplementedError(f"data layout class {cls} (version {ver})")
        return np.frombuffer(raw, dtype=dtype, count=count).reshape(shape).astype(
            dtype.newbyteorder("="))

    def _btree_chunks(self, addr, rank, chunk, dtype, out):
        b, p = self.buf, self._addr(addr)
        if b[p:p + 4] != b"TREE" or b[p + 4] != 1:
            raise ValueError("bad chunk B-tree node")
        level, n = b[p + 5], self._u(p + 6, 2)
        ksz = 8 + 8 * (rank + 1)
        q = p + 8 + 2 * self.so
        for _ in range(n):
            size, mask = self._u(q, 4), self._u(q + 4, 4)
            offs = [self._u(q + 8 + 8 * k, 8) for k in range(rank)]
            child = self._off(q + ksz)
            q += ksz + self.so
            if level > 0:
                self._btree_chunks(child, rank, chunk, dtype, out)
                continue
            if mask:
                raise NotImplementedError("filtered chunks")
            c = self._addr(child)
            blk = np.frombuffer(b[c:c + size], dtype=dtype,
                                count=int(np.prod(chunk))).reshape(chunk)
            sl = tuple(slice(o, min(o + k, s)) for o, k, s in zip(offs, chunk, out.shape))
            out[sl] = blk[tuple(slice(0, x.stop - x.start) for x in sl)]


def _data_item(item, base_dir, files):
    fmt = item.get("Format", "XML")
    dims = tuple(int(x) for x in item.get("Dimensions", "").split())
    text = (item.text or "").strip()
    if fmt == "HDF":
        fname, dset = text.split(":", 1)
        fpath = os.path.join(base_dir, fname)
        if fpath not in files:
            files[fpath] = H5File(fpath)
        a = files[fpath][dset]
    elif fmt == "XML":
        kind = item.get("DataType", "Float")
        a = np.array(text.split(), dtype=np.float64 if kind == "Float" else np.int64)
    else:
        raise NotImplementedError(f"XDMF DataItem format {fmt}")
    if dims and int(np.prod(dims)) == a.size:
        a = a.reshape(dims)
    return a


def read_xdmf(path):
    """(conn int32 C x 4, xyz float64 V x 3, cell attribute int32 C or None) of the first grid
    of a tetrahedral XDMF file."""
    root = ET.parse(path).getroot()
    grid = root.find(".//Grid")
    if grid is None:
        raise ValueError(f"{path}: no Grid")
    base_dir = os.path.dirname(os.path.abspath(path))
    files = {}
    topo = grid.find("Topology")
    if topo is None or topo.get("TopologyType", "").lower() != "tetrahedron":
        raise NotImplementedError("only Tetrahedron topologies")
    geo = grid.find("Geometry")
    if geo is None or geo.get("GeometryType", "XYZ") != "XYZ":
        raise NotImplementedError("only XYZ geometries")
    conn = _data_item(topo.find("DataItem"), base_dir, files).reshape(-1, 4)
    xyz = _data_item(geo.find("DataItem"), base_dir, files).reshape(-1, 3).astype(np.float64)
    tags = None
    for at in grid.findall("Attribute"):
        if at.get("Center") == "Cell":
            tags = _data_item(at.find("DataItem"), base_dir, files).reshape(-1).astype(np.int32)
            break
    if conn.min() < 0 or conn.max() >= len(xyz):
        raise ValueError(f"{path}: connectivity out of range")
    return np.ascontiguousarray(conn, dtype=np.int32), np.ascontiguousarray(xyz), tags


def load_mesh(xdmf_path: str, comm=None):
    """(dsv.core.Mesh on the current GPU, cell tags or None), the reference's load_mesh
    (`domain/mesh_utils.py:197-267`) for a single process."""
    from .core import Mesh
    conn, xyz, tags = read_xdmf(xdmf_path)
    return Mesh(conn, xyz), tags
