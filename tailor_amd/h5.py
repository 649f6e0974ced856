"""Minimal HDF5 reader and writer for the case cache and the eigenmode output.

The image has no libhdf5 binding for the product interpreter, so this module
implements the subset of the HDF5 file format (version 0 superblock, version 1
object headers, symbol-table groups) that the two files of the path use:

* reading: groups (symbol tables, and link messages of newer files), datasets
  with contiguous, compact or unfiltered chunked storage, attributes with
  numeric, fixed-length string and variable-length string types; that is what
  PETSc's HDF5 viewer and h5py write for the schema-v2 cache
  (Prepare.cpp:443-668 reads it);
* writing: nested groups, contiguous datasets and scalar attributes (64-bit
  integers, doubles, fixed-length null-terminated strings), which is the
  layout of the schema-v1 eigenmode file of EigenOutput.cpp:85-304.

Layouts follow the published HDF5 File Format Specification (version 2.0 of
the document, sections II-IV).  Complex vectors use PETSc's convention: a
trailing axis of length 2 holding (re, im) and an integer attribute
``complex = 1`` on the dataset.
"""
from __future__ import annotations

import os
import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(RuntimeError):
    pass


# --------------------------------------------------------------------------
# reading
# --------------------------------------------------------------------------

@dataclass
class Dataset:
    name: str
    data: np.ndarray
    attrs: dict = field(default_factory=dict)


@dataclass
class Group:
    name: str
    members: dict = field(default_factory=dict)
    attrs: dict = field(default_factory=dict)

    def __getitem__(self, path: str):
        node = self
        for part in [p for p in path.split("/") if p]:
            if not isinstance(node, Group) or part not in node.members:
                raise KeyError(path)
            node = node.members[part]
        return node

    def __contains__(self, path: str) -> bool:
        try:
            self[path]
            return True
        except KeyError:
            return False


class _Reader:
    def __init__(self, buf: bytes):
        self.b = buf
        if buf[:8] != SIGNATURE:
            raise H5Error("not an HDF5 file (signature missing)")
        version = buf[8]
        if version not in (0, 1):
            raise H5Error(f"unsupported HDF5 superblock version {version}")
        if buf[13] != 8 or buf[14] != 8:
            raise H5Error("only 8-byte offsets and lengths are supported")
        pos = 24 + (4 if version == 1 else 0)
        self.base = self.u64(pos)
        root_entry = pos + 32
        self.root_header = self.u64(root_entry + 8)

    def u8(self, p):
        return self.b[p]

    def u16(self, p):
        return struct.unpack_from("<H", self.b, p)[0]

    def u32(self, p):
        return struct.unpack_from("<I", self.b, p)[0]

    def u64(self, p):
        return struct.unpack_from("<Q", self.b, p)[0]

    # object headers ------------------------------------------------------
    def messages(self, addr: int) -> list[tuple[int, bytes]]:
        b = self.b
        if b[addr:addr + 4] == b"OHDR":
            return self._messages_v2(addr)
        if b[addr] != 1:
            raise H5Error(f"unsupported object header version {b[addr]} at {addr}")
        nmsg = self.u16(addr + 2)
        size = self.u32(addr + 8)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(out) < nmsg:
                mtype, msize, flags = self.u16(p), self.u16(p + 2), b[p + 4]
                data = b[p + 8:p + 8 + msize]
                if flags & 0x02:
                    raise H5Error("shared object header messages are not supported")
                if mtype == 0x10:
                    blocks.append((struct.unpack_from("<Q", data, 0)[0],
                                   struct.unpack_from("<Q", data, 8)[0]))
                out.append((mtype, data))
                p += 8 + msize
        return out

    def _messages_v2(self, addr: int) -> list[tuple[int, bytes]]:
        b = self.b
        flags = b[addr + 5]
        p = addr + 6
        if flags & 0x20:
            p += 16
        if flags & 0x10:
            p += 4
        width = 1 << (flags & 0x03)
        size = int.from_bytes(b[p:p + width], "little")
        p += width
        out = []
        blocks = [(p, size, False)]
        while blocks:
            start, length, is_cont = blocks.pop(0)
            q = start + (4 if is_cont else 0)
            end = start + length - (4 if is_cont else 0)
            while q + 4 <= end:
                mtype, msize, mflags = b[q], struct.unpack_from("<H", b, q + 1)[0], b[q + 3]
                q += 4
                if flags & 0x04:
                    q += 2
                data = b[q:q + msize]
                if mflags & 0x02:
                    raise H5Error("shared object header messages are not supported")
                if mtype == 0x10:
                    blocks.append((struct.unpack_from("<Q", data, 0)[0],
                                   struct.unpack_from("<Q", data, 8)[0], True))
                out.append((mtype, data))
                q += msize
        return out

    # datatypes and dataspaces --------------------------------------------
    @staticmethod
    def datatype(d: bytes) -> tuple[dict, int]:
        cls, bits, size = d[0] & 0x0F, d[1:4], struct.unpack_from("<I", d, 4)[0]
        if cls == 0:
            endian = ">" if bits[0] & 1 else "<"
            kind = "i" if bits[0] & 0x08 else "u"
            return {"kind": "num", "dtype": np.dtype(f"{endian}{kind}{size}")}, 12
        if cls == 1:
            endian = ">" if bits[0] & 1 else "<"
            return {"kind": "num", "dtype": np.dtype(f"{endian}f{size}")}, 20
        if cls == 3:
            return {"kind": "str", "size": size}, 8
        if cls == 9:
            base, blen = _Reader.datatype(d[8:])
            if bits[0] & 0x0F == 1:
                return {"kind": "vlenstr", "size": 16}, 8 + blen
            raise H5Error("variable-length sequences are not supported")
        raise H5Error(f"unsupported HDF5 datatype class {cls}")

    @staticmethod
    def dataspace(d: bytes) -> tuple[int, ...]:
        version, rank, flags = d[0], d[1], d[2]
        if version == 1:
            p = 8
        elif version == 2:
            if d[3] == 2:
                return ()
            p = 4
        else:
            raise H5Error(f"unsupported dataspace version {version}")
        return tuple(struct.unpack_from("<Q", d, p + 8 * k)[0] for k in range(rank))

    def decode(self, raw: bytes, dt: dict, shape: tuple[int, ...]):
        count = int(np.prod(shape)) if shape else 1
        if dt["kind"] == "num":
            arr = np.frombuffer(raw, dtype=dt["dtype"], count=count).astype(
                dt["dtype"].newbyteorder("="))
            return arr.reshape(shape) if shape else arr[0].item()
        if dt["kind"] == "str":
            vals = [raw[k * dt["size"]:(k + 1) * dt["size"]].split(b"\0", 1)[0].decode()
                    for k in range(count)]
            return np.array(vals).reshape(shape) if shape else vals[0]
        vals = []
        for k in range(count):
            length, gaddr, index = struct.unpack_from("<IQI", raw, 16 * k)
            vals.append(self.global_heap_object(gaddr, index)[:length].decode())
        return np.array(vals).reshape(shape) if shape else vals[0]

    def global_heap_object(self, addr: int, index: int) -> bytes:
        b = self.b
        if b[addr:addr + 4] != b"GCOL":
            raise H5Error("bad global heap collection")
        end = addr + self.u64(addr + 8)
        p = addr + 16
        while p + 16 <= end:
            idx, size = self.u16(p), self.u64(p + 8)
            if idx == index:
                return bytes(b[p + 16:p + 16 + size])
            if idx == 0:
                break
            p += 16 + ((size + 7) & ~7)
        raise H5Error(f"global heap object {index} not found")

    def attribute(self, d: bytes) -> tuple[str, object]:
        version = d[0]
        nsize, tsize, ssize = struct.unpack_from("<HHH", d, 2)
        if version == 1:
            pad = lambda n: (n + 7) & ~7
            p = 8
            name = d[p:p + nsize].split(b"\0", 1)[0].decode()
            p += pad(nsize)
            dt, _ = self.datatype(d[p:p + tsize])
            p += pad(tsize)
            shape = self.dataspace(d[p:p + ssize])
            p += pad(ssize)
        elif version in (2, 3):
            p = 8 + (1 if version == 3 else 0)
            name = d[p:p + nsize].split(b"\0", 1)[0].decode()
            p += nsize
            dt, _ = self.datatype(d[p:p + tsize])
            p += tsize
            shape = self.dataspace(d[p:p + ssize])
            p += ssize
        else:
            raise H5Error(f"unsupported attribute message version {version}")
        return name, self.decode(bytes(d[p:]), dt, shape)

    # storage ---------------------------------------------------------------
    def chunked(self, btree: int, shape, chunk, esize: int) -> bytes:
        out = np.zeros(int(np.prod(shape)) * esize, dtype=np.uint8)
        full = np.asarray(shape, dtype=np.int64)
        ch = np.asarray(chunk, dtype=np.int64)
        rank = len(shape)
        view = out.reshape(tuple(shape) + (esize,))

        def walk(addr):
            b = self.b
            if b[addr:addr + 4] != b"TREE" or b[addr + 4] != 1:
                raise H5Error("bad chunk B-tree node")
            level, used = b[addr + 5], self.u16(addr + 6)
            p = addr + 24
            ksize = 8 + 8 * (rank + 1)
            for _ in range(used):
                csize, mask = self.u32(p), self.u32(p + 4)
                offs = [self.u64(p + 8 + 8 * k) for k in range(rank)]
                child = self.u64(p + ksize)
                if level > 0:
                    walk(child)
                else:
                    if mask != 0:
                        raise H5Error("filtered chunks are not supported")
                    raw = np.frombuffer(self.b, dtype=np.uint8, count=int(np.prod(ch)) * esize,
                                        offset=child).reshape(tuple(ch) + (esize,))
                    sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, ch, full))
                    sub = tuple(slice(0, s.stop - s.start) for s in sl)
                    view[sl] = raw[sub]
                p += ksize + 8
        walk(btree)
        return out.tobytes()

    # objects ---------------------------------------------------------------
    def read_object(self, addr: int, name: str):
        msgs = self.messages(addr)
        types = {t for t, _ in msgs}
        attrs = {}
        for t, d in msgs:
            if t == 0x0C:
                k, v = self.attribute(d)
                attrs[k] = v
            elif t == 0x15:
                raise H5Error("dense attribute storage is not supported")
            elif t == 0x0B:
                raise H5Error("filtered datasets are not supported")
        if 0x11 in types or 0x06 in types or 0x02 in types:
            group = Group(name, attrs=attrs)
            for t, d in msgs:
                if t == 0x11:
                    btree, heap = struct.unpack_from("<QQ", d, 0)
                    for child_name, child_addr in self.symbol_table(btree, heap):
                        group.members[child_name] = self.read_object(child_addr, child_name)
                elif t == 0x06:
                    child_name, child_addr = self.link(d)
                    group.members[child_name] = self.read_object(child_addr, child_name)
                elif t == 0x02 and len(d) > 10 and struct.unpack_from("<Q", d, 2 + (8 if d[1] & 1 else 0))[0] != UNDEF:
                    raise H5Error("dense link storage is not supported")
            return group
        space = dtype = layout = None
        for t, d in msgs:
            if t == 0x01:
                space = self.dataspace(d)
            elif t == 0x03:
                dtype, _ = self.datatype(d)
            elif t == 0x08:
                layout = d
        if space is None or dtype is None or layout is None:
            raise H5Error(f"object {name!r} is neither a group nor a dataset")
        esize = dtype["dtype"].itemsize if dtype["kind"] == "num" else dtype["size"]
        nbytes = int(np.prod(space)) * esize if space else esize
        version, cls = layout[0], layout[1]
        if version != 3:
            raise H5Error(f"unsupported data layout version {version}")
        if cls == 0:
            size = struct.unpack_from("<H", layout, 2)[0]
            raw = bytes(layout[4:4 + size])
        elif cls == 1:
            a = struct.unpack_from("<Q", layout, 2)[0]
            raw = bytes(nbytes) if a == UNDEF else bytes(self.b[a:a + nbytes])
        elif cls == 2:
            rank1 = layout[2]
            btree = struct.unpack_from("<Q", layout, 3)[0]
            chunk = struct.unpack_from(f"<{rank1}I", layout, 11)[:-1]
            raw = bytes(nbytes) if btree == UNDEF else self.chunked(btree, space, chunk, esize)
        else:
            raise H5Error(f"unsupported layout class {cls}")
        return Dataset(name, np.asarray(self.decode(raw, dtype, space)), attrs)

    def symbol_table(self, btree: int, heap: int):
        b = self.b
        if b[heap:heap + 4] != b"HEAP":
            raise H5Error("bad local heap")
        seg = self.u64(heap + 24)
        names = []

        def walk(addr):
            if b[addr:addr + 4] != b"TREE" or b[addr + 4] != 0:
                raise H5Error("bad group B-tree node")
            level, used = b[addr + 5], self.u16(addr + 6)
            for k in range(used):
                child = self.u64(addr + 24 + 8 + 16 * k)
                if level > 0:
                    walk(child)
                    continue
                if b[child:child + 4] != b"SNOD":
                    raise H5Error("bad symbol table node")
                for e in range(self.u16(child + 6)):
                    p = child + 8 + 40 * e
                    off, hdr = self.u64(p), self.u64(p + 8)
                    end = b.index(b"\0", seg + off)
                    names.append((b[seg + off:end].decode(), hdr))
        walk(btree)
        return names

    def link(self, d: bytes) -> tuple[str, int]:
        flags = d[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = d[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        width = 1 << (flags & 0x03)
        nlen = int.from_bytes(d[p:p + width], "little")
        p += width
        name = bytes(d[p:p + nlen]).decode()
        p += nlen
        if ltype != 0:
            raise H5Error(f"link {name!r}: only hard links are supported")
        return name, struct.unpack_from("<Q", d, p)[0]


def read(path: str | os.PathLike) -> Group:
    """Read a whole (small) HDF5 file into a tree of Group/Dataset objects."""
    buf = Path(path).read_bytes()
    r = _Reader(buf)
    root = r.read_object(r.root_header, "/")
    if not isinstance(root, Group):
        raise H5Error("root object is not a group")
    return root


# --------------------------------------------------------------------------
# writing
# --------------------------------------------------------------------------

def _pad8(n: int) -> int:
    return (n + 7) & ~7


def _msg(mtype: int, data: bytes) -> bytes:
    data = data + b"\0" * (_pad8(len(data)) - len(data))
    return struct.pack("<HHB3x", mtype, len(data), 0) + data


def _dtype_bytes(value) -> tuple[bytes, bytes, int]:
    """(datatype message, raw data of one element or array, element size)."""
    if isinstance(value, str):
        raw = value.encode() + b"\0"
        return struct.pack("<B3sI", 0x13, b"\0\0\0", len(raw)), raw, len(raw)
    arr = np.asarray(value)
    if arr.dtype.kind in "iub":
        arr = arr.astype("<i8")
        return struct.pack("<B3sIHH", 0x10, b"\x08\0\0", 8, 0, 64), arr.tobytes(), 8
    if arr.dtype.kind == "f":
        arr = arr.astype("<f8")
        return (struct.pack("<B3sIHHBBBBI", 0x11, b"\x20\x3f\0", 8, 0, 64, 52, 11, 0, 52, 1023),
                arr.tobytes(), 8)
    raise H5Error(f"cannot write values of type {arr.dtype}")


def _space_bytes(shape: tuple[int, ...]) -> bytes:
    return struct.pack("<BBBx4x", 1, len(shape), 0) + b"".join(
        struct.pack("<Q", int(s)) for s in shape)


def _attr_msg(name: str, value) -> bytes:
    dt, raw, _ = _dtype_bytes(value)
    shape = () if isinstance(value, str) or np.ndim(value) == 0 else np.shape(value)
    sp = _space_bytes(shape)
    nm = name.encode() + b"\0"
    body = struct.pack("<BxHHH", 1, len(nm), len(dt), len(sp))
    body += nm + b"\0" * (_pad8(len(nm)) - len(nm))
    body += dt + b"\0" * (_pad8(len(dt)) - len(dt))
    body += sp + b"\0" * (_pad8(len(sp)) - len(sp))
    return _msg(0x0C, body + raw)


class _Writer:
    SYM_LEAF_K = 4       # symbol table nodes hold up to 2K = 8 entries
    BTREE_K = 16         # group B-tree nodes hold up to 2K = 32 children

    def __init__(self):
        self.buf = bytearray(96)

    def alloc(self, data: bytes) -> int:
        addr = len(self.buf)
        self.buf += data
        self.buf += b"\0" * (_pad8(len(self.buf)) - len(self.buf))
        return addr

    def header(self, msgs: list[bytes]) -> int:
        body = b"".join(msgs)
        return self.alloc(struct.pack("<BxHII4x", 1, len(msgs), 1, len(body)) + body)

    def dataset(self, data: np.ndarray, attrs: dict) -> int:
        arr = np.ascontiguousarray(data)
        dt, raw, _ = _dtype_bytes(arr)
        addr = self.alloc(raw) if raw else UNDEF
        msgs = [_msg(0x01, _space_bytes(arr.shape)), _msg(0x03, dt),
                _msg(0x05, bytes([2, 1, 2, 0])),
                _msg(0x08, struct.pack("<BBQQ", 3, 1, addr, len(raw)))]
        msgs += [_attr_msg(k, v) for k, v in attrs.items()]
        return self.header(msgs)

    def group(self, node: Group) -> tuple[int, int, int]:
        children = []
        for name in sorted(node.members, key=lambda s: s.encode()):
            child = node.members[name]
            if isinstance(child, Group):
                hdr, bt, hp = self.group(child)
                children.append((name, hdr, (bt, hp)))
            else:
                children.append((name, self.dataset(child.data, child.attrs), None))
        if len(children) > 2 * self.SYM_LEAF_K * 2 * self.BTREE_K:
            raise H5Error("too many members in one group")
        # local heap: offset 0 is the empty name
        heap_data = bytearray(8)
        offsets = []
        for name, _, _ in children:
            offsets.append(len(heap_data))
            nm = name.encode() + b"\0"
            heap_data += nm + b"\0" * (_pad8(len(nm)) - len(nm))
        heap_data += b"\0" * 8
        seg = self.alloc(bytes(heap_data))
        heap = self.alloc(b"HEAP" + bytes([0, 0, 0, 0]) +
                          struct.pack("<QQQ", len(heap_data), 1, seg))
        # symbol table nodes of up to 2K entries each
        per = 2 * self.SYM_LEAF_K
        snods, keys = [], [0]
        for s in range(0, max(len(children), 1), per):
            part = children[s:s + per]
            ent = b""
            for (name, hdr, scratch), off in zip(part, offsets[s:s + per]):
                if scratch is None:
                    ent += struct.pack("<QQI4x16x", off, hdr, 0)
                else:
                    ent += struct.pack("<QQI4xQQ", off, hdr, 1, *scratch)
            ent += b"\0" * (40 * per - len(ent))
            snods.append(self.alloc(b"SNOD" + bytes([1, 0]) + struct.pack("<H", len(part)) + ent))
            keys.append(offsets[s + len(part) - 1] if part else 0)
        if not children:
            snods, keys = [], [0]
        tree = b"TREE" + bytes([0, 0]) + struct.pack("<HQQ", len(snods), UNDEF, UNDEF)
        body = struct.pack("<Q", keys[0])
        for child, key in zip(snods, keys[1:]):
            body += struct.pack("<QQ", child, key)
        cap = 8 * (2 * self.BTREE_K + 1) + 8 * 2 * self.BTREE_K
        tree += body + b"\0" * (cap - len(body))
        btree = self.alloc(tree)
        msgs = [_msg(0x11, struct.pack("<QQ", btree, heap))]
        msgs += [_attr_msg(k, v) for k, v in node.attrs.items()]
        return self.header(msgs), btree, heap

    def finish(self, root: Group) -> bytes:
        hdr, btree, heap = self.group(root)
        sb = bytearray(SIGNATURE)
        sb += bytes([0, 0, 0, 0, 0, 8, 8, 0])
        sb += struct.pack("<HHI", self.SYM_LEAF_K, self.BTREE_K, 0)
        sb += struct.pack("<QQQQ", 0, UNDEF, len(self.buf), UNDEF)
        sb += struct.pack("<QQI4xQQ", 0, hdr, 1, btree, heap)
        assert len(sb) == 96
        self.buf[:96] = sb
        return bytes(self.buf)


def write(path: str | os.PathLike, root: Group) -> None:
    """Write a tree of Group/Dataset objects as an HDF5 file."""
    Path(path).write_bytes(_Writer().finish(root))


def complex_dataset(name: str, values: np.ndarray) -> Dataset:
    """PETSc's complex-vector layout: trailing (re, im) axis, ``complex = 1``."""
    v = np.asarray(values, dtype=np.complex128)
    return Dataset(name, np.stack([v.real, v.imag], axis=-1), {"complex": 1})


def as_complex(ds: Dataset) -> np.ndarray:
    """Inverse of :func:`complex_dataset`."""
    a = np.asarray(ds.data, dtype=np.float64)
    if a.shape[-1] != 2:
        raise H5Error(f"dataset {ds.name!r} has no trailing (re, im) axis")
    return a[..., 0] + 1j * a[..., 1]
