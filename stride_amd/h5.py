"""Mosaic's HDF5 file protocol 0.1 (mosaic/file_manipulation/h5.py:14-330), without h5py.

Stride dumps and loads Traces, ScalarField and wavefield data through Mosaic's ``HDF5``
class: a nested description (dicts become groups, lists of dicts become groups of
``<name>_<index:08d>`` groups, everything else a dataset) with the attributes ``protocol``,
``datetime`` and ``is_array`` on groups and ``is_ndarray``, ``is_list``, ``is_tuple``,
``is_str``, ``is_bytes``, ``is_none`` on datasets (h5.py:34-114). This module reads and writes
that layout with numpy and the standard library only (h5py is not part of this image).

HDF5 coverage, from the published HDF5 file format specification (version 3.0):
- read: superblock versions 0-3; object headers v1 and v2 (with continuation blocks);
  groups as symbol tables (v1 B-tree + local heap, the format of files written with the
  default library bounds, e.g. the breast2D fixture) or as compact link messages (files written
  with ``libver='latest'``, as Mosaic does, h5.py:305); contiguous and compact datasets;
  fixed-point, floating-point, fixed and variable-length string and enum (h5py's bool)
  datatypes; links and attributes stored compact or dense (fractal heap + v2 B-tree, one
  level).
- write: superblock v2, v2 object headers, compact links (any number: the group info message
  raises the compact threshold) and attributes, contiguous datasets; strings as fixed-length
  UTF-8, booleans as h5py's enum.
- not supported (raises): chunked or filtered (compressed) datasets, dense storage whose
  B-tree has more than one level, compound and reference types.
"""

import datetime as _dt
import mmap
import os
import struct

import numpy as np

__all__ = ['read', 'write', 'HDF5', 'file_exists', 'read_file', 'write_file', 'lookup3']

PROTOCOL = '0.1'
_SIG = b'\x89HDF\r\n\x1a\n'
_UNDEF = 0xFFFFFFFFFFFFFFFF


class HDF5FormatError(RuntimeError):
    pass


# ----------------------------------------------------------------------------- checksum ----

def _rot(x, k):
    return ((x << k) | (x >> (32 - k))) & 0xFFFFFFFF


def lookup3(data, initval=0):
    """Bob Jenkins' lookup3 ``hashlittle`` (byte-wise form), the checksum of HDF5 v2 metadata."""
    M = 0xFFFFFFFF
    n = len(data)
    a = b = c = (0xDEADBEEF + n + initval) & M
    k = bytes(data)
    i = 0
    while n > 12:
        a = (a + k[i] + (k[i + 1] << 8) + (k[i + 2] << 16) + (k[i + 3] << 24)) & M
        b = (b + k[i + 4] + (k[i + 5] << 8) + (k[i + 6] << 16) + (k[i + 7] << 24)) & M
        c = (c + k[i + 8] + (k[i + 9] << 8) + (k[i + 10] << 16) + (k[i + 11] << 24)) & M
        a = (a - c) & M; a ^= _rot(c, 4); c = (c + b) & M
        b = (b - a) & M; b ^= _rot(a, 6); a = (a + c) & M
        c = (c - b) & M; c ^= _rot(b, 8); b = (b + a) & M
        a = (a - c) & M; a ^= _rot(c, 16); c = (c + b) & M
        b = (b - a) & M; b ^= _rot(a, 19); a = (a + c) & M
        c = (c - b) & M; c ^= _rot(b, 4); b = (b + a) & M
        n -= 12
        i += 12
    if n == 0:
        return c
    tail = k[i:i + n] + bytes(12 - n)
    a = (a + int.from_bytes(tail[0:4], 'little')) & M
    b = (b + int.from_bytes(tail[4:8], 'little')) & M
    c = (c + int.from_bytes(tail[8:12], 'little')) & M
    c ^= b; c = (c - _rot(b, 14)) & M
    a ^= c; a = (a - _rot(c, 11)) & M
    b ^= a; b = (b - _rot(a, 25)) & M
    c ^= b; c = (c - _rot(b, 16)) & M
    a ^= c; a = (a - _rot(c, 4)) & M
    b ^= a; b = (b - _rot(a, 14)) & M
    c ^= b; c = (c - _rot(b, 24)) & M
    return c


# ------------------------------------------------------------------------------ reading ----

class _Dtype:
    """A parsed datatype message: numpy dtype, or a string / vlen-string / bool marker."""

    def __init__(self, kind, np_dtype=None, size=0, base=None, enum=None):
        self.kind = kind          # 'num', 'str', 'vlen_str', 'bool', 'enum'
        self.np = np_dtype
        self.size = size
        self.base = base
        self.enum = enum


class _Reader:
    def __init__(self, buf):
        self.b = buf
        base = 0
        while base < len(buf) and buf[base:base + 8] != _SIG:
            base = 512 if base == 0 else base * 2
        if base >= len(buf):
            raise HDF5FormatError('not an HDF5 file')
        self.base = base
        v = buf[base + 8]
        if v in (0, 1):
            self.so, self.sl = buf[base + 13], buf[base + 14]
            p = base + 24 + (4 if v == 1 else 0)
            p += 4 * self.so  # base address, free-space info, EOF, driver info
            self.root = self._entry(p)[1]
        elif v in (2, 3):
            self.so, self.sl = buf[base + 9], buf[base + 10]
            p = base + 12
            end = p + 4 * self.so
            if lookup3(buf[base:end]) != struct.unpack_from('<I', buf, end)[0]:
                raise HDF5FormatError('superblock checksum mismatch')
            self.root = self._off(p + 3 * self.so)
        else:
            raise HDF5FormatError('superblock version %d' % v)
        if self.so != 8 or self.sl != 8:
            raise HDF5FormatError('only 8-byte offsets and lengths are supported')

    def _u(self, p, n):
        return int.from_bytes(self.b[p:p + n], 'little')

    def _off(self, p):
        return self._u(p, self.so)

    def _nul(self, p):
        end = self.b.find(b'\0', p)
        if end < 0:
            raise HDF5FormatError('unterminated string')
        return end

    def _entry(self, p):
        """symbol table entry -> (link name offset, object header address)"""
        return self._off(p), self._off(p + self.so)

    # object headers ---------------------------------------------------------------------
    def messages(self, addr):
        b = self.b
        out = []
        if b[addr:addr + 4] == b'OHDR':
            flags = b[addr + 5]
            p = addr + 6
            if flags & 0x20:
                p += 16
            if flags & 0x10:
                p += 4
            wid = 1 << (flags & 3)
            size = self._u(p, wid)
            p += wid
            blocks = [(p, p + size, flags)]
            while blocks:
                start, stop, fl = blocks.pop(0)
                q = start
                while q + 4 <= stop:
                    mtype, msize, mflags = b[q], self._u(q + 1, 2), b[q + 3]
                    q += 4 + (2 if fl & 0x04 else 0)
                    if mtype == 0x10:
                        cont, clen = self._off(q), self._u(q + self.so, self.sl)
                        if b[cont:cont + 4] != b'OCHK':
                            raise HDF5FormatError('bad continuation block')
                        blocks.append((cont + 4, cont + clen - 4, fl))
                    else:
                        out.append((mtype, q, msize, mflags))
                    q += msize
        elif b[addr] == 1:
            nmsg = self._u(addr + 2, 2)
            size = self._u(addr + 8, 4)
            blocks = [(addr + 16, addr + 16 + size)]
            while blocks and len(out) < nmsg:
                start, stop = blocks.pop(0)
                q = start
                while q + 8 <= stop:
                    mtype, msize, mflags = self._u(q, 2), self._u(q + 2, 2), b[q + 4]
                    q += 8
                    if mtype == 0x10:
                        blocks.append((self._off(q), self._off(q) + self._u(q + self.so, self.sl)))
                    elif mtype != 0:
                        out.append((mtype, q, msize, mflags))
                    q += msize
        else:
            raise HDF5FormatError('object header at %d' % addr)
        for m in out:
            if m[3] & 0x02:
                raise HDF5FormatError('shared messages are not supported')
        return out

    # datatypes and dataspaces -----------------------------------------------------------
    def datatype(self, p):
        b = self.b
        cls, ver = b[p] & 0x0F, b[p] >> 4
        bits = self._u(p + 1, 3)
        size = self._u(p + 4, 4)
        q = p + 8
        order = '>' if bits & 1 else '<'
        if cls == 0:
            kind = ('i' if bits & 0x08 else 'u') + str(size)
            return _Dtype('num', np.dtype(order + kind), size), q + 4
        if cls == 1:
            if size not in (2, 4, 8):
                raise HDF5FormatError('float of %d bytes' % size)
            return _Dtype('num', np.dtype(order + 'f' + str(size)), size), q + 12
        if cls == 3:
            return _Dtype('str', np.dtype('S%d' % size), size), q
        if cls == 8:
            nmem = bits & 0xFFFF
            base, q = self.datatype(q)
            names = []
            for _ in range(nmem):  # null-terminated; padded to a multiple of 8 bytes before version 3
                end = self._nul(q)
                names.append(bytes(b[q:end]).decode())
                q = end + 1 if ver >= 3 else q + (end - q + 1 + 7) // 8 * 8
            vals = [self._u(q + i * base.size, base.size) for i in range(nmem)]
            q += nmem * base.size
            enum = dict(zip(vals, names))
            if sorted(names) == ['FALSE', 'TRUE'] and base.size == 1:
                return _Dtype('bool', np.dtype(bool), size, base, enum), q
            return _Dtype('enum', base.np, size, base, enum), q
        if cls == 9:
            base, q = self.datatype(q)
            if bits & 0x0F == 1:
                return _Dtype('vlen_str', None, size, base), q
            raise HDF5FormatError('variable-length sequences are not supported')
        raise HDF5FormatError('datatype class %d is not supported' % cls)

    def dataspace(self, p):
        b = self.b
        ver, rank, flags = b[p], b[p + 1], b[p + 2]
        if ver == 1:
            q = p + 8
            kind = 1 if rank else 0
        else:
            kind = b[p + 3]
            q = p + 4
        dims = tuple(self._u(q + i * self.sl, self.sl) for i in range(rank))
        if kind == 2:
            return None
        return dims

    # values -----------------------------------------------------------------------------
    def _gheap(self, coll, index):
        b = self.b
        if b[coll:coll + 4] != b'GCOL':
            raise HDF5FormatError('bad global heap collection')
        size = self._u(coll + 8, self.sl)
        q = coll + 8 + self.sl
        while q < coll + size:
            idx = self._u(q, 2)
            osize = self._u(q + 8, self.sl)
            if idx == 0:
                break
            if idx == index:
                return bytes(b[q + 8 + self.sl:q + 8 + self.sl + osize])
            q += 8 + self.sl + (osize + 7) // 8 * 8
        raise HDF5FormatError('global heap object %d not found' % index)

    def values(self, dt, dims, raw):
        count = int(np.prod(dims)) if dims else 1
        if dt.kind == 'vlen_str':
            out = []
            for i in range(count):
                rec = raw[i * 16:(i + 1) * 16]
                length = int.from_bytes(rec[0:4], 'little')
                coll = int.from_bytes(rec[4:12], 'little')
                idx = int.from_bytes(rec[12:16], 'little')
                out.append(self._gheap(coll, idx)[:length] if length else b'')
            arr = np.empty(count, dtype=object)
            arr[:] = out
        elif dt.kind == 'bool':
            arr = np.frombuffer(raw, dtype=np.uint8, count=count).astype(bool)
        elif dt.kind == 'enum':
            arr = np.frombuffer(raw, dtype=dt.np, count=count).copy()
        else:
            arr = np.frombuffer(raw, dtype=dt.np, count=count).copy()
            if dt.np.byteorder == '>':
                arr = arr.astype(dt.np.newbyteorder('='))
        return arr.reshape(dims) if dims else arr.reshape(())

    def attribute(self, p):
        b = self.b
        ver = b[p]
        nlen, tlen, slen = self._u(p + 2, 2), self._u(p + 4, 2), self._u(p + 6, 2)
        q = p + 8 + (1 if ver == 3 else 0)
        pad = (lambda n: (n + 7) // 8 * 8) if ver == 1 else (lambda n: n)
        name = bytes(b[q:q + nlen]).split(b'\0')[0].decode()
        q += pad(nlen)
        dt, _ = self.datatype(q)
        q += pad(tlen)
        dims = self.dataspace(q)
        q += pad(slen)
        if dims is None:
            return name, None
        count = int(np.prod(dims)) if dims else 1
        return name, self.values(dt, dims, b[q:q + count * dt.size])

    # objects ----------------------------------------------------------------------------
    def obj(self, addr):
        """('group', {name: address}, attrs) or ('dataset', array, attrs)"""
        msgs = self.messages(addr)
        attrs, links = {}, {}
        dt = dims = layout = None
        is_group = False
        for mtype, p, size, _ in msgs:
            if mtype == 0x0C:
                k, v = self.attribute(p)
                attrs[k] = v
            elif mtype == 0x11:
                is_group = True
                links.update(self._symbol_table(self._off(p), self._off(p + self.so)))
            elif mtype == 0x02:
                is_group = True
                flags = self.b[p + 1]
                q = p + 2 + (8 if flags & 1 else 0)
                if self._off(q) != _UNDEF:  # dense storage: links in a fractal heap
                    for o in self._heap_objects(self._off(q), self._off(q + self.so)):
                        name, target = self._link(o)
                        if target is not None:
                            links[name] = target
            elif mtype == 0x06:
                is_group = True
                name, target = self._link(p)
                if target is not None:
                    links[name] = target
            elif mtype == 0x01:
                dims = self.dataspace(p)
            elif mtype == 0x03:
                dt, _ = self.datatype(p)
            elif mtype == 0x08:
                layout = p
            elif mtype == 0x0B:
                raise HDF5FormatError('filtered (compressed) datasets are not supported')
            elif mtype == 0x15:
                flags = self.b[p + 1]
                q = p + 2 + (2 if flags & 1 else 0)
                if self._off(q) != _UNDEF:  # dense storage: attributes in a fractal heap
                    for o in self._heap_objects(self._off(q), self._off(q + self.so)):
                        k, v = self.attribute(o)
                        attrs[k] = v
        if is_group:
            return 'group', links, attrs
        if dt is None or layout is None:
            raise HDF5FormatError('object at %d is neither a group nor a dataset' % addr)
        count = int(np.prod(dims)) if dims else 1
        nbytes = count * dt.size
        lver, lcls = self.b[layout], self.b[layout + 1]
        if lver < 3:
            raise HDF5FormatError('data layout version %d is not supported' % lver)
        mv = memoryview(self.b)
        if lcls == 0:
            raw = mv[layout + 4:layout + 4 + nbytes]
        elif lcls == 1:
            a = self._off(layout + 2)
            raw = bytes(nbytes) if a == _UNDEF else mv[self.base + a:self.base + a + nbytes]
        else:
            raise HDF5FormatError('chunked datasets are not supported')
        if dims is None:
            return 'dataset', None, attrs
        return 'dataset', self.values(dt, dims, raw), attrs

    def _nbytes(self, v):
        return (int(v).bit_length() + 7) // 8

    def _heap_objects(self, heap, btree):
        """Objects of a fractal heap named by the records of a v2 B-tree (link name index,
        type 5, or attribute name index, type 8): managed objects in direct blocks of a root
        direct block or of a root indirect block's direct rows."""
        b = self.b
        if b[heap:heap + 4] != b'FRHP' or b[btree:btree + 4] != b'BTHD':
            raise HDF5FormatError('bad dense storage headers')
        so, sl = self.so, self.sl
        idlen, flags = self._u(heap + 5, 2), b[heap + 9]
        if self._u(heap + 7, 2):
            raise HDF5FormatError('filtered fractal heaps are not supported')
        max_obj = self._u(heap + 10, 4)
        q = heap + 14 + sl + so + sl + so + 4 * sl + 4 * sl
        width = self._u(q, 2)
        start = self._u(q + 2, sl)
        max_direct = self._u(q + 2 + sl, sl)
        max_heap_bits = self._u(q + 2 + 2 * sl, 2)
        root = self._off(q + 6 + 2 * sl)
        rows = self._u(q + 6 + 2 * sl + so, 2)
        off_bytes = (max_heap_bits + 7) // 8
        len_bytes = self._nbytes(min(max_direct, max_obj))
        csum = 4 if flags & 0x02 else 0
        blocks = []  # (heap offset of block start, file address, size)

        def direct(addr, size):
            if b[addr:addr + 4] != b'FHDB':
                raise HDF5FormatError('bad fractal heap direct block')
            blocks.append((self._u(addr + 5 + so, off_bytes), addr, size))
        if rows == 0:
            direct(root, start)
        else:
            if b[root:root + 4] != b'FHIB':
                raise HDF5FormatError('bad fractal heap indirect block')
            p = root + 5 + so + off_bytes
            for r in range(rows):
                size = start if r < 2 else start << (r - 1)
                if size > max_direct:
                    raise HDF5FormatError('nested fractal heap blocks are not supported')
                for _ in range(width):
                    a = self._off(p)
                    p += so
                    if a != _UNDEF:
                        direct(a, size)
        btype, rec_size, depth = b[btree + 5], self._u(btree + 10, 2), self._u(btree + 12, 2)
        leaf, nrec = self._off(btree + 16), self._u(btree + 16 + so, 2)
        if depth != 0:
            raise HDF5FormatError('dense storage with a multi-level B-tree is not supported')
        if b[leaf:leaf + 4] != b'BTLF':
            raise HDF5FormatError('bad v2 B-tree leaf')
        out = []
        for i in range(nrec):
            r = leaf + 6 + i * rec_size
            hid = r + (4 if btype == 5 else 0)
            if (b[hid] >> 4) & 3:
                raise HDF5FormatError('huge or tiny heap objects are not supported')
            off = self._u(hid + 1, off_bytes)
            length = self._u(hid + 1 + off_bytes, len_bytes)
            for bo, addr, size in blocks:
                if bo <= off < bo + size:
                    out.append(addr + off - bo)
                    break
            else:
                raise HDF5FormatError('heap object outside the managed blocks')
            del length
        del csum
        return out

    def _link(self, p):
        b = self.b
        flags = b[p + 1]
        q = p + 2
        ltype = 0
        if flags & 0x08:
            ltype = b[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        wid = 1 << (flags & 3)
        n = self._u(q, wid)
        q += wid
        name = bytes(b[q:q + n]).decode('utf-8')
        q += n
        return name, (self._off(q) if ltype == 0 else None)

    def _symbol_table(self, btree, heap):
        b = self.b
        if b[heap:heap + 4] != b'HEAP':
            raise HDF5FormatError('bad local heap')
        data = self._off(heap + 8 + 2 * self.sl)
        out = {}

        def walk(node):
            if b[node:node + 4] != b'TREE':
                raise HDF5FormatError('bad group B-tree node')
            level, used = b[node + 5], self._u(node + 6, 2)
            q = node + 8 + 2 * self.so + self.sl  # first key
            for i in range(used):
                child = self._off(q)
                if level > 0:
                    walk(child)
                else:
                    if b[child:child + 4] != b'SNOD':
                        raise HDF5FormatError('bad symbol table node')
                    nsym = self._u(child + 6, 2)
                    e = child + 8
                    for _ in range(nsym):
                        noff, oh = self._entry(e)
                        start = data + noff
                        out[bytes(b[start:self._nul(start)]).decode('utf-8')] = oh
                        e += 2 * self.so + 24
                q += self.so + self.sl
        walk(btree)
        return out


def read_file(filename):
    """The whole file as nested dicts: groups -> {'attrs': {...}, 'items': {name: ...}},
    datasets -> {'attrs': {...}, 'data': array}."""
    with open(filename, 'rb') as f:
        buf = mmap.mmap(f.fileno(), 0, access=mmap.ACCESS_READ)  # pages read on first touch
    try:
        r = _Reader(buf)

        def load(addr):
            kind, content, attrs = r.obj(addr)
            if kind == 'group':
                return {'attrs': attrs, 'items': {k: load(v) for k, v in content.items()}}
            return {'attrs': attrs, 'data': content}
        return load(r.root)
    finally:
        r = None
        try:
            buf.close()
        except BufferError:  # a view is still referenced: closed with the map
            pass


# ------------------------------------------------------------------------------ writing ----

class _Writer:
    """Appends objects to a byte buffer: superblock v2, then every object header and its
    raw data, children before parents (an object header needs its children's addresses)."""

    def __init__(self):
        self.parts = [bytes(48)]  # superblock v2, filled in last
        self.size = 48

    def _append(self, data):
        """data: bytes, or a C-contiguous array written as it is"""
        addr = self.size
        self.parts.append(data)
        self.size += data.nbytes if isinstance(data, np.ndarray) else len(data)
        return addr

    @staticmethod
    def _msg(mtype, body):
        return struct.pack('<BHB', mtype, len(body), 0) + body

    def _ohdr(self, messages):
        body = b''.join(messages)
        head = b'OHDR' + bytes([2, 0x02]) + struct.pack('<I', len(body))
        blob = head + body
        return self._append(blob + struct.pack('<I', lookup3(blob)))

    @staticmethod
    def _datatype(arr):
        dt = arr.dtype
        if dt.kind == 'b':  # h5py's bool: enum FALSE = 0, TRUE = 1 over int8
            base = bytes([0x10, 0x08, 0, 0]) + struct.pack('<I', 1) + struct.pack('<HH', 0, 8)
            names = b''
            for n in (b'FALSE', b'TRUE'):
                names += n + b'\0' + bytes((-(len(n) + 1)) % 8)
            return bytes([0x18, 2, 0, 0]) + struct.pack('<I', 1) + base + names + bytes([0, 1])
        if dt.kind in 'iu':
            sign = 0x08 if dt.kind == 'i' else 0
            return bytes([0x10, sign, 0, 0]) + struct.pack('<I', dt.itemsize) + struct.pack('<HH', 0,
                                                                                             8 * dt.itemsize)
        if dt.kind == 'f':
            if dt.itemsize == 4:
                props = struct.pack('<HHBBBBI', 0, 32, 23, 8, 0, 23, 127)
                return bytes([0x11, 0x20, 31, 0]) + struct.pack('<I', 4) + props
            if dt.itemsize == 8:
                props = struct.pack('<HHBBBBI', 0, 64, 52, 11, 0, 52, 1023)
                return bytes([0x11, 0x20, 63, 0]) + struct.pack('<I', 8) + props
        if dt.kind == 'S':  # fixed-length, null-padded, UTF-8
            return bytes([0x13, 0x11, 0, 0]) + struct.pack('<I', max(dt.itemsize, 1))
        raise TypeError('cannot store arrays of dtype %s' % dt)

    @staticmethod
    def _dataspace(arr):
        if arr.ndim == 0:
            return bytes([2, 0, 0, 0])
        return bytes([2, arr.ndim, 0, 1]) + b''.join(struct.pack('<Q', n) for n in arr.shape)

    @staticmethod
    def _prep(value):
        """numpy array for a dataset or attribute value (strings -> UTF-8 fixed-length bytes)."""
        if isinstance(value, str):
            return np.array(value.encode('utf-8'))
        arr = np.asarray(value)
        if arr.dtype.kind == 'U':
            flat = [s.encode('utf-8') for s in arr.ravel().tolist()]
            width = max([len(s) for s in flat] + [1])
            arr = np.array(flat, dtype='S%d' % width).reshape(arr.shape)
        if arr.dtype.kind == 'S' and arr.dtype.itemsize == 0:
            arr = arr.astype('S1')
        if arr.dtype.byteorder == '>':
            arr = arr.astype(arr.dtype.newbyteorder('<'))
        return arr if arr.flags.c_contiguous else arr.copy(order='C')  # keeps 0-d arrays 0-d

    def _attribute(self, name, value):
        arr = self._prep(value)
        nm = name.encode('utf-8') + b'\0'
        dt, ds = self._datatype(arr), self._dataspace(arr)
        body = bytes([3, 0]) + struct.pack('<HHH', len(nm), len(dt), len(ds)) + bytes([1]) + nm + dt + ds
        return self._msg(0x0C, body + arr.tobytes())

    def dataset(self, value, attrs):
        arr = self._prep(value)
        addr = self._append(arr) if arr.nbytes else _UNDEF
        msgs = [self._msg(0x01, self._dataspace(arr)), self._msg(0x03, self._datatype(arr)),
                self._msg(0x05, bytes([3, 0x09])),
                self._msg(0x08, bytes([3, 1]) + struct.pack('<QQ', addr, arr.nbytes))]
        msgs += [self._attribute(k, v) for k, v in attrs.items()]
        return self._ohdr(msgs)

    def group(self, children, attrs):
        """children: {name: object header address}"""
        # compact links only: the group info message raises the compact-to-dense threshold to
        # the number of links (the readers' default is 8)
        msgs = [self._msg(0x02, bytes([0, 0]) + struct.pack('<QQ', _UNDEF, _UNDEF)),
                self._msg(0x0A, bytes([0, 1]) + struct.pack('<HH', max(8, len(children)), 6))]
        for name, addr in children.items():
            nm = name.encode('utf-8')
            if len(nm) > 255:
                raise ValueError('link name too long: %s' % name)
            msgs.append(self._msg(0x06, bytes([1, 0x10, 1, len(nm)]) + nm + struct.pack('<Q', addr)))
        msgs += [self._attribute(k, v) for k, v in attrs.items()]
        return self._ohdr(msgs)

    def finish(self, root):
        sb = _SIG + bytes([2, 8, 8, 0]) + struct.pack('<QQQQ', 0, _UNDEF, self.size, root)
        self.parts[0] = sb + struct.pack('<I', lookup3(sb))
        return self.parts


def write_file(filename, tree):
    """Write nested dicts in the form ``read_file`` returns."""
    w = _Writer()

    def put(node):
        if 'items' in node:
            return w.group({k: put(v) for k, v in node['items'].items()}, node.get('attrs', {}))
        return w.dataset(node['data'], node.get('attrs', {}))
    root = put(tree)
    with open(filename, 'wb') as f:
        for part in w.finish(root):
            f.write(memoryview(part).cast('B') if isinstance(part, np.ndarray) else part)


# ------------------------------------------------------------------- Mosaic protocol 0.1 ----

def _describe(name, obj, root=False):
    """h5.py:34-114: a description -> the group / dataset tree with Mosaic's attributes."""
    if isinstance(obj, dict):
        attrs = {'protocol': PROTOCOL, 'datetime': str(_dt.datetime.now())} if root else {}
        attrs['is_array'] = np.bool_(False)
        return {'attrs': attrs, 'items': {k: _describe(k, v) for k, v in obj.items()}}
    if isinstance(obj, list) and len(obj) > 0 and isinstance(obj[0], dict):
        items = {'%s_%08d' % (name, i): _describe('%s_%08d' % (name, i), v) for i, v in enumerate(obj)}
        return {'attrs': {'is_array': np.bool_(True), 'len': np.int64(len(obj))}, 'items': items}
    is_bytes, is_none = isinstance(obj, bytes), obj is None
    data = np.frombuffer(obj, dtype=np.uint8) if is_bytes else ('None' if is_none else obj)
    attrs = {'is_ndarray': np.bool_(isinstance(obj, np.ndarray)), 'is_list': np.bool_(isinstance(obj, list)),
             'is_tuple': np.bool_(isinstance(obj, tuple)), 'is_str': np.bool_(isinstance(obj, str)),
             'is_bytes': np.bool_(is_bytes), 'is_none': np.bool_(is_none)}
    if not isinstance(obj, np.ndarray) and not is_bytes:
        flat = np.asarray(data).flatten().tolist()
        attrs['is_str'] = np.bool_(bool(flat) and isinstance(flat[0], str))
    return {'attrs': attrs, 'data': data}


def _decode(x):
    if isinstance(x, list):
        return [_decode(v) for v in x]
    return x.decode('utf-8') if isinstance(x, bytes) else x


def _restore(node):
    """h5.py:128-221: the tree -> the description (non-lazy)."""
    attrs = node['attrs']
    if 'items' in node:
        if bool(attrs.get('is_array', False)):
            return [_restore(node['items'][k]) for k in sorted(node['items'])]
        return {k: _restore(node['items'][k]) for k in sorted(node['items'])}  # h5py's name order
    data = node['data']
    if bool(attrs.get('is_none', False)):
        return None
    if bool(attrs.get('is_ndarray', False)):
        return data
    if bool(attrs.get('is_bytes', False)):
        return np.asarray(data).tobytes()
    if data.dtype == object:
        data = np.vectorize(lambda v: v, otypes=[object])(data)
    if bool(attrs.get('is_str', False)) and not bool(attrs.get('is_list', False)):
        v = data.item() if data.shape == () else data.ravel()[0]
        return v.decode('utf-8') if isinstance(v, bytes) else v
    if bool(attrs.get('is_tuple', False)):
        return tuple(data.tolist())
    if bool(attrs.get('is_list', False)):
        out = list(data.tolist())
        return _decode(out) if bool(attrs.get('is_str', False)) else out
    return data.item()


def write(filename, description):
    """Mosaic ``HDF5.dump``: write a description (dict) as protocol 0.1."""
    write_file(filename, _describe('/', description, root=True))


def read(filename):
    """Mosaic ``HDF5.load`` (non-lazy): the description stored in a protocol 0.1 file."""
    return _restore(read_file(filename))


def _camel_case(name):
    parts = name.replace('-', '_').split('_')
    return parts[0] + ''.join(p[:1].upper() + p[1:] for p in parts[1:])


def _filename(path, project_name, parameter, version, extension):
    stem = '%s-%s' % (project_name, _camel_case(parameter))
    name = stem + ('-%05d' % version if version > 0 else '') + extension
    return os.path.join(path, name)


class HDF5:
    """Mosaic's ``HDF5`` file interface (h5.py:224-337): ``filename=`` or ``project_name`` +
    ``parameter`` (+ ``version``, ``path``, ``extension``) name the file as
    ``<project>-<parameter in camelCase>[-<version:05d>]<ext>``; without a version the next free
    one (write) or the last existing one (read) is used. ``dump`` writes, ``load`` reads."""

    def __init__(self, *args, **kwargs):
        self._mode = kwargs.pop('mode')
        filename = args[0] if args else kwargs.pop('filename', None)
        path = kwargs.pop('path', None) or os.getcwd()
        if filename is None:
            project_name = kwargs.pop('project_name', None)
            parameter = kwargs.pop('parameter', None)
            if project_name is None or parameter is None:
                raise RuntimeError('Either filename or project_name and parameter are needed to generate a filename')
            version = kwargs.pop('version', None)
            extension = kwargs.pop('extension', '.h5')
            if version is None or version < 0:
                version = kwargs.pop('version_start', 0)
                while os.path.exists(_filename(path, project_name, parameter, version, extension)):
                    version += 1
                if self._mode.startswith('r'):
                    version -= 1
            filename = _filename(path, project_name, parameter, version, extension)
        elif not os.path.isabs(filename):
            filename = os.path.join(path, filename)
        self._filename = filename

    @property
    def mode(self):
        return self._mode

    @property
    def filename(self):
        return self._filename

    def dump(self, description):
        if self._mode.startswith('r'):
            raise RuntimeError('file opened for reading')
        write(self._filename, description)

    def load(self):
        return read(self._filename)

    def close(self):
        pass

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        self.close()


def file_exists(*args, **kwargs):
    """Mosaic's ``file_exists`` (h5.py:340-381)."""
    filename = args[0] if args else kwargs.pop('filename', None)
    path = kwargs.pop('path', None) or os.getcwd()
    if filename is None:
        filename = _filename(path, kwargs.pop('project_name'), kwargs.pop('parameter'), kwargs.pop('version', 0),
                             kwargs.pop('extension', '.h5'))
    elif not os.path.isabs(filename):
        filename = os.path.join(path, filename)
    return os.path.exists(filename)
