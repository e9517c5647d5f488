"""A second, independent reader of OP2 HDF5 mesh files (numpy, for the tests only).

It walks the same subset of the HDF5 format as the library's C++ reader
(superblock 0/1, version-1 object headers, a symbol-table root group, contiguous
or compact datasets) but shares no code with it, so the two can be compared.
"""
import struct

import numpy as np


class H5:
    def __init__(self, path):
        self.d = open(path, "rb").read()
        d = self.d
        assert d[:8] == b"\x89HDF\r\n\x1a\n", "not an HDF5 file"
        ver = d[8]
        root = (24 if ver == 0 else 28) + 32
        ohdr = self._q(root + 8)
        st = [b for t, b in self._messages(ohdr) if t == 0x11][0]
        self.items = self._group(self._q(st), self._q(st + 8))

    def _q(self, off):
        return struct.unpack_from("<Q", self.d, off)[0]

    def _messages(self, addr):
        d = self.d
        nmsg = struct.unpack_from("<H", d, addr + 2)[0]
        blocks = [(addr + 16, struct.unpack_from("<I", d, addr + 8)[0])]
        out = []
        while blocks and len(out) < nmsg:
            start, size = blocks.pop(0)
            p = start
            while p + 8 <= start + size and len(out) < nmsg:
                t, sz = struct.unpack_from("<HH", d, p)
                if t == 0x10:
                    blocks.append((self._q(p + 8), self._q(p + 16)))
                out.append((t, p + 8))
                p += 8 + sz
        return out

    def _group(self, btree, heap):
        d = self.d
        hdata = self._q(heap + 24)
        res = {}

        def name(off):
            e = d.index(b"\0", hdata + off)
            return d[hdata + off:e].decode()

        def walk(a):
            level, used = d[a + 5], struct.unpack_from("<H", d, a + 6)[0]
            for i in range(used):
                child = self._q(a + 24 + 16 * i + 8)
                if level:
                    walk(child)
                    continue
                n = struct.unpack_from("<H", d, child + 6)[0]
                for j in range(n):
                    e = child + 8 + 40 * j
                    res[name(self._q(e))] = self._q(e + 8)
        walk(btree)
        return res

    def read(self, key):
        dims, off, dtype = None, None, None
        for t, b in self._messages(self.items[key]):
            if t == 1:
                ver, rank = self.d[b], self.d[b + 1]
                o = b + (8 if ver == 1 else 4)
                dims = struct.unpack_from("<%dQ" % rank, self.d, o)
            elif t == 3:
                dtype = "<i4" if self.d[b] & 15 == 0 else "<f8"
            elif t == 8:
                if self.d[b] == 3 and self.d[b + 1] == 1:
                    off = self._q(b + 2)
                elif self.d[b] == 3 and self.d[b + 1] == 0:
                    off = b + 4
        return np.frombuffer(self.d, dtype=dtype, count=int(np.prod(dims)), offset=off).reshape(dims)
