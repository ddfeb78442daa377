"""Writes tests/golden/ns_amr_lgf_ref.npz from the reference's own fixture
tests/ns_amr_lgf/configs/ref.hdf5 (the velocity and pressure of the
vortex-ring IF-HERK test, read by NS_AMR_LGF::run through
simulation.read_h5, tests/ns_amr_lgf/ns_amr_lgf.hpp:374-378).

The file is HDF5 (superblock 0, old-style groups) in the Chombo layout: one
level, 36 boxes over the cells [-12, 11]^3 at dx 0.25, eight components
(cell_aux, edge_aux_0..2, u_0..2, p), each box's components one after the
other, x fastest.  No HDF5 library is available here, so this script reads
the few structures it needs itself (superblock, version-1 object headers,
symbol-table groups, contiguous datasets, attributes); it executes nothing
from the file.  Output: u (3, 24, 24, 24) and p (24, 24, 24), indexed
[z][y][x] from cell -12, FP64.

Usage: python tests/golden/make_ns_amr_lgf_golden.py [path to ref.hdf5]
"""
import os
import struct
import sys

import numpy as np

REF = "/root/reference/tests/ns_amr_lgf/configs/ref.hdf5"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ns_amr_lgf_ref.npz")


class Hdf5Reader:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.d = fh.read()
        d = self.d
        if d[:8] != b"\x89HDF\r\n\x1a\n" or d[8] != 0 or d[13] != 8 or d[14] != 8:
            raise ValueError("not an HDF5 file with superblock 0 and 8-byte offsets")
        self.root = struct.unpack_from("<Q", d, 56 + 8)[0]  # root symbol-table entry's object header

    def messages(self, addr):
        d = self.d
        ver, _, nmsg, _, hsize = struct.unpack_from("<BBHII", d, addr)
        if ver != 1:
            raise ValueError("object header version %d" % ver)
        out, blocks = [], [(addr + 16, hsize)]
        while blocks:
            start, size = blocks.pop(0)
            p = start
            while p + 8 <= start + size and len(out) < nmsg:
                mtype, msize = struct.unpack_from("<HH", d, p)
                data = d[p + 8:p + 8 + msize]
                if mtype == 0x10:  # continuation
                    blocks.append(struct.unpack_from("<QQ", data, 0))
                out.append((mtype, data))
                p += 8 + msize
        return out

    def group(self, addr):
        st = [m for t, m in self.messages(addr) if t == 0x11]
        if not st:
            raise ValueError("not a symbol-table group")
        btree, heap = struct.unpack_from("<QQ", st[0], 0)
        d = self.d
        if d[heap:heap + 4] != b"HEAP":
            raise ValueError("local heap expected")
        heap_data = struct.unpack_from("<Q", d, heap + 24)[0]
        out = {}
        self._btree(btree, heap_data, out)
        return out

    def _btree(self, addr, heap_data, out):
        d = self.d
        if d[addr:addr + 4] != b"TREE":
            raise ValueError("B-tree node expected")
        _, level, used = struct.unpack_from("<BBH", d, addr + 4)
        p = addr + 8 + 16 + 8
        for _ in range(used):
            child = struct.unpack_from("<Q", d, p)[0]
            p += 16
            if level > 0:
                self._btree(child, heap_data, out)
            else:
                if d[child:child + 4] != b"SNOD":
                    raise ValueError("symbol-table node expected")
                n = struct.unpack_from("<H", d, child + 6)[0]
                for i in range(n):
                    noff, oh = struct.unpack_from("<QQ", d, child + 8 + 40 * i)
                    e = d.index(b"\0", heap_data + noff)
                    out[d[heap_data + noff:e].decode()] = oh

    def attrs(self, addr):
        out = {}
        for t, data in self.messages(addr):
            if t != 0x0C or data[0] != 1:
                continue
            nsz, tsz, ssz = struct.unpack_from("<HHH", data, 2)
            p = 8
            name = data[p:p + nsz].split(b"\0")[0].decode()
            p += (nsz + 7) // 8 * 8
            dt = data[p:p + tsz]
            p += (tsz + 7) // 8 * 8 + (ssz + 7) // 8 * 8
            cls, size = dt[0] & 0xF, struct.unpack_from("<I", dt, 4)[0]
            val = data[p:p + size]
            if cls == 3:
                out[name] = val.split(b"\0")[0].decode()
            elif cls == 0 and size == 4:
                out[name] = struct.unpack_from("<i", val)[0]
            elif cls == 1 and size == 8:
                out[name] = struct.unpack_from("<d", val)[0]
            else:
                out[name] = val
        return out

    def contiguous(self, addr):
        """(address, bytes, element size) of a contiguous dataset."""
        info = {t: m for t, m in self.messages(addr)}
        lay = info[8]
        if lay[0] != 3 or lay[1] != 1:
            raise ValueError("contiguous layout (version 3) expected")
        daddr, size = struct.unpack_from("<QQ", lay, 2)
        esize = struct.unpack_from("<I", info[3], 4)[0]
        return daddr, size, esize


def main(path=REF):
    f = Hdf5Reader(path)
    root = f.group(f.root)
    ra = f.attrs(f.root)
    ncomp = ra["num_components"]
    names = [ra["component_%d" % i] for i in range(ncomp)]
    lv = f.group(root["level_0"])
    la = f.attrs(root["level_0"])
    assert abs(la["dx"] - 0.25) < 1e-15, la["dx"]
    baddr, bsize, besize = f.contiguous(lv["boxes"])
    nbox = bsize // besize
    boxes = [struct.unpack_from("<6i", f.d, baddr + i * besize) for i in range(nbox)]
    daddr, dsize, desize = f.contiguous(lv["data:datatype=0"])
    assert desize == 8
    data = np.frombuffer(f.d[daddr:daddr + dsize], dtype="<f8")
    lo = min(min(b[:3]) for b in boxes)
    hi = max(max(b[3:]) for b in boxes)
    m = hi - lo + 1
    field = np.full((ncomp, m, m, m), np.nan)
    p = 0
    for b in boxes:
        ext = [b[3 + d] - b[d] + 1 for d in range(3)]
        cnt = ext[0] * ext[1] * ext[2]
        for c in range(ncomp):
            blk = data[p:p + cnt].reshape(ext[2], ext[1], ext[0])  # x fastest
            field[c, b[2] - lo:b[5] - lo + 1, b[1] - lo:b[4] - lo + 1, b[0] - lo:b[3] - lo + 1] = blk
            p += cnt
    assert p == data.size and not np.isnan(field).any()
    u = np.stack([field[names.index("u_%d" % c)] for c in range(3)])
    pr = field[names.index("p")]
    np.savez_compressed(OUT, u=u, p=pr, lo=lo, dx=la["dx"])
    print(f"wrote {OUT}: {nbox} boxes, cells [{lo}, {hi}]^3, |u| max {np.abs(u).max():.4g}")


if __name__ == "__main__":
    main(*sys.argv[1:])
