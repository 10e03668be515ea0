"""A reader of NanoVDB 32.4 FloatGrid buffers and .nvdb files, written for
the tests from the published layout (NanoVDB.h / IO.h of that version) and
kept apart from the library's writer and reader (csrc/nm_nvdb.cpp): it
decodes the structs field by field with `struct`, checks the layout rules
(32-byte alignment, node sizes, breadth-first child order, counts, bounding
boxes, min/max) and looks voxels up the way NanoVDB's accessor does: root
tile by key, then the upper, lower and leaf node by child slot."""
import struct

import numpy as np

MAGIC = 0x304244566F6E614E  # "NanoVDB0"
GRID, TREE, ROOT, TILE = 672, 64, 64, 32
UPPER, LOWER, LEAF = 270400, 33856, 2144


def root_key(x, y, z):
    u = lambda c: (c & 0xFFFFFFFF) >> 12  # noqa: E731
    return u(z) | (u(y) << 21) | (u(x) << 42)


def slot(x, y, z, log2dim, child_total):
    m = (1 << (log2dim + child_total)) - 1
    return ((((x & m) >> child_total) << (2 * log2dim)) |
            (((y & m) >> child_total) << log2dim) |
            ((z & m) >> child_total))


def string_hash(name):
    h = 0
    for ch in name.encode():
        overflow = h >> 56
        h = (h * 67 + ch + overflow) & 0xFFFFFFFFFFFFFFFF
    return h


class Grid:
    """One FloatGrid buffer."""

    def __init__(self, buf):
        self.b = bytes(buf)
        b = self.b
        (self.magic, self.checksum, self.version, self.flags, self.index,
         self.count, self.size) = struct.unpack_from("<QQIIIIQ", b, 0)
        self.name = b[40:296].split(b"\0", 1)[0].decode()
        self.mat_f = struct.unpack_from("<9f", b, 296)
        self.inv_f = struct.unpack_from("<9f", b, 332)
        self.vec_f = struct.unpack_from("<3f", b, 368)
        self.taper_f = struct.unpack_from("<f", b, 380)[0]
        self.mat_d = struct.unpack_from("<9d", b, 384)
        self.inv_d = struct.unpack_from("<9d", b, 456)
        self.vec_d = struct.unpack_from("<3d", b, 528)
        self.taper_d = struct.unpack_from("<d", b, 552)[0]
        self.world_bbox = struct.unpack_from("<6d", b, 560)
        self.voxel_size = struct.unpack_from("<3d", b, 608)
        self.grid_class, self.grid_type = struct.unpack_from("<II", b, 632)
        t = GRID
        self.node_offset = struct.unpack_from("<4Q", b, t)
        self.node_count = struct.unpack_from("<3I", b, t + 32)
        self.tile_count = struct.unpack_from("<3I", b, t + 44)
        self.voxel_count = struct.unpack_from("<Q", b, t + 56)[0]
        self.root = t + self.node_offset[3]
        r = self.root
        self.root_bbox = struct.unpack_from("<6i", b, r)
        self.table_size, = struct.unpack_from("<I", b, r + 24)
        self.background, self.vmin, self.vmax = struct.unpack_from(
            "<3f", b, r + 28)
        self.tiles = []
        for i in range(self.table_size):
            key, child, state, value = struct.unpack_from(
                "<QqIf", b, r + ROOT + TILE * i)
            self.tiles.append((key, child, state, value))

    # -- NanoVDB's lookup path --------------------------------------------
    def _bit(self, at, n):
        return (self.b[at + (n >> 3)] >> (n & 7)) & 1

    def probe(self, x, y, z):
        """(value, active, level) of voxel (x, y, z): level 0 leaf, 1 lower
        tile, 2 upper tile, 3 root tile or background."""
        key = root_key(x, y, z)
        for k, child, state, value in self.tiles:
            if k == key:
                break
        else:
            return self.background, False, 3
        if child == 0:
            return value, bool(state), 3
        up = self.root + child
        n = slot(x, y, z, 5, 7)
        e = up + 8256 + 8 * n
        if not self._bit(up + 4128, n):
            return struct.unpack_from("<f", self.b, e)[0], bool(
                self._bit(up + 32, n)), 2
        lo = up + struct.unpack_from("<q", self.b, e)[0]
        n = slot(x, y, z, 4, 3)
        e = lo + 1088 + 8 * n
        if not self._bit(lo + 544, n):
            return struct.unpack_from("<f", self.b, e)[0], bool(
                self._bit(lo + 32, n)), 1
        lf = lo + struct.unpack_from("<q", self.b, e)[0]
        n = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7)
        v = struct.unpack_from("<f", self.b, lf + 96 + 4 * n)[0]
        return v, bool(self._bit(lf + 16, n)), 0

    # -- walk of the whole tree, with the layout checks --------------------
    def leaves(self):
        """[(origin, values[512] float32, mask bits[512] bool)] in tree
        order, after checking the layout rules."""
        b = self.b
        assert self.magic == MAGIC
        assert self.version >> 21 == 32
        assert self.grid_type == 1  # Float
        assert self.size == len(b) and self.size % 32 == 0
        uppers, lowers, leaves = [], [], []
        for key, child, _, _ in self.tiles:
            if child:
                uppers.append((self.root + child, key))
        out = []
        active_total = 0
        for up, key in uppers:
            ox = ((key >> 42) & 0x1FFFFF) << 12
            oy = ((key >> 21) & 0x1FFFFF) << 12
            oz = (key & 0x1FFFFF) << 12
            origin_u = [c - (1 << 32) if c >= 1 << 31 else c
                        for c in (ox, oy, oz)]
            for n in range(32768):
                if not self._bit(up + 4128, n):
                    continue
                lo = up + struct.unpack_from("<q", b, up + 8256 + 8 * n)[0]
                lorg = (origin_u[0] + ((n >> 10) << 7),
                        origin_u[1] + (((n >> 5) & 31) << 7),
                        origin_u[2] + ((n & 31) << 7))
                lowers.append(lo)
                for m in range(4096):
                    if not self._bit(lo + 544, m):
                        continue
                    lf = lo + struct.unpack_from("<q", b,
                                                 lo + 1088 + 8 * m)[0]
                    org = (lorg[0] + ((m >> 8) << 3),
                           lorg[1] + (((m >> 4) & 15) << 3),
                           lorg[2] + ((m & 15) << 3))
                    leaves.append(lf)
                    bmin = struct.unpack_from("<3i", b, lf)
                    assert tuple(c & ~7 for c in bmin) == org
                    vals = np.frombuffer(b, np.float32, 512, lf + 96).copy()
                    words = np.frombuffer(b, np.uint64, 8, lf + 16)
                    bits = ((words[np.arange(512) >> 6] >>
                             (np.arange(512, dtype=np.uint64) & 63)) & 1
                            ).astype(bool)
                    active_total += int(bits.sum())
                    # leaf bbox: min and extent of its active voxels, flag
                    # bit 1 set when it has any
                    dif = struct.unpack_from("<3B", b, lf + 12)
                    flags = b[lf + 15]
                    if bits.any():
                        q = np.nonzero(bits)[0]
                        c = np.stack([q >> 6, (q >> 3) & 7, q & 7], 1)
                        assert tuple(bmin) == tuple(
                            int(a) for a in np.add(org, c.min(0)))
                        assert dif == tuple(int(a) for a in
                                            c.max(0) - c.min(0))
                        assert flags & 2
                        assert struct.unpack_from("<2f", b, lf + 80) == (
                            vals[bits].min(), vals[bits].max())
                    else:
                        assert not flags & 2
                    out.append((org, vals, bits))
        # breadth-first: each level contiguous, in the order its parents
        # and their child slots list them, at the published node sizes
        for nodes, size, first in ((uppers, UPPER, 2), (lowers, LOWER, 1),
                                   (leaves, LEAF, 0)):
            addrs = [a if isinstance(a, int) else a[0] for a in nodes]
            assert self.node_count[first] == len(addrs)
            for i, a in enumerate(addrs):
                assert a % 32 == 0
                assert a == addrs[0] + i * size
            if addrs:
                assert GRID + self.node_offset[first] == addrs[0]
        assert self.voxel_count == active_total
        return out


def read_file(data):
    """(header fields, [(record fields, name, Grid)]) of a .nvdb file."""
    magic, version, count, codec = struct.unpack_from("<QIHH", data, 0)
    at = 16
    grids = []
    for _ in range(count):
        rec = struct.unpack_from("<QQQQII6d6i3dI4I3IHHI", data, at)
        at += 176
        keys = ("grid_size", "file_size", "name_key", "voxel_count",
                "grid_type", "grid_class", "world_bbox", "index_bbox",
                "voxel_size", "name_size", "node_count", "tile_count",
                "codec", "padding", "version")
        r = dict(zip(keys[:6], rec[:6]))
        r["world_bbox"] = rec[6:12]
        r["index_bbox"] = rec[12:18]
        r["voxel_size"] = rec[18:21]
        r["name_size"] = rec[21]
        r["node_count"] = rec[22:26]
        r["tile_count"] = rec[26:29]
        r["codec"], r["padding"], r["version"] = rec[29:32]
        name = data[at:at + r["name_size"] - 1].decode()
        assert data[at + r["name_size"] - 1] == 0
        at += r["name_size"]
        grids.append((r, name, Grid(data[at:at + r["grid_size"]])))
        at += r["grid_size"]
    return dict(magic=magic, version=version, count=count,
                codec=codec), grids
