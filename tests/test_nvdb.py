"""NanoVDB export and import (csrc/nm_nvdb.cpp), on the CPU.

NanoVDB is not in this image, so the grids are checked against what its
published 32.4 layout fixes, through tests/nvdb_layout.py (a separate
reader of that layout): the node sizes (672 + 64 + 64 + 32 per upper node,
270400, 33856 and 2144 bytes), which nodes a leaf shares with another
(closed-form buffer sizes), the root key of known coordinates, NanoVDB's
lookup of every voxel, the breadth-first order and counts, the file header
and grid record.  Whether NanoVDB's own reader accepts the files is not
tested here (parity unpinned against the library itself)."""
import struct

import numpy as np
import pytest

import nanomap_amd as nm
from nanomap_amd import _capi

import nvdb_layout as L


def _leaves(rng, n, span=20000):
    """n distinct leaves with origins in [-span, span) (several upper
    nodes, negative coordinates), random values and masks, plus one leaf
    with no active voxel and one with all."""
    orgs = set()
    while len(orgs) < n:
        orgs.add(tuple(int(c) * 8 for c in rng.integers(-span // 8, span // 8, 3)))
    o = np.array(sorted(orgs), np.int32)
    v = rng.normal(0, 1.5, (n, 512)).astype(np.float32)
    k = rng.integers(0, 2**32, (n, 16), dtype=np.uint64).astype(np.uint32)
    k &= rng.integers(0, 2**32, (n, 16), dtype=np.uint64).astype(np.uint32)
    if n >= 2:
        k[0] = 0
        k[1] = 0xFFFFFFFF
    return o, v, k


def _bits(k):
    return ((k[:, np.arange(512) >> 5] >> (np.arange(512) & 31)) & 1).astype(bool)


def _sorted(o, v, bits):
    idx = np.lexsort((o[:, 2], o[:, 1], o[:, 0]))
    return o[idx], v[idx], bits[idx]


SIZE_EMPTY = L.GRID + L.TREE + L.ROOT


def test_node_sizes_closed_form():
    v = np.zeros((1, 512), np.float32)
    k = np.zeros((1, 16), np.uint32)

    def size(origins):
        o = np.array(origins, np.int32)
        return len(nm.nvdb_build(o, np.repeat(v, len(o), 0),
                                 np.repeat(k, len(o), 0), 0.1))
    one = SIZE_EMPTY + L.TILE + L.UPPER + L.LOWER + L.LEAF
    assert one == 672 + 64 + 64 + 32 + 270400 + 33856 + 2144
    assert size([[0, 0, 0]]) == one
    # same lower node (128^3): one more leaf
    assert size([[0, 0, 0], [120, 0, 8]]) == one + L.LEAF
    # same upper node (4096^3), another lower node
    assert size([[0, 0, 0], [128, 0, 0]]) == one + L.LOWER + L.LEAF
    # another upper node: a root tile and a whole branch
    assert size([[0, 0, 0], [4096, 0, 0]]) == 2 * one - SIZE_EMPTY
    # -8 lies in the upper node below 0 on x
    assert size([[0, 0, 0], [-8, 0, 0]]) == 2 * one - SIZE_EMPTY
    assert len(nm.nvdb_build(np.zeros((0, 3), np.int32),
                             np.zeros((0, 512), np.float32),
                             np.zeros((0, 16), np.uint32), 0.1)) == SIZE_EMPTY


def test_root_keys_and_header_fields():
    o = np.array([[-8, 4096, 8192], [0, 0, 0]], np.int32)
    v = np.full((2, 512), 0.5, np.float32)
    k = np.zeros((2, 16), np.uint32)
    k[:, 0] = 1  # voxel 0 of each leaf
    g = L.Grid(nm.nvdb_build(o, v, k, 0.05, "occupancy"))
    # (-8 as u32) >> 12 = 0xfffff, 4096 >> 12 = 1, 8192 >> 12 = 2
    assert L.root_key(-8, 4096, 8192) == (2 | (1 << 21) | (0xFFFFF << 42))
    keys = sorted(t[0] for t in g.tiles)
    assert keys == sorted([L.root_key(-8, 4096, 8192), L.root_key(0, 0, 0)])
    assert g.magic == L.MAGIC and g.version == (32 << 21) | (4 << 10) | 2
    assert g.checksum == 2**64 - 1 and g.count == 1 and g.index == 0
    assert g.name == "occupancy" and g.grid_type == 1 and g.grid_class == 0
    assert g.flags == (1 << 1) | (1 << 2) | (1 << 5)
    assert g.mat_d[0] == g.mat_d[4] == g.mat_d[8] == 0.05
    assert g.inv_d[0] == 1 / 0.05 and g.mat_d[1] == 0 and g.vec_d == (0, 0, 0)
    assert g.mat_f[0] == np.float32(0.05) and g.taper_d == 1.0
    assert g.voxel_size == (0.05, 0.05, 0.05)
    assert g.root_bbox == (-8, 0, 0, 0, 4096, 8192)
    assert g.world_bbox[0] == -8 * 0.05 and g.world_bbox[5] == 8192 * 0.05
    assert g.vmin == g.vmax == np.float32(0.5) and g.background == 0.0
    assert g.node_count == (2, 2, 2) and g.voxel_count == 2


@pytest.mark.parametrize("seed", [0, 1])
def test_every_voxel_through_the_lookup_path(seed):
    rng = np.random.default_rng(seed)
    o, v, k = _leaves(rng, 60)
    g = L.Grid(nm.nvdb_build(o, v, k, 0.1))
    bits = _bits(k)
    got = g.leaves()  # layout checks
    assert len(got) == len(o)
    go = np.array([x[0] for x in got], np.int32)
    gv = np.stack([x[1] for x in got])
    gb = np.stack([x[2] for x in got])
    so, sv, sb = _sorted(o, v, bits)
    to, tv, tb = _sorted(go, gv, gb)
    assert np.array_equal(so, to)
    assert np.array_equal(sv.view(np.uint32), tv.view(np.uint32))
    assert np.array_equal(sb, tb)
    # point lookups (NanoVDB's accessor path) in a few leaves, every voxel
    for i in rng.choice(len(o), 6, replace=False):
        for q in range(0, 512, 7):
            x, y, z = o[i] + np.array([q >> 6, (q >> 3) & 7, q & 7])
            val, act, level = g.probe(int(x), int(y), int(z))
            assert level == 0
            assert np.float32(val).view(np.uint32) == v[i, q].view(np.uint32)
            assert act == bits[i, q]
    # voxels outside every leaf: background, inactive
    have = set(map(tuple, o.tolist()))
    for _ in range(200):
        c = rng.integers(-21000, 21000, 3)
        if tuple(int(a) & ~7 for a in c) in have:
            continue
        val, act, _ = g.probe(*(int(a) for a in c))
        assert val == 0.0 and not act
    # stats: min/max over active values; the leaf bbox spans its active
    # voxels; node count and voxel count
    assert g.voxel_count == int(bits.sum())
    act_vals = v[bits]
    assert g.vmin == act_vals.min() and g.vmax == act_vals.max()


def test_file_round_trip(tmp_path):
    rng = np.random.default_rng(7)
    o, v, k = _leaves(rng, 40)
    path = tmp_path / "map.nvdb"
    nm.nvdb_write(path, o, v, k, 0.2, "log_odds")
    data = path.read_bytes()
    head, grids = L.read_file(data)
    assert head == dict(magic=L.MAGIC, version=(32 << 21) | (4 << 10) | 2,
                        count=1, codec=0)
    rec, name, g = grids[0]
    assert name == "log_odds" and rec["name_key"] == L.string_hash("log_odds")
    assert rec["grid_size"] == rec["file_size"] == g.size
    assert rec["voxel_count"] == g.voxel_count == int(_bits(k).sum())
    assert rec["grid_type"] == 1 and rec["codec"] == 0
    assert rec["node_count"] == g.node_count + (1,)
    assert rec["index_bbox"] == g.root_bbox
    assert rec["voxel_size"] == (0.2, 0.2, 0.2)
    assert len(data) == 16 + 176 + len("log_odds") + 1 + g.size
    ro, rv, rk, dx = nm.nvdb_read(path)
    assert dx == 0.2
    a = _sorted(o, v, _bits(k))
    b = _sorted(ro, rv, _bits(rk))
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint8),
                              np.asarray(y).view(np.uint8))
    # by name, and a name the file does not hold
    assert len(nm.nvdb_read(path, "log_odds")[0]) == len(o)
    with pytest.raises(_capi.NanoMapError):
        nm.nvdb_read(path, "density")


def test_reader_expands_active_tiles(tmp_path):
    # turn one lower-node child of a built grid into an active tile: the
    # reader gives that 8^3 block as a leaf of active voxels at the tile's
    # value
    o = np.array([[0, 0, 0], [8, 0, 0]], np.int32)
    v = np.ones((2, 512), np.float32)
    k = np.full((2, 16), 0xFFFFFFFF, np.uint32)
    buf = bytearray(nm.nvdb_build(o, v, k, 0.1, "g"))
    g = L.Grid(buf)
    up = g.root + g.tiles[0][1]
    lo = up + struct.unpack_from("<q", buf, up + 8256)[0]
    n = L.slot(8, 0, 0, 4, 3)
    buf[lo + 544 + (n >> 3)] &= ~(1 << (n & 7)) & 0xFF   # not a child
    buf[lo + 32 + (n >> 3)] |= 1 << (n & 7)              # an active tile
    struct.pack_into("<f", buf, lo + 1088 + 8 * n, 2.5)
    assert L.Grid(buf).probe(9, 1, 2) == (2.5, True, 1)
    path = tmp_path / "tile.nvdb"
    head = struct.pack("<QIHH", L.MAGIC, (32 << 21) | (4 << 10) | 2, 1, 0)
    rec = bytearray(176)
    struct.pack_into("<QQ", rec, 0, len(buf), len(buf))
    struct.pack_into("<I", rec, 32, 1)
    struct.pack_into("<I", rec, 136, 2)
    path.write_bytes(head + bytes(rec) + b"g\0" + bytes(buf))
    ro, rv, rk, _ = nm.nvdb_read(path)
    assert sorted(map(tuple, ro.tolist())) == [(0, 0, 0), (8, 0, 0)]
    t = [i for i in range(2) if ro[i, 0] == 8][0]
    assert np.all(rv[t] == 2.5) and np.all(rk[t] == 0xFFFFFFFF)


def test_rejects_bad_input(tmp_path):
    v = np.zeros((2, 512), np.float32)
    k = np.zeros((2, 16), np.uint32)
    with pytest.raises(_capi.NanoMapError):   # the same leaf twice
        nm.nvdb_build(np.zeros((2, 3), np.int32), v, k, 0.1)
    with pytest.raises(_capi.NanoMapError):   # origin not on the leaf grid
        nm.nvdb_build(np.array([[4, 0, 0]], np.int32), v[:1], k[:1], 0.1)
    with pytest.raises(_capi.NanoMapError):
        nm.nvdb_build(np.zeros((1, 3), np.int32), v[:1], k[:1], 0.1, "x" * 256)
    with pytest.raises(_capi.NanoMapError):
        nm.nvdb_build(np.zeros((1, 3), np.int32), v[:1], k[:1], 0.0)
    path = tmp_path / "m.nvdb"
    nm.nvdb_write(path, np.zeros((1, 3), np.int32), v[:1], k[:1], 0.1)
    data = path.read_bytes()
    cut = tmp_path / "cut.nvdb"
    cut.write_bytes(data[:-100])
    with pytest.raises(_capi.NanoMapError):
        nm.nvdb_read(cut)
    zipped = tmp_path / "zip.nvdb"
    zipped.write_bytes(data[:14] + struct.pack("<H", 1) + data[16:])
    with pytest.raises(_capi.NanoMapError):
        nm.nvdb_read(zipped)
    # a child offset pointing past the buffer
    bad = bytearray(data)
    g0 = 16 + 176 + len("occupancy") + 1
    grid = L.Grid(bytes(bad[g0:]))
    struct.pack_into("<q", bad, g0 + grid.root + L.ROOT + 8, 1 << 40)
    broken = tmp_path / "broken.nvdb"
    broken.write_bytes(bytes(bad))
    with pytest.raises(_capi.NanoMapError):
        nm.nvdb_read(broken)
