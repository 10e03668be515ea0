"""GPU tests of whole-map operations (csrc/nm_map.hip): growth, batched
point queries, snapshot and restore.  Each is checked bit for bit against
the oracle's map (Handler::integrateTempGrid restated in oracle/) or
against the map it was derived from."""
import numpy as np
import pytest

import oracle
from nanomap_amd import workloads as W

pytestmark = pytest.mark.gpu

RD = 1


def _same_map(a, b):
    ao, av, am = a
    bo, bv, bm = b
    assert ao.shape == bo.shape, (ao.shape, bo.shape)
    assert np.array_equal(ao, bo)
    assert np.array_equal(av.view(np.uint32), bv.view(np.uint32))
    assert np.array_equal(am, bm)


def _hall(nm, n):
    _, cam, frames = W.hall_walk_frames(nm, n_frames=n)
    return cam, frames


def _feed(nm, ctx, gmap, omap, frames):
    for pos, quat, cloud in frames:
        f = W.frame(nm, pos, quat, 0.05, 10.0, RD)
        ctx.submit(gmap, f, cloud)
        if omap is not None:
            oracle.process_frame(omap, f, cloud)
    return ctx.stats()


def test_auto_grow_keeps_every_update(gpu):
    # a 64-leaf map that grows as the hall frames need: same map as the
    # oracle, no overflow; without auto-grow the same map overflows
    nm = gpu
    cam, frames = _hall(nm, 6)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP, leaf_capacity=64)
    gmap.set_auto_grow(True)
    omap = oracle.Map(*W.CLAMP)
    print("auto-grow: 6 hall frames from 64 leaves", flush=True)
    for k, (pos, quat, cloud) in enumerate(frames):
        f = W.frame(nm, pos, quat, 0.05, 10.0, RD)
        ctx.submit(gmap, f, cloud)
        oracle.process_frame(omap, f, cloud)
        st = ctx.stats()
        assert st["overflow"] == 0 and st["errors"] == 0, (k, st)
    assert gmap.capacity() >= gmap.leaf_count() > 64
    _same_map(gmap.download(), omap.export())

    small = nm.Map(0.05, *W.CLAMP, leaf_capacity=64)
    st = _feed(nm, ctx, small, None, frames[:1])
    assert st["overflow"] > 0
    assert small.leaf_count() <= 64


def test_reserve_and_queries(gpu):
    nm = gpu
    cam, frames = _hall(nm, 4)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    print("reserve + query: 4 hall frames", flush=True)
    _feed(nm, ctx, gmap, omap, frames[:2])
    before = gmap.download()
    cap0 = gmap.capacity()
    gmap.reserve(4 * cap0)
    assert gmap.capacity() >= 4 * cap0
    _same_map(gmap.download(), before)
    # the grown map keeps integrating exactly
    _feed(nm, ctx, gmap, omap, frames[2:])
    o, v, m = gmap.download()
    _same_map((o, v, m), omap.export())
    # point queries: every voxel of 200 leaves, plus voxels of absent
    # leaves (background 0, inactive), from host and device arrays
    rng = np.random.default_rng(7)
    pick = rng.choice(len(o), min(200, len(o)), replace=False)
    li = np.arange(512)
    off = np.stack([li >> 6, (li >> 3) & 7, li & 7], 1)
    ijk = (o[pick][:, None, :] + off[None]).reshape(-1, 3).astype(np.int32)
    far = (rng.integers(-10, 10, size=(500, 3)) * 8 + 4000).astype(np.int32)
    q = np.concatenate([ijk, far])
    vals, act = gmap.query(q)
    want_v = v[pick].reshape(-1)
    bits = ((m[pick][:, li >> 5] >> (li & 31)) & 1).reshape(-1).astype(bool)
    assert np.array_equal(vals[:len(ijk)].view(np.uint32), want_v.view(np.uint32))
    assert np.array_equal(act[:len(ijk)], bits)
    assert np.all(vals[len(ijk):] == 0) and not act[len(ijk):].any()
    import torch
    dq = torch.from_numpy(q).cuda()
    dv, da = gmap.query(dq)
    torch.cuda.synchronize()
    assert np.array_equal(dv.cpu().numpy().view(np.uint32), vals.view(np.uint32))
    assert np.array_equal(da.cpu().numpy(), act)


def test_snapshot_restore_continues_exactly(gpu, tmp_path):
    # save after 5 of 10 frames, restore into a new map (on a fresh
    # context), integrate the other 5: the oracle's 10-frame map, bit for
    # bit; the restored map equals the saved one before any new frame
    nm = gpu
    cam, frames = _hall(nm, 10)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP)
    omap = oracle.Map(*W.CLAMP)
    print("snapshot: 10 hall frames, restore at 5", flush=True)
    _feed(nm, ctx, gmap, omap, frames[:5])
    path = tmp_path / "hall5.nmap"
    gmap.save(path)
    saved = gmap.download()
    gmap.close()
    restored = nm.Map.load(path)
    assert restored.grid_res == np.float32(0.05)
    assert restored.capacity() >= 2 * len(saved[0])
    _same_map(restored.download(), saved)
    exact = nm.Map.load(path, leaf_capacity=1)  # room for the file only
    assert exact.capacity() == len(saved[0])
    _same_map(exact.download(), saved)
    exact.close()
    ctx2 = nm.Context(cam.width * cam.height, 200.0)
    st = _feed(nm, ctx2, restored, omap, frames[5:])
    assert st["overflow"] == 0
    _same_map(restored.download(), omap.export())
    # a truncated file is rejected
    data = path.read_bytes()
    bad = tmp_path / "bad.nmap"
    bad.write_bytes(data[:len(data) // 2])
    with pytest.raises(nm.NanoMapError):
        nm.Map.load(bad)


def test_nanovdb_export_and_load(gpu, tmp_path):
    # the map of 8 hall frames written as a NanoVDB FloatGrid: every leaf
    # (values, active masks) read back through the layout reader equals the
    # device map; a map loaded from the file equals it too, and 4 more
    # frames on both maps give the same map
    import nvdb_layout as L
    nm = gpu
    cam, frames = _hall(nm, 12)
    ctx = nm.Context(cam.width * cam.height, 200.0)
    gmap = nm.Map(0.05, *W.CLAMP, leaf_capacity=1 << 16)
    _feed(nm, ctx, gmap, None, frames[:8])
    path = tmp_path / "hall.nvdb"
    gmap.write_nvdb(path, "hall")
    o, v, m = gmap.download()
    _, grids = L.read_file(path.read_bytes())
    rec, name, grid = grids[0]
    assert name == "hall" and rec["voxel_size"] == (np.float32(0.05),) * 3
    got = grid.leaves()
    go = np.array([x[0] for x in got], np.int32)
    gv = np.stack([x[1] for x in got])
    gb = np.stack([x[2] for x in got])
    bits = ((m[:, np.arange(512) >> 5] >> (np.arange(512) & 31)) & 1).astype(bool)
    i = np.lexsort((o[:, 2], o[:, 1], o[:, 0]))
    j = np.lexsort((go[:, 2], go[:, 1], go[:, 0]))
    assert np.array_equal(o[i], go[j])
    assert np.array_equal(v[i].view(np.uint32), gv[j].view(np.uint32))
    assert np.array_equal(bits[i], gb[j])
    loaded = nm.Map.load_nvdb(path, "hall", *W.CLAMP)
    lo, lv, lm = loaded.download()
    k = np.lexsort((lo[:, 2], lo[:, 1], lo[:, 0]))
    assert np.array_equal(o[i], lo[k])
    assert np.array_equal(v[i].view(np.uint32), lv[k].view(np.uint32))
    assert np.array_equal(m[i], lm[k])
    _feed(nm, ctx, gmap, None, frames[8:])
    _feed(nm, ctx, loaded, None, frames[8:])
    a, b = gmap.download(), loaded.download()
    i = np.lexsort((a[0][:, 2], a[0][:, 1], a[0][:, 0]))
    k = np.lexsort((b[0][:, 2], b[0][:, 1], b[0][:, 0]))
    _same_map((a[0][i], a[1][i], a[2][i]), (b[0][k], b[1][k], b[2][k]))
