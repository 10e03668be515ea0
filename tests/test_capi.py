"""CPU tests of the C ABI: the library loads, exports every symbol the
public headers declare, and its host-only functions agree with the oracle.
No device calls here."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(nm_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_symbols()
    for n in ("nm_ctx_create", "nm_cast", "nm_update_map", "nm_frame_submit",
              "nm_map_create", "nm_map_download", "nm_scene_render"):
        assert n in names


def test_library_exports_every_declared_symbol(nm):
    L = ctypes.CDLL(nm._capi.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each of them
    from nanomap_amd._capi import SIGNATURES
    assert not [n for n in declared_symbols() if n not in SIGNATURES]


def test_fold_host_matches_oracle_sequential_sums(nm, oracle_lib):
    rng = np.random.default_rng(21)
    cases = [(0.0, -0.4, k) for k in (0, 1, 2, 7, 100, 4096, 300000, 921600)]
    cases += [(0.0, 0.85, k) for k in (1, 5, 3000, 100000)]
    for _ in range(300):
        v = float(np.float32(rng.uniform(-200, 200)))
        a = float(np.float32(rng.choice([-0.4, 0.85, rng.uniform(-2, 2)])))
        cases.append((v, a, int(rng.integers(0, 20000))))
    for mode in (0, 1):
        for v, a, k in cases:
            g = np.float32(nm.fold(v, a, k, mode)).view(np.uint32)
            o = np.float32(oracle.fold(v, a, k, mode)).view(np.uint32)
            assert g == o, (mode, v, a, k)


def _pow2_neighbours():
    out = []
    for e in (-20, -3, 0, 1, 5, 23, 40):
        p = np.float32(2.0 ** e)
        for f in (p, np.nextafter(p, np.float32(np.inf)),
                  np.nextafter(np.nextafter(p, np.float32(np.inf)),
                               np.float32(np.inf)),
                  np.float32(p * np.float32(1.0 + 3 * 2.0 ** -23)),
                  np.nextafter(p, np.float32(-np.inf))):
            out += [float(f), -float(f)]
    return out


def test_fold_host_tiny_increments_at_binade_edges(nm, oracle_lib):
    """Increments far below the double ulp of v, with v on or next to a
    power of two: a jump must stop where the exact sum first leaves the
    binade, also when the step would land exactly on 2^e."""
    rng = np.random.default_rng(22)
    tiny = [1e-20, 2.0 ** -60, 3e-30, 1e-12, 2.0 ** -40 * 3, 7e-9]
    cases = [(1.0, -1e-20, 30), (float(np.float32(1.0 + 3 * 2.0 ** -23)),
                                 -2.0 ** -60, 100)]
    for v in _pow2_neighbours():
        for t in tiny:
            for sgn in (-1.0, 1.0):
                for k in (25, 26, 31, 64, 200, 5000):
                    cases.append((v, float(np.float32(sgn * t)), k))
    # increments a little below/above one ulp of v, and random mantissas
    for _ in range(400):
        e = int(rng.integers(-10, 30))
        v = float(np.float32(rng.choice([-1, 1]) * 2.0 ** e *
                             rng.choice([1.0, 1.0 + 2.0 ** -23,
                                         rng.uniform(1, 2)])))
        ulp = 2.0 ** (e - 23)
        a = float(np.float32(rng.choice([-1, 1]) * ulp *
                             rng.choice([0.25, 0.5, 0.75, 1.0, 1.5,
                                         rng.uniform(0, 3)])))
        cases.append((v, a, int(rng.integers(25, 3000))))
    for mode in (0, 1):
        for v, a, k in cases:
            g = np.float32(nm.fold(v, a, k, mode)).view(np.uint32)
            o = np.float32(oracle.fold(v, a, k, mode)).view(np.uint32)
            assert g == o, (mode, v, a, k, hex(int(g)), hex(int(o)))
    # the reported case, against the exact rational round-down sum
    from fractions import Fraction
    acc = np.float32(1.0)
    for _ in range(30):
        x = Fraction(float(acc)) + Fraction(float(np.float32(-1e-20)))
        f = np.float32(float(x))
        while Fraction(float(f)) > x:
            f = np.nextafter(f, np.float32(-np.inf))
        acc = f
    assert np.float32(nm.fold(1.0, -1e-20, 30, 1)) == acc
    assert np.float32(acc).view(np.uint32) == np.float32(
        float.fromhex("0x1.ffffc4p-1")).view(np.uint32)


def test_fold_host_overflow_and_non_finite(nm, oracle_lib):
    big = float(np.finfo(np.float32).max)
    # sums that overflow: round-down keeps FLT_MAX going up, reaches -inf
    # going down; an infinite or NaN start stays put; every case returns
    # at once even for 2^32 - 1 steps
    cases = [(big, 1e38, 5), (-big, -1e38, 5), (3e38, 2e37, 40),
             (float("inf"), -0.4, 7), (float("-inf"), 0.85, 7)]
    for mode in (0, 1):
        for v, a, k in cases:
            g = np.float32(nm.fold(v, a, k, mode)).view(np.uint32)
            o = np.float32(oracle.fold(v, a, k, mode)).view(np.uint32)
            assert g == o, (mode, v, a, k)
        for v in (float("inf"), float("-inf")):
            assert nm.fold(v, -0.4, 2 ** 32 - 1, mode) == v
        assert np.isnan(nm.fold(float("nan"), 0.85, 2 ** 32 - 1, mode))
        assert nm.fold(-big, -1e38, 2 ** 32 - 1, mode) == float("-inf")


def test_frame_from_pose_closed_forms(nm):
    # identity quaternion, identity frame transform -> R = I, eye = p / res
    I = np.eye(3, dtype=np.float32)
    f = nm.frame_from_pose([1.0, -2.0, 0.5], [1, 0, 0, 0], I, 0.1, 5.0,
                           0.85, -0.4, nm.ROUND_NEAREST)
    assert list(f.rotation) == [1, 0, 0, 0, 1, 0, 0, 0, 1]
    inv = 1.0 / float(np.float32(0.1))
    assert list(f.eye) == [float(np.float32(1.0)) * inv,
                           float(np.float32(-2.0)) * inv, 0.5 * inv]
    assert f.max_time == np.float32(np.float32(5.0) / np.float32(0.1))
    # 90 degrees about z: exact permutation matrix
    s = np.float32(np.sqrt(0.5))
    f = nm.frame_from_pose([0, 0, 0], [s, 0, 0, s], I, 0.1, 5.0, 0.85, -0.4, 0)
    R = np.array(f.rotation).reshape(3, 3)
    assert np.allclose(R, [[0, -1, 0], [1, 0, 0], [0, 0, 1]], atol=1e-6)
    # sample config frame transform maps optical z to body x
    f = nm.frame_from_pose([0, 0, 0], [1, 0, 0, 0], nm.DEFAULT_F, 0.1, 5.0,
                           0.85, -0.4, 0)
    R = np.array(f.rotation).reshape(3, 3)
    assert np.array_equal(R @ [0, 0, 1], [1, 0, 0])
    # round-down mode: max_time = RD(5 / 0.1f) < RN(5 / 0.1f)
    f_rd = nm.frame_from_pose([0, 0, 0], [1, 0, 0, 0], I, 0.1, 5.0, 0.85, -0.4, 1)
    f_rn = nm.frame_from_pose([0, 0, 0], [1, 0, 0, 0], I, 0.1, 5.0, 0.85, -0.4, 0)
    assert f_rd.max_time < f_rn.max_time


def test_scene_is_deterministic_and_plausible(nm):
    from nanomap_amd import workloads
    _, cam, frames = workloads.box_room_frame(nm, seed=1)
    _, _, frames2 = workloads.box_room_frame(nm, seed=1)
    a, b = frames[0][2], frames2[0][2]
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    z = a[:, 2]
    assert np.all(z > 0) and np.all(z <= 5.0)
    # a good share of pixels sees a surface inside the 5 m cap
    assert np.mean(z < 5.0) > 0.1
    _, _, frames3 = workloads.box_room_frame(nm, seed=2)
    assert not np.array_equal(a, frames3[0][2])


def test_dropouts_and_noise(nm):
    scene = nm.Scene(3, (10.0, 10.0, 4.0), 5)
    pos, quat = scene.sample_pose(3)
    cam = nm.camera(64, 48, 80.0, 60.0, 8.0, noise_k=0.0, dropout=0.25)
    c = scene.render(cam, pos, quat, 7)
    frac = np.mean(np.isnan(c[:, 0]))
    assert 0.15 < frac < 0.35
    cam0 = nm.camera(64, 48, 80.0, 60.0, 8.0, noise_k=0.0, dropout=0.0)
    camn = nm.camera(64, 48, 80.0, 60.0, 8.0, noise_k=0.002, dropout=0.0)
    c0 = scene.render(cam0, pos, quat, 7)
    cn = scene.render(camn, pos, quat, 7)
    d = np.abs(cn[:, 2] - c0[:, 2])
    ok = c0[:, 2] < 7.9
    assert np.all(d[ok] < 0.002 * c0[ok, 2] ** 2 * 6 + 1e-6)
    assert np.mean(d[ok] > 0) > 0.9


def test_context_needs_device(nm):
    import torch
    if torch.cuda.is_available():
        pytest.skip("device present: covered by gpu tests")
    with pytest.raises(nm.NanoMapError):
        nm.Context(100, 50.0)


def test_frustum_box_closed_form(nm):
    # FrustumData::calculateSensorBounds: half height = range tan(vfov/2),
    # half width = aspect * half height, z from 0 (the origin corner) to
    # the range; the half extents are positive products, so round-down
    # shrinks the box on both sides.
    for w, h, vfov, rmin, rmax in ((640, 480, 58.0, 0.0, 10.0),
                                   (160, 120, 45.0, 0.3, 5.0),
                                   (848, 480, 60.0, 0.1, 15.0)):
        lo_n, hi_n = nm.frustum_box(w, h, vfov, rmin, rmax, nm.ROUND_NEAREST)
        lo_d, hi_d = nm.frustum_box(w, h, vfov, rmin, rmax, nm.ROUND_DOWN)
        hh = rmax * np.tan(np.radians(vfov) / 2)
        hw = w / h * hh
        assert np.allclose(hi_n, [hw, hh, rmax], rtol=2e-7)
        assert np.allclose(lo_n, [-hw, -hh, 0.0], rtol=2e-7)
        assert np.all(hi_d <= hi_n) and np.all(lo_d >= lo_n)
        assert np.all(np.abs(hi_d - hi_n) <= 2 * np.spacing(hi_n))
    # 87x58 deg at 640x480: the box is the aspect-derived 72.9 deg, not 87
    lo, hi = nm.frustum_box(640, 480, 58.0, 0.0, 10.0)
    assert abs(np.degrees(2 * np.arctan(hi[0] / 10.0)) - 72.9) < 0.1
    with pytest.raises(nm.NanoMapError):
        nm.frustum_box(640, 480, 190.0, 0.0, 10.0)


def test_map_load_rejects_bad_files_before_any_device_work(nm, tmp_path):
    # the snapshot is parsed and checked on the host first
    p = tmp_path / "junk.nmap"
    p.write_bytes(b"NOTAMAP!" + bytes(100))
    with pytest.raises(nm.NanoMapError):
        nm.Map.load(p)
    with pytest.raises(nm.NanoMapError):
        nm.Map.load(tmp_path / "missing.nmap")


def test_depth_back_projection_closed_forms(nm):
    # (d ax[u], d ay[v], d): z is the depth itself; columns u and W-1-u
    # mirror exactly (cx = W/2); the corner ray's angle is the half FOV
    # less half a pixel; NaN depth gives a NaN point
    W_, H_ = 640, 480
    cam = nm.depth_camera(W_, H_, 87.0, 58.0)
    assert cam.cx == 320.0 and cam.cy == 240.0
    d = np.full(W_ * H_, 2.5, np.float32)
    d[7] = np.nan
    xyz = nm.depth_to_cloud(cam, d).reshape(H_, W_, 3)
    assert np.all(xyz[..., 2].reshape(-1)[np.arange(W_ * H_) != 7] == 2.5)
    assert np.isnan(xyz[0, 7]).all()
    assert np.array_equal(xyz[5, :, 0], -xyz[5, ::-1, 0])
    assert np.array_equal(xyz[:, 9, 1], -xyz[::-1, 9, 1])
    ang = np.degrees(np.arctan2(xyz[100, W_ - 1, 0], 2.5))
    want = np.degrees(np.arctan((W_ / 2 - 0.5) / cam.fx))
    assert abs(ang - want) < 1e-5 and abs(want - 43.5) < 0.1
    assert xyz[0, 0, 1] > 0 > xyz[H_ - 1, 0, 1]  # row 0 is the top: y up
    # the scene renderer's cloud is the same pinhole: its z is the depth
    from nanomap_amd import workloads
    _, cam2, frames = workloads.box_room_frame(nm)
    cloud = frames[0][2]
    c = nm.depth_camera(cam2.width, cam2.height, cam2.hfov_deg, cam2.vfov_deg)
    back = nm.depth_to_cloud(c, cloud[:, 2].copy())
    ok = ~np.isnan(cloud[:, 2])
    assert np.allclose(back[ok], cloud[ok], rtol=0, atol=2e-6)
    with pytest.raises(nm.NanoMapError):
        nm.depth_camera(0, 480, 87.0, 58.0)
