"""Pins of the CPU oracle (oracle/nm_oracle.cpp), on the CPU.

The reference ships no tests, fixtures or golden vectors for this path and
cannot be built here (it needs CUDA, OpenVDB, NanoVDB and Eigen), so the
oracle is pinned by what the mathematics fixes:

  ray set-up + DDA (Handler.cpp:58-85, openvdb::math::DDA):
    - closed forms: axis-aligned rays, the exact three-way tie (1,1,1),
      the max-range hit rule, skipped rays (NaN, too long);
    - invariants on random rays: unit steps, monotone direction,
      updates = 1 + |end - start|_1;
    - brute force: exact rational traversal (fractions) of random rays,
      near-ties excluded, with the ray set-up recomputed in numpy (IEEE
      round-to-nearest), so a wrong tie rule, a dropped first voxel or an
      off-by-one end voxel fails;
  FP32 fold in a rounding mode: numpy sequential sums (nearest) and exact
    rational rounding toward -inf, plus the bit patterns measured in the
    survey (7 x -0.4f: 0xC0333334 nearest, 0xC0333335 down);
  temp accumulate + integrate (Handler.cpp:88-122): closed forms for one
    ray, repeated frames, clamp saturation, activity thresholds;
  voxel filter restatement: an independent pure-numpy binning.

Parity unpinned (no check can reach them without the reference build):
Eigen's quaternion rounding, OpenVDB's compiled worldToIndex, the
reference GPU filter's atomic order.  The harness computes R and the eye
once on the host and feeds identical values to oracle and kernels.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle

RN, RD = 0, 1


def f32bits(x):
    return int(np.float32(x).view(np.uint32))


# ---------------------------------------------------------------- DDA ----

def test_axis_aligned_rays(oracle_lib):
    res = 0.1
    for axis in range(3):
        for sign in (1, -1):
            eye = [3.25, -2.5, 7.75]
            d = [0.0, 0.0, 0.0]
            d[axis] = sign * 1.234  # metres
            out = oracle.dda_trace(eye, d, res, 50.0, RN)
            assert out is not None
            vox, hit = out
            L = np.float64(np.float32(1.234)) / np.float64(np.float32(res))
            end = math.floor(eye[axis] + sign * L)
            start = math.floor(eye[axis])
            n = abs(end - start) + 1
            assert len(vox) == n and hit
            for k, v in enumerate(vox):
                exp = [math.floor(e) for e in eye]
                exp[axis] = start + sign * k
                assert list(v) == exp


def test_exact_three_way_tie_order(oracle_lib):
    # (1,1,1) from a voxel centre: all crossing times tie exactly; the
    # DDA resolves ties toward the higher axis: z, then y, then x.
    vox, hit = oracle.dda_trace([0.5, 0.5, 0.5], [0.1, 0.1, 0.1], 0.1, 50.0, RN)
    exp = [(0, 0, 0), (0, 0, 1), (0, 1, 1), (1, 1, 1)]
    assert [tuple(v) for v in vox[:4]] == exp
    assert hit


def test_hit_rule_exit_time(oracle_lib):
    # eye x = 0.5, ray +x of 0.99 m at 0.1 m: end voxel 10 exits at t = 10.5
    vox, hit = oracle.dda_trace([0.5, 0.5, 0.5], [0.99, 0, 0], 0.1, 10.2, RN)
    assert [v[0] for v in vox] == list(range(10)) and not hit
    vox, hit = oracle.dda_trace([0.5, 0.5, 0.5], [0.99, 0, 0], 0.1, 10.6, RN)
    assert [v[0] for v in vox] == list(range(11)) and hit


def test_skipped_rays(oracle_lib):
    assert oracle.dda_trace([0.5, 0.5, 0.5], [np.nan, 0, 0], 0.1, 50.0, RN) is None
    assert oracle.dda_trace([0.5, 0.5, 0.5], [5.01, 0, 0], 0.1, 50.0, RN) is None
    # 5.0 / double(0.1f) = 49.99999925 < 50: valid, but the end voxel 49
    # exits at t = 49.5 + 1 > 50, so no hit.
    vox, hit = oracle.dda_trace([0.5, 0.5, 0.5], [5.0, 0, 0], 0.1, 50.0, RN)
    assert len(vox) == 50 and not hit


def _setup_rn(p, res):
    """Ray set-up of Handler.cpp:65-75 in numpy (IEEE nearest)."""
    x, y, z = (np.float64(np.float32(c)) for c in p)
    ln = np.sqrt(x * x + y * y + z * z)
    L = ln / np.float64(np.float32(res))
    s = np.float64(1.0) / ln
    return (float(x * s), float(y * s), float(z * s)), float(L)


def _exact_traverse(eye, d, L):
    """Exact rational voxel traversal of eye + t d over t in [0, L]."""
    v = [math.floor(e) for e in eye]
    ev = []
    for a in range(3):
        if d[a] == 0:
            continue
        s = 1 if d[a] > 0 else -1
        k = v[a] + 1 if s > 0 else v[a]
        while True:
            t = (Fraction(k) - Fraction(eye[a])) / Fraction(d[a])
            if t > L:
                break
            ev.append((t, a, s))
            k += s
        # next crossing beyond L (exit time of the end voxel on this axis)
        ev.append((t, a, None))
    inside = sorted([e for e in ev if e[2] is not None], key=lambda e: (e[0], -e[1]))
    exits = [e[0] for e in ev if e[2] is None]
    seq = [tuple(v)]
    for t, a, s in inside:
        v[a] += s
        seq.append(tuple(v))
    times = [e[0] for e in inside]
    return seq, times, min(exits)


def test_brute_force_rational_traversal(oracle_lib):
    rng = np.random.default_rng(12345)
    res = np.float32(0.1)
    checked = 0
    for _ in range(300):
        eye = list(rng.uniform(-20, 20, 3))
        p = rng.normal(size=3)
        p = p / np.linalg.norm(p) * rng.uniform(0.05, 1.5)
        p = p.astype(np.float32)
        d, L = _setup_rn(p, res)
        seq, times, t_exit = _exact_traverse(eye, d, L)
        # exclude near-ties and near-boundary ends
        ts = [float(t) for t in times] + [float(t_exit)]
        gaps = np.diff(sorted([0.0] + ts))
        if len(gaps) and gaps.min() < 1e-9:
            continue
        if abs(float(t_exit) - L) < 1e-9 or (times and abs(float(times[-1]) - L) < 1e-9):
            continue
        max_time = 50.0
        out = oracle.dda_trace(eye, p, float(res), max_time, RN)
        assert out is not None
        vox, hit = out
        assert [tuple(v) for v in vox] == seq
        assert hit == (float(t_exit) < max_time)
        checked += 1
    assert checked > 250


def test_random_ray_invariants_both_modes(oracle_lib):
    rng = np.random.default_rng(7)
    for mode in (RN, RD):
        for _ in range(200):
            eye = rng.uniform(-100, 100, 3)
            p = (rng.normal(size=3) * rng.uniform(0.1, 4.0)).astype(np.float32)
            out = oracle.dda_trace(eye, p, 0.05, 200.0, mode)
            if out is None:
                continue
            vox, hit = out
            steps = np.abs(np.diff(vox, axis=0)).sum(axis=1)
            assert np.all(steps == 1)
            sgn = np.sign(p.astype(np.float64))
            dv = np.diff(vox, axis=0)
            assert np.all(dv * sgn[None, :] >= 0)
            assert len(vox) == 1 + np.abs(vox[-1] - vox[0]).sum()


def test_rn_and_rd_agree_away_from_ties(oracle_lib):
    rng = np.random.default_rng(99)
    same = 0
    for _ in range(200):
        eye = rng.uniform(-30, 30, 3)
        p = (rng.normal(size=3) * 2).astype(np.float32)
        a = oracle.dda_trace(eye, p, 0.05, 400.0, RN)
        b = oracle.dda_trace(eye, p, 0.05, 400.0, RD)
        assert (a is None) == (b is None)
        if a is not None and len(a[0]) == len(b[0]) and np.all(a[0] == b[0]):
            same += 1
    assert same >= 190


# --------------------------------------------------------------- fold ----

def _round_down_f32(x):
    """Largest float32 <= x (x a Fraction)."""
    f = np.float32(float(x))
    while Fraction(float(f)) > x:
        f = np.nextafter(f, np.float32(-np.inf))
    while Fraction(float(np.nextafter(f, np.float32(np.inf)))) <= x:
        f = np.nextafter(f, np.float32(np.inf))
    return f


def test_fold_bit_patterns(oracle_lib):
    assert f32bits(oracle.fold(0.0, -0.4, 7, RN)) == 0xC0333334
    assert f32bits(oracle.fold(0.0, -0.4, 7, RD)) == 0xC0333335
    assert f32bits(-0.4) == 0xBECCCCCD and f32bits(0.85) == 0x3F59999A


def test_fold_round_nearest_matches_numpy(oracle_lib):
    rng = np.random.default_rng(3)
    for _ in range(50):
        v = np.float32(rng.uniform(-50, 50))
        a = np.float32(rng.choice([-0.4, 0.85, rng.uniform(-3, 3)]))
        k = int(rng.integers(0, 3000))
        acc = v
        for _ in range(k):
            acc = np.float32(acc + a)
        assert f32bits(oracle.fold(v, a, k, RN)) == f32bits(acc)


def test_fold_round_down_matches_exact_rational(oracle_lib):
    rng = np.random.default_rng(4)
    for _ in range(40):
        v = np.float32(rng.uniform(-20, 20))
        a = np.float32(rng.choice([-0.4, 0.85, rng.uniform(-3, 3)]))
        k = int(rng.integers(0, 400))
        acc = v
        for _ in range(k):
            acc = _round_down_f32(Fraction(float(acc)) + Fraction(float(a)))
        assert f32bits(oracle.fold(v, a, k, RD)) == f32bits(acc)


# ------------------------------------------------- temp + integrate ------

class _F:
    """Minimal frame record for the oracle binding."""

    def __init__(self, eye, res=0.1, max_time=50.0, rounding=RN):
        self.rotation = [1, 0, 0, 0, 1, 0, 0, 0, 1]
        self.eye = eye
        self.grid_res = res
        self.max_time = max_time
        self.logodds_hit = 0.85
        self.logodds_miss = -0.4
        self.rounding = rounding


def test_single_ray_temp_and_map(oracle_lib):
    f = _F([0.5, 0.5, 0.5])
    cloud = np.array([[0.95, 0.0, 0.0]], np.float32)  # 9.5 voxels along x
    om = oracle.Map()
    t = oracle.process_frame(om, f, cloud)
    ijk, miss, hit, val = t.export()
    assert [tuple(v) for v in ijk] == [(i, 0, 0) for i in range(10)]
    assert list(miss) == [1] * 9 + [0] and list(hit) == [0] * 9 + [1]
    assert np.all(val[:9] == np.float32(-0.4)) and val[9] == np.float32(0.85)
    ok, v, act = om.get(3, 0, 0)
    assert ok and v == np.float32(-0.4) and not act
    ok, v, act = om.get(9, 0, 0)
    assert ok and v == np.float32(0.85) and act
    assert om.get(100, 0, 0)[0] is False


def test_clamp_saturation_and_thresholds(oracle_lib):
    f = _F([0.5, 0.5, 0.5])
    cloud = np.array([[0.95, 0.0, 0.0]], np.float32)
    om = oracle.Map()
    for k in range(1, 7):
        oracle.process_frame(om, f, cloud)
        _, v, act = om.get(0, 0, 0)
        exp = max(-2.0, float(oracle.fold(0.0, -0.4, k, RN)))
        assert v == np.float32(exp)
        _, v, act = om.get(9, 0, 0)
        assert v == np.float32(min(3.5, float(oracle.fold(0.0, 0.85, k, RN))))
        assert act
    _, v, _ = om.get(0, 0, 0)
    assert v == np.float32(-2.0)
    _, v, _ = om.get(9, 0, 0)
    assert v == np.float32(3.5)


def test_duplicate_rays_sum_in_pixel_order(oracle_lib):
    # hit of ray 0 and miss of ray 1 on voxel 9: temp = fl(0.85 - 0.4)
    f = _F([0.5, 0.5, 0.5])
    cloud = np.array([[0.95, 0, 0], [1.15, 0, 0]], np.float32)
    t = oracle.Temp()
    t.cast_cloud(f, cloud)
    ijk, miss, hit, val = t.export()
    i9 = [k for k, v in enumerate(ijk) if tuple(v) == (9, 0, 0)][0]
    assert miss[i9] == 1 and hit[i9] == 1
    assert val[i9] == np.float32(np.float32(0.85) + np.float32(-0.4))


# ------------------------------------------------------------ filter ----

def _filter_numpy(cloud, R, eye, res, max_time, precision=0):
    """Ascending-key rays; precision 1 sums FP16-rounded fractions and the
    count in numpy's float16 (correctly rounded: float32 holds the exact
    sum of two halves' 11-bit significands to 24 bits, so rounding twice
    is innocuous), 2 and 3 give voxel centres, 3 with an 8-bit count."""
    R = np.asarray(R, np.float32).reshape(3, 3)
    origin = np.array([np.float32(e - math.floor(e)) for e in eye], np.float32)
    bins = {}
    max_range = np.float32(np.float32(max_time) * np.float32(res))
    for p in cloud:
        p = p.astype(np.float32)
        if np.any(np.isnan(p)):
            continue
        nrm = np.sqrt(np.float32(np.float32(p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]))
        if not (nrm > 0 and nrm < max_range):
            continue
        r = [np.float32(np.float32(R[i, 0] * p[0] + R[i, 1] * p[1]) + R[i, 2] * p[2]) for i in range(3)]
        key, fr = [], []
        for a in range(3):
            q = np.float32(r[a] / np.float32(res))
            key.append(int(math.floor(np.float32(q + origin[a]))))
            fr.append(np.float32(q - np.float32(math.floor(q))))
        k = tuple(key)
        if k not in bins:
            bins[k] = [np.float32(0)] * 3 + [0] + [np.float16(0)] * 4
        b = bins[k]
        for a in range(3):
            b[a] = np.float32(b[a] + fr[a])
            b[4 + a] = np.float16(b[4 + a] + np.float16(fr[a]))
        b[3] += 1
        b[7] = np.float16(b[7] + np.float16(1))
    out = []
    for k in sorted(bins):
        b = bins[k]
        if precision == 0:
            c = np.float32(b[3])
            pt = [np.float32(np.float32(b[a] / c) + np.float32(k[a])) for a in range(3)]
        elif precision == 1:
            c = np.float32(b[7])
            pt = [np.float32(np.float32(np.float32(b[4 + a]) / c) + np.float32(k[a]))
                  for a in range(3)]
        else:
            if precision == 3 and b[3] % 256 == 0:
                continue
            pt = [np.float32(np.float32(k[a]) + np.float32(0.5)) for a in range(3)]
        ln = np.sqrt(np.float32(np.float32(pt[0] * pt[0] + pt[1] * pt[1]) + pt[2] * pt[2]))
        if ln > 0 and ln < np.float32(max_time):
            out.append([pt[0] / ln, pt[1] / ln, pt[2] / ln, ln])
    return np.array(out, np.float32).reshape(-1, 4)


def test_voxel_filter_matches_numpy_restatement(oracle_lib):
    rng = np.random.default_rng(5)
    cloud = rng.normal(size=(600, 3)).astype(np.float32) * np.float32(0.6)
    cloud[::37] = np.nan
    th = 0.3
    R = [math.cos(th), -math.sin(th), 0, math.sin(th), math.cos(th), 0, 0, 0, 1]
    f = _F([12.3, -4.7, 2.2], res=0.1, max_time=20.0)
    f.rotation = [float(np.float32(x)) for x in R]
    got = oracle.voxel_filter(f, cloud)
    exp = _filter_numpy(cloud, f.rotation, f.eye, f.grid_res, f.max_time)
    assert got.shape == exp.shape and got.shape[0] > 50
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))


@pytest.mark.parametrize("precision", [0, 1, 2, 3])
def test_voxel_filter_precision_modes(oracle_lib, precision):
    # Scattered points plus dense voxels: 256 points (the 8-bit count
    # wraps to 0), 300, and 2100 (the FP16 count stops at 2048).
    rng = np.random.default_rng(17 + precision)
    f = _F([12.3, -4.7, 2.2], res=0.1, max_time=20.0)
    origin = np.array([e - math.floor(e) for e in f.eye])
    parts = [rng.normal(size=(400, 3)).astype(np.float32) * np.float32(0.6)]
    for n, k in ((256, (14, 8, 5)), (300, (-3, 4, 3)), (2100, (2, -4, -1))):
        c = (np.asarray(k) + 0.5 - origin) * 0.1  # that voxel's centre
        parts.append((c + rng.uniform(-0.03, 0.03, size=(n, 3))).astype(np.float32))
    cloud = np.concatenate(parts)
    got = oracle.voxel_filter(f, cloud, precision=precision)
    exp = _filter_numpy(cloud, f.rotation, f.eye, f.grid_res, f.max_time,
                        precision)
    assert got.shape == exp.shape and got.shape[0] > 50
    assert np.array_equal(got.view(np.uint32), exp.view(np.uint32))
    if precision == 1:  # FP16 sums differ from FP32 ones in some bins
        ref = oracle.voxel_filter(f, cloud, precision=0)
        assert ref.shape == got.shape and not np.array_equal(ref, got)
    if precision == 3:  # exactly the 256-point voxel is dropped
        assert got.shape[0] == oracle.voxel_filter(f, cloud,
                                                   precision=2).shape[0] - 1


def test_half_rounding_matches_numpy_float16(oracle_lib):
    # Subnormal, tie and saturating-count cases of the FP16 sums, checked
    # through single-point and repeated-point voxels.
    f = _F([0.5, 0.5, 0.5], res=1.0, max_time=20.0)
    for x in (1.0 + 2.0 ** -12, 1.0 + 3 * 2.0 ** -12, 1.0 + 2.0 ** -20,
              1.5 + 2.0 ** -11, 2.0 + 1e-7):
        cloud = np.array([[x, 0.0, 0.0]] * 3, np.float32)
        got = oracle.voxel_filter(f, cloud, precision=1)
        exp = _filter_numpy(cloud, f.rotation, f.eye, f.grid_res, f.max_time, 1)
        assert np.array_equal(got.view(np.uint32), exp.view(np.uint32)), x


def test_frame_cast_totals_match_ray_traces(oracle_lib):
    rng = np.random.default_rng(11)
    cloud = (rng.normal(size=(200, 3)) * 1.5).astype(np.float32)
    f = _F([10.3, 20.6, 3.1], res=0.1, max_time=40.0)
    t = oracle.Temp()
    nrays = t.cast_cloud(f, cloud)
    _, miss, hit, _ = t.export()
    tot_m = tot_h = valid = 0
    for p in cloud:
        out = oracle.dda_trace(f.eye, p, 0.1, 40.0, RN)
        if out is None:
            continue
        valid += 1
        vox, h = out
        tot_h += int(h)
        tot_m += len(vox) - int(h)
    assert nrays == valid and int(miss.sum()) == tot_m and int(hit.sum()) == tot_h
    assert t.visits() == tot_m + tot_h
