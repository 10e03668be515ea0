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
    rational rounding toward -inf, plus known bit patterns (7 x -0.4f
    summed in order: 0xC0333334 nearest, 0xC0333335 down);
  temp accumulate + integrate (Handler.cpp:88-122): closed forms for one
    ray, repeated frames, clamp saturation, activity thresholds;
  voxel filter: closed forms on exact inputs (integer eye: the endpoint
    is the bin's centroid and lies inside it; one point per bin: that
    point, or the point + 1 on axes where the origin offset carries; the
    voxel-centre modes emit bin + 0.5 and the 8-bit count drops exactly
    bins of 256 k points), brute-force bin counts on tiny clouds, the
    frustum box faces at the 640x480 87x58 deg sensor (on the face kept,
    one ulp outside dropped), the strict range cap, and FP16 sums against
    numpy's float16.

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
# The voxel filter (filterCloud.cu:84-133, 279-318) is pinned by what the
# mathematics fixes, on inputs where every quantity is exact: an identity
# or signed-permutation R (r = R p has no rounding), a power-of-two voxel
# size (q = r / res is exact) and points on a 1/64 m grid, so bins and
# centroids have closed forms to compare with.

DEFAULT_F = [0.0, 0.0, 1.0, -1.0, 0.0, 0.0, 0.0, 1.0, 0.0]  # r = (z, -x, y)


def _rays_endpoints(rays):
    """Endpoints (dir * len) of filter rays, in voxel units (float64)."""
    r = np.asarray(rays, np.float64)
    return r[:, :3] * r[:, 3:4]


def _clustered(rng, n_bins, per_bin, res, span=6):
    """Points on a 1/64 m grid, per_bin(k) of them in each of n_bins
    distinct voxels of size res (metres) around the origin."""
    pts, seen = [], set()
    while len(seen) < n_bins:
        b = tuple(int(x) for x in rng.integers(-span, span, size=3))
        if b in seen or b == (0, 0, 0):
            continue
        seen.add(b)
        k = per_bin(len(seen))
        cells = rng.integers(0, int(64 * res), size=(k, 3))
        pts.append((np.asarray(b) * res + cells / 64.0).astype(np.float32))
    return np.concatenate(pts)


@pytest.mark.parametrize("rot", ["identity", "frame"])
def test_voxel_filter_integer_eye_gives_bin_centroids(oracle_lib, rot):
    # Integer eye: voxelOriginOffset = 0, so a point's bin is floor(q) and
    # the emitted endpoint frac-mean + bin is the mean of the bin's q: the
    # centroid, which lies inside the bin.  One ray per occupied bin, in
    # ascending (x, y, z) bin order.
    rng = np.random.default_rng(51)
    res = 0.25
    cloud = _clustered(rng, 40, lambda k: 1 + k % 7, res)
    cloud = cloud[rng.permutation(len(cloud))]
    f = _F([12.0, -5.0, 2.0], res=res, max_time=60.0)
    f.rotation = [1, 0, 0, 0, 1, 0, 0, 0, 1] if rot == "identity" else DEFAULT_F
    Rm = np.asarray(f.rotation, np.float64).reshape(3, 3)
    rays = oracle.voxel_filter(f, cloud)
    q = (cloud.astype(np.float64) @ Rm.T) / res  # exact
    bins = {}
    for qi in q:
        bins.setdefault(tuple(int(math.floor(x)) for x in qi), []).append(qi)
    keys = sorted(bins)
    # (R negates x for "frame": points on a voxel's lower x face then fall
    # into the neighbouring bin, so there can be more than 40 bins)
    assert len(rays) == len(keys) >= 40
    ends = _rays_endpoints(rays)
    for k, e in zip(keys, ends):
        mean = [sum(Fraction(float(x[a])) for x in bins[k]) / len(bins[k])
                for a in range(3)]
        for a in range(3):
            assert abs(e[a] - float(mean[a])) < 2e-5, (k, a, e, mean)
            assert k[a] - 1e-5 <= e[a] <= k[a] + 1 + 1e-5


def test_voxel_filter_one_point_per_bin_is_that_point(oracle_lib):
    # One point per bin: the endpoint is frac(q) + floor(q + o) with the
    # eye's fractional part o (the bin carries the offset, the fraction
    # does not, filterCloud.cu:119-128): q itself, or q + 1 on an axis
    # where frac(q) + o >= 1.
    rng = np.random.default_rng(52)
    res = 0.5
    cloud = _clustered(rng, 60, lambda k: 1, res, span=10)
    f = _F([3.25, 7.5, -1.75], res=res, max_time=40.0)
    o = np.array([e - math.floor(e) for e in f.eye])
    rays = oracle.voxel_filter(f, cloud)
    q = cloud.astype(np.float64) / res
    expect = {}
    for qi in q:
        key = tuple(int(math.floor(qi[a] + o[a])) for a in range(3))
        expect[key] = np.array([qi[a] - math.floor(qi[a]) + key[a]
                                for a in range(3)])
    assert len(rays) == len(expect) == 60
    for k, e in zip(sorted(expect), _rays_endpoints(rays)):
        assert np.allclose(e, expect[k], rtol=0, atol=2e-5), (k, e, expect[k])
    # the offset carries (endpoint = q + 1) on some axis of some bins here
    carried = sum(int(np.any(np.floor(qi + o) > np.floor(qi))) for qi in q)
    assert 0 < carried < 60


def test_voxel_filter_ray_count_is_distinct_bins(oracle_lib):
    # Brute force on tiny clouds: the number of rays equals the number of
    # distinct bins floor(q + o) (exact here: q and o have few bits), for
    # every precision mode that emits a ray per bin.
    rng = np.random.default_rng(53)
    for trial in range(20):
        res = 0.25
        cloud = (rng.integers(-300, 300, size=(int(rng.integers(1, 40)), 3))
                 / 64.0).astype(np.float32)
        f = _F([5.25 + trial, -2.5, 0.75], res=res, max_time=60.0)
        o = [e - math.floor(e) for e in f.eye]
        keys = set()
        for p in cloud.astype(np.float64):
            if not 0 < np.linalg.norm(p) < 60.0 * res:
                continue
            keys.add(tuple(int(math.floor(p[a] / res + o[a]))
                           for a in range(3)))
        for prec in (0, 1, 2):
            got = oracle.voxel_filter(f, cloud, precision=prec)
            assert len(got) == len(keys), (trial, prec)


def _box_for(width, height, vfov, rmin, rmax):
    """FrustumData::calculateSensorBounds in closed form (float64)."""
    hh = rmax * math.tan(math.radians(vfov) / 2)
    hw = width / height * hh
    return np.array([-hw, -hh, 0.0]), np.array([hw, hh, rmax])


@pytest.mark.parametrize("precision", [0, 1, 2, 3])
def test_voxel_filter_box_faces(oracle_lib, precision):
    # The 640x480, 58 deg vertical FOV, 10 m sensor: box x within +-7.39 m
    # (the aspect-derived half width), y within +-5.54 m, z in [0, 10].  A
    # point exactly on a face is kept (the reference drops x < min and
    # x > max), one ulp outside is dropped, one ulp inside kept.  Each
    # probe point sits alone in its voxel, so the ray count says which.
    lo, hi = _box_for(640, 480, 58.0, 0.0, 10.0)
    lo = lo.astype(np.float32)
    hi = hi.astype(np.float32)
    res = 0.05
    f = _F([0.5, 0.5, 0.5], res=res, max_time=400.0)  # range cap 20 m
    f.box_min, f.box_max = list(lo), list(hi)
    centre = np.float32([0.0, 0.0, 5.0])
    kept, dropped = [], []
    for a in range(3):
        for face, out_dir in ((lo[a], -np.inf), (hi[a], np.inf)):
            if a == 2 and face == 0.0:
                continue  # z = 0 with x = y = 0 is the origin: |p| = 0
            base = centre.copy()
            if a == 2:
                base[0] = np.float32(1.0)  # z-faces probed off the axis
            p_on, p_out, p_in = base.copy(), base.copy(), base.copy()
            p_on[a] = face
            p_out[a] = np.nextafter(np.float32(face), np.float32(out_dir))
            p_in[a] = np.nextafter(np.float32(face), np.float32(-out_dir))
            kept += [p_on, p_in]
            dropped += [p_out]
    kept = np.array(kept, np.float32)
    dropped = np.array(dropped, np.float32)
    n_kept = len(oracle.voxel_filter(f, kept, precision=precision))
    # face / inside pairs share a voxel: one ray per pair
    assert n_kept == len(kept) // 2
    assert len(oracle.voxel_filter(f, dropped, precision=precision)) == 0
    both = np.concatenate([kept, dropped])
    assert len(oracle.voxel_filter(f, both, precision=precision)) == n_kept
    # without the box the outside probes count too (one voxel each)
    g = _F(f.eye, res=res, max_time=400.0)
    assert len(oracle.voxel_filter(g, dropped, precision=precision)) == \
        len(dropped) == 5


def test_voxel_filter_range_cap_is_strict(oracle_lib):
    # 0 < |p| < max_range (metres): a point at exactly max_range is
    # dropped, the float below it kept; the origin is dropped.
    f = _F([0.5, 0.5, 0.5], res=0.25, max_time=50.0)  # 12.5 m of rays
    f.max_range = 10.0
    on = np.array([[0.0, 0.0, 10.0]], np.float32)
    below = np.array([[0.0, 0.0, np.nextafter(np.float32(10.0),
                                               np.float32(0))]], np.float32)
    zero = np.zeros((1, 3), np.float32)
    assert len(oracle.voxel_filter(f, on)) == 0
    assert len(oracle.voxel_filter(f, below)) == 1
    assert len(oracle.voxel_filter(f, zero)) == 0


def test_voxel_filter_centre_modes(oracle_lib):
    # PrecisionType 2 and 3 emit the voxel centre bin + 0.5; the 8-bit
    # count of type 3 wraps, so exactly the bins with 256 k points vanish.
    rng = np.random.default_rng(54)
    res = 0.25
    sizes = {1: 256, 2: 512, 3: 255, 4: 257, 5: 1, 6: 300}
    cloud = _clustered(rng, 6, lambda k: sizes[k], res)
    f = _F([4.0, 4.0, 4.0], res=res, max_time=60.0)
    r2 = oracle.voxel_filter(f, cloud, precision=2)
    r3 = oracle.voxel_filter(f, cloud, precision=3)
    assert len(r2) == 6 and len(r3) == 4
    for e in _rays_endpoints(r2):
        assert np.allclose(e - np.floor(e), 0.5, atol=2e-5)


def test_half_rounding_matches_numpy_float16(oracle_lib):
    # PrecisionType 1: each fraction rounded to FP16 and every sum, the
    # count included, rounded to FP16 (half2 atomics in pixel order).
    # numpy's float16 is an independent correctly rounded FP16: three equal
    # points in one voxel give ((h + h) + h) / 3h + bin, with h = fp16(f);
    # ties, a subnormal-free fraction and one that rounds up are covered.
    f = _F([0.0, 0.0, 0.0], res=1.0, max_time=20.0)
    for x in (1.0 + 2.0 ** -12, 1.0 + 3 * 2.0 ** -12, 1.0 + 2.0 ** -20,
              1.5 + 2.0 ** -11, 2.0 + 1e-7, 3.0 + 0.3333):
        cloud = np.array([[x, 0.5, 0.5]] * 3, np.float32)
        got = oracle.voxel_filter(f, cloud, precision=1)
        fr = np.float32(np.float32(x) - np.float32(math.floor(x)))
        h = np.float16(fr)
        s = np.float16(np.float16(h + h) + h)
        c = np.float16(np.float16(np.float16(1) + np.float16(1)) + np.float16(1))
        ex = np.float32(np.float32(np.float32(s) / np.float32(c)) +
                        np.float32(math.floor(x)))
        e = _rays_endpoints(got)[0]
        assert abs(e[0] - float(ex)) < 1e-6, (x, e[0], ex)
    # the FP16 count stops at 2048: 2100 points count as 2048
    cloud = np.array([[0.25, 0.5, 0.5]] * 2100, np.float32)
    got = oracle.voxel_filter(f, cloud, precision=1)
    h = np.float16(0)
    for _ in range(2100):
        h = np.float16(h + np.float16(1))
    assert float(h) == 2048.0
    e = _rays_endpoints(got)[0]
    sx = np.float16(0)
    for _ in range(2100):
        sx = np.float16(sx + np.float16(0.25))
    assert abs(e[0] - float(sx) / 2048.0) < 1e-6


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
