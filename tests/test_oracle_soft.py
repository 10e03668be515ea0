"""CPU pins of the soft-march oracle (oracle/nm_soft_oracle.cpp) and of the
host-side SO(3) helpers of nanomap_amd.soft.

The soft march has no counterpart in the reference; the oracle restates its
model (include/nanomap_amd_soft.h) in FP64.  Pinned here by:
  - a closed form: constant log-odds give a geometric series;
  - the hard limit: a +-50 log-odds half space gives the distance of the
    first sample past the surface (p -> {0, 1});
  - invariance: moving the camera and the grid origin together;
  - FP64 central finite differences of the oracle's own forward for its
    analytic adjoint (pose and log-odds).
"""
import numpy as np
import pytest
import torch

import oracle
from nanomap_amd import soft


def _look(axis_to):
    """Camera rotation (optical -> world) with the optical z axis along
    axis_to, y down (-world z where possible)."""
    z = np.asarray(axis_to, np.float64)
    z = z / np.linalg.norm(z)
    up = np.array([0.0, 0.0, 1.0]) if abs(z[2]) < 0.9 else np.array([1.0, 0, 0])
    x = np.cross(z, up)
    x /= np.linalg.norm(x)
    y = np.cross(z, x)
    return np.stack([x, y, z], 1)  # columns = optical axes in world


def _cam(R, p):
    return np.concatenate([np.asarray(R, np.float64).reshape(9),
                           np.asarray(p, np.float64)])


def test_constant_grid_closed_form():
    l0, h, N = -0.7, 0.03, 50
    g = np.full((6, 6, 6), l0)
    sc = oracle.SoftScene(g, [0, 0, 0], 0.1, l0, 3, 3, 60.0, 40.0, h, N)
    cams = np.stack([_cam(_look([1, 0.2, 0.1]), [0.31, 0.27, 0.29])])
    D = sc.forward(cams)
    p = 1.0 / (1.0 + np.exp(-l0))
    q = 1.0 - p
    # sum_{i<N} p q^i (i + 1/2) h, by the closed forms of sum q^i, i q^i
    s0 = (1 - q ** N) / (1 - q)
    s1 = q * (1 - N * q ** (N - 1) + (N - 1) * q ** N) / (1 - q) ** 2
    expect = p * h * (s1 + 0.5 * s0)
    assert np.allclose(D, expect, rtol=1e-13, atol=0)


def test_hard_limit_is_first_sample_past_the_surface():
    res = 0.1
    g = np.full((40, 9, 9), -50.0)
    g[23:, :, :] = 50.0  # occupied for x >= 2.3 m
    sc = oracle.SoftScene(g, [0, 0, 0], res, -50.0, 5, 5, 30.0, 30.0, res, 35)
    # camera on a voxel face looking along +x: the centre pixel samples the
    # voxel centres 0.45, 0.55, ... m, where the spline is -37.5 (last free
    # voxel) or +37.5 (first occupied voxel); the first past 2.3 m is 2.35 m
    eye = [0.4, 0.45, 0.45]
    cams = np.stack([_cam(_look([1, 0, 0]), eye)])
    D = sc.forward(cams)
    d_first = 2.35 - eye[0]
    assert abs(D[0, 2, 2] - d_first) < 1e-12
    # moving the eye by half a voxel moves every sample onto a face, where
    # the spline is 0 (p = 1/2) at the surface: the depth is then the mean
    # of the face sample and the next one
    cams = np.stack([_cam(_look([1, 0, 0]), [0.45, 0.45, 0.45])])
    D = sc.forward(cams)
    assert abs(D[0, 2, 2] - 0.5 * ((2.3 - 0.45) + (2.4 - 0.45))) < 1e-9


def test_translation_invariance():
    rng = np.random.RandomState(0)
    g = rng.normal(0.0, 1.5, (12, 11, 10))
    cams = np.stack([_cam(_look([1, 0.3, -0.2]), [0.35, 0.42, 0.51])])
    a = oracle.SoftScene(g, [0.0, 0.0, 0.0], 0.1, 0.3, 6, 4, 70, 50, 0.037, 40)
    b = oracle.SoftScene(g, [1.25, -2.5, 0.75], 0.1, 0.3, 6, 4, 70, 50, 0.037, 40)
    shifted = cams.copy()
    shifted[0, 9:] += [1.25, -2.5, 0.75]
    assert np.allclose(a.forward(cams), b.forward(shifted), rtol=1e-12, atol=1e-12)


def _smooth_grid(shape, seed):
    rng = np.random.RandomState(seed)
    g = rng.normal(0.0, 2.0, shape)
    for ax in range(3):
        g = (np.roll(g, 1, ax) + 2 * g + np.roll(g, -1, ax)) / 4
    return g


def test_oracle_adjoint_matches_central_differences():
    g = _smooth_grid((24, 22, 20), 1)
    sc = oracle.SoftScene(g, [0, 0, 0], 0.1, 0.2, 7, 5, 60.0, 45.0, 0.043, 36)
    cams = np.stack([_cam(_look([1, 0.4, 0.15]), [0.52, 0.61, 0.83]),
                     _cam(_look([0.3, 1, -0.2]), [0.77, 0.31, 1.05])])
    rng = np.random.RandomState(2)
    c = rng.normal(size=(2, 5, 7))
    gp, gg = sc.backward(cams, c)
    eps = 1e-6

    def loss(cs):
        return float((sc.forward(cs) * c).sum())
    for cam in range(2):
        fd = np.zeros(6)
        for k in range(6):
            xi = np.zeros(6)
            xi[k] = eps
            cp, cm = cams.copy(), cams.copy()
            cp[cam] = oracle.perturb_camera(cams[cam], xi)
            cm[cam] = oracle.perturb_camera(cams[cam], -xi)
            fd[k] = (loss(cp) - loss(cm)) / (2 * eps)
        assert np.linalg.norm(gp[cam] - fd) <= 1e-6 * np.linalg.norm(fd), (gp[cam], fd)
    # log-odds: the largest-gradient voxels and a few others
    idx = np.argsort(-np.abs(gg).ravel())[:6]
    for flat in idx:
        gplus, gminus = g.copy(), g.copy()
        gplus.ravel()[flat] += eps
        gminus.ravel()[flat] -= eps
        sp = oracle.SoftScene(gplus, [0, 0, 0], 0.1, 0.2, 7, 5, 60.0, 45.0, 0.043, 36)
        sm = oracle.SoftScene(gminus, [0, 0, 0], 0.1, 0.2, 7, 5, 60.0, 45.0, 0.043, 36)
        fd = ((sp.forward(cams) * c).sum() - (sm.forward(cams) * c).sum()) / (2 * eps)
        assert abs(gg.ravel()[flat] - fd) <= 1e-6 * abs(fd) + 1e-12


def test_so3_helpers():
    torch.manual_seed(0)
    w = torch.randn(5, 3, dtype=torch.float64) * 0.7
    R = soft.so3_exp(w)
    assert torch.allclose(R @ R.transpose(1, 2),
                          torch.eye(3, dtype=torch.float64).expand(5, 3, 3),
                          atol=1e-14)
    # exp((w + d)^) ~ exp((J_l(w) d)^) exp(w^), error O(|d|^2)
    d = torch.randn(5, 3, dtype=torch.float64) * 1e-6
    lhs = soft.so3_exp(w + d)
    rhs = soft.so3_exp((soft.so3_left_jacobian(w) @ d.unsqueeze(-1)).squeeze(-1)) @ R
    assert (lhs - rhs).abs().max() < 1e-11
    # a rotation by pi/2 about z maps x to y
    Rz = soft.so3_exp(torch.tensor([[0.0, 0.0, np.pi / 2]], dtype=torch.float64))
    assert torch.allclose(Rz[0] @ torch.tensor([1.0, 0, 0], dtype=torch.float64),
                          torch.tensor([0.0, 1.0, 0], dtype=torch.float64),
                          atol=1e-15)
