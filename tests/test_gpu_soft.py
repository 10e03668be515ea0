"""GPU checks of the differentiable soft march (include/nanomap_amd_soft.h).

The box-room world mapped at 0.05 m, 16 cameras of 320x240 with
an 8 m range (160 samples).  Checked:
  - the FP64 kernels against the FP64 oracle (forward and adjoint);
  - the FP32 forward against the FP64 one;
  - FP32 adjoint pose and log-odds gradients against FP64 central finite
    differences (eps = 1e-4), within 1e-4 relative;
  - the torch.autograd.Function, at zero and nonzero pose perturbations;
  - one batched launch against per-camera launches (bit-identical);
  - 20 Gauss-Newton steps of scan-to-map pose refinement from
    perturbations of up to 0.1 m and 3 deg.
"""
import numpy as np
import pytest
import torch

import oracle
from nanomap_amd import soft
from nanomap_amd import workloads as W

pytestmark = pytest.mark.gpu

RES = 0.05


@pytest.fixture(scope="module")
def soft_walk(gpu):
    nm = gpu
    _, cam, pos, quat, clouds = W.soft_walk(nm)
    ctx = nm.Context(cam.width * cam.height, 8.0 / RES + 1)
    m = nm.Map(RES, *W.CLAMP, leaf_capacity=1 << 19)
    frames = []
    for k in range(len(clouds)):
        f = W.frame(nm, pos[k], quat[k], RES, 8.0, nm.ROUND_DOWN)
        ctx.submit(m, f, clouds[k])
        frames.append(f)
    o, v, _ = m.download()
    grid, lo = W.dense_grid(o, v)
    print("soft_walk grid", grid.shape, "leaves", len(o), flush=True)
    # the same grid gathered on the device from the map (nm_map_dense)
    dev = m.dense(lo, grid.shape)
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy().view(np.uint32),
                          grid.view(np.uint32))
    sel = list(range(0, 48, 3))[:16]
    R = torch.tensor(np.stack([np.array(list(frames[k].rotation),
                                        np.float64).reshape(3, 3) for k in sel]))
    p = torch.tensor(np.stack([np.asarray(pos[k], np.float64) for k in sel]))
    march = soft.SoftMarch(grid.shape, lo * RES, RES, 320, 240, 87.0, 58.0,
                           RES, 160)
    return nm, march, dev, R.cuda(), p.cuda(), grid, lo


def _small(soft_walk, w=16, h=12, ns=60):
    """A cropped grid around camera 0 and a small image (oracle-sized).  The
    oracle gets the FP32 values of the voxel size and step that the kernel
    parameters hold."""
    _, march, g32, R, p, grid, lo = soft_walk
    c = np.floor(p[0].cpu().numpy() / RES).astype(int) - lo
    a = np.maximum(c - 40, 0)
    b = np.minimum(c + 40, grid.shape)
    sub = np.ascontiguousarray(grid[a[0]:b[0], a[1]:b[1], a[2]:b[2]])
    origin = (lo + a) * RES
    m = soft.SoftMarch(sub.shape, origin, RES, w, h, 87.0, 58.0, 0.037, ns,
                       background=0.0)
    cams = soft.cameras(R[:2], p[:2])
    sc = oracle.SoftScene(sub.astype(np.float64), origin,
                          float(np.float32(RES)), 0.0, w, h, 87.0, 58.0,
                          float(np.float32(0.037)), ns)
    return m, sub, cams, sc


def test_fp64_kernels_match_oracle(soft_walk):
    m, sub, cams, sc = _small(soft_walk)
    g64 = torch.from_numpy(sub.astype(np.float64)).cuda()
    d64 = m.forward64(g64, cams)
    od = sc.forward(cams.cpu().numpy())
    assert np.allclose(d64.cpu().numpy(), od, rtol=1e-12, atol=1e-12)
    # FP64 adjoint vs the oracle's suffix-sum adjoint
    c = torch.randn(d64.shape, dtype=torch.float64,
                    generator=torch.Generator().manual_seed(3)).cuda()
    _, lt64 = m.forward64(g64, cams, with_logt=True)
    gp, gg = m.backward64(g64, cams, lt64, c, grid=True)
    ogp, ogg = sc.backward(cams.cpu().numpy(), c.cpu().numpy())
    assert np.allclose(gp.cpu().numpy(), ogp, rtol=1e-9, atol=1e-9)
    assert np.allclose(gg.cpu().numpy(), ogg, rtol=1e-9, atol=1e-11)
    # the FP32 kernels are within FP32 accuracy of the oracle
    d32 = m.forward(torch.from_numpy(sub).cuda(), cams)
    assert np.abs(d32.cpu().numpy() - od).max() < 1e-4


def test_batched_launch_equals_per_camera(soft_walk):
    _, march, g32, R, p, _, _ = soft_walk
    cams = soft.cameras(R, p)
    d = march.forward(g32, cams)
    for k in (0, 7, 15):
        dk = march.forward(g32, cams[k:k + 1])
        assert torch.equal(d[k], dk[0])
    assert torch.isfinite(d).all() and float(d.min()) > 0


def test_pose_gradient_matches_fp64_central_differences(soft_walk):
    _, march, g32, R, p, grid, _ = soft_walk
    g64 = g32.double()
    cams = soft.cameras(R, p)
    depth, logt = march.forward(g32, cams, with_logt=True)
    c = torch.randn(depth.shape, generator=torch.Generator().manual_seed(5)).cuda()
    gp, _, _ = march.backward(g32, cams, logt, c, pose=True)
    # the FP64 adjoint, to tell FP32 rounding from finite-difference error
    _, lt64 = march.forward64(g64, cams, with_logt=True)
    gp64, _ = march.backward64(g64, cams, lt64, c)
    # eps = 1e-5: the differences' truncation comes from samples that cross
    # a knot of the C1 spline within +-eps, so it shrinks with eps, while
    # FP64 round-off of the sums stays ~1e-9 relative
    eps = 1e-5
    fd = torch.zeros_like(gp)
    for k in range(6):
        xi = torch.zeros((16, 6), dtype=torch.float64, device="cuda")
        xi[:, k] = eps
        lp = (march.forward64(g64, soft.cameras(R, p, xi)) * c).sum((1, 2))
        lm = (march.forward64(g64, soft.cameras(R, p, -xi)) * c).sum((1, 2))
        fd[:, k] = (lp - lm) / (2 * eps)
    err = (gp - fd).norm(dim=1) / fd.norm(dim=1)
    err64 = (gp64 - fd).norm(dim=1) / fd.norm(dim=1)
    err32 = (gp - gp64).norm(dim=1) / gp64.norm(dim=1)
    print("pose gradient rel err per camera: FP32 adjoint vs FD",
          err.cpu().numpy(), "FP64 adjoint vs FD", err64.cpu().numpy(),
          "FP32 vs FP64 adjoint", err32.cpu().numpy(), flush=True)
    # the adjoint's mathematics: the FP64 adjoint against FP64 central
    # differences
    assert float(err64.max()) <= 2e-5
    # the FP32 product against the same adjoint in FP64: FP32 rounding only
    # (measured up to 1.7e-5 on these 16 cameras)
    assert float(err32.max()) <= 5e-5
    # and end to end
    assert float(err.max()) <= 1e-4


def test_logodds_gradient_matches_fp64_central_differences(soft_walk):
    _, march, g32, R, p, _, _ = soft_walk
    cams = soft.cameras(R[:4], p[:4])
    depth, logt = march.forward(g32, cams, with_logt=True)
    c = torch.randn(depth.shape, generator=torch.Generator().manual_seed(6)).cuda()
    _, gg, _ = march.backward(g32, cams, logt, c, pose=False, grid=True)
    flat = gg.reshape(-1)
    top = torch.topk(flat.abs(), 12).indices
    g64 = g32.double()
    eps = 1e-4
    for i in top.tolist():
        gpl, gmi = g64.clone(), g64.clone()
        gpl.view(-1)[i] += eps
        gmi.view(-1)[i] -= eps
        fd = ((march.forward64(gpl, cams) - march.forward64(gmi, cams)) * c).sum() / (2 * eps)
        assert abs(float(flat[i]) - float(fd)) <= 1e-4 * abs(float(fd)) + 1e-7, (i, float(flat[i]), float(fd))


def test_autograd_function_at_nonzero_perturbation(soft_walk):
    _, march, g32, R, p, _, _ = soft_walk
    R4, p4 = R[:4], p[:4]
    xi = torch.tensor([[0.02, -0.01, 0.015, 0.01, -0.02, 0.03]] * 4,
                      dtype=torch.float64, device="cuda", requires_grad=True)
    depth = soft.soft_depth(xi, g32, R4, p4, march)
    c = torch.randn(depth.shape, generator=torch.Generator().manual_seed(7)).cuda()
    (depth * c).sum().backward()
    g = xi.grad.clone()
    eps = 1e-4
    g64 = g32.double()
    fd = torch.zeros_like(g)
    for k in range(6):
        d = torch.zeros_like(g)
        d[:, k] = eps
        lp = (march.forward64(g64, soft.cameras(R4, p4, xi.detach() + d)) * c).sum((1, 2))
        lm = (march.forward64(g64, soft.cameras(R4, p4, xi.detach() - d)) * c).sum((1, 2))
        fd[:, k] = (lp - lm) / (2 * eps)
    err = (g - fd).norm(dim=1) / fd.norm(dim=1)
    print("autograd rel err", err.cpu().numpy(), flush=True)
    assert float(err.max()) <= 1e-4
    # gradient w.r.t. the grid through autograd equals the raw adjoint
    # (float atomics: equal up to summation order)
    gl = g32.clone().requires_grad_(True)
    z = torch.zeros((4, 6), dtype=torch.float64, device="cuda")
    d2 = soft.soft_depth(z, gl, R4, p4, march)
    (d2 * c).sum().backward()
    cams = soft.cameras(R4, p4)
    _, lt = march.forward(g32, cams, with_logt=True)
    _, raw, _ = march.backward(g32, cams, lt, c, pose=False, grid=True)
    assert float((gl.grad - raw).norm()) <= 1e-5 * float(raw.norm())


def test_pose_refinement_recovers_perturbed_poses(soft_walk):
    _, march, g32, R, p, _, _ = soft_walk
    target = march.forward(g32, soft.cameras(R, p))
    gen = torch.Generator().manual_seed(11)
    dt = (torch.rand((16, 3), generator=gen, dtype=torch.float64) * 2 - 1) * 0.1
    axis = torch.randn((16, 3), generator=gen, dtype=torch.float64)
    axis = axis / axis.norm(dim=1, keepdim=True)
    ang = (torch.rand((16, 1), generator=gen, dtype=torch.float64) * 2 - 1) * np.radians(3)
    R0 = soft.so3_exp((axis * ang).cuda()) @ R
    p0 = p + dt.cuda()
    Rf, pf, hist = soft.refine_poses(march, g32, R0, p0, target, iters=20)
    t_err0 = (p0 - p).norm(dim=1)
    t_err = (pf - p).norm(dim=1)
    cosang = ((Rf.transpose(1, 2) @ R).diagonal(dim1=1, dim2=2).sum(1) - 1) / 2
    r_err = torch.rad2deg(torch.arccos(cosang.clamp(-1, 1)))
    print("refinement: translation error %s -> %s m, rotation error -> %s deg"
          % (t_err0.cpu().numpy().round(3), t_err.cpu().numpy().round(4),
             r_err.cpu().numpy().round(3)), flush=True)
    print("rms residual per step (median over cameras):",
          hist.median(1).values.cpu().numpy().round(4), flush=True)
    assert float(t_err.median()) < 0.01 and float(r_err.median()) < 0.3
    assert float(hist[-1].median()) < 0.1 * float(hist[0].median())
