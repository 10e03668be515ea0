"""Differentiable expected-depth ray march (include/nanomap_amd_soft.h)
for PyTorch: a torch.autograd.Function over the hand-written HIP forward
and adjoint kernels, and Gauss-Newton scan-to-map pose refinement.

A camera is 12 float64 numbers: the row-major optical->world rotation R
and the position p (metres).  Pose gradients are taken with respect to
xi = (rho, omega) in world axes: R(xi) = exp(omega^) R0, p(xi) = p0 + rho.
"""
import ctypes

import torch

from ._capi import SoftGrid, SoftParams, check, lib


class SoftMarch:
    """Grid geometry and camera model of a soft march.

    logodds: dense [nx, ny, nz] tensor (float32 for the product kernels,
    float64 for the FP64 reference); voxel (i, j, k) covers
    origin + res * [i, i+1) x [j, j+1) x [k, k+1)."""

    def __init__(self, shape, origin, res, width, height, hfov_deg, vfov_deg,
                 step, n_samples, background=0.0):
        self.grid = SoftGrid(int(shape[0]), int(shape[1]), int(shape[2]),
                             (ctypes.c_double * 3)(*[float(x) for x in origin]),
                             float(res), float(background))
        self.params = SoftParams(int(width), int(height), float(hfov_deg),
                                 float(vfov_deg), float(step), int(n_samples))
        self.shape = tuple(int(x) for x in shape)
        self.width, self.height = int(width), int(height)

    @staticmethod
    def _stream():
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _check_grid(self, logodds, dtype):
        assert logodds.is_cuda and logodds.dtype == dtype
        assert tuple(logodds.shape) == self.shape and logodds.is_contiguous()

    def forward(self, logodds, cams, with_logt=False):
        """Expected depth [n, height, width] (float32) of cams [n, 12]; with
        with_logt also the per-ray log-transmittance (float64) that the
        adjoint takes."""
        self._check_grid(logodds, torch.float32)
        cams = cams.to(torch.float64).contiguous()
        n = cams.shape[0]
        depth = torch.empty((n, self.height, self.width), dtype=torch.float32,
                            device=logodds.device)
        logt = (torch.empty((n, self.height, self.width), dtype=torch.float64,
                            device=logodds.device) if with_logt else None)
        check(lib().nm_soft_forward(ctypes.byref(self.grid),
                                    ctypes.c_void_p(logodds.data_ptr()),
                                    ctypes.c_void_p(cams.data_ptr()), n,
                                    ctypes.byref(self.params),
                                    ctypes.c_void_p(depth.data_ptr()),
                                    None if logt is None else
                                    ctypes.c_void_p(logt.data_ptr()),
                                    self._stream()))
        return (depth, logt) if with_logt else depth

    def backward(self, logodds, cams, logt, grad_depth=None, pose=True,
                 grid=False, per_ray=False):
        """(dL/dxi [n, 6] float64 or None, dL/dlogodds or None, per-ray
        dD/dxi [n, height, width, 6] or None)."""
        self._check_grid(logodds, torch.float32)
        cams = cams.to(torch.float64).contiguous()
        n = cams.shape[0]
        dev = logodds.device
        gd = None if grad_depth is None else grad_depth.float().contiguous()
        gp = torch.zeros((n, 6), dtype=torch.float64, device=dev) if pose else None
        gg = torch.zeros_like(logodds) if grid else None
        gr = (torch.zeros((n, self.height, self.width, 6), dtype=torch.float32,
                          device=dev) if per_ray else None)

        def ptr(t):
            return None if t is None else ctypes.c_void_p(t.data_ptr())
        check(lib().nm_soft_backward(ctypes.byref(self.grid), ptr(logodds),
                                     ptr(cams), n, ctypes.byref(self.params),
                                     ptr(logt.contiguous()), ptr(gd), ptr(gp),
                                     ptr(gr), ptr(gg), self._stream()))
        return gp, gg, gr

    def forward64(self, logodds64, cams, with_logt=False):
        """FP64 expected depth (the finite-difference reference)."""
        self._check_grid(logodds64, torch.float64)
        cams = cams.to(torch.float64).contiguous()
        n = cams.shape[0]
        depth = torch.empty((n, self.height, self.width), dtype=torch.float64,
                            device=logodds64.device)
        logt = torch.empty_like(depth) if with_logt else None
        check(lib().nm_soft_forward_f64(ctypes.byref(self.grid),
                                        ctypes.c_void_p(logodds64.data_ptr()),
                                        ctypes.c_void_p(cams.data_ptr()), n,
                                        ctypes.byref(self.params),
                                        ctypes.c_void_p(depth.data_ptr()),
                                        None if logt is None else
                                        ctypes.c_void_p(logt.data_ptr()),
                                        self._stream()))
        return (depth, logt) if with_logt else depth

    def backward64(self, logodds64, cams, logt, grad_depth, grid=False):
        """FP64 adjoint: (dL/dxi [n, 6], dL/dlogodds or None)."""
        self._check_grid(logodds64, torch.float64)
        cams = cams.to(torch.float64).contiguous()
        n = cams.shape[0]
        gp = torch.zeros((n, 6), dtype=torch.float64, device=logodds64.device)
        gg = torch.zeros_like(logodds64) if grid else None
        check(lib().nm_soft_backward_f64(
            ctypes.byref(self.grid), ctypes.c_void_p(logodds64.data_ptr()),
            ctypes.c_void_p(cams.data_ptr()), n, ctypes.byref(self.params),
            ctypes.c_void_p(logt.data_ptr()),
            ctypes.c_void_p(grad_depth.to(torch.float64).contiguous().data_ptr()),
            ctypes.c_void_p(gp.data_ptr()),
            None if gg is None else ctypes.c_void_p(gg.data_ptr()),
            self._stream()))
        return gp, gg


def hat(w):
    """[..., 3] -> skew matrices [..., 3, 3]."""
    z = torch.zeros_like(w[..., 0])
    return torch.stack([torch.stack([z, -w[..., 2], w[..., 1]], -1),
                        torch.stack([w[..., 2], z, -w[..., 0]], -1),
                        torch.stack([-w[..., 1], w[..., 0], z], -1)], -2)


def so3_exp(w):
    """Rodrigues: [..., 3] -> [..., 3, 3] (float64)."""
    w = w.to(torch.float64)
    th = w.norm(dim=-1, keepdim=True).unsqueeze(-1)
    K = hat(w)
    small = th < 1e-8
    ths = torch.where(small, torch.ones_like(th), th)
    a = torch.where(small, 1.0 - th * th / 6.0, torch.sin(ths) / ths)
    b = torch.where(small, 0.5 - th * th / 24.0, (1.0 - torch.cos(ths)) / (ths * ths))
    eye = torch.eye(3, dtype=torch.float64, device=w.device).expand_as(K)
    return eye + a * K + b * (K @ K)


def so3_left_jacobian(w):
    """J_l(w): exp((w + d)^) = exp((J_l(w) d)^) exp(w^) to first order."""
    w = w.to(torch.float64)
    th = w.norm(dim=-1, keepdim=True).unsqueeze(-1)
    K = hat(w)
    small = th < 1e-6
    ths = torch.where(small, torch.ones_like(th), th)
    a = torch.where(small, 0.5 - th * th / 24.0, (1.0 - torch.cos(ths)) / (ths * ths))
    b = torch.where(small, 1.0 / 6.0 - th * th / 120.0,
                    (ths - torch.sin(ths)) / (ths ** 3))
    eye = torch.eye(3, dtype=torch.float64, device=w.device).expand_as(K)
    return eye + a * K + b * (K @ K)


def cameras(R, p, xi=None):
    """Camera table [n, 12] float64 from rotations [n, 3, 3], positions
    [n, 3] and an optional perturbation xi [n, 6]."""
    R = R.to(torch.float64)
    p = p.to(torch.float64)
    if xi is not None:
        xi = xi.to(torch.float64)
        R = so3_exp(xi[:, 3:]) @ R
        p = p + xi[:, :3]
    return torch.cat([R.reshape(-1, 9), p.reshape(-1, 3)], 1).contiguous()


class _SoftDepth(torch.autograd.Function):
    @staticmethod
    def forward(ctx, xi, logodds, R0, p0, march):
        cams = cameras(R0, p0, xi.detach())
        depth, logt = march.forward(logodds.detach(), cams, with_logt=True)
        ctx.march = march
        ctx.save_for_backward(xi, logodds, cams, logt)
        return depth

    @staticmethod
    def backward(ctx, grad_depth):
        xi, logodds, cams, logt = ctx.saved_tensors
        want_pose, want_grid = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        gp, gg, _ = ctx.march.backward(logodds.detach(), cams, logt,
                                       grad_depth, pose=want_pose,
                                       grid=want_grid)
        gxi = None
        if want_pose:
            # gradient w.r.t. a perturbation at the current pose -> w.r.t. xi
            J = so3_left_jacobian(xi.detach()[:, 3:])
            gw = (J.transpose(1, 2) @ gp[:, 3:].unsqueeze(-1)).squeeze(-1)
            gxi = torch.cat([gp[:, :3], gw], 1).to(xi.dtype)
        return gxi, gg, None, None, None


def soft_depth(xi, logodds, R0, p0, march):
    """Expected depth [n, H, W] of cameras (R0, p0) moved by xi [n, 6];
    differentiable in xi and logodds."""
    return _SoftDepth.apply(xi, logodds, R0, p0, march)


def refine_poses(march, logodds, R, p, target, iters=20, damping=1e-3,
                 weight=None):
    """Gauss-Newton scan-to-map refinement of every camera independently:
    minimise sum over pixels of (D(pose) - target)^2, using the adjoint's
    per-ray Jacobians.  Returns refined (R, p) and the per-step RMS
    residuals [iters + 1, n]."""
    R = R.to(torch.float64).clone()
    p = p.to(torch.float64).clone()
    n = R.shape[0]
    hist = []
    for it in range(iters + 1):
        cams = cameras(R, p)
        depth, logt = march.forward(logodds, cams, with_logt=True)
        r = (depth - target).reshape(n, -1).to(torch.float64)
        w = None if weight is None else weight.reshape(n, -1).to(torch.float64)
        hist.append(torch.sqrt((r * r if w is None else w * r * r).mean(1)))
        if it == iters:
            break
        _, _, J = march.backward(logodds, cams, logt, None, pose=False,
                                 per_ray=True)
        J = J.reshape(n, -1, 6).to(torch.float64)
        Jw = J if w is None else J * w.unsqueeze(-1)
        # 6x6 normal equations as column reductions over the rays: a batched
        # product with a 76800-long inner dimension and a 6x6 result runs
        # ~100x slower in the library GEMM than these 7 reductions
        H = torch.stack([(Jw[..., a:a + 1] * J).sum(1) for a in range(6)], 1)
        g = (Jw * r.unsqueeze(-1)).sum(1)
        lam = damping * torch.diagonal(H, dim1=1, dim2=2).mean(1)
        H = H + lam[:, None, None] * torch.eye(6, dtype=torch.float64,
                                               device=H.device)
        delta = -torch.linalg.solve(H, g.unsqueeze(-1)).squeeze(-1)
        R = so3_exp(delta[:, 3:]) @ R
        p = p + delta[:, :3]
    return R, p, torch.stack(hist)
