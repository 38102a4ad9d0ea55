"""Off-grid refinement of docked poses: a torch.autograd.Function around the companion
score of libmdk (mdk_score_companion) and a gradient-ascent loop over the rigid pose.

The pose of decoy b is a rotation R_b = exp([w_b]x) Q_b (Q_b the docked rotation, w_b a
rotation vector, radians) and a continuous translation t_b in voxels (the units of a
result's index[1..3]).  The device returns the score with its gradient for a rotation
applied after R_b (dS/dd at exp([d]x) R_b, d = 0); dS/dw follows with the left Jacobian
of SO(3): exp([w + dw]x) = exp([J_l(w) dw]x) exp([w]x), so dS/dw = J_l(w)^T dS/dd.

PyTorch is the harness here (autograd bookkeeping, small 3x3 algebra in float64); the
scores and gradients come from the HIP kernels through the C ABI.  Refined poses are
re-judged by the hard-voxel score at the refined pose (mdk_score_spectral).

Two refinements:
  refine()        ascent on the companion score alone (monotone in the companion);
  refine_hard()   ascent whose steps are accepted only when the hard-voxel score rises:
                  translation along the exact gradient of the band-limited hard score
                  (SpectralScore), rotation along the companion's torque (the hard
                  voxelizer has no rotation gradient), so the hard-voxel score, by which
                  refined poses are judged, never drops.
"""
import numpy as np
import torch


def skew(w):
    z = torch.zeros_like(w[..., 0])
    return torch.stack([torch.stack([z, -w[..., 2], w[..., 1]], -1),
                        torch.stack([w[..., 2], z, -w[..., 0]], -1),
                        torch.stack([-w[..., 1], w[..., 0], z], -1)], -2)


def rotvec_matrix(w):
    """exp([w]x) for a batch of rotation vectors (B, 3), float64."""
    th = torch.linalg.norm(w, dim=-1, keepdim=True)[..., None]
    K = skew(w)
    eye = torch.eye(3, dtype=w.dtype).expand(K.shape)
    small = th < 1e-8
    a = torch.where(small, torch.ones_like(th), torch.sin(th) / torch.where(small, torch.ones_like(th), th))
    b = torch.where(small, 0.5 * torch.ones_like(th), (1 - torch.cos(th)) / torch.where(small, torch.ones_like(th), th) ** 2)
    return eye + a * K + b * (K @ K)


def left_jacobian(w):
    """J_l(w) = I + (1 - cos th)/th^2 [w]x + (th - sin th)/th^3 [w]x^2, (B, 3, 3)."""
    th = torch.linalg.norm(w, dim=-1, keepdim=True)[..., None]
    K = skew(w)
    eye = torch.eye(3, dtype=w.dtype).expand(K.shape)
    small = th < 1e-6
    ths = torch.where(small, torch.ones_like(th), th)
    a = torch.where(small, 0.5 * torch.ones_like(th), (1 - torch.cos(ths)) / ths ** 2)
    b = torch.where(small, torch.ones_like(th) / 6.0, (ths - torch.sin(ths)) / ths ** 3)
    return eye + a * K + b * (K @ K)


class CompanionScore(torch.autograd.Function):
    """S(w, t) for a batch of poses; gradients with respect to w (B, 3) and t (B, 3)."""

    @staticmethod
    def forward(ctx, w, t, base, dock):
        R = rotvec_matrix(w) @ base
        out = dock.score_companion(R.reshape(-1, 9).detach().numpy().astype(np.float32),
                                   t.detach().numpy().astype(np.float32))
        out = torch.from_numpy(out.astype(np.float64))
        ctx.save_for_backward(w, out[:, 1:4], out[:, 4:7])
        return out[:, 0]

    @staticmethod
    def backward(ctx, gs):
        w, dt, dd = ctx.saved_tensors
        dw = torch.einsum("bji,bj->bi", left_jacobian(w), dd)
        return gs[:, None] * dw, gs[:, None] * dt, None, None


class SpectralScore(torch.autograd.Function):
    """Hard-voxel score of a batch of poses (rotations R (B, 3, 3) fixed) at continuous
    translations t (B, 3), band-limited between grid points; gradient with respect to t."""

    @staticmethod
    def forward(ctx, t, R, dock):
        out = dock.score_spectral(R.reshape(-1, 9).detach().numpy().astype(np.float32),
                                  t.detach().numpy().astype(np.float32))
        out = torch.from_numpy(out.astype(np.float64))
        ctx.save_for_backward(out[:, 1:4])
        return out[:, 0]

    @staticmethod
    def backward(ctx, gs):
        (dt,) = ctx.saved_tensors
        return gs[:, None] * dt, None, None


def refine_hard(dock, rotcoef, t0, steps=20, step_t=0.25, step_w=0.02, min_step=1e-3):
    """Refine poses (rotcoef (B, 9), t0 (B, 3)) on the hard-voxel score: each step moves
    the translation along the hard score's gradient and the rotation along the companion
    torque, with normalised lengths, and is kept only if the hard score rises (length x1.2,
    else x0.5 and retried).  Returns (rotcoef (B, 9) float32, t (B, 3) float32, history
    (steps + 1, B) of hard scores, non-decreasing per decoy)."""
    R = torch.from_numpy(np.asarray(rotcoef, np.float64).reshape(-1, 3, 3).copy())
    B = R.shape[0]
    t = torch.from_numpy(np.asarray(t0, np.float64).reshape(-1, 3).copy())
    st = torch.full((B,), float(step_t), dtype=torch.float64)
    sw = torch.full((B,), float(step_w), dtype=torch.float64)
    zero = torch.zeros(B, 3, dtype=torch.float64)
    history = []
    for it in range(steps + 1):
        tv = t.clone().requires_grad_(True)
        s = SpectralScore.apply(tv, R, dock)
        s.sum().backward()
        history.append(s.detach().numpy().copy())
        if it == steps:
            break
        wv = zero.clone().requires_grad_(True)
        CompanionScore.apply(wv, t, R, dock).sum().backward()
        gt, gw = tv.grad, wv.grad
        nt = torch.linalg.norm(gt, dim=1).clamp_min(1e-30)
        nw = torch.linalg.norm(gw, dim=1).clamp_min(1e-30)
        cur = s.detach().clone()
        idx = torch.arange(B)  # decoys still trying this step; only they are evaluated
        while idx.numel() > 0:
            t_try = t[idx] + (st[idx] / nt[idx])[:, None] * gt[idx]
            R_try = rotvec_matrix((sw[idx] / nw[idx])[:, None] * gw[idx]) @ R[idx]
            with torch.no_grad():
                s_try = SpectralScore.apply(t_try, R_try, dock)
            ok = s_try > cur[idx]
            acc, rej = idx[ok], idx[~ok]
            t[acc], R[acc], cur[acc] = t_try[ok], R_try[ok], s_try[ok]
            st[acc] *= 1.2
            sw[acc] *= 1.2
            st[rej] *= 0.5
            sw[rej] *= 0.5
            idx = rej[st[rej] >= min_step]
    return R.reshape(-1, 9).numpy().astype(np.float32), t.numpy().astype(np.float32), np.array(history)


def refine(dock, rotcoef, t0, steps=30, step_t=0.25, step_w=0.02, min_step=1e-4):
    """Gradient ascent on the companion score from docked poses (rotcoef (B, 9), t0 (B, 3)).

    Each decoy takes normalised steps along its translation and rotation gradients; a
    step is kept only if the score rises (then its length grows by 1.2), otherwise it is
    halved, so every decoy's score is non-decreasing.  Returns (rotcoef (B, 9) float32,
    t (B, 3) float32, history (steps + 1, B) of companion scores)."""
    base = torch.from_numpy(np.asarray(rotcoef, np.float64).reshape(-1, 3, 3))
    B = base.shape[0]
    w = torch.zeros(B, 3, dtype=torch.float64)
    t = torch.from_numpy(np.asarray(t0, np.float64).reshape(-1, 3).copy())
    st = torch.full((B,), float(step_t), dtype=torch.float64)
    sw = torch.full((B,), float(step_w), dtype=torch.float64)
    history = []
    for _ in range(steps + 1):
        wv = w.clone().requires_grad_(True)
        tv = t.clone().requires_grad_(True)
        s = CompanionScore.apply(wv, tv, base, dock)
        s.sum().backward()
        history.append(s.detach().numpy().copy())
        if len(history) == steps + 1:
            break
        gt, gw = tv.grad, wv.grad
        nt = torch.linalg.norm(gt, dim=1).clamp_min(1e-30)
        nw = torch.linalg.norm(gw, dim=1).clamp_min(1e-30)
        cur = s.detach().clone()
        idx = torch.arange(B)  # backtracking, per decoy: only the decoys still trying are evaluated
        while idx.numel() > 0:
            t_try = t[idx] + (st[idx] / nt[idx])[:, None] * gt[idx]
            w_try = w[idx] + (sw[idx] / nw[idx])[:, None] * gw[idx]
            with torch.no_grad():
                s_try = CompanionScore.apply(w_try, t_try, base[idx], dock)
            ok = s_try > cur[idx]
            acc, rej = idx[ok], idx[~ok]
            t[acc], w[acc], cur[acc] = t_try[ok], w_try[ok], s_try[ok]
            st[acc] *= 1.2
            sw[acc] *= 1.2
            st[rej] *= 0.5
            sw[rej] *= 0.5
            idx = rej[st[rej] >= min_step]
    R = (rotvec_matrix(w) @ base).reshape(-1, 9).numpy().astype(np.float32)
    return R, t.numpy().astype(np.float32), np.array(history)
