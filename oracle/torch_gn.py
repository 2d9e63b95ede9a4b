"""ORACLE — test infrastructure only.

Pure-PyTorch restatement of the unrolled Gauss-Newton forward of one frame
pair with fixed correspondences, differentiable by torch autograd. It is the
checker of the hand-written adjoint (slamtb.registration.diff): the tests
compare the adjoint's gradients with autograd of this function, and this
function's gradients with FP64 central finite differences.

Per step k (slamtb/registration/icp.py:210-256, icp_jacobian_geometric.cu:
54-85, icp_jacobian_feature.cu:58-109, feature_map.hpp:8-85):
  p' = R_k p + t_k; geometric r = (q - p') . n, J = [n; p' x n];
  feature (last channel) r = I(u, v) - f_s, J = [g; p' x g] with
  g = (du fx/z, dv fy/z, -(du x fx + dv y fy)/z^2) and du, dv the h = 0.05
  central differences of the bilinear sample;
  A = gw^2 sum Jg Jg^T + fw^2 sum Jf Jf^T, b = gw sum Jg rg - fw sum Jf rf;
  x = A^-1 b; T_{k+1} = exp_rt(x) T_k.
Owners (source pixel -> target pixel) and the feature gate of each step are
inputs, taken from the decisions the GPU forward made. So, optionally, are
the bilinear cells (the truncated sample positions) of the five feature
samples: they are piecewise constant too, and a checker evaluating u, v in
FP64 would otherwise pick the neighbouring cell wherever the FP32 forward's
u +- h lies within rounding of a pixel edge.
"""
import torch


def exp_rt(x):
    """[6] -> [4, 4]: Rodrigues rotation of x[3:], raw translation x[:3]."""
    w = x[3:]
    th = torch.sqrt((w * w).sum())
    a = w / th
    K = torch.zeros(3, 3, dtype=x.dtype, device=x.device)
    K[0, 1], K[0, 2], K[1, 0] = -a[2], a[1], a[2]
    K[1, 2], K[2, 0], K[2, 1] = -a[0], -a[1], a[0]
    R = (torch.eye(3, dtype=x.dtype, device=x.device) + torch.sin(th) * K +
         (1 - torch.cos(th)) * (K @ K))
    T = torch.eye(4, dtype=x.dtype, device=x.device)
    T = T.clone()
    T[:3, :3] = R
    T[:3, 3] = x[:3]
    return T


def bilinear(ch, u, v, cell=None):
    """Bilinear sample of ch [H, W] at (u, v) with truncating cell index and
    zero padding past the right/bottom edge (feature_map.hpp:8-43); `cell`
    = (ui, vi) gives the cell instead of truncating (u, v)."""
    H, W = ch.shape
    if cell is None:
        ui = torch.trunc(u.detach()).long()
        vi = torch.trunc(v.detach()).long()
    else:
        ui, vi = cell
    ur = u - ui.to(u.dtype)
    vr = v - vi.to(v.dtype)
    rx = ui + 1 < W
    ry = vi + 1 < H
    flat = ch.reshape(-1)
    def tap(r, c, ok):
        idx = (r.clamp(max=H - 1) * W + c.clamp(max=W - 1))
        return torch.where(ok, flat[idx], torch.zeros_like(ur))
    v00 = flat[vi * W + ui]
    v10 = tap(vi, ui + 1, rx)
    v01 = tap(vi + 1, ui, ry)
    v11 = tap(vi + 1, ui + 1, rx & ry)
    a = v00 * (1 - ur) + v10 * ur
    c = v01 * (1 - ur) + v11 * ur
    return a * (1 - vr) + c * vr


def pin(T, value):
    """T with the value `value` and T's derivatives (straight through)."""
    return value + (T - T.detach())


def unrolled(src_points, tgt_points, tgt_normals, src_feats, tgt_feats, kcam,
             x0, steps, geom_weight, feat_weight, h=0.05, poses=None):
    """src_points [N, 3]; tgt_points/normals [H, W, 3]; src_feats [C, N];
    tgt_feats [C, H, W]; x0 [6]; steps: list of (src_idx, tgt_pos, feat_ok)
    long/long/bool tensors of the matched pixels of each step, optionally
    with a fourth entry: the cells [5, 2, M] (ui, vi of the samples at
    (u, v), (u + h, v), (max(u - h, 0), v), (u, v + h), (u, max(v - h, 0))
    of the M feature pixels). With `poses` (the values T_0 .. T_steps of a
    forward run) each step is evaluated at that run's pose, so the
    derivatives are those of the same unrolled map taken along that run's
    trajectory instead of this function's own FP64 one."""
    fx, fy, cx, cy = kcam[0, 0], kcam[1, 1], kcam[0, 2], kcam[1, 2]
    T = exp_rt(x0)
    if poses is not None:
        T = pin(T, poses[0])
    tp = tgt_points.reshape(-1, 3)
    tn = tgt_normals.reshape(-1, 3)
    for k, step in enumerate(steps):
        src_idx, tgt_pos, feat_ok = step[:3]
        cells = step[3] if len(step) > 3 else None
        cell = (lambda i: None) if cells is None else (lambda i: (cells[i, 0], cells[i, 1]))
        p = src_points[src_idx] @ T[:3, :3].T + T[:3, 3]
        A = torch.zeros(6, 6, dtype=p.dtype, device=p.device)
        b = torch.zeros(6, dtype=p.dtype, device=p.device)
        if geom_weight > 0:
            q, n = tp[tgt_pos], tn[tgt_pos]
            r = ((q - p) * n).sum(1)
            J = torch.cat([n, torch.cross(p, n, dim=1)], 1)
            A = A + geom_weight ** 2 * J.T @ J
            b = b + geom_weight * J.T @ r
        if feat_weight > 0 and src_feats is not None:
            pf = p[feat_ok]
            x, y, z = pf[:, 0], pf[:, 1], pf[:, 2]
            u = fx * x / z + cx
            v = fy * y / z + cy
            ch = tgt_feats[-1]
            rf = bilinear(ch, u, v, cell(0)) - src_feats[-1][src_idx[feat_ok]]
            du = (bilinear(ch, u + h, v, cell(1)) -
                  bilinear(ch, torch.clamp(u - h, min=0), v, cell(2))) / (2 * h)
            dv = (bilinear(ch, u, v + h, cell(3)) -
                  bilinear(ch, u, torch.clamp(v - h, min=0), cell(4))) / (2 * h)
            g = torch.stack([du * fx / z, dv * fy / z,
                             -(du * x * fx + dv * y * fy) / z ** 2], 1)
            J = torch.cat([g, torch.cross(pf, g, dim=1)], 1)
            A = A + feat_weight ** 2 * J.T @ J
            b = b - feat_weight * J.T @ rf
        xk = torch.linalg.solve(A, b)
        T = exp_rt(xk) @ T
        if poses is not None:
            T = pin(T, poses[k + 1])
    return T
