"""Differentiable ICP: unrolled Gauss-Newton steps with a hand-written adjoint.

The forward pass runs ``num_steps`` fused GN iterations for a batch of frame
pairs (the batched kernels of libslamtb_hip.so) and records, per step, only
the pose T_k and whether the step was accepted. The backward pass walks the
steps in reverse:

* through the pose update T_{k+1} = exp_rt(x_k) T_k: dL/dU_k = G T_k^T,
  dL/dT_k += U_k^T G and dL/dx_k by the exp_rt_to_matrix backward
  (src/exp_rt.hpp:87-141);
* through the solve x_k = A_k^-1 b_k by its implicit adjoint:
  lambda = A_k^-1 dL/dx_k, with (A_k, b_k) recomputed at T_k in FP64 per
  pixel with the forward's decisions (stb_icp_batch_system64);
* through the per-pixel Jacobians and residuals, which the adjoint kernel
  recomputes at T_k (stb_icp_batch_adjoint) instead of storing them, into
  source points (hence source depth), target points and normals, source and
  target features, and the part of dL/dT_k that flows through the
  transformed source points.

Correspondences and the feature-residual gate are piecewise constant: they
are treated as fixed, exactly as automatic differentiation through the same
forward would. The initial pose enters as a 6-vector x0 with
T_0 = exp_rt_to_matrix(x0). Source normals only gate correspondences and get
no gradient.
"""
import ctypes
import math

import torch

from slamtb import _lib
from slamtb._cslamtb import ExpRtToMatrixOp

from ._fused import Level, STATE_BYTES, make_params

# Slot of JtJ[a][c] among the 21 upper-triangle slots, for every (a, c).
_TRI = torch.tensor([min(a, c) * (11 - min(a, c)) // 2 + max(a, c)
                     for a in range(6) for c in range(6)])

_OFF = {name: getattr(_lib.StbIcpState, name).offset
        for name in ("transform", "last_jtj", "last_jr", "iterations")}


def _state_f64(states, name, count):
    off = _OFF[name]
    return states[:, off:off + 8 * count].clone().view(torch.float64)


def _iterations(states):
    off = _OFF["iterations"]
    return states[:, off:off + 4].clone().view(torch.int32).view(-1)


def exp_rt(x):
    """[B, 6] FP64 -> [B, 4, 4] with the library's exp_rt_to_matrix."""
    m = torch.empty(x.size(0), 3, 4, dtype=torch.float64, device=x.device)
    ExpRtToMatrixOp.forward(x.contiguous(), m)
    T = torch.zeros(x.size(0), 4, 4, dtype=torch.float64, device=x.device)
    T[:, :3] = m
    T[:, 3, 3] = 1.0
    return T


def exp_rt_vjp(d_matrix, x, matrix):
    out = torch.empty(x.size(0), 6, dtype=torch.float64, device=x.device)
    ExpRtToMatrixOp.backward(d_matrix.contiguous(), x.contiguous(),
                             matrix[:, :3].contiguous(), out)
    return out


def depth_to_points(depth, kcam):
    """[B, H, W] depth -> [B, H*W, 3] camera points (float32)."""
    B, H, W = depth.shape
    pts = torch.empty(B, H * W, 3, dtype=torch.float32, device=depth.device)
    d = depth.contiguous().float()
    fx, fy, cx, cy = (float(kcam[0, 0]), float(kcam[1, 1]), float(kcam[0, 2]),
                      float(kcam[1, 2]))
    for b in range(B):
        _lib.check(_lib.lib().stb_depth_to_points(
            _lib.STB_F32, _lib.ptr(d[b]), H, W, fx, fy, cx, cy, 1.0, 0.0,
            _lib.ptr(pts[b]), _lib.stream()), "depth_to_points")
    return pts


def pose_loss(T, T_gt):
    """Squared geodesic distance on SE(3) as rotation angle^2 + translation^2
    of T_gt^-1 T (per pair), differentiable by torch."""
    D = torch.linalg.solve(T_gt, T)
    R = D[..., :3, :3]
    c = (R.diagonal(dim1=-2, dim2=-1).sum(-1) - 1.0) / 2.0
    v = torch.stack([R[..., 2, 1] - R[..., 1, 2], R[..., 0, 2] - R[..., 2, 0],
                     R[..., 1, 0] - R[..., 0, 1]], -1) / 2.0
    theta = torch.atan2(v.norm(dim=-1), c)
    return theta ** 2 + (D[..., :3, 3] ** 2).sum(-1)


class _Batch:
    """Batched plan, states and scratch for B pairs at one resolution."""

    def __init__(self, kcam, src_points, src_normals, src_mask, tgt_points,
                 tgt_normals, tgt_mask, src_feats, tgt_feats, params, tile_px):
        B = src_points.size(0)
        self.levels = [Level(kcam, src_points[b], src_normals[b], src_mask[b],
                             tgt_points[b], tgt_normals[b], tgt_mask[b],
                             None if src_feats is None else src_feats[b],
                             None if tgt_feats is None else tgt_feats[b])
                       for b in range(B)]
        arr = (_lib.StbIcpLevel * B)(*[L.struct for L in self.levels])
        self.sizes = _lib.StbIcpBatchSizes()
        lib = _lib.lib()
        _lib.check(lib.stb_icp_batch_plan(B, arr, tile_px, None, 0,
                                          ctypes.byref(self.sizes)), "plan")
        host = (ctypes.c_uint8 * self.sizes.plan_bytes)()
        _lib.check(lib.stb_icp_batch_plan(B, arr, tile_px, host,
                                          self.sizes.plan_bytes,
                                          ctypes.byref(self.sizes)), "plan")
        dev = src_points.device
        self.plan = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(dev)
        self.states = torch.zeros(B, STATE_BYTES, dtype=torch.uint8, device=dev)
        self.scratch = torch.empty(self.sizes.scratch_bytes, dtype=torch.uint8,
                                   device=dev)
        self.params = params
        self.B = B

    def _common(self):
        return (_lib.ptr(self.plan), ctypes.byref(self.sizes))

    def begin(self, T):
        _lib.check(_lib.lib().stb_icp_batch_begin(
            *self._common(), _lib.ptr(self.states), _lib.ptr(T.contiguous()), 0,
            _lib.ptr(self.scratch), self.scratch.numel(), _lib.stream()),
            "stb_icp_batch_begin")

    def iterate(self, n):
        _lib.check(_lib.lib().stb_icp_batch_iterate(
            *self._common(), ctypes.byref(self.params), _lib.ptr(self.states),
            _lib.ptr(self.scratch), self.scratch.numel(), n, _lib.stream()),
            "stb_icp_batch_iterate")

    def system64(self):
        """Every pair's normal equations at its current pose recomputed in
        FP64 per pixel with the forward's decisions: A [B, 6, 6], b [B, 6]."""
        out = torch.empty(self.B, 27, dtype=torch.float64, device=self.states.device)
        _lib.check(_lib.lib().stb_icp_batch_system64(
            *self._common(), ctypes.byref(self.params), _lib.ptr(self.states),
            _lib.ptr(out), _lib.ptr(self.scratch), self.scratch.numel(),
            _lib.stream()), "stb_icp_batch_system64")
        A = out[:, _TRI].view(self.B, 6, 6)
        return A, out[:, 21:27].contiguous()

    def adjoint(self, grads_dev, coef, dT):
        _lib.check(_lib.lib().stb_icp_batch_adjoint(
            *self._common(), ctypes.byref(self.params), _lib.ptr(self.states),
            ctypes.c_void_p(grads_dev.data_ptr()), _lib.ptr(coef), _lib.ptr(dT),
            _lib.ptr(self.scratch), self.scratch.numel(), _lib.stream()),
            "stb_icp_batch_adjoint")


class _UnrolledGN(torch.autograd.Function):

    @staticmethod
    def forward(ctx, src_depth, tgt_points, tgt_normals, src_feats, tgt_feats,
                x0, kcam, src_normals, src_mask, tgt_mask, cfg):
        B, H, W = src_depth.shape
        src_points = depth_to_points(src_depth, kcam)
        use_feat = src_feats is not None and cfg["feat_weight"] > 0
        params = make_params(cfg["geom_weight"], cfg["feat_weight"],
                             cfg["distance_threshold"],
                             cfg["normals_angle_thresh"],
                             cfg["feat_residual_thresh"], use_feat)
        bt = _Batch(kcam, src_points, src_normals.reshape(B, -1, 3),
                    src_mask.reshape(B, -1), tgt_points, tgt_normals, tgt_mask,
                    src_feats.reshape(B, src_feats.size(1), -1) if use_feat else None,
                    tgt_feats if use_feat else None, params, cfg["tile_px"])
        T = exp_rt(x0.double())
        bt.begin(T)
        Ts, acc = [], []
        prev = _iterations(bt.states)
        for _ in range(cfg["num_steps"]):
            Ts.append(_state_f64(bt.states, "transform", 16).view(B, 4, 4))
            bt.iterate(1)
            it = _iterations(bt.states)
            acc.append(it > prev)
            prev = it
        T_out = _state_f64(bt.states, "transform", 16).view(B, 4, 4)
        ctx.bt = bt
        ctx.cfg = cfg
        ctx.kcam = kcam
        ctx.hw = (B, H, W)
        ctx.use_feat = use_feat
        ctx.steps = (Ts, acc)
        ctx.save_for_backward(x0, src_points)
        return T_out

    @staticmethod
    def backward(ctx, grad_T):
        x0, src_points = ctx.saved_tensors
        bt, (B, H, W) = ctx.bt, ctx.hw
        Ts, acc = ctx.steps
        dev = src_points.device
        C = bt.levels[0].channels if ctx.use_feat else 0
        n = H * W
        g_sp = torch.zeros(B, n, 3, dtype=torch.float32, device=dev)
        g_tp = torch.zeros(B, n, 3, dtype=torch.float32, device=dev)
        g_tn = torch.zeros(B, n, 3, dtype=torch.float32, device=dev)
        g_sf = torch.zeros(B, max(C, 1), n, dtype=torch.float32, device=dev)
        g_tf = torch.zeros(B, max(C, 1), n, dtype=torch.float32, device=dev)
        ptrs = []
        for b in range(B):
            ptrs += [g_sp[b].data_ptr(), g_sf[b].data_ptr(), g_tp[b].data_ptr(),
                     g_tn[b].data_ptr(), g_tf[b].data_ptr()]
        grads_dev = torch.tensor(ptrs, dtype=torch.int64).to(dev)
        G = grad_T.double().clone()
        G[:, 3] = 0.0
        dT = torch.empty(B, 12, dtype=torch.float64, device=dev)
        eye = torch.eye(6, dtype=torch.float64, device=dev)
        for k in reversed(range(len(Ts))):
            ok = acc[k]
            # The step's system recomputed in FP64 per pixel at T_k with the
            # forward's decisions: the forward's own sums carry the FP32
            # rounding of u and v (up to 3e-5 px at 640x480, ~1e-4 in the
            # sampled feature derivatives), and dL/dT_k, the difference of
            # two nearly cancelling terms, would grow that error about 2.5x
            # per step.
            bt.begin(Ts[k])
            A64, b64 = bt.system64()
            A = torch.where(ok[:, None, None], A64, eye)
            # A is symmetric positive definite for every accepted step (its
            # device Cholesky succeeded); one factorisation serves both
            # solves, with no host synchronisation.
            Lc = torch.linalg.cholesky_ex(A)[0]
            x = torch.cholesky_solve(torch.where(ok[:, None], b64, 0.0)[..., None],
                                     Lc)[..., 0]
            U = exp_rt(x)
            dU = G[:, :3] @ Ts[k].transpose(1, 2)
            G_direct = U.transpose(1, 2) @ G
            xbar = exp_rt_vjp(dU, x, U)
            lam = torch.cholesky_solve(xbar[..., None], Lc)[..., 0]
            coef = torch.cat([lam, x], 1)
            coef = torch.where(ok[:, None], coef, 0.0).contiguous()
            bt.adjoint(grads_dev, coef, dT)
            dTk = torch.zeros_like(G)
            dTk[:, :3] = dT.view(B, 3, 4)
            G = torch.where(ok[:, None, None], G_direct + dTk, G)
            G[:, 3] = 0.0
        T0 = Ts[0]
        g_x0 = exp_rt_vjp(G[:, :3], x0.double(), T0).to(x0.dtype)
        kc = ctx.kcam
        v, u = torch.meshgrid(torch.arange(H, device=dev, dtype=torch.float32),
                              torch.arange(W, device=dev, dtype=torch.float32),
                              indexing="ij")
        dpdz = torch.stack([(u - float(kc[0, 2])) / float(kc[0, 0]),
                            (v - float(kc[1, 2])) / float(kc[1, 1]),
                            torch.ones_like(u)], -1).view(1, n, 3)
        g_depth = (g_sp * dpdz).sum(-1).view(B, H, W)
        g_tf_out = g_tf.view(B, -1, H, W) if ctx.use_feat else None
        g_sf_out = g_sf.view(B, -1, H, W) if ctx.use_feat else None
        return (g_depth, g_tp.view(B, H, W, 3), g_tn.view(B, H, W, 3), g_sf_out,
                g_tf_out, g_x0, None, None, None, None, None)


class DifferentiableICP:
    """Pose after ``num_steps`` GN steps, differentiable with respect to the
    source depth, target points and normals, source and target features and
    the initial 6-DoF pose x0 ([t; w], T_0 = exp_rt_to_matrix(x0)).

    Inputs are batched: depth [B, H, W], normals/points [B, H, W, 3], masks
    [B, H, W], features [B, C, H, W], x0 [B, 6]; kcam is the shared [3, 3]
    intrinsics. Returns T [B, 4, 4] (FP64).
    """

    def __init__(self, num_steps=5, geom_weight=1.0, feat_weight=0.1,
                 distance_threshold=0.1, normals_angle_thresh=math.radians(20),
                 feat_residual_thresh=2.75, tile_px=1024):
        self.cfg = dict(num_steps=num_steps, geom_weight=geom_weight,
                        feat_weight=feat_weight,
                        distance_threshold=distance_threshold,
                        normals_angle_thresh=normals_angle_thresh,
                        feat_residual_thresh=feat_residual_thresh,
                        tile_px=tile_px)

    def __call__(self, kcam, src_depth, src_normals, src_mask, tgt_points,
                 tgt_normals, tgt_mask, x0, src_feats=None, tgt_feats=None):
        kcam = torch.as_tensor(kcam).detach().to("cpu", torch.float32)
        return _UnrolledGN.apply(src_depth, tgt_points, tgt_normals, src_feats,
                                 tgt_feats, x0, kcam, src_normals, src_mask,
                                 tgt_mask, self.cfg)
