"""Pose estimation by automatic differentiation
(slamtb/registration/autogradicp.py).

The pose is the 6-vector [t; w] (translation, exponential-map rotation).
Every cost evaluation transforms the source points, matches them to the
target frame with the differentiable HIP matcher, forms the point-to-plane
and feature costs with torch, and back-propagates through the HIP matcher
and the HIP exponential map; scipy's BFGS drives the first stage and plain
SGD the refinement, as in the reference.
"""
import math

import numpy as np
import scipy.optimize
import torch

from slamtb.camera import RigidTransform
from slamtb.frame import Frame, FramePointCloud
from slamtb.spatial.matching import FramePointCloudMatcher

from .result import RegistrationResult
from .se3 import ExpRtToMatrix, MatrixToExpRt


def _to_4x4(mtx):
    if mtx.size(0) == 4 and mtx.size(1) == 4:
        return mtx
    ret = torch.eye(4, device=mtx.device, dtype=mtx.dtype)
    ret[:3, :4] = mtx[:3, :4]
    return ret


class _ClosureBox:
    """Adapts a torch closure to scipy: ``func(x)`` writes x into the
    optimised tensor and returns (cost, gradient) (autogradicp.py:29-58)."""

    def __init__(self, closure, x):
        self.closure = closure
        self.x = x

    def func(self, x):
        with torch.no_grad():
            self.x.copy_(torch.as_tensor(np.asarray(x, dtype=np.float64),
                                         dtype=self.x.dtype))
        loss = self.closure()
        if not torch.is_tensor(loss):
            # No term was formed (nothing to differentiate): a zero cost
            # with a zero gradient. The reference returned a bare 0 here,
            # which scipy cannot unpack.
            return 0.0, np.zeros(6)
        value = loss.detach().cpu().item()
        loss.backward()
        return value, self.x.grad.detach().cpu().double().numpy()


class _HuberLoss(torch.nn.Module):
    """autogradicp.py:61-90: mean of the residuals below ``alpha`` (taken
    as they are) plus the mean of ``alpha*|r| - alpha`` over the rest."""

    def __init__(self, alpha):
        super().__init__()
        self.alpha = alpha

    def forward(self, residual):
        loss1 = residual[residual < self.alpha]
        loss2 = residual[residual >= self.alpha]
        loss2 = self.alpha * torch.abs(loss2) - self.alpha
        loss = 0
        if loss1.numel() > 0:
            loss = loss1.mean()
        if loss2.numel() > 0:
            loss += loss2.mean()
        return loss


def _compacted_loss(tsrc, tpoints, tnormals, tfeatures, smask, source_feats,
                    icp, huber_loss):
    """The reference's cost on compacted matches (autogradicp.py:183-201):
    returns (geometric loss, feature loss); 0 for a disabled term."""
    geom_loss = 0
    feat_loss = 0
    if icp.geom_weight > 0:
        diff = tpoints - tsrc[smask]
        cost = (tnormals * diff).sum(1)
        geom_loss = torch.pow(cost, 2).mean()
    if icp.feat_weight > 0 and source_feats is not None and tfeatures is not None:
        feat_diff = torch.norm(tfeatures - source_feats[:, smask], 2, dim=0)
        if icp.feat_residual_thresh > 0:
            thr2 = icp.feat_residual_thresh * icp.feat_residual_thresh
            feat_diff = feat_diff[feat_diff.pow(2) < thr2]
        if icp.huber_loss_alpha > 0:
            feat_loss = huber_loss(feat_diff)
        else:
            feat_loss = feat_diff.mean()
    return geom_loss, feat_loss


def _masked_mean(values, keep):
    """Mean of values[keep] as a masked sum over the count; NaN when nothing
    is kept, as the mean of an empty selection."""
    return torch.where(keep, values, 0).sum() / keep.sum()


def _row_loss(tsrc, match, source_feats, icp):
    """The same cost on uncompacted match rows, as masked sums: no host
    synchronisation. Equal to :func:`_compacted_loss` up to summation
    order; returns (geometric loss, feature loss)."""
    mask = match.mask
    geom_loss = 0
    feat_loss = 0
    if icp.geom_weight > 0:
        cost = ((match.points - tsrc) * match.normals).sum(1)
        geom_loss = _masked_mean(cost.pow(2), mask)
    if icp.feat_weight > 0 and source_feats is not None:
        d = torch.norm(match.features - source_feats, 2, dim=0)
        keep = mask
        if icp.feat_residual_thresh > 0:
            thr2 = icp.feat_residual_thresh * icp.feat_residual_thresh
            keep = keep & (d.detach().pow(2) < thr2)
        alpha = icp.huber_loss_alpha
        if alpha > 0:
            # Each Huber part counts only when it selected something
            # (autogradicp.py:84-88); with nothing selected its sum is 0.
            lo = keep & (d.detach() < alpha)
            hi = keep & (d.detach() >= alpha)
            feat_loss = (torch.where(lo, d, 0).sum() / lo.sum().clamp(min=1)
                         + torch.where(hi, alpha * torch.abs(d) - alpha, 0).sum()
                         / hi.sum().clamp(min=1))
        else:
            feat_loss = _masked_mean(d, keep)
    return geom_loss, feat_loss


class AutogradICP:
    """ICP whose 6-D gradient comes from autograd (autogradicp.py:95-374).

    Attributes mirror the reference: ``num_iters`` BFGS iterations,
    ``learning_rate`` scales the cost handed to BFGS, ``geom_weight`` and
    ``feat_weight`` weight the point-to-plane and feature terms,
    ``distance_threshold`` / ``normals_angle_thresh`` gate matches,
    ``feat_residual_thresh`` drops feature residuals at or above it,
    ``huber_loss_alpha`` (0 disables Huber) and ``gradient_descent_iters``
    SGD refinement steps (None skips them).

    With a frame matcher each cost evaluation runs on the device without a
    host round trip (masked sums over the uncompacted match rows) until
    scipy or the SGD loop reads the cost; another matcher is used through
    ``find_correspondences`` and the reference's compacted cost."""

    def __init__(self, num_iters, learning_rate=0.01, geom_weight=0.5,
                 feat_weight=0.5, distance_threshold=0.1,
                 normals_angle_thresh=math.pi / 8.0, feat_residual_thresh=0.5,
                 huber_loss_alpha=4.5, gradient_descent_iters=100):
        self.num_iters = num_iters
        self.geom_weight = geom_weight
        self.feat_weight = feat_weight
        self.learning_rate = learning_rate
        self.distance_threshold = distance_threshold
        self.normals_angle_thresh = normals_angle_thresh
        self.feat_residual_thresh = feat_residual_thresh
        self.huber_loss_alpha = huber_loss_alpha
        self.gradient_descent_iters = gradient_descent_iters

    def _estimate(self, source_points, source_normals, source_feats,
                  target_matcher, transform=None):
        device = source_points.device
        dtype = source_points.dtype

        if transform is None:
            exp_rt = torch.zeros(6, device=device, dtype=dtype)
        else:
            exp_rt = MatrixToExpRt.apply(
                transform[:3, :4].to(device, dtype)).detach().clone()
        exp_rt.requires_grad_(True)

        loss = math.inf
        if self.feat_weight <= 0:
            source_feats = None
        huber_loss = _HuberLoss(self.huber_loss_alpha)
        rows = hasattr(target_matcher, "match_rows")

        def _closure():
            nonlocal loss
            if exp_rt.grad is not None:
                exp_rt.grad.zero_()

            rigid = RigidTransform(ExpRtToMatrix.apply(exp_rt))
            tsrc = (rigid @ source_points).view(-1, 3)
            with torch.no_grad():
                tnormals = rigid.transform_normals(source_normals)
            if rows:
                match = target_matcher.match_rows(tsrc, tnormals)
                geom_loss, feat_loss = _row_loss(tsrc, match, source_feats, self)
            else:
                geom_loss, feat_loss = _compacted_loss(
                    tsrc, *target_matcher.find_correspondences(tsrc, tnormals),
                    source_feats, self, huber_loss)

            loss = geom_loss * self.geom_weight + feat_loss * self.feat_weight
            if not torch.is_tensor(loss):
                return loss
            return loss * self.learning_rate

        box = _ClosureBox(_closure, exp_rt)
        opt_res = scipy.optimize.minimize(
            box.func, exp_rt.detach().cpu().double().numpy(), method="BFGS",
            jac=True, tol=1e-9, options=dict(disp=False, maxiter=self.num_iters))

        best_exp_rt = exp_rt.detach().clone()
        best_residual = 5555

        if self.gradient_descent_iters is not None:
            opt = torch.optim.SGD([exp_rt], 0.1)
            opt_res = None
            for _ in range(self.gradient_descent_iters):
                opt.zero_grad()
                _closure()
                if not torch.is_tensor(loss):
                    break
                loss.backward()
                tmp = loss.detach().cpu().item()
                if tmp < best_residual:
                    best_residual = tmp
                    best_exp_rt = exp_rt.detach().clone()
                opt.step()

        with torch.no_grad():
            transform = _to_4x4(ExpRtToMatrix.apply(best_exp_rt).cpu())

        return RegistrationResult(
            transform,
            (torch.from_numpy(opt_res.hess_inv).inverse().float()
             if opt_res is not None else None),
            loss, 1.0)

    def estimate(self, kcam, source_points, source_normals, source_mask,
                 target_points=None, target_mask=None, target_normals=None,
                 target_feats=None, source_feats=None, transform=None):
        """autogradicp.py:241-295. ``kcam`` is a ``KCamera``; source tensors
        are organised [H, W, 3] (or flat) with a [H, W] mask, features
        [C, H, W]; targets are organised [H, W, ...]."""
        source_points = source_points[source_mask].view(-1, 3)
        source_normals = source_normals[source_mask].view(-1, 3)
        if source_feats is not None:
            source_feats = source_feats[:, source_mask].view(
                source_feats.size(0), -1)
        matcher = FramePointCloudMatcher(
            target_points, target_normals, target_mask, target_feats, kcam,
            self.distance_threshold, self.normals_angle_thresh)
        return self._estimate(source_points, source_normals, source_feats,
                              matcher, transform)

    def estimate_pcl(self, source_pcl, target_pcl, transform=None,
                     source_feats=None, target_feats=None):
        """Align two point (or surfel) clouds (autogradicp.py:297-330):
        8 nearest neighbours within ``distance_threshold`` per source point,
        the nearest one's point and normal for the geometric term and their
        interpolated features for the feature term."""
        from slamtb.spatial.matching import PointCloudMatcher
        matcher = PointCloudMatcher(
            target_pcl.points, target_pcl.normals, target_feats,
            num_neighbors=8, distance_upper_bound=self.distance_threshold,
            normals_angle_thresh=self.normals_angle_thresh)
        return self._estimate(source_pcl.points, source_pcl.normals,
                              source_feats, matcher, transform)

    def estimate_frame(self, source_frame, target_frame, source_feats=None,
                       target_feats=None, transform=None, device="cpu"):
        """autogradicp.py:332-374."""
        if isinstance(source_frame, Frame):
            source_frame = FramePointCloud.from_frame(source_frame).to(device)
        if isinstance(target_frame, Frame):
            target_frame = FramePointCloud.from_frame(target_frame).to(device)
        return self.estimate(source_frame.kcam, source_frame.points,
                             source_frame.normals, source_frame.mask,
                             source_feats=source_feats,
                             target_points=target_frame.points,
                             target_mask=target_frame.mask,
                             target_normals=target_frame.normals,
                             target_feats=target_feats, transform=transform)
