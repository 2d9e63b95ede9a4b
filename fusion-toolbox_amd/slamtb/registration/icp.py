"""Point-to-plane (+ feature) ICP odometry with Gauss-Newton
(slamtb/registration/icp.py of the reference).

``ICPOdometry.estimate`` keeps the reference signature and result. Instead of
the reference's per-iteration sequence -- two correspondence maps, per-pixel
[N,6,6] partials, torch sums, a host Cholesky and a host SE3 update
(icp.py:210-256) -- it enqueues the fused device loop of libslamtb_hip.so:
one correspondence kernel and one Jacobian/reduction/solve kernel per
iteration, with the pose kept on the device and a single host read at the
end.

Semantics kept from the reference:
* JtJ = fw^2 JtJ_feat + gw^2 JtJ_geom, Jr = -fw Jr_feat + gw Jr_geom
  (icp.py:221-234), residual = (fw r2_feat + gw r2_geom) / match_count,
  match_count = max(feature count, geometric count);
* the feature Jacobian keeps only the last channel (icp_jacobian_common.hpp:
  50-60) while the squared residual sums every channel;
* a failed Cholesky stops the loop and the result keeps the last accepted
  iteration (icp.py:236-242);
* the update is exp_rt_to_matrix(x) left-multiplied into the pose, with
  x = [t; w] (src/exp_rt_to_matrix.cu:22-36);
* the information matrix is sum G^T G at the final pose (icp.py:258-261).
Deviation: with geom_weight == 0 the reference crashes computing the
information matrix; here ``information`` is None.
"""
import math

import torch

from slamtb._cslamtb import ICPJacobian as _ICPJacobian
from slamtb._utils import empty_ensured_size
from slamtb.frame import Frame, FramePointCloud

from ._fused import Level, make_params, runner_for, state_matrix
from .correspondence_map import CorrespondenceMap
from .result import RegistrationResult


class _JacobianStep:
    """One linearisation of one term with the drop-in per-pixel kernels
    (icp.py:19-108): correspondence map, per-pixel partials, sums."""

    def __init__(self, geom, so3, distance_threshold, normals_angle_thresh,
                 feat_residual_thresh):
        self.JtJ = None
        self.Jtr = None
        self.residual = None
        self.geom = geom
        self.so3 = so3
        self.distance_threshold = distance_threshold
        self.normals_angle_thresh = normals_angle_thresh
        self.feat_residual_thresh = feat_residual_thresh

    def __call__(self, target_points, target_normals, target_feats,
                 target_mask, source_points, source_normals, source_feats,
                 source_mask, kcam, transform):
        if self.so3:
            raise NotImplementedError("so3 ICP is not provided (its reference "
                                      "path cannot run: icp.py:71-76)")
        device = target_points.device
        dtype = source_points.dtype
        n = source_points.size(0)
        self.JtJ = empty_ensured_size(self.JtJ, n, 6, 6, device=device, dtype=dtype)
        self.Jtr = empty_ensured_size(self.Jtr, n, 6, device=device, dtype=dtype)
        self.residual = empty_ensured_size(self.residual, n, device=device,
                                           dtype=dtype)
        corresp_map = CorrespondenceMap(self.distance_threshold,
                                        self.normals_angle_thresh)(
            source_points, source_normals, source_mask, transform.to(dtype),
            target_points, target_normals, target_mask, kcam)
        if self.geom:
            match_count = _ICPJacobian.estimate_geometric(
                source_points, source_mask, transform.to(dtype), target_points,
                target_normals, kcam, corresp_map, self.JtJ, self.Jtr,
                self.residual)
        else:
            match_count = _ICPJacobian.estimate_feature(
                source_points, source_feats, source_mask, transform.to(dtype),
                target_feats, kcam, corresp_map, self.feat_residual_thresh,
                self.JtJ, self.Jtr, self.residual)
        return self.JtJ.sum(0), self.Jtr.sum(0), self.residual.sum(), match_count

    def get_information_matrix(self, target_points, target_normals,
                               target_mask, source_points, source_normals,
                               source_mask, kcam, transform):
        dtype = source_points.dtype
        GtG = torch.empty(source_points.size(0), 6, 6, device=source_points.device,
                          dtype=dtype)
        corresp_map = CorrespondenceMap(self.distance_threshold,
                                        self.normals_angle_thresh)(
            source_points, source_normals, source_mask, transform.to(dtype),
            target_points, target_normals, target_mask, kcam)
        _ICPJacobian.get_information_matrix(source_points, source_mask,
                                            transform.to(dtype), target_points,
                                            kcam, corresp_map, GtG)
        return GtG.sum(0).cpu()


class ICPOdometry:
    """Iterative closest points with the point-to-plane error and
    Gauss-Newton (icp.py:111-310).

    Args: num_iters, geom_weight, feat_weight, so3 (not provided),
    distance_threshold (m), normals_angle_thresh (rad), feat_residual_thresh.
    """

    def __init__(self, num_iters, geom_weight=1.0, feat_weight=1.0, so3=False,
                 distance_threshold=.5, normals_angle_thresh=math.pi / 4,
                 feat_residual_thresh=2.75):
        if so3:
            raise NotImplementedError("so3 ICP is not provided (its reference "
                                      "path cannot run: icp.py:71-76)")
        self.num_iters = num_iters
        self.geom_weight = geom_weight
        self.feat_weight = feat_weight
        self.so3 = so3
        self.distance_threshold = distance_threshold
        self.normals_angle_thresh = normals_angle_thresh
        self.feat_residual_thresh = feat_residual_thresh
        self._geom_step = (_JacobianStep(True, so3, distance_threshold,
                                         normals_angle_thresh, feat_residual_thresh)
                           if geom_weight > 0 else None)
        self._feature_step = (_JacobianStep(False, so3, distance_threshold,
                                            normals_angle_thresh, feat_residual_thresh)
                              if feat_weight > 0 else None)

    def params(self, has_features):
        return make_params(self.geom_weight, self.feat_weight,
                           self.distance_threshold, self.normals_angle_thresh,
                           self.feat_residual_thresh,
                           has_features and self._feature_step is not None)

    def level(self, kcam, source_points, source_normals, source_mask,
              source_feats, target_points, target_mask, target_normals,
              target_feats, storage=None, feat_layout="planar"):
        has_features = not None in (source_feats, target_feats,
                                    self._feature_step)
        return Level(kcam, source_points, source_normals, source_mask,
                     target_points, target_normals, target_mask,
                     source_feats if has_features else None,
                     target_feats if has_features else None, storage=storage,
                     feat_layout=feat_layout)

    def enqueue(self, runner, level, transform=None, chain=False,
                information=True):
        """Enqueue begin + iterations (+ information) without reading back."""
        params = self.params(level.has_features)
        runner.begin(level, transform, chain=chain)
        runner.iterate(level, params, self.num_iters)
        if information and self._geom_step is not None:
            runner.information(level, params)

    def estimate(self, kcam, source_points, source_normals, source_mask,
                 source_feats=None, target_points=None, target_mask=None,
                 target_normals=None, target_feats=None, transform=None):
        """Estimate the transform that aligns the source onto the target
        (icp.py:152-262). Tensors must be on a HIP device."""
        level = self.level(kcam, source_points, source_normals, source_mask,
                           source_feats, target_points, target_mask,
                           target_normals, target_feats)
        runner = runner_for(level.device)
        self.enqueue(runner, level, transform)
        return result_from_state(runner.read_state(), level.n,
                                 self._geom_step is not None)

    def estimate_frame(self, source_frame, target_frame, source_feats=None,
                       target_feats=None, transform=None, device="cuda:0"):
        """icp.py:264-310."""
        if isinstance(source_frame, Frame):
            source_frame = FramePointCloud.from_frame(source_frame, device=device)
        if isinstance(target_frame, Frame):
            target_frame = FramePointCloud.from_frame(target_frame, device=device)
        return self.estimate(source_frame.kcam, source_frame.points,
                             source_frame.normals, source_frame.mask,
                             source_feats=source_feats,
                             target_points=target_frame.points,
                             target_normals=target_frame.normals,
                             target_mask=target_frame.mask,
                             target_feats=target_feats, transform=transform)


def result_from_state(st, n_src, with_information):
    """RegistrationResult from the device state (icp.py:255-262)."""
    if st.iterations == 0:
        result = RegistrationResult()
    else:
        result = RegistrationResult(
            state_matrix(st.best_transform), state_matrix(st.best_jtj, 6),
            float(st.best_residual), float(st.best_match_count) / n_src)
    if with_information:
        result.information = state_matrix(st.information, 6).float()
    result.iterations = int(st.iterations)
    return result
