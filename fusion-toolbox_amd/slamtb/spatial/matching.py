"""Differentiable frame-to-point-cloud matching (slamtb/spatial/matching.py).

``FramePointCloudMatcher`` pairs each source point with the target pixel the
projective correspondence map gives it and returns the matched target
points, normals and bilinearly sampled target features; the features carry
a gradient back to the source points. Both passes are HIP kernels
(``_cslamtb.FPCLMatcherOp``).
"""
import math

import torch

from slamtb._cslamtb import FPCLMatcherOp as _FPCLMatcherOp


class FPCLMatcherOp(torch.autograd.Function):
    """matching.py:10-85. ``forward(source_points, source_normals, target,
    match, grad_precision)`` fills ``match`` and returns the [C, N] matched
    features; the backward pass returns d loss / d source_points."""

    class Target:
        def __init__(self, points, normals, mask, features, kcam,
                     distance_threshold, normals_angle_thresh):
            self.points = points
            self.normals = normals
            self.mask = mask
            self.features = features
            self.kcam = kcam
            self.distance_threshold = distance_threshold
            self.normals_angle_thresh = normals_angle_thresh
            self._kcam_dev = {}

        def kcam_on(self, device, dtype):
            """The intrinsic matrix on `device` in `dtype`, copied once: a
            forward/backward pass issues no host-to-device copy, so it can be
            captured in a hipGraph."""
            key = (str(device), dtype)
            if key not in self._kcam_dev:
                self._kcam_dev[key] = torch.as_tensor(
                    self.kcam.matrix).to(device=device, dtype=dtype)
            return self._kcam_dev[key]

    class Match:
        def __init__(self):
            self.points = None
            self.normals = None
            self.features = None
            self.mask = None

    @staticmethod
    def forward(ctx, source_points, source_normals, target, match,
                grad_precision=0.005):
        from slamtb.registration.correspondence_map import CorrespondenceMap

        device = source_points.device
        dtype = source_points.dtype
        size = source_points.size(0)
        kcam = target.kcam_on(device, target.kcam.matrix.dtype)

        # The source points arrive already transformed: identity pose and a
        # one-element mask (every source point takes part, matching.py:44-49).
        corresp_map = CorrespondenceMap(
            target.distance_threshold, target.normals_angle_thresh)(
                source_points, source_normals,
                torch.ones(1, dtype=torch.bool, device=device),
                torch.eye(4, dtype=dtype, device=device),
                target.points, target.normals, target.mask, kcam)

        features = target.features
        if features is None:
            # Geometry only: a zero-channel feature map (the reference
            # requires a feature tensor here).
            features = torch.empty(0, target.points.size(0),
                                   target.points.size(1), device=device,
                                   dtype=dtype)
        match.points = torch.empty(size, 3, device=device, dtype=dtype)
        match.normals = torch.empty(size, 3, device=device, dtype=dtype)
        match.features = torch.empty(features.size(0), size, device=device,
                                     dtype=dtype)
        match.mask = torch.empty(size, device=device, dtype=torch.bool)
        _FPCLMatcherOp.forward(target.points, target.normals, features,
                               source_points, kcam, corresp_map, match.points,
                               match.normals, match.features, match.mask)

        ctx.target = target
        ctx.source_points = source_points
        ctx.match_mask = match.mask
        ctx.grad_precision = grad_precision
        return match.features

    @staticmethod
    def backward(ctx, dl_features):
        if ctx.target.features is None:
            return None, None, None, None, None
        dx_points = torch.empty(dl_features.size(1), 3,
                                device=dl_features.device,
                                dtype=dl_features.dtype)
        _FPCLMatcherOp.backward(ctx.target.features, ctx.source_points,
                                ctx.match_mask, dl_features.contiguous(),
                                ctx.target.kcam_on(dl_features.device,
                                                   ctx.target.kcam.matrix.dtype),
                                ctx.grad_precision, dx_points)
        return dx_points, None, None, None, None


class FramePointCloudMatcher:
    """Differentiable correspondence layer against an organised target
    (matching.py:88-124)."""

    def __init__(self, target_points, target_normals, target_mask,
                 target_features, kcam, distance_threshold=1.0,
                 normals_angle_thresh=math.pi):
        self.target = FPCLMatcherOp.Target(
            target_points, target_normals, target_mask, target_features, kcam,
            distance_threshold, normals_angle_thresh)

    @classmethod
    def from_frame_pcl(cls, target_fpcl, features):
        return cls(target_fpcl.points, target_fpcl.normals, target_fpcl.mask,
                   features, target_fpcl.kcam)

    def find_correspondences(self, source_points, source_normals):
        """Returns (points [M, 3], normals [M, 3], features [C, M], mask [N])
        of the M matched source points."""
        match = self.match_rows(source_points, source_normals)
        return (match.points[match.mask, :], match.normals[match.mask, :],
                match.features[:, match.mask], match.mask)

    def match_rows(self, source_points, source_normals):
        """The uncompacted match: a ``FPCLMatcherOp.Match`` whose points and
        normals are [N, 3] and features [C, N] (zero rows where ``mask`` is
        false), the features differentiable in ``source_points``. Selecting
        rows by the mask needs the match count on the host; callers that
        can work with masked sums avoid that synchronisation."""
        match = FPCLMatcherOp.Match()
        match.features = FPCLMatcherOp.apply(source_points, source_normals,
                                             self.target, match)
        return match


class PointCloudMatcher:
    """Correspondences against an unorganised point cloud through the
    nearest-neighbour layer (matching.py:127-172): the nearest target point
    and normal of every source point with a neighbour within
    ``distance_upper_bound``, and the target features interpolated from its
    ``num_neighbors`` nearest neighbours (differentiable in the source
    points). The reference's normal-angle filter is disabled code
    (matching.py:160-170) and is not applied here either."""

    def __init__(self, target_points, target_normals, target_features,
                 num_neighbors=5, distance_upper_bound=math.inf,
                 normals_angle_thresh=1):
        from .kdtree_layer import KDTreeLayer
        self.target_points = target_points
        self.target_normals = target_normals
        self.target_features = target_features
        self.num_neighbors = num_neighbors
        self.distance_upper_bound = distance_upper_bound
        self.normals_angle_thresh = normals_angle_thresh
        self._layer = KDTreeLayer
        self.kdtree_op = KDTreeLayer.setup(target_points)

    @classmethod
    def from_point_cloud(cls, pcl, features, num_neighbors=5,
                         distance_upper_bound=math.inf):
        return cls(pcl.points, pcl.normals, features, num_neighbors=num_neighbors,
                   distance_upper_bound=distance_upper_bound)

    def find_correspondences(self, source_points, source_normals=None):
        """Returns (points [M, 3], normals [M, 3], features [C, M] or None,
        mask [N]) of the M source points with a neighbour."""
        # The layer's state is per class, as in the reference: re-register
        # this matcher's target before querying.
        self.kdtree_op = self._layer.setup(self.target_points)
        match_mask = self._layer.query(
            source_points, k=self.num_neighbors,
            distance_upper_bound=self.distance_upper_bound)
        top1 = self._layer.last_query[1][match_mask, 0]
        matched_points = self.target_points[top1]
        matched_normals = self.target_normals[top1]
        matched_features = None
        if self.target_features is not None:
            matched_features = self.kdtree_op(self.target_features,
                                              source_points)[:, match_mask]
        return matched_points, matched_normals, matched_features, match_mask
