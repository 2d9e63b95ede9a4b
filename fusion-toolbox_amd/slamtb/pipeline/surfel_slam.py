"""Frame-to-frame camera tracking, the front end of the reference's
``SurfelSLAM.step`` (slamtb/pipeline/surfel_slam.py:31-143).

Per frame: bilateral-filtered depth (the reference's call with the frame's
depth scale), its point cloud and normals, registration of the new frame
onto the previous one, the verifier, and the camera pose update
``rt_camera = rt_camera.transform(relative)``. Surfel fusion (``EFFusion``,
``SurfelModel``, the OpenGL rasterisers of ``tenviz``) is not built here:
``step`` returns the tracking statistics instead of fusion statistics.

With the default multiscale Gauss-Newton registration and fixed image size
the whole step after the first frame is one captured hipGraph: the new depth
(and features) are copied into static device buffers, the graph filters the
depth, builds the point cloud, normals and pyramid of the new frame, runs
every pyramid level of the fused GN loop against the previous frame's
pyramid, and copies the new pyramid into the previous-frame buffers; the
host then reads the 1.3 KB state once. The reference's step is a sequence
of eager launches, tensor allocations and host synchronisations (and a
scipy optimiser: its live configuration is ``AutogradICP(50, 0.01, 1, 1)``,
surfel_slam.py:47; pass that as ``registration`` to reproduce it, eagerly).
"""
import math
import time

import torch

from slamtb.camera import RTCamera
from slamtb.frame import Frame, FramePointCloud
from slamtb.processing import bilateral_depth_filter, feature_pyramid
from slamtb.registration import (ICPOdometry, MultiscaleRegistration,
                                 RegistrationVerifier)
from slamtb.registration._fused import GNRunner
from slamtb.registration.icp import result_from_state


def default_registration():
    """The multiscale GN preset of the reference's tracking branch
    (surfel_slam.py:38-44): 3 levels of 10 iterations, geometry 10 :
    features 1."""
    return MultiscaleRegistration([
        (1.0, ICPOdometry(10, geom_weight=10, feat_weight=1)),
        (0.5, ICPOdometry(10, geom_weight=10, feat_weight=1)),
        (0.5, ICPOdometry(10, geom_weight=10, feat_weight=1))])


def _rotation_angle(R):
    c = (float(torch.trace(R[:3, :3])) - 1.0) / 2.0
    v = torch.stack([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    return math.atan2(0.5 * float(v.norm()), c)


def estimate_confidence_weight(prev_rt_cam, curr_rt_cam):
    """Fusion confidence from the inter-frame motion (surfel_slam.py:17-25):
    max(rotation angle, translation) clamped to 0.01, mapped to
    [0.5, 1]. The reference takes the angle from ``RigidTransform.rodrigues``,
    whose native code returns a unit axis (so its weight is always 0.5);
    this uses the rotation angle that the formula intends."""
    rel = torch.linalg.inv(prev_rt_cam.matrix) @ curr_rt_cam.matrix
    conf = max(_rotation_angle(rel), float(rel[:3, 3].norm()))
    conf = min(conf, 0.01)
    return max(1 - conf / 0.01, 0.5)


class TrackingStats:
    """Outcome of one ``SurfelSLAM.step``: the camera pose after the frame,
    the registration result (None for the first frame), whether the
    verifier accepted it, the fusion confidence weight and the wall time."""

    def __init__(self, rt_camera, result, accepted, confidence_weight, ms):
        self.rt_camera = rt_camera
        self.result = result
        self.accepted = accepted
        self.confidence_weight = confidence_weight
        self.ms = ms

    def __repr__(self):
        return (f"TrackingStats(accepted={self.accepted}, "
                f"confidence_weight={self.confidence_weight:.3f}, "
                f"ms={self.ms:.3f})")


def _frame_pyramid(depth, info, feats, scales, method):
    """Filtered depth -> coarse-first list of (kcam, points, normals, mask,
    feats) for every scale (the preprocessing of surfel_slam.py:81-91 and
    multiscale.py:38-53)."""
    d = bilateral_depth_filter(depth, depth > 0, depth_scale=info.depth_scale)
    fpcl = FramePointCloud.from_frame(Frame(info, d))
    fpcl.normals  # full-resolution normals, carried into the coarse levels
    pyr = fpcl.pyramid(scales, method)
    fpyr = feature_pyramid(feats, scales) if feats is not None else [None] * len(scales)
    return [(p.kcam, p.points, p.normals, p.mask, f) for p, f in zip(pyr, fpyr)]


class _GraphStep:
    """The captured tracking step for one image size (see the module
    docstring). ``prev`` holds the previous frame's pyramid in static
    buffers; ``depth``/``feats`` are the static inputs of the graph."""

    def __init__(self, registration, info, depth, feats, device):
        self.reg = registration
        self.info = info
        self.device = device
        self.scales = [s for s, _ in registration.estimators]
        self.method = registration.downsample_xyz_method
        self.depth = torch.empty_like(depth, device=device)
        self.feats = None if feats is None else torch.empty_like(feats, device=device)
        self.runner = GNRunner(device)
        self.graph = None
        self.prev = None

    def _load(self, depth, feats):
        self.depth.copy_(depth, non_blocking=True)
        if self.feats is not None:
            self.feats.copy_(feats, non_blocking=True)

    def first(self, depth, feats):
        """Preprocess the first frame into the previous-frame buffers."""
        self._load(depth, feats)
        levels = _frame_pyramid(self.depth, self.info, self.feats, self.scales,
                                self.method)
        self.prev = [tuple(None if t is None or not torch.is_tensor(t)
                           else t.clone() for t in lv[1:]) for lv in levels]
        for lv, (pts, nrm, msk, f) in zip(levels, self.prev):
            self.runner.reserve(pts.size(0) * pts.size(1), pts.size(0), pts.size(1))

    def _body(self, hand_over=True):
        cur = _frame_pyramid(self.depth, self.info, self.feats, self.scales,
                             self.method)
        levels = []
        for (_, est), (kc, pts, nrm, msk, f), (ppts, pnrm, pmsk, pf) in zip(
                self.reg.estimators[::-1], cur, self.prev):
            levels.append((est, est.level(kc, pts, nrm, msk, f, ppts, pmsk,
                                          pnrm, pf)))
        MultiscaleRegistration.enqueue_levels(self.runner, levels)
        if hand_over:
            for (_, pts, nrm, msk, f), (ppts, pnrm, pmsk, pf) in zip(cur, self.prev):
                ppts.copy_(pts)
                pnrm.copy_(nrm)
                pmsk.copy_(msk)
                if pf is not None:
                    pf.copy_(f)
        self._levels = levels  # keeps the level structs alive with the graph

    def step(self, depth, feats):
        self._load(depth, feats)
        if self.graph is None:
            # One eager pass (no hand-over, so the previous frame stays in
            # place) initialises everything the captured launches use.
            side = torch.cuda.Stream(self.device)
            side.wait_stream(torch.cuda.current_stream(self.device))
            with torch.cuda.stream(side):
                self._body(hand_over=False)
                self.graph = torch.cuda.CUDAGraph()
                with torch.cuda.graph(self.graph, stream=side):
                    self._body()
            torch.cuda.current_stream(self.device).wait_stream(side)
        self.graph.replay()
        n = self._levels[-1][1].n
        return result_from_state(self.runner.read_state(), n, True)


class SurfelSLAM:
    """Tracking front end of surfel_slam.py:31-143 (no surfel fusion).

    Args: device; tracking_mode (only 'frame-to-frame': frame-to-model
    tracking rasterises the surfel model); registration (default: the
    multiscale GN preset of the reference's tracking branch); verifier
    (default ``RegistrationVerifier()``, the reference's ``ICPVerifier``);
    use_graph (capture the step in a hipGraph when the registration is a
    multiscale GN). The fusion arguments of the reference's constructor are
    accepted and unused.
    """

    def __init__(self, max_surfels=1024 * 1024 * 15, device="cuda:0",
                 tracking_mode='frame-to-frame', stable_conf_thresh=10,
                 stable_time_thresh=20, search_size=4, indexmap_scale=4,
                 feature_size=3, registration=None, verifier=None,
                 use_graph=True):
        if tracking_mode != 'frame-to-frame':
            raise NotImplementedError(
                "frame-to-model tracking needs the surfel model rasteriser, "
                "which is not provided")
        self.device = torch.device(device)
        self.rt_camera = RTCamera(torch.eye(4, dtype=torch.float64))
        self.icp = registration if registration is not None else default_registration()
        self.icp_verifier = verifier if verifier is not None else RegistrationVerifier()
        self.use_graph = use_graph and isinstance(self.icp, MultiscaleRegistration) \
            and all(isinstance(e, ICPOdometry) for _, e in self.icp.estimators)
        self._previous_fpcl = None
        self._previous_features = None
        self._graph = None
        self._graph_key = None

    def _key(self, frame, features):
        k = frame.info.kcam.matrix
        return (tuple(frame.depth_image.shape), str(frame.depth_image.dtype),
                tuple(k.flatten().tolist()), float(frame.info.depth_scale),
                float(frame.info.depth_bias),
                None if features is None else tuple(features.shape))

    def step(self, frame, features=None):
        """Track one frame; returns :class:`TrackingStats`."""
        t0 = time.perf_counter()
        depth = torch.as_tensor(frame.depth_image)
        if features is not None:
            features = features.to(self.device)
        result = None
        if self.use_graph:
            key = self._key(frame, features)
            if self._graph is None or key != self._graph_key:
                self._graph = _GraphStep(self.icp, frame.info, depth, features,
                                         self.device)
                self._graph_key = key
                self._graph.first(depth, features)
            else:
                result = self._graph.step(depth, features)
        else:
            d = depth.to(self.device)
            d = bilateral_depth_filter(d, d > 0, depth_scale=frame.info.depth_scale)
            live = FramePointCloud.from_frame(Frame(frame.info, d))
            live.normals  # full-resolution normals, carried into the pyramid
            if self._previous_fpcl is not None:
                prev = self._previous_fpcl
                if isinstance(self.icp, MultiscaleRegistration):
                    result = self.icp.estimate_frame(
                        live, prev, source_feats=features,
                        target_feats=self._previous_features, device=self.device)
                else:
                    result = self.icp.estimate(
                        frame.info.kcam, live.points, live.normals, live.mask,
                        source_feats=features, target_points=prev.points,
                        target_mask=prev.mask, target_normals=prev.normals,
                        target_feats=self._previous_features)
            self._previous_fpcl = live
            self._previous_features = features
        accepted, weight = False, 1.0
        if result is not None:
            accepted = bool(self.icp_verifier(result))
            relative = result.transform
            if relative is None:
                relative = torch.eye(4, dtype=torch.float64)
            rt_camera = self.rt_camera.transform(relative.cpu())
            weight = estimate_confidence_weight(self.rt_camera, rt_camera)
            self.rt_camera = rt_camera
        return TrackingStats(self.rt_camera, result, accepted, weight,
                             (time.perf_counter() - t0) * 1e3)
