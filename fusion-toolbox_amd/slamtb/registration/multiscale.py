"""Coarse-to-fine registration over an image pyramid
(slamtb/registration/multiscale.py).

When every estimator is an ``ICPOdometry`` the levels are chained on the
device: each level starts from the previous level's accepted pose (identity
if it accepted none, as ``transform = curr_result.transform`` does with a
result whose transform is None), and the host reads the state once at the
end. The returned result is the last level's, as in the reference.
"""
from slamtb.frame import Frame, FramePointCloud
from slamtb.processing import DownsampleXYZMethod, feature_pyramid

from ._fused import runner_for
from .icp import ICPOdometry, result_from_state


class MultiscaleRegistration:
    def __init__(self, estimators,
                 downsample_xyz_method=DownsampleXYZMethod.Nearest):
        """estimators: [(scale, estimator), ...] listed fine to coarse with
        cumulative scales; they run coarse to fine."""
        self.estimators = estimators
        self.downsample_xyz_method = downsample_xyz_method

    def pyramids(self, source_frame, target_frame, source_feats=None,
                 target_feats=None):
        scales = [scale for scale, _ in self.estimators]
        source_pyr = source_frame.pyramid(scales, self.downsample_xyz_method)
        target_pyr = target_frame.pyramid(scales, self.downsample_xyz_method)
        src_fp = tgt_fp = [None] * len(scales)
        if source_feats is not None:
            src_fp = feature_pyramid(source_feats, scales)
        if target_feats is not None:
            tgt_fp = feature_pyramid(target_feats, scales)
        return source_pyr, src_fp, target_pyr, tgt_fp

    def levels(self, source_frame, target_frame, source_feats=None,
               target_feats=None, storage=None, feat_layout="planar"):
        """Per-level fused inputs, coarse first (estimators must be ICP);
        storage and feat_layout as in ``_fused.Level``."""
        src_pyr, src_fp, tgt_pyr, tgt_fp = self.pyramids(
            source_frame, target_frame, source_feats, target_feats)
        out = []
        for (_, est), src, sf, tgt, tf in zip(self.estimators[::-1], src_pyr,
                                              src_fp, tgt_pyr, tgt_fp):
            out.append((est, est.level(src.kcam, src.points, src.normals,
                                       src.mask, sf, tgt.points, tgt.mask,
                                       tgt.normals, tf, storage=storage,
                                       feat_layout=feat_layout)))
        return out

    @staticmethod
    def enqueue_levels(runner, levels, transform=None):
        for i, (est, level) in enumerate(levels):
            last = i == len(levels) - 1
            est.enqueue(runner, level, transform if i == 0 else None,
                        chain=i > 0, information=last)

    def estimate_frame(self, source_frame, target_frame, source_feats=None,
                       target_feats=None, transform=None, device="cuda:0"):
        if isinstance(source_frame, Frame):
            source_frame = FramePointCloud.from_frame(source_frame, device=device)
        if isinstance(target_frame, Frame):
            target_frame = FramePointCloud.from_frame(target_frame, device=device)
        if all(isinstance(est, ICPOdometry) for _, est in self.estimators):
            levels = self.levels(source_frame, target_frame, source_feats,
                                 target_feats)
            runner = runner_for(levels[-1][1].device)
            for _, level in levels:
                runner.reserve(level.n, level.height, level.width)
            self.enqueue_levels(runner, levels, transform)
            est, level = levels[-1]
            return result_from_state(runner.read_state(), level.n,
                                     est._geom_step is not None)
        src_pyr, src_fp, tgt_pyr, tgt_fp = self.pyramids(
            source_frame, target_frame, source_feats, target_feats)
        curr_result = None
        for (_, est), src, sf, tgt, tf in zip(self.estimators[::-1], src_pyr,
                                              src_fp, tgt_pyr, tgt_fp):
            curr_result = est.estimate_frame(src, tgt, source_feats=sf,
                                             target_feats=tf,
                                             transform=transform, device=device)
            transform = curr_result.transform
        return curr_result

    def estimate_pcl(self, source_pcl, target_pcl, transform=None, device=None):
        """Coarse-to-fine registration of point or surfel clouds
        (multiscale.py:69-91): per estimator, both clouds voxel-merged at
        its scale (a voxel size; 0 keeps the cloud), then the estimator's
        ``estimate_pcl`` from the previous level's transform. The
        estimators must provide ``estimate_pcl`` (``AutogradICP`` does; the
        reference's ``ICPOdometry`` has none). Voxel merging runs on the
        device instead of the reference's host round trip."""
        if device is None:
            device = source_pcl.device
        if transform is not None:
            transform = transform.cpu()
        result = None
        for scale, estimator in self.estimators[::-1]:
            if scale > 0:
                src = source_pcl.to(device).downsample(scale)
                tgt = target_pcl.to(device).downsample(scale)
            else:
                src, tgt = source_pcl, target_pcl
            result = estimator.estimate_pcl(src, tgt, transform=transform,
                                            source_feats=src.features,
                                            target_feats=tgt.features)
            transform = result.transform
        return result
