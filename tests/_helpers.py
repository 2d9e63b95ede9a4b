"""Shared inputs for the tests: seeded synthetic frame pairs prepared with
the oracle's preprocessing, so both sides see identical inputs."""
import numpy as np

from oracle import oracle as orc
from slamtb import synthetic


def prepared_pair(seed, width, height, kcam=None, channels=3, noise=True,
                  bilateral=False, descriptor=False):
    k = kcam if kcam is not None else synthetic.intrinsics(width, height)
    pair = synthetic.make_pair(seed, width, height, kcam=k, channels=channels,
                               noise=noise, descriptor=descriptor)
    out = dict(kcam=k, T_gt=pair["T_gt"])
    for side in ("src", "tgt"):
        depth = pair[f"{side}_depth"]
        if bilateral:
            depth = orc.bilateral(depth, depth > 0, 6, 0.03, 4.5, 1.0)
        mask = orc.erode(depth > 0)
        pts = orc.backproject(depth, k)
        out[f"{side}_depth"] = depth
        out[f"{side}_mask"] = mask
        out[f"{side}_pts"] = pts
        out[f"{side}_nrm"] = orc.normals(pts, mask)
        out[f"{side}_feats"] = pair[f"{side}_feats"]
    return out


def to_dev(d, device):
    import torch
    return {k: (torch.from_numpy(np.ascontiguousarray(v)).to(device)
                if isinstance(v, np.ndarray) else v) for k, v in d.items()}


def level_args(level):
    """The oracle's positional inputs (kcam and the six frame arrays) of a
    fused ``Level`` (device tensors), as FP32 numpy arrays."""
    k = np.array([[level.fx, 0, level.cx], [0, level.fy, level.cy], [0, 0, 1]],
                 np.float32)
    c = lambda t: t.detach().float().cpu().numpy()
    return (k, c(level.src_points), c(level.src_normals),
            level.src_mask.cpu().numpy(), c(level.tgt_points),
            c(level.tgt_normals), level.tgt_mask.cpu().numpy())


def level_kwargs(est, level):
    c = lambda t: t.detach().float().cpu().numpy()
    return dict(src_feats=c(level.src_feats) if level.has_features else None,
                tgt_feats=c(level.tgt_feats) if level.has_features else None,
                geom_weight=est.geom_weight, feat_weight=est.feat_weight,
                distance_threshold=est.distance_threshold,
                normals_angle_thresh=est.normals_angle_thresh,
                feat_residual_thresh=est.feat_residual_thresh)


def state_pose(st):
    return np.array(list(st.transform), np.float64).reshape(4, 4)
