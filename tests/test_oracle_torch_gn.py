"""Pins oracle/torch_gn.py, the pure-PyTorch FP64 unrolled Gauss-Newton
forward that checks the hand-written adjoint:

* its forward, driven with the oracle's own correspondences and feature gate
  at every step, reproduces the oracle's GN poses (slamtb/registration/
  icp.py:208-256 restated in oracle/icp_oracle.cpp) in FP64;
* its autograd gradients of the pose loss equal FP64 central finite
  differences with respect to the source depth (64 random pixels) and the
  initial 6-DoF pose.

The correspondences and the feature gate are piecewise constant in the
inputs, so both checks hold them fixed, as the adjoint does.
"""
import math

import numpy as np
import torch

from oracle import oracle as orc
from oracle import torch_gn
from slamtb.registration.diff import pose_loss

from _helpers import prepared_pair

CFG = dict(geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
           normals_angle_thresh=math.radians(20), feat_residual_thresh=2.75)
STEPS = 5
NONE = 0x0FFFFFFF7F800000


def f64(P):
    return {k: (np.asarray(v, np.float64) if k.endswith(("pts", "nrm", "feats"))
                else v) for k, v in P.items()}


def oracle_decisions(Q, T):
    """Owners (source index, target position) and the feature gate of one
    step at pose T, from the oracle's FP64 kernels."""
    rt = T[:3]
    cmap = orc.corresp_map(Q["src_pts"], Q["src_nrm"], Q["src_mask"], rt,
                           Q["tgt_pts"], Q["tgt_nrm"], Q["tgt_mask"], Q["kcam"],
                           CFG["distance_threshold"], CFG["normals_angle_thresh"])
    flat = cmap.reshape(-1)
    pos = np.flatnonzero(flat != NONE)
    src = flat[pos] >> 32
    _, _, r2, _ = orc.feature(Q["src_pts"], Q["src_feats"], Q["src_mask"], rt,
                              Q["tgt_feats"], Q["kcam"], cmap,
                              CFG["feat_residual_thresh"])
    return (torch.from_numpy(src), torch.from_numpy(pos),
            torch.from_numpy(r2[src] > 0))


def oracle_trajectory(Q, x0):
    """The oracle's GN poses T_0..T_STEPS from T_0 = exp_rt(x0), and the
    decisions it made at each step."""
    T = orc.exp_rt_4x4(x0)
    poses, steps = [T], []
    for _ in range(STEPS):
        steps.append(oracle_decisions(Q, T))
        r = orc.icp_estimate(Q["kcam"], Q["src_pts"], Q["src_nrm"], Q["src_mask"],
                             Q["tgt_pts"], Q["tgt_nrm"], Q["tgt_mask"], 1,
                             src_feats=Q["src_feats"], tgt_feats=Q["tgt_feats"],
                             transform=T, dtype=np.float64,
                             with_information=False, **CFG)
        T = r["transform"]
        poses.append(T)
    return poses, steps


def backproject(depth, kcam):
    H, W = depth.shape
    v, u = torch.meshgrid(torch.arange(H, dtype=depth.dtype),
                          torch.arange(W, dtype=depth.dtype), indexing="ij")
    x = (u - kcam[0, 2]) * depth / kcam[0, 0]
    y = (v - kcam[1, 2]) * depth / kcam[1, 1]
    return torch.stack([x, y, depth], -1).view(-1, 3)


def unrolled(R, steps, x0, depth=None):
    return torch_gn.unrolled(backproject(R["depth"] if depth is None else depth, R["k"]),
                             R["tp"], R["tn"], R["sf"], R["tf"], R["k"], x0, steps,
                             CFG["geom_weight"], CFG["feat_weight"])


def torch_inputs(P):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).double()
    return dict(depth=d(P["src_depth"]), tp=d(P["tgt_pts"]), tn=d(P["tgt_nrm"]),
                sf=d(P["src_feats"]).view(P["src_feats"].shape[0], -1),
                tf=d(P["tgt_feats"]), k=d(P["kcam"]))


def case(seed):
    P = prepared_pair(seed, 160, 120, channels=3)
    # The oracle's FP64 path is driven with the same FP64 points torch
    # backprojects from the depth.
    Q = f64(P)
    Q["src_pts"] = backproject(torch.from_numpy(P["src_depth"]).double(),
                               torch.from_numpy(P["kcam"]).double()).view(
                                   P["src_pts"].shape).numpy()
    rng = np.random.default_rng(seed)
    x0 = rng.normal(size=6) * 0.01
    return P, Q, x0


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_reproduces_the_oracle_poses():
    P, Q, x0 = case(71)
    poses, steps = oracle_trajectory(Q, x0)
    R = torch_inputs(P)
    x = torch.from_numpy(x0)
    for k in range(1, STEPS + 1):
        T = unrolled(R, steps[:k], x).numpy()
        err = orc.pose_error(T, poses[k])
        assert err[0] < 1e-9 and err[1] < 1e-9, (k, err)
    # The steps move the pose: the check is not of a fixed point.
    assert orc.pose_error(poses[0], poses[-1])[0] > 1e-3


def test_gradients_match_central_differences():
    P, Q, x0 = case(72)
    _, steps = oracle_trajectory(Q, x0)
    R = torch_inputs(P)
    T_gt = torch.from_numpy(P["T_gt"])
    depth = R["depth"].clone().requires_grad_()
    xg = torch.from_numpy(x0).clone().requires_grad_()
    pose_loss(unrolled(R, steps, xg, depth), T_gt).backward()
    x = torch.from_numpy(x0)
    loss = lambda xx, dd=None: pose_loss(unrolled(R, steps, xx, dd), T_gt).item()
    rng = np.random.default_rng(0)
    pix = rng.choice(np.flatnonzero(P["src_mask"].reshape(-1)), 64, replace=False)
    fd = []
    eps = 1e-5
    with torch.no_grad():
        for p in pix:
            dp, dm = R["depth"].clone(), R["depth"].clone()
            dp.view(-1)[p] += eps
            dm.view(-1)[p] -= eps
            fd.append((loss(x, dp) - loss(x, dm)) / (2 * eps))
        # A small step for the pose: every pixel moves, and the bilinear
        # samples' derivatives jump where a tap crosses a pixel edge.
        ex = 1e-7
        fdx = [(loss(x + ex * e) - loss(x - ex * e)) / (2 * ex)
               for e in torch.eye(6, dtype=torch.float64)]
    ad = depth.grad.view(-1)[pix].numpy()
    assert rel(ad, fd) < 1e-4, rel(ad, fd)
    assert rel(xg.grad.numpy(), fdx) < 1e-4, (xg.grad.numpy(), fdx)
