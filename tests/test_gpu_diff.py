"""Differentiable ICP (unrolled GN with the hand-written adjoint) against
torch autograd through the pure-PyTorch FP64 restatement
(oracle/torch_gn.py, itself pinned on the CPU against the oracle's GN poses
and central finite differences).

For a batch of two pairs and 5 geometric + 3-channel feature steps, the
restatement is driven with the correspondences and feature gate the GPU
forward made at its own poses (the same drop-in kernels at the same FP32
pose), so both sides differentiate the same piecewise-smooth function. The
adjoint runs in FP32 per pixel: gradients with respect to the source depth,
target points and normals, source and target features and the initial 6-DoF
pose must agree with autograd to a norm-wise relative error of 2e-3.
"""
import math

import numpy as np
import pytest
import torch

from oracle import torch_gn
from slamtb._cslamtb import ICPJacobian
from slamtb.registration.correspondence_map import CorrespondenceMap
from slamtb.registration.diff import DifferentiableICP, pose_loss

from _helpers import prepared_pair

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
CFG = dict(geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
           normals_angle_thresh=math.radians(20), feat_residual_thresh=2.75)
STEPS = 5
NONE = 0x0FFFFFFF7F800000


def decisions(P, T):
    """Owners (source index, target position) and feature gate of one step at
    pose T, through the drop-in kernels."""
    g = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    k = g(P["kcam"])
    rt = T.float().to(DEV)
    cmap = CorrespondenceMap(CFG["distance_threshold"], CFG["normals_angle_thresh"])(
        g(P["src_pts"]).view(-1, 3), g(P["src_nrm"]).view(-1, 3),
        g(P["src_mask"]).view(-1), rt, g(P["tgt_pts"]), g(P["tgt_nrm"]),
        g(P["tgt_mask"]), k)
    flat = cmap.view(-1)
    pos = torch.nonzero(flat != NONE).view(-1)
    src = flat[pos] >> 32
    n = flat.numel()
    jtj = torch.empty(n, 6, 6, device=DEV)
    jr = torch.empty(n, 6, device=DEV)
    r2 = torch.empty(n, device=DEV)
    ICPJacobian.estimate_feature(g(P["src_pts"]).view(-1, 3), g(P["src_feats"]).view(3, -1),
                                 g(P["src_mask"]).view(-1), rt, g(P["tgt_feats"]), k,
                                 cmap, CFG["feat_residual_thresh"], jtj, jr, r2)
    return src.cpu(), pos.cpu(), (r2[src] > 0).cpu()


def backproject(depth, kcam):
    H, W = depth.shape
    v, u = torch.meshgrid(torch.arange(H, dtype=depth.dtype),
                          torch.arange(W, dtype=depth.dtype), indexing="ij")
    x = (u - kcam[0, 2]) * depth / kcam[0, 0]
    y = (v - kcam[1, 2]) * depth / kcam[1, 1]
    return torch.stack([x, y, depth], -1).view(-1, 3)


def reference_inputs(P):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).double()
    return dict(depth=d(P["src_depth"]), tp=d(P["tgt_pts"]), tn=d(P["tgt_nrm"]),
                sf=d(P["src_feats"]).view(3, -1), tf=d(P["tgt_feats"]),
                k=d(P["kcam"]))


def reference_pose(R, steps, x0, depth=None, tp=None, tn=None, sf=None, tf=None):
    pick = lambda v, key: R[key] if v is None else v
    return torch_gn.unrolled(backproject(pick(depth, "depth"), R["k"]),
                             pick(tp, "tp"), pick(tn, "tn"), pick(sf, "sf"),
                             pick(tf, "tf"), R["k"], x0, steps,
                             CFG["geom_weight"], CFG["feat_weight"])


@pytest.fixture(scope="module")
def batch():
    pairs = [prepared_pair(70 + i, 160, 120, channels=3) for i in range(2)]
    rng = np.random.default_rng(5)
    x0 = rng.normal(size=(2, 6)) * 0.01
    return pairs, x0


def stacked(pairs, key, dt=torch.float32):
    return torch.stack([torch.from_numpy(np.ascontiguousarray(p[key]))
                        for p in pairs]).to(DEV, dt)


def run_gpu(pairs, x0, steps, grad=False):
    depth = stacked(pairs, "src_depth").requires_grad_(grad)
    tp = stacked(pairs, "tgt_pts").requires_grad_(grad)
    tn = stacked(pairs, "tgt_nrm").requires_grad_(grad)
    sf = stacked(pairs, "src_feats").requires_grad_(grad)
    tf = stacked(pairs, "tgt_feats").requires_grad_(grad)
    x = torch.from_numpy(x0).to(DEV).requires_grad_(grad)
    icp = DifferentiableICP(steps, **CFG)
    T = icp(pairs[0]["kcam"], depth, stacked(pairs, "src_nrm"),
            stacked(pairs, "src_mask", torch.bool), tp, tn,
            stacked(pairs, "tgt_mask", torch.bool), x, src_feats=sf, tgt_feats=tf)
    if not grad:
        return T.detach().cpu(), None
    T_gt = torch.from_numpy(np.stack([p["T_gt"] for p in pairs])).to(DEV)
    pose_loss(T, T_gt).sum().backward()
    return T.detach().cpu(), dict(depth=depth.grad.cpu(), tp=tp.grad.cpu(),
                                  tn=tn.grad.cpu(), sf=sf.grad.cpu(),
                                  tf=tf.grad.cpu(), x0=x.grad.cpu())


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_adjoint_matches_autograd_of_the_reference(batch):
    pairs, x0 = batch
    # The GPU forward's own poses T_0 .. T_4 and the decisions taken there.
    poses = [run_gpu(pairs, x0, k)[0] for k in range(STEPS)]
    T_gpu, G = run_gpu(pairs, x0, STEPS, grad=True)
    errors = []
    for b, P in enumerate(pairs):
        steps = [decisions(P, poses[k][b]) for k in range(STEPS)]
        R = reference_inputs(P)
        leaves = {k: R[k].clone().requires_grad_() for k in ("depth", "tp", "tn", "sf", "tf")}
        xg = torch.from_numpy(x0[b]).clone().requires_grad_()
        T_ref = reference_pose(R, steps, xg, **leaves)
        fwd = float((T_ref.detach() - T_gpu[b]).abs().max())
        print(f"pair {b}: matched {[len(s[0]) for s in steps]}, |T_gpu - T_ref| {fwd:.2e}")
        assert fwd < 1e-5
        pose_loss(T_ref, torch.from_numpy(P["T_gt"])).backward()
        ref = {k: v.grad.numpy() for k, v in leaves.items()}
        ref["x0"] = xg.grad.numpy()
        got = {"depth": G["depth"][b].numpy(), "tp": G["tp"][b].numpy(),
               "tn": G["tn"][b].numpy(), "sf": G["sf"][b].reshape(3, -1).numpy(),
               "tf": G["tf"][b].numpy(), "x0": G["x0"][b].numpy()}
        for key in ref:
            e = rel(got[key].reshape(-1), ref[key].reshape(-1))
            print(f"pair {b} d/d{key}: rel err {e:.2e} (|ref| {np.linalg.norm(ref[key]):.3e})")
            errors.append((b, key, e))
    bad = [e for e in errors if not e[2] < 2e-3]
    assert not bad, bad
