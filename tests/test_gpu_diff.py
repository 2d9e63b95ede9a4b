"""Differentiable ICP (unrolled GN with the hand-written adjoint).

Checks, for a batch of pairs and 5 geometric + 3-channel feature steps:
* the adjoint's gradients of the pose loss with respect to the source depth,
  target points and normals, source and target features and the initial
  6-DoF pose equal torch autograd through the pure-PyTorch FP64 restatement
  (oracle/torch_gn.py) driven with the same correspondences (norm-wise
  relative error <= 2e-3: the adjoint runs in FP32);
* that restatement's gradients equal FP64 central finite differences
  (eps = 1e-4) on 64 random source pixels and all 6 pose dimensions.
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


def decisions(P, T, dev=DEV):
    """Owners and feature gate of one step at pose T (the GPU forward's own
    rules, through the drop-in kernels)."""
    g = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    k = g(P["kcam"])
    rt = T.float().to(dev)
    cmap = CorrespondenceMap(CFG["distance_threshold"], CFG["normals_angle_thresh"])(
        g(P["src_pts"]).view(-1, 3), g(P["src_nrm"]).view(-1, 3),
        g(P["src_mask"]).view(-1), rt, g(P["tgt_pts"]), g(P["tgt_nrm"]),
        g(P["tgt_mask"]), k)
    flat = cmap.view(-1)
    pos = torch.nonzero(flat != 0x0FFFFFFF7F800000).view(-1)
    src = (flat[pos] >> 32)
    n = flat.numel()
    jtj = torch.empty(n, 6, 6, device=dev)
    jr = torch.empty(n, 6, device=dev)
    r2 = torch.empty(n, device=dev)
    ICPJacobian.estimate_feature(g(P["src_pts"]).view(-1, 3), g(P["src_feats"]).view(3, -1),
                                 g(P["src_mask"]).view(-1), rt, g(P["tgt_feats"]), k,
                                 cmap, CFG["feat_residual_thresh"], jtj, jr, r2)
    ok = r2[src] > 0
    return src.cpu(), pos.cpu(), ok.cpu()


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


def reference_steps(P, R, x0):
    steps = []
    with torch.no_grad():
        for _ in range(STEPS):
            T = torch_gn.unrolled(backproject(R["depth"], R["k"]), R["tp"], R["tn"],
                                  R["sf"], R["tf"], R["k"], x0, steps,
                                  CFG["geom_weight"], CFG["feat_weight"])
            steps.append(decisions(P, T))
    return steps


def reference_loss(R, steps, x0, T_gt, depth=None, tp=None, tn=None, sf=None,
                   tf=None):
    depth = R["depth"] if depth is None else depth
    T = torch_gn.unrolled(backproject(depth, R["k"]),
                          R["tp"] if tp is None else tp,
                          R["tn"] if tn is None else tn,
                          R["sf"] if sf is None else sf,
                          R["tf"] if tf is None else tf, R["k"], x0, steps,
                          CFG["geom_weight"], CFG["feat_weight"])
    return pose_loss(T, T_gt)


@pytest.fixture(scope="module")
def batch():
    pairs = [prepared_pair(70 + i, 160, 120, channels=3) for i in range(2)]
    rng = np.random.default_rng(5)
    x0 = rng.normal(size=(2, 6)) * np.array([0.01] * 3 + [0.01] * 3)
    return pairs, x0


def gpu_grads(pairs, x0):
    st = lambda key, dt=torch.float32: torch.stack(
        [torch.from_numpy(np.ascontiguousarray(p[key])) for p in pairs]).to(DEV, dt)
    depth = st("src_depth").requires_grad_()
    tp = st("tgt_pts").requires_grad_()
    tn = st("tgt_nrm").requires_grad_()
    sf = st("src_feats").requires_grad_()
    tf = st("tgt_feats").requires_grad_()
    x = torch.from_numpy(x0).to(DEV).requires_grad_()
    T_gt = torch.from_numpy(np.stack([p["T_gt"] for p in pairs])).to(DEV)
    icp = DifferentiableICP(STEPS, **CFG)
    T = icp(pairs[0]["kcam"], depth, st("src_nrm"), st("src_mask", torch.bool), tp, tn,
            st("tgt_mask", torch.bool), x, src_feats=sf, tgt_feats=tf)
    loss = pose_loss(T, T_gt).sum()
    loss.backward()
    return T.detach().cpu(), dict(depth=depth.grad.cpu(), tp=tp.grad.cpu(),
                                  tn=tn.grad.cpu(), sf=sf.grad.cpu(),
                                  tf=tf.grad.cpu(), x0=x.grad.cpu())


def rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_adjoint_matches_autograd_of_the_reference(batch):
    pairs, x0 = batch
    T_gpu, G = gpu_grads(pairs, x0)
    for b, P in enumerate(pairs):
        R = reference_inputs(P)
        xr = torch.from_numpy(x0[b]).clone()
        steps = reference_steps(P, R, xr)
        leaves = {k: R[k].clone().requires_grad_() for k in ("depth", "tp", "tn", "sf", "tf")}
        xg = xr.clone().requires_grad_()
        L = reference_loss(R, steps, xg, torch.from_numpy(P["T_gt"]),
                           **{k: v for k, v in leaves.items()})
        L.backward()
        ref = {k: v.grad.numpy() for k, v in leaves.items()}
        ref["x0"] = xg.grad.numpy()
        got = {"depth": G["depth"][b].numpy(), "tp": G["tp"][b].numpy(),
               "tn": G["tn"][b].numpy(), "sf": G["sf"][b].reshape(3, -1).numpy(),
               "tf": G["tf"][b].numpy(), "x0": G["x0"][b].numpy()}
        for key in ref:
            e = rel(got[key].reshape(-1), ref[key].reshape(-1))
            print(f"pair {b} d/d{key}: rel err {e:.2e} (|ref| {np.linalg.norm(ref[key]):.3e})")
            assert e < 2e-3, (b, key, e)


def test_reference_gradients_match_central_differences(batch):
    pairs, x0 = batch
    P = pairs[0]
    R = reference_inputs(P)
    xr = torch.from_numpy(x0[0]).clone()
    steps = reference_steps(P, R, xr)
    T_gt = torch.from_numpy(P["T_gt"])
    depth = R["depth"].clone().requires_grad_()
    xg = xr.clone().requires_grad_()
    reference_loss(R, steps, xg, T_gt, depth=depth).backward()
    eps = 1e-4
    rng = np.random.default_rng(0)
    valid = np.flatnonzero(P["src_mask"].reshape(-1))
    pix = rng.choice(valid, 64, replace=False)
    fd = []
    with torch.no_grad():
        for p in pix:
            dp, dm = R["depth"].clone(), R["depth"].clone()
            dp.view(-1)[p] += eps
            dm.view(-1)[p] -= eps
            fd.append((reference_loss(R, steps, xr, T_gt, depth=dp) -
                       reference_loss(R, steps, xr, T_gt, depth=dm)).item() / (2 * eps))
        fdx = []
        for i in range(6):
            e = torch.zeros(6, dtype=torch.float64)
            e[i] = eps
            fdx.append((reference_loss(R, steps, xr + e, T_gt) -
                        reference_loss(R, steps, xr - e, T_gt)).item() / (2 * eps))
    ad = depth.grad.view(-1)[pix].numpy()
    assert rel(ad, fd) < 1e-3, rel(ad, fd)
    assert rel(xg.grad.numpy(), fdx) < 1e-3, rel(xg.grad.numpy(), fdx)
