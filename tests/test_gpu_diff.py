"""Differentiable ICP (unrolled GN with the hand-written adjoint) against
torch autograd through the pure-PyTorch FP64 restatement
(oracle/torch_gn.py, itself pinned on the CPU against the oracle's GN poses
and central finite differences).

For a batch of two pairs and 5 geometric + 3-channel feature steps, the
restatement is driven with the correspondences, the feature gate and the
bilinear cells of the feature samples that the GPU forward used at its own
poses (the same drop-in kernels at the same FP32 pose, and the cells from
the kernels' FP32 arithmetic), and with the GPU's FP32 source points as
values (moving along each pixel's ray with the depth), so both sides
differentiate the same piecewise-smooth function of the same inputs. The
restatement's steps are evaluated at the GPU forward's poses (derivatives
straight through), so both differentiate along the same trajectory.
Gradients with respect to the source depth, target points and normals,
source and target features and the initial 6-DoF pose must agree with
autograd to a norm-wise relative error of 1e-4.
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
    src, ok = src.cpu(), (r2[src] > 0).cpu()
    return src, pos.cpu(), ok, fp32_cells(P, T, src[ok])


def fp32_cells(P, T, src):
    """The cells (truncated u, v) of the five feature samples of source
    pixels `src` at pose T, evaluated as the kernels evaluate them: the
    float pose, R p + t summed left to right, u = fx x / z + cx, each
    operation rounded to FP32 (no contraction), h = 0.05f."""
    rt = T[:3].to(torch.float32)
    sp = torch.from_numpy(np.ascontiguousarray(P["src_pts"])).view(-1, 3)[src]
    row = lambda i: ((rt[i, 0] * sp[:, 0] + rt[i, 1] * sp[:, 1]) +
                     rt[i, 2] * sp[:, 2]) + rt[i, 3]
    x, y, z = row(0), row(1), row(2)
    k = torch.from_numpy(np.ascontiguousarray(P["kcam"])).to(torch.float32)
    u = k[0, 0] * x / z + k[0, 2]
    v = k[1, 1] * y / z + k[1, 2]
    h = torch.tensor(0.05, dtype=torch.float32)
    zero = torch.zeros((), dtype=torch.float32)
    samples = [(u, v), (u + h, v), (torch.maximum(u - h, zero), v), (u, v + h),
               (u, torch.maximum(v - h, zero))]
    return torch.stack([torch.stack([torch.trunc(a).long(), torch.trunc(b).long()])
                        for a, b in samples])


def backproject(depth, kcam):
    H, W = depth.shape
    v, u = torch.meshgrid(torch.arange(H, dtype=depth.dtype, device=depth.device),
                          torch.arange(W, dtype=depth.dtype, device=depth.device),
                          indexing="ij")
    x = (u - kcam[0, 2]) * depth / kcam[0, 0]
    y = (v - kcam[1, 2]) * depth / kcam[1, 1]
    return torch.stack([x, y, depth], -1).view(-1, 3)


def reference_inputs(P):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).double()
    return dict(depth=d(P["src_depth"]), tp=d(P["tgt_pts"]), tn=d(P["tgt_nrm"]),
                sf=d(P["src_feats"]).view(3, -1), tf=d(P["tgt_feats"]),
                k=d(P["kcam"]), sp=d(P["src_pts"]).view(-1, 3))


def source_points(R, depth):
    """The source points as the GPU holds them (the FP32 back-projection,
    widened), moving with the depth along each pixel's ray: the values are
    the GPU's inputs and the derivative with respect to the depth is exact."""
    ray = backproject(torch.ones_like(depth), R["k"])
    return R["sp"] + (depth - R["depth"]).reshape(-1, 1) * ray


def reference_pose(R, steps, x0, depth=None, tp=None, tn=None, sf=None, tf=None,
                   poses=None):
    pick = lambda v, key: R[key] if v is None else v
    return torch_gn.unrolled(source_points(R, pick(depth, "depth")),
                             pick(tp, "tp"), pick(tn, "tn"), pick(sf, "sf"),
                             pick(tf, "tf"), R["k"], x0, steps,
                             CFG["geom_weight"], CFG["feat_weight"], poses=poses)


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
    poses = [run_gpu(pairs, x0, k)[0] for k in range(STEPS + 1)]
    T_gpu, G = run_gpu(pairs, x0, STEPS, grad=True)
    errors = []
    for b, P in enumerate(pairs):
        steps = [decisions(P, poses[k][b]) for k in range(STEPS)]
        R = reference_inputs(P)
        leaves = {k: R[k].clone().requires_grad_() for k in ("depth", "tp", "tn", "sf", "tf")}
        xg = torch.from_numpy(x0[b]).clone().requires_grad_()
        with torch.no_grad():
            fwd = float((reference_pose(R, steps, xg) - T_gpu[b]).abs().max())
        print(f"pair {b}: matched {[len(s[0]) for s in steps]}, |T_gpu - T_ref| {fwd:.2e}")
        assert fwd < 1e-5
        # Evaluated along the GPU forward's trajectory (see the 640x480 test).
        T_ref = reference_pose(R, steps, xg, **leaves,
                               poses=[poses[k][b] for k in range(STEPS + 1)])
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
    bad = [e for e in errors if not e[2] < 1e-4]
    assert not bad, bad


def to_dev64(R):
    return {k: v.to(DEV) for k, v in R.items()}


def test_adjoint_at_640x480_batch_of_eight():
    """Config-size check: 8 pairs at 640x480 (bilateral-filtered depth), 5
    geometric + 3-channel feature steps. The 8-pair forward gives pairs 0
    and 1 the poses a 2-pair batch gives them (each pair's tiles and sums are
    its own), and the same gradients.

    For pairs 0 and 1 the hand-written adjoint's gradients (depth, target
    points and normals, both feature maps, initial pose) equal autograd of
    the FP64 PyTorch restatement evaluated along the GPU forward's trajectory
    (each step's pose value taken from the forward, derivatives straight
    through) to 1e-4 norm-wise (seen: 1e-8, 3e-6 for the depth). Along its
    own FP64 trajectory the restatement's gradients differ by up to ~1e-2:
    at this size the forward is still moving after 5 steps, the FP32 and
    FP64 trajectories part by ~1e-6, and the derivative of the unrolled map
    changes that much with the points at which the feature samples are
    taken -- printed, not asserted. That restatement in turn equals FP64
    central differences of itself (pair 0): 64 random depth pixels with a
    step of 1e-4 (a depth step moves one pixel), and the 6 pose dimensions
    with a step of 1e-7 to 1e-3 (a pose step moves every pixel, and where
    one of the ~10^6 bilinear taps crosses a pixel edge the derivative
    jumps; at 1e-4 a few percent of them do, so the quotient at 1e-4 --
    printed -- is no longer the derivative at the point)."""
    pairs = [prepared_pair(80 + i, 640, 480, channels=3, bilateral=True)
             for i in range(8)]
    rng = np.random.default_rng(6)
    x0 = rng.normal(size=(8, 6)) * 0.01
    T8, G8 = run_gpu(pairs, x0, STEPS, grad=True)
    T2, G2 = run_gpu(pairs[:2], x0[:2], STEPS, grad=True)
    assert torch.equal(T8[:2], T2)
    for key in G2:
        assert rel(G8[key][:2].numpy(), G2[key].numpy()) < 1e-5, key
    poses = [run_gpu(pairs[:2], x0[:2], k)[0] for k in range(STEPS + 1)]
    for b in range(2):
        P = pairs[b]
        steps = [tuple(t.to(DEV) for t in decisions(P, poses[k][b]))
                 for k in range(STEPS)]
        R = to_dev64(reference_inputs(P))
        T_gt = torch.from_numpy(P["T_gt"]).to(DEV)
        got = {"depth": G2["depth"][b], "tp": G2["tp"][b], "tn": G2["tn"][b],
               "sf": G2["sf"][b].reshape(3, -1), "tf": G2["tf"][b], "x0": G2["x0"][b]}
        refs = {}
        for pinned in (True, False):
            leaves = {k: R[k].clone().requires_grad_()
                      for k in ("depth", "tp", "tn", "sf", "tf")}
            xg = torch.from_numpy(x0[b]).to(DEV).requires_grad_()
            T_ref = reference_pose(
                R, steps, xg, **leaves,
                poses=[p[b].to(DEV) for p in poses] if pinned else None)
            if not pinned:
                fwd = float((T_ref.detach().cpu() - T2[b]).abs().max())
                print(f"640x480 pair {b}: matched {[len(s[0]) for s in steps]}, "
                      f"FP64 trajectory ends {fwd:.2e} from the GPU's")
                assert fwd < 1e-5
            pose_loss(T_ref, T_gt).backward()
            ref = {k: v.grad.cpu().numpy() for k, v in leaves.items()}
            ref["x0"] = xg.grad.cpu().numpy()
            refs[pinned] = ref
        for key in refs[True]:
            e = rel(got[key].numpy().reshape(-1), refs[True][key].reshape(-1))
            e_own = rel(got[key].numpy().reshape(-1), refs[False][key].reshape(-1))
            print(f"640x480 pair {b} d/d{key}: rel err {e:.2e} along the GPU "
                  f"trajectory ({e_own:.2e} along the FP64 one)")
            assert e < 1e-4, (b, key, e)
        if b:
            continue
        ref = refs[False]
        loss = lambda xx, dd=None: pose_loss(
            reference_pose(R, steps, xx, depth=dd), T_gt).item()
        x = torch.from_numpy(x0[b]).to(DEV)
        pick = np.random.default_rng(0).choice(
            np.flatnonzero(P["src_mask"].reshape(-1)), 64, replace=False)
        eps = 1e-4
        with torch.no_grad():
            fd = []
            for p in pick:
                dp, dm = R["depth"].clone(), R["depth"].clone()
                dp.view(-1)[p] += eps
                dm.view(-1)[p] -= eps
                fd.append((loss(x, dp) - loss(x, dm)) / (2 * eps))
            eye = torch.eye(6, dtype=torch.float64, device=DEV)
            fdx = {ex: [(loss(x + ex * e) - loss(x - ex * e)) / (2 * ex) for e in eye]
                   for ex in (1e-7, 1e-4)}
        e_d = rel(ref["depth"].reshape(-1)[pick], fd)
        e_x = rel(ref["x0"], fdx[1e-7])
        e_wide = rel(ref["x0"], fdx[1e-4])
        print(f"restatement vs central differences: depth (1e-4) {e_d:.2e}; "
              f"pose (1e-7) {e_x:.2e}; pose (1e-4) {e_wide:.2e}")
        assert e_d < 1e-4 and e_x < 1e-3
