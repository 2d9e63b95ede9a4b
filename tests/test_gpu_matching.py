"""GPU parity of the differentiable matcher and projection (FPCLMatcherOp,
ProjectOp) against the oracle, autograd checks of both, and AutogradICP on a
synthetic pair.

Bar: match masks, gathered points/normals and feature samples bit-exact;
backward gradients bit-exact (the kernels and the oracle perform the same
IEEE operations in the same order, contraction off); gradcheck at the
reference's own settings (eps 1e-6, atol 1e-4).
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from slamtb import synthetic
from slamtb._cslamtb import FPCLMatcherOp as _Op, ProjectOp, RigidTransformOp
from slamtb.camera import KCamera, Project, RigidTransform
from slamtb.frame import Frame, FrameInfo, FramePointCloud
from slamtb.registration import ICPOdometry
from slamtb.registration.autogradicp import AutogradICP
from slamtb.registration.correspondence_map import CorrespondenceMap
from slamtb.spatial.matching import FPCLMatcherOp, FramePointCloudMatcher

from _helpers import prepared_pair

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
NP = {torch.float32: np.float32, torch.float64: np.float64}


@pytest.fixture(scope="module")
def pair():
    k = synthetic.intrinsics(160, 120).astype(np.float32)
    return prepared_pair(21, 160, 120, kcam=k, channels=3)


def moved_source(P, T, dtype):
    """Source points and normals moved by T (computed in float64 on the
    host), valid pixels only."""
    m = P["src_mask"].reshape(-1)
    p = P["src_pts"].reshape(-1, 3)[m].astype(np.float64)
    n = P["src_nrm"].reshape(-1, 3)[m].astype(np.float64)
    p = p @ T[:3, :3].T + T[:3, 3]
    n = n @ T[:3, :3].T
    return p.astype(NP[dtype]), n.astype(NP[dtype])


def t(a, dtype=None):
    x = torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return x.to(dtype) if dtype is not None else x


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fpcl_forward_and_backward_bit_exact(pair, dtype):
    P = pair
    T = synthetic.random_se3(np.random.default_rng(2), 2.0, 0.02) @ P["T_gt"]
    sp, sn = moved_source(P, T, dtype)
    npdt = NP[dtype]
    tp, tn = P["tgt_pts"].astype(npdt), P["tgt_nrm"].astype(npdt)
    tf = P["tgt_feats"].astype(npdt)
    k = P["kcam"].astype(npdt)
    # The matcher's own map: identity pose, one-element mask.
    cmap = CorrespondenceMap(0.1, math.radians(30))(
        t(sp), t(sn), torch.ones(1, dtype=torch.bool, device=DEV),
        torch.eye(4, dtype=dtype, device=DEV), t(tp), t(tn), t(P["tgt_mask"]), t(k))
    ref_map = orc.corresp_map(sp, sn, np.ones(1, np.uint8), np.eye(4), tp, tn,
                              P["tgt_mask"], k, 0.1, math.radians(30))
    assert np.array_equal(cmap.cpu().numpy(), ref_map)

    n, C = sp.shape[0], tf.shape[0]
    op = torch.zeros(n, 3, dtype=dtype, device=DEV)
    on = torch.zeros(n, 3, dtype=dtype, device=DEV)
    of = torch.zeros(C, n, dtype=dtype, device=DEV)
    mask = torch.empty(n, dtype=torch.bool, device=DEV)
    _Op.forward(t(tp), t(tn), t(tf), t(sp), t(k), cmap, op, on, of, mask)
    r_op, r_on, r_of, r_mask = orc.fpcl_forward(tp, tn, tf, sp, k, ref_map)
    print(f"{dtype}: {int(r_mask.sum())} of {n} matched")
    assert r_mask.sum() > 0.5 * n
    assert np.array_equal(mask.cpu().numpy(), r_mask)
    assert np.array_equal(op.cpu().numpy(), r_op)
    assert np.array_equal(on.cpu().numpy(), r_on)
    assert np.array_equal(of.cpu().numpy(), r_of)

    dl = np.random.default_rng(4).normal(size=(C, n)).astype(npdt)
    dx = torch.empty(n, 3, dtype=dtype, device=DEV)
    _Op.backward(t(tf), t(sp), mask, t(dl), t(k), 0.005, dx)
    ref_dx = orc.fpcl_backward(tf, sp, r_mask, dl, k, 0.005)
    assert np.array_equal(dx.cpu().numpy(), ref_dx)
    assert np.abs(ref_dx).max() > 0


def test_fpcl_backward_rejects_large_steps(pair):
    tf = t(pair["tgt_feats"])
    sp = torch.zeros(4, 3, device=DEV)
    with pytest.raises(RuntimeError, match="grad_precision"):
        _Op.backward(tf, sp, torch.ones(4, dtype=torch.bool, device=DEV),
                     torch.zeros(tf.size(0), 4, device=DEV), t(pair["kcam"]), 0.5,
                     torch.empty(4, 3, device=DEV))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_project_bit_exact(dtype):
    rng = np.random.default_rng(6)
    npdt = NP[dtype]
    p = np.stack([rng.uniform(-1, 1, 5000), rng.uniform(-1, 1, 5000),
                  rng.uniform(0.3, 4, 5000)], 1).astype(npdt)
    k = synthetic.intrinsics(640, 480).astype(npdt)
    dy = rng.normal(size=(5000, 2)).astype(npdt)
    uv = ProjectOp.forward(t(p), t(k))
    assert np.array_equal(uv.cpu().numpy(), orc.project_forward(p, k))
    dx = ProjectOp.backward(t(dy), t(p), t(k))
    assert np.array_equal(dx.cpu().numpy(), orc.project_backward(dy, p, k))
    # A [2, 3] intrinsic matrix (the reference's own test uses one).
    uv2 = ProjectOp.forward(t(p), t(k[:2]))
    assert torch.equal(uv, uv2)


def test_project_gradcheck():
    """slamtb/_test/test_camera.py:29-38 on the HIP device."""
    torch.manual_seed(10)
    inputs = (torch.rand(3, dtype=torch.double, device=DEV, requires_grad=True),
              torch.tensor([[45.0, 0, 24], [0, 45, 24]], dtype=torch.double,
                           device=DEV))
    assert torch.autograd.gradcheck(Project.apply, inputs, eps=1e-6, atol=1e-4,
                                    raise_exception=True)


def test_fpcl_matcher_gradcheck(pair):
    """spatial/_test/test_fpcl_matcher.py:13-38 on the HIP device: a 10x10
    patch of target points nudged by up to 1 cm as the source. Target
    features are linear in (u, v) (so the central-difference backward is
    the exact slope); points whose projection lies within 2e-3 px of a
    rounding boundary are left out so that the eps-perturbed forward
    passes take the same match decisions."""
    P = pair
    H, W = P["tgt_mask"].shape
    v, u = np.mgrid[0:H, 0:W].astype(np.float64)
    feats = np.stack([0.02 * u - 0.01 * v + 1.0, 0.005 * u + 0.03 * v])
    kc = KCamera(torch.from_numpy(P["kcam"].astype(np.float64)))
    matcher = FramePointCloudMatcher(
        t(P["tgt_pts"], torch.float64), t(P["tgt_nrm"], torch.float64),
        t(P["tgt_mask"]), t(feats), kc)
    target = matcher.target
    gen = torch.Generator().manual_seed(10)
    patch = P["tgt_pts"][50:60, 70:80].reshape(-1, 3).astype(np.float64)
    patch = patch + torch.rand(patch.shape, generator=gen, dtype=torch.float64).numpy() * 0.01
    uv = orc.project_forward(patch, P["kcam"].astype(np.float64))
    frac = np.abs(uv - np.floor(uv) - 0.5)
    keep = (frac > 2e-3).all(1) & (patch[:, 2] > 0)
    patch = patch[keep]
    normals = P["tgt_nrm"][50:60, 70:80].reshape(-1, 3)[keep]
    src = t(patch).requires_grad_(True)
    match = FPCLMatcherOp.Match()
    inputs = (src, t(normals, torch.float64), target, match)
    assert torch.autograd.gradcheck(FPCLMatcherOp.apply, inputs, eps=1e-6,
                                    atol=1e-4, raise_exception=True)
    assert match.mask.sum() > 0.5 * len(patch)


def test_find_correspondences_and_rigid_transform(pair):
    P = pair
    T = torch.from_numpy(P["T_gt"]).to(DEV, torch.float32)
    m = torch.from_numpy(P["src_mask"]).to(DEV)
    sp = t(P["src_pts"])[m]
    sn = t(P["src_nrm"])[m]
    rig = RigidTransform(T)
    tsp = rig @ sp
    ref = (T[:3, :3] @ sp.T).T + T[:3, 3]
    torch.testing.assert_close(tsp, ref, rtol=1e-6, atol=1e-6)
    matcher = FramePointCloudMatcher(t(P["tgt_pts"]), t(P["tgt_nrm"]),
                                     t(P["tgt_mask"]), t(P["tgt_feats"]),
                                     KCamera(torch.from_numpy(P["kcam"])), 0.1,
                                     math.radians(30))
    pts, nrm, feats, mask = matcher.find_correspondences(
        tsp, rig.transform_normals(sn))
    assert pts.shape[0] == nrm.shape[0] == feats.shape[1] == int(mask.sum())
    # At the true motion the matched target points sit on the moved source
    # surface: point-to-plane residuals at the depth-noise level.
    res = ((pts - tsp[mask]) * nrm).sum(1).abs()
    print("matched", int(mask.sum()), "of", sp.shape[0], "median |r|",
          float(res.median()))
    assert mask.sum() > 0.5 * sp.shape[0]
    assert float(res.median()) < 1e-2


def pose_err(A, B):
    A, B = np.asarray(A, np.float64), np.asarray(B, np.float64)
    D = np.linalg.inv(B) @ A
    ang = math.degrees(math.acos(min(1.0, max(-1.0, (np.trace(D[:3, :3]) - 1) / 2))))
    return ang, float(np.linalg.norm(A[:3, 3] - B[:3, 3]))


@pytest.mark.parametrize("use_feats", [False, True])
def test_autogradicp_recovers_the_synthetic_motion(use_feats):
    k = synthetic.intrinsics(160, 120).astype(np.float32)
    P = prepared_pair(33, 160, 120, kcam=k, channels=3, noise=False)
    kc = KCamera(torch.from_numpy(k))
    kwargs = dict(geom_weight=1.0, feat_weight=0.0, distance_threshold=0.1,
                  normals_angle_thresh=math.pi / 4)
    if use_feats:
        kwargs.update(geom_weight=0.5, feat_weight=0.5, feat_residual_thresh=2.0)
    icp = AutogradICP(40, learning_rate=0.1, gradient_descent_iters=20, **kwargs)
    g = {k_: t(v) for k_, v in P.items() if isinstance(v, np.ndarray) and k_ != "kcam"}
    res = icp.estimate(kc, g["src_pts"], g["src_nrm"], g["src_mask"],
                       target_points=g["tgt_pts"], target_mask=g["tgt_mask"],
                       target_normals=g["tgt_nrm"],
                       target_feats=g["tgt_feats"] if use_feats else None,
                       source_feats=g["src_feats"] if use_feats else None)
    err = pose_err(res.transform.numpy(), P["T_gt"])
    start = pose_err(np.eye(4), P["T_gt"])
    gn = ICPOdometry(20, geom_weight=1.0, feat_weight=0.0, distance_threshold=0.1,
                     normals_angle_thresh=math.pi / 4).estimate(
        torch.from_numpy(k), g["src_pts"], g["src_nrm"], g["src_mask"],
        target_points=g["tgt_pts"], target_mask=g["tgt_mask"],
        target_normals=g["tgt_nrm"])
    print("AutogradICP", err, "start", start, "GN", pose_err(gn.transform.numpy(), P["T_gt"]))
    assert res.transform.shape == (4, 4)
    assert err[0] < 0.25 * start[0] and err[1] < 0.25 * start[1]


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_rigid_transform_kernels_bit_exact(dtype):
    rng = np.random.default_rng(12)
    npdt = NP[dtype]
    T = synthetic.random_se3(rng, 20.0, 0.5)
    T[:3, :3] *= 1.3  # not orthonormal: the normal matrix is a true inverse
    M = T.astype(npdt)
    p = rng.normal(size=(70001, 3)).astype(npdt)
    dy = rng.normal(size=(70001, 3)).astype(npdt)
    out = torch.empty(70001, 3, dtype=dtype, device=DEV)
    RigidTransformOp.transform(t(M), t(p), out)
    assert np.array_equal(out.cpu().numpy(), orc.rigid_points(M, p))
    RigidTransformOp.transform_normals(t(M), t(p), out)
    assert np.array_equal(out.cpu().numpy(), orc.rigid_points(M, p, normals=True))
    inplace = t(p)
    RigidTransformOp.transform_inplace(t(M[:3]), inplace)
    assert np.array_equal(inplace.cpu().numpy(), orc.rigid_points(M, p))
    dx = torch.empty_like(out)
    dM = torch.empty(3, 4, dtype=dtype, device=DEV)
    RigidTransformOp.transform_backward(t(M), t(p), t(dy), dx, dM)
    rdx, rdM = orc.rigid_backward(M, p, dy)
    assert np.array_equal(dx.cpu().numpy(), rdx)
    # FP64 sums of 70001 terms in another order: equal to the FP64
    # accumulation error (n eps sum|terms|), then rounded to dtype.
    bound = 70001 * np.finfo(np.float64).eps * np.abs(p).max() * np.abs(dy).max() * 4
    np.testing.assert_allclose(dM.cpu().numpy(), rdM,
                               rtol=2 * np.finfo(npdt).eps, atol=bound)


def test_rigid_transform_autograd():
    torch.manual_seed(3)
    M = torch.randn(4, 4, dtype=torch.float64, device=DEV, requires_grad=True)
    p = torch.randn(40, 3, dtype=torch.float64, device=DEV, requires_grad=True)
    assert torch.autograd.gradcheck(lambda m, q: RigidTransform(m) @ q, (M, p),
                                    eps=1e-6, atol=1e-8)
    M3 = torch.randn(3, 4, dtype=torch.float64, device=DEV, requires_grad=True)
    assert torch.autograd.gradcheck(lambda m, q: RigidTransform(m) @ q, (M3, p),
                                    eps=1e-6, atol=1e-8)
    # The reference's expression (camera.py:333-337) on the CPU.
    Mc, pc = M.detach().cpu(), p.detach().cpu()
    ref = (Mc[:3, :3].matmul(pc.view(-1, 3, 1)) + Mc[:3, 3].view(3, 1)).squeeze()
    torch.testing.assert_close((RigidTransform(M.detach()) @ p.detach()).cpu(), ref,
                               rtol=1e-12, atol=1e-12)


def test_autogradicp_frame_objective_and_gradient_match_the_oracle(pair):
    """The cost AutogradICP.estimate hands its optimiser (reference
    autogradicp.py:183-201 over the FPCL matcher of matching.py:29-125), at
    the true motion and at a perturbed pose: the matcher's rows against the
    oracle's correspondence map and FPCL forward on the same FP32 points,
    both cost terms against an FP64 restatement on those rows, and the
    geometric term's autograd gradient against FP64 central differences
    with the matches held."""
    from slamtb.registration.autogradicp import _row_loss
    from slamtb.registration.se3 import ExpRtToMatrix
    P = pair
    m = P["src_mask"].reshape(-1)
    sp = P["src_pts"].reshape(-1, 3)[m]
    sn = P["src_nrm"].reshape(-1, 3)[m]
    sf = P["src_feats"].reshape(P["src_feats"].shape[0], -1)[:, m]
    tp, tn, tf, k = P["tgt_pts"], P["tgt_nrm"], P["tgt_feats"], P["kcam"]
    athr = math.pi / 4
    icp = AutogradICP(10, geom_weight=1.0, feat_weight=0.5, distance_threshold=0.1,
                      normals_angle_thresh=athr, feat_residual_thresh=2.0,
                      huber_loss_alpha=0.3)
    matcher = FramePointCloudMatcher(t(tp), t(tn), t(P["tgt_mask"]), t(tf),
                                     KCamera(torch.from_numpy(k)), 0.1, athr)
    x_true = orc.matrix_to_exp_rt(P["T_gt"])[0]
    rng = np.random.default_rng(6)
    for x in (x_true, x_true + np.concatenate([rng.normal(0, 0.01, 3),
                                               rng.normal(0, 0.01, 3)])):
        xt = torch.tensor(x, dtype=torch.float32, device=DEV, requires_grad=True)
        rigid = RigidTransform(ExpRtToMatrix.apply(xt))
        tsrc = (rigid @ t(sp)).view(-1, 3)
        with torch.no_grad():
            tnrm = rigid.transform_normals(t(sn))
        match = matcher.match_rows(tsrc, tnrm)
        geom, feat = _row_loss(tsrc, match, t(sf), icp)
        geom.backward()
        grad = xt.grad.double().cpu().numpy()
        gv, fv = geom.item(), feat.item()
        ts32, tn32 = tsrc.detach().cpu().numpy(), tnrm.cpu().numpy()

        cmap = orc.corresp_map(ts32, tn32, np.ones(1, np.uint8), np.eye(4), tp, tn,
                               P["tgt_mask"], k, 0.1, athr)
        rp, rn, rf, rmask = orc.fpcl_forward(tp, tn, tf, ts32, k, cmap)
        assert np.array_equal(match.mask.cpu().numpy(), rmask)
        assert rmask.sum() > 0.5 * len(rmask)
        cost = (rn[rmask].astype(np.float64)
                * (rp[rmask].astype(np.float64) - ts32[rmask])).sum(1)
        og = float(np.mean(cost ** 2))
        d = np.linalg.norm(rf[:, rmask].astype(np.float64) - sf[:, rmask], axis=0)
        d = d[d ** 2 < 2.0 ** 2]
        lo, hi = d[d < 0.3], d[d >= 0.3]
        of = (lo.mean() if lo.size else 0.0) + ((0.3 * hi - 0.3).mean() if hi.size else 0.0)
        print(f"geom {gv:.9g} vs {og:.9g}; feat {fv:.9g} vs {of:.9g}")
        assert abs(gv - og) <= 1e-5 * og
        assert abs(fv - of) <= 1e-5 * abs(of)

        sel = np.nonzero(rmask)[0]
        q, nq = rp[sel].astype(np.float64), rn[sel].astype(np.float64)

        def g64(xx):
            M = orc.exp_rt_to_matrix(xx)
            s = sp[sel].astype(np.float64) @ M[:, :3].T + M[:, 3]
            return float(np.mean(((q - s) * nq).sum(1) ** 2))
        fd = np.empty(6)
        for i in range(6):
            e = np.zeros(6)
            e[i] = 1e-6
            fd[i] = (g64(x + e) - g64(x - e)) / 2e-6
        rel = np.linalg.norm(grad - fd) / np.linalg.norm(fd)
        print(f"  gradient {grad} vs {fd}: {rel:.2e}")
        # At the true motion the residuals are ~1e-3 m and the gradient
        # ~1e-4: the FP32 rounding of the transformed points (~1e-7 m) then
        # moves it by ~5e-4 of itself (seen 4.8e-4; 2e-6 at the other pose).
        assert rel < 3e-3
