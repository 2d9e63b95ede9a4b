"""GPU parity of the point-cloud path (exact k nearest neighbours, the
nearest-neighbour feature layer and its backward, the surfel voxel merge)
against the oracle, the reference's KDTreeLayer gradcheck, and point-cloud
registration (AutogradICP.estimate_pcl, MultiscaleRegistration.estimate_pcl)
on a synthetic pair.

Bar: neighbour indices, squared distances, interpolated features, epsilon
weights, gradients and merged surfels bit-exact (same IEEE operations in the
same order); gradcheck at the reference's settings.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from slamtb import synthetic
from slamtb._cslamtb import NearestNeighborsOp, knn
from slamtb.camera import KCamera
from slamtb.frame import Frame, FrameInfo, FramePointCloud
from slamtb.registration import MultiscaleRegistration
from slamtb.registration.autogradicp import AutogradICP
from slamtb.spatial.kdtree_layer import KDTreeLayer
from slamtb.surfel import SurfelCloud

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
NP = {torch.float32: np.float32, torch.float64: np.float64}


def t(a, dtype=None):
    x = torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return x if dtype is None else x.to(dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("k,bound", [(1, math.inf), (3, 0.08), (8, math.inf), (16, 0.12),
                                     (8, 0.5), (4, 0.013)])
def test_knn_bit_exact(dtype, k, bound):
    rng = np.random.default_rng(k)
    npdt = NP[dtype]
    tg = rng.uniform(0, 1, size=(1000, 3)).astype(npdt)
    tg[100] = tg[200] = tg[300]
    q = rng.uniform(0, 1, size=(777, 3)).astype(npdt)
    q[5] = tg[100]
    rd, ri = orc.knn(tg, q, k, bound)
    # A finite bound searches the uniform grid by default; both searches
    # give the oracle's result.
    for method in ("auto", "brute"):
        d, i = knn(t(tg), t(q), k, bound, method=method)
        assert torch.equal(i.cpu(), torch.from_numpy(ri)), method
        assert np.array_equal(d.cpu().numpy(), rd), method
    assert ri[5, 0] == 100


@pytest.mark.parametrize("k,bound", [(1, 0.08), (8, 0.12), (16, 0.3)])
def test_knn_grid_one_lane_form_bit_exact(k, bound):
    """The one-lane grid search (taken above 2^20 queries per call) forced on
    a small call: the same indices and distances as the oracle's brute
    force, as the four-lane form gives."""
    from slamtb import _lib
    rng = np.random.default_rng(100 + k)
    tg = rng.uniform(0, 1, size=(3000, 3)).astype(np.float32)
    tg[7] = tg[8] = tg[9]
    q = rng.uniform(0, 1, size=(1500, 3)).astype(np.float32)
    q[3] = tg[8]
    rd, ri = orc.knn(tg, q, k, bound)
    lib = _lib.lib()
    prev = lib.stb_knn_grid_set_lane_threshold(0)
    try:
        d, i = knn(t(tg), t(q), k, bound)
    finally:
        lib.stb_knn_grid_set_lane_threshold(prev)
    assert torch.equal(i.cpu(), torch.from_numpy(ri))
    assert np.array_equal(d.cpu().numpy(), rd)
    d4, i4 = knn(t(tg), t(q), k, bound)
    assert torch.equal(i4.cpu(), i.cpu()) and torch.equal(d4.cpu(), d.cpu())


def test_kdtree_layer_sees_in_place_target_updates():
    """A target written in place between two queries is searched as it is
    now (the cached search grid is rebuilt): the neighbours equal the
    brute-force search over the new points."""
    rng = np.random.default_rng(5)
    target = t(rng.uniform(0, 1, size=(2000, 3)).astype(np.float32))
    src = t(rng.uniform(0, 1, size=(500, 3)).astype(np.float32))
    KDTreeLayer.setup(target)
    KDTreeLayer.query(src, 8, 0.1)
    first = KDTreeLayer.last_query[1].clone()
    with torch.no_grad():
        target.mul_(0.5).add_(0.25)  # same tensor, new contents
    KDTreeLayer.setup(target)
    KDTreeLayer.query(src, 8, 0.1)
    dist, index = KDTreeLayer.last_query
    rd, ri = orc.knn(target.cpu().numpy(), src.cpu().numpy(), 8, 0.1)
    assert torch.equal(index.cpu(), torch.from_numpy(ri))
    assert np.array_equal(dist.cpu().numpy(), rd)
    assert not torch.equal(first, index)
    # Without a new setup the query still sees a later in-place write.
    with torch.no_grad():
        target.add_(0.01)
    KDTreeLayer.query(src, 8, 0.1)
    rd, ri = orc.knn(target.cpu().numpy(), src.cpu().numpy(), 8, 0.1)
    assert torch.equal(KDTreeLayer.last_query[1].cpu(), torch.from_numpy(ri))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_nn_feature_ops_bit_exact(dtype):
    rng = np.random.default_rng(7)
    npdt = NP[dtype]
    tg = rng.uniform(-1, 1, size=(600, 3)).astype(npdt)
    src = rng.uniform(-1, 1, size=(300, 3)).astype(npdt)
    feats = rng.normal(size=(4, 600)).astype(npdt)
    dl = rng.normal(size=(4, 300)).astype(npdt)
    rd, ri = orc.knn(tg, src, 8, 0.25)  # some rows have missing neighbours
    assert (ri == 600).any()
    out = torch.empty(4, 300, dtype=dtype, device=DEV)
    NearestNeighborsOp.forward(t(rd), t(ri), t(feats), out)
    assert np.array_equal(out.cpu().numpy(), orc.nn_forward(rd, ri, feats))
    eps = torch.zeros(300, 6, 8, dtype=dtype, device=DEV)
    NearestNeighborsOp.compute_epsilon_distances(t(tg), t(src), t(ri), eps)
    reps = orc.nn_epsilon(tg, src, ri)  # slots past the neighbours stay 0 on both
    assert np.array_equal(eps.cpu().numpy(), reps)
    dx = torch.empty(300, 3, dtype=dtype, device=DEV)
    NearestNeighborsOp.backward(t(reps), t(ri), t(feats), t(dl), dx)
    assert np.array_equal(dx.cpu().numpy(), orc.nn_backward(reps, ri, feats, dl))


def test_kdtree_layer_gradcheck():
    """slamtb/spatial/_test/test_kdtree_layer.py:16-34 on the HIP device."""
    torch.manual_seed(10)
    target = torch.rand(100, 3, dtype=torch.double, device=DEV)
    features = torch.rand(5, 100, dtype=torch.double, device=DEV)
    layer = KDTreeLayer.setup(target)
    KDTreeLayer._gradcheck = True
    try:
        source = torch.rand(20, 3, dtype=torch.double, device=DEV, requires_grad=True)
        assert torch.autograd.gradcheck(layer, (features, source), eps=1e-6,
                                        atol=1e-4, raise_exception=True)
    finally:
        KDTreeLayer._gradcheck = False


def random_surfels(rng, n, C=3):
    pts = rng.uniform(-1, 1, size=(n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3)).astype(np.float32)
    col = rng.integers(0, 256, size=(n, 3)).astype(np.uint8)
    conf = rng.uniform(0, 1, n).astype(np.float32)
    rad = rng.uniform(0.001, 0.01, n).astype(np.float32)
    feats = rng.normal(size=(C, n)).astype(np.float32)
    return pts, nrm, col, conf, rad, feats


@pytest.mark.parametrize("voxel", [0.05, 0.3])
def test_voxel_merge_bit_exact(voxel):
    rng = np.random.default_rng(11)
    pts, nrm, col, conf, rad, feats = random_surfels(rng, 20000)
    cloud = SurfelCloud(t(pts), t(conf), t(nrm), t(rad), t(col),
                        torch.zeros(20000, dtype=torch.int32, device=DEV), t(feats))
    out = cloud.downsample(voxel)
    ref = orc.voxel_merge(pts, nrm, col, conf, rad, feats, voxel)
    print(f"voxel {voxel}: {out.size} surfels from 20000")
    assert out.size == len(ref["points"]) and out.size < 20000
    assert np.array_equal(out.points.cpu().numpy(), ref["points"])
    assert np.array_equal(out.normals.cpu().numpy(), ref["normals"])
    assert np.array_equal(out.colors.cpu().numpy(), ref["colors"])
    assert np.array_equal(out.confidences.cpu().numpy(), ref["confidences"])
    assert np.array_equal(out.radii.cpu().numpy(), ref["radii"])
    assert np.array_equal(out.features.cpu().numpy(), ref["features"])
    assert not out.times.any()


def pair_clouds(seed=3, W=160, H=120):
    k = synthetic.intrinsics(W, H).astype(np.float32)
    pair = synthetic.make_pair(seed, W, H, kcam=k, channels=3, noise=False,
                               max_angle_deg=3.0, max_trans=0.03)
    kc = KCamera(torch.from_numpy(k))
    clouds = []
    for side in ("src", "tgt"):
        d = torch.from_numpy(pair[f"{side}_depth"]).to(DEV)
        fp = FramePointCloud.from_frame(Frame(FrameInfo(kc), depth_image=d), device=DEV)
        clouds.append(SurfelCloud.from_frame_pcl(
            fp, features=torch.from_numpy(pair[f"{side}_feats"]).to(DEV)))
    return clouds[0], clouds[1], pair["T_gt"]


def pose_err(A, B):
    D = np.linalg.inv(np.asarray(B, np.float64)) @ np.asarray(A, np.float64)
    ang = math.acos(min(1.0, max(-1.0, (np.trace(D[:3, :3]) - 1) / 2)))
    return ang, float(np.linalg.norm(D[:3, 3]))


def test_autogradicp_estimate_pcl_recovers_the_motion():
    src, tgt, T_gt = pair_clouds()
    src, tgt = src.downsample(0.01), tgt.downsample(0.01)
    icp = AutogradICP(40, learning_rate=0.1, geom_weight=1.0, feat_weight=0.0,
                      distance_threshold=0.1, gradient_descent_iters=20)
    res = icp.estimate_pcl(src, tgt)
    err, start = pose_err(res.transform.numpy(), T_gt), pose_err(np.eye(4), T_gt)
    print(f"estimate_pcl: {src.size} / {tgt.size} surfels; error {err} from {start}")
    assert err[0] < 0.25 * start[0] and err[1] < 0.25 * start[1]


def test_multiscale_estimate_pcl():
    src, tgt, T_gt = pair_clouds(5)
    kw = dict(learning_rate=0.1, geom_weight=1.0, feat_weight=0.0,
              distance_threshold=0.1, gradient_descent_iters=10)
    ms = MultiscaleRegistration([(0.01, AutogradICP(20, **kw)),
                                 (0.03, AutogradICP(20, **kw))])
    res = ms.estimate_pcl(src, tgt)
    err, start = pose_err(res.transform.numpy(), T_gt), pose_err(np.eye(4), T_gt)
    print(f"multiscale estimate_pcl: error {err} from {start}")
    assert err[0] < 0.25 * start[0] and err[1] < 0.25 * start[1]


def test_knn_grid_edge_cases():
    """Targets on the bounding-box faces, queries outside the box, non-finite
    targets and a bound wider than the cloud: grid search equals brute
    force."""
    rng = np.random.default_rng(21)
    tg = rng.uniform(0, 1, size=(3000, 3)).astype(np.float32)
    tg[:8] = [[0, 0, 0], [1, 1, 1], [0, 1, 0], [1, 0, 1], [1, 1, 0], [0, 0, 1],
              [1, 0, 0], [0, 1, 1]]
    tg[8] = np.nan
    tg[9] = np.inf
    q = rng.uniform(-0.3, 1.3, size=(2000, 3)).astype(np.float32)
    q[:8] = tg[:8] + 0.01
    for bound in (0.05, 0.2, 5.0):
        d0, i0 = knn(t(tg), t(q), 8, bound, method="brute")
        d1, i1 = knn(t(tg), t(q), 8, bound)
        assert torch.equal(i0, i1) and torch.equal(d0, d1), bound
        assert not bool(((i1 == 8) | (i1 == 9)).any())


def _oracle_pcl_losses(tsrc, src_feats, tgt_pts, tgt_nrm, tgt_feats, bound,
                       feat_thresh, alpha):
    """FP64 restatement of the estimate_pcl cost (reference
    autogradicp.py:183-201 on the correspondences of matching.py:127-172):
    the nearest target within `bound` of each transformed source point (the
    oracle's exact k-nearest search, k = 8), point-to-plane cost against its
    point and normal, and the Huber cost of the distance between the source
    features and the target features interpolated from the 8 neighbours."""
    d, i = orc.knn(tgt_pts.astype(tsrc.dtype), tsrc, 8, bound)
    m = tgt_pts.shape[0]
    mask = i[:, 0] < m
    top1 = i[mask, 0]
    s = tsrc[mask].astype(np.float64)
    cost = (tgt_nrm[top1].astype(np.float64)
            * (tgt_pts[top1].astype(np.float64) - s)).sum(1)
    geom = float(np.mean(cost ** 2))
    f = orc.nn_forward(d.astype(np.float64), i, tgt_feats.astype(np.float64))
    fd = np.linalg.norm(f[:, mask] - src_feats[:, mask].astype(np.float64), axis=0)
    fd = fd[fd ** 2 < feat_thresh ** 2]
    lo, hi = fd[fd < alpha], fd[fd >= alpha]
    feat = (lo.mean() if lo.size else 0.0) + (
        (alpha * np.abs(hi) - alpha).mean() if hi.size else 0.0)
    return geom, feat, mask, top1


def test_estimate_pcl_objective_and_gradient_match_the_oracle():
    """The cost that AutogradICP.estimate_pcl hands its optimiser, at the
    identity and at a perturbed pose: the transformed source points against
    the oracle's FP64 rigid transform, the matches and both cost terms
    against the FP64 restatement above (its neighbour search run on the same
    FP32 points, so that the decisions are shared), and the autograd
    gradient of the geometric term against FP64 central differences of the
    restated cost with the matches held."""
    from slamtb.camera import RigidTransform
    from slamtb.registration.autogradicp import _compacted_loss, _HuberLoss
    from slamtb.registration.se3 import ExpRtToMatrix
    from slamtb.spatial.matching import PointCloudMatcher
    src, tgt, _ = pair_clouds()
    src, tgt = src.downsample(0.01), tgt.downsample(0.01)
    icp = AutogradICP(10, geom_weight=1.0, feat_weight=0.5, distance_threshold=0.1,
                      feat_residual_thresh=2.75, huber_loss_alpha=0.3)
    sp = src.points.cpu().numpy()
    tp, tn = tgt.points.cpu().numpy(), tgt.normals.cpu().numpy()
    sf, tf = src.features.cpu().numpy(), tgt.features.cpu().numpy()
    rng = np.random.default_rng(8)
    for x in (np.zeros(6), np.concatenate([rng.normal(0, 0.01, 3),
                                           rng.normal(0, 0.01, 3)])):
        xt = torch.tensor(x, dtype=torch.float32, device=DEV, requires_grad=True)
        rigid = RigidTransform(ExpRtToMatrix.apply(xt))
        tsrc = (rigid @ src.points).view(-1, 3)
        with torch.no_grad():
            tnormals = rigid.transform_normals(src.normals)
        matcher = PointCloudMatcher(tgt.points, tgt.normals, tgt.features,
                                    num_neighbors=8, distance_upper_bound=0.1)
        found = matcher.find_correspondences(tsrc, tnormals)
        geom, feat = _compacted_loss(tsrc, *found, src.features, icp,
                                     _HuberLoss(icp.huber_loss_alpha))
        geom.backward()
        grad = xt.grad.double().cpu().numpy()
        geom, feat = geom.detach(), feat.detach()
        ts32 = tsrc.detach().cpu().numpy()

        M = orc.exp_rt_to_matrix(x.astype(np.float64))
        ts64 = sp.astype(np.float64) @ M[:, :3].T + M[:, 3]
        assert np.abs(ts32 - ts64).max() <= 1e-5 * max(1.0, np.abs(ts64).max())
        og, of, mask, top1 = _oracle_pcl_losses(ts32, sf, tp, tn, tf, 0.1, 2.75, 0.3)
        assert np.array_equal(found[3].cpu().numpy(), mask)
        assert mask.sum() > 0.5 * len(mask)
        gv, fv = geom.item(), feat.item()
        print(f"pose {x}: geom {gv:.9g} vs {og:.9g}, feat {fv:.9g} vs {of:.9g}")
        assert abs(gv - og) <= 1e-5 * og
        # (the Huber part alpha |d| - alpha is negative for d < 1)
        assert abs(fv - of) <= 1e-5 * abs(of)

        def g64(xx):
            Mx = orc.exp_rt_to_matrix(xx)
            s = sp[mask].astype(np.float64) @ Mx[:, :3].T + Mx[:, 3]
            c = (tn[top1].astype(np.float64) * (tp[top1].astype(np.float64) - s)).sum(1)
            return float(np.mean(c ** 2))
        fd = np.empty(6)
        for k in range(6):
            e = np.zeros(6)
            e[k] = 1e-6
            fd[k] = (g64(x + e) - g64(x - e)) / 2e-6
        rel = np.linalg.norm(grad - fd) / np.linalg.norm(fd)
        print(f"  gradient {grad} vs {fd}: {rel:.2e}")
        assert rel < 1e-3
