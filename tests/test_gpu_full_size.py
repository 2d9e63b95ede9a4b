"""The hot path at its full workload sizes, and at edge shapes.

* 640x480, 3 levels (10/5/3 iterations), geometric + 3-channel features,
  bilateral-filtered depth: pose within 1e-6 of the oracle fed the same level
  inputs.
* 1280x720 with 32-channel descriptors in BF16 and in FP16 storage (top
  level, 3 iterations): pose within 1e-6 of the oracle on the same
  (rounded) values.
* 1280x720 fused linearisation against the drop-in per-pixel partials summed
  in FP64 on the host (the drop-in partials are pinned bit-exact to the
  oracle at 320x240): 1e-5 relative. No oracle run at this size.
* 3840x2160 self-registration, a size-independent property: a frame
  registered to itself from the identity meets every valid pixel at its own
  position with a zero point-to-plane residual, so the map is the identity
  permutation, every Gauss-Newton update is exactly zero and the pose stays
  the identity bit for bit.
* Ragged and tiny levels (161x97, 8x6) against the oracle, and an
  all-invalid source mask (no match: the loop's `break`).
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from slamtb import synthetic
from slamtb._cslamtb import ICPJacobian
from slamtb.camera import KCamera
from slamtb.frame import Frame, FrameInfo, FramePointCloud
from slamtb.processing import bilateral_depth_filter
from slamtb.registration import ICPOdometry, MultiscaleRegistration
from slamtb.registration._fused import Level, make_params, runner_for
from slamtb.registration.correspondence_map import CorrespondenceMap

from _helpers import prepared_pair, to_dev

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
KW = dict(geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
          normals_angle_thresh=math.radians(20), feat_residual_thresh=2.75)


def gpu_levels(ms, pair, storage=None, bilateral=True):
    kcam = KCamera(torch.from_numpy(pair["kcam"]))
    fps = []
    for side in ("src", "tgt"):
        d = torch.from_numpy(pair[f"{side}_depth"]).to(DEV)
        if bilateral:
            d = bilateral_depth_filter(d, d > 0, None, 6, 0.03, 4.5, 1.0)
        fps.append(FramePointCloud.from_frame(Frame(FrameInfo(kcam), depth_image=d),
                                              device=DEV))
    feats = [torch.from_numpy(pair[f"{s}_feats"]).to(DEV) for s in ("src", "tgt")]
    return ms.levels(fps[0], fps[1], feats[0], feats[1], storage=storage)


def oracle_chain(levels):
    T = np.eye(4)
    c = lambda t: t.detach().float().cpu().numpy()
    for est, L in levels:
        k = np.array([[L.fx, 0, L.cx], [0, L.fy, L.cy], [0, 0, 1]], np.float32)
        r = orc.icp_estimate(k, c(L.src_points), c(L.src_normals),
                             L.src_mask.cpu().numpy(), c(L.tgt_points),
                             c(L.tgt_normals), L.tgt_mask.cpu().numpy(),
                             est.num_iters, src_feats=c(L.src_feats),
                             tgt_feats=c(L.tgt_feats), transform=T,
                             with_information=False, **KW)
        T = r["transform"] if r["transform"] is not None else np.eye(4)
    return T


def run_levels(levels):
    from slamtb.registration.icp import result_from_state
    run = runner_for(DEV)
    for _, L in levels:
        run.reserve(L.n, L.height, L.width)
    MultiscaleRegistration.enqueue_levels(run, levels)
    return result_from_state(run.read_state(), levels[-1][1].n, True)


def test_640x480_three_levels_match_the_oracle():
    k = synthetic.intrinsics(640, 480).astype(np.float32)
    pair = synthetic.make_pair(2, 640, 480, kcam=k, channels=3)
    ms = MultiscaleRegistration([(1.0, ICPOdometry(10, **KW)),
                                 (0.5, ICPOdometry(5, **KW)),
                                 (0.5, ICPOdometry(3, **KW))])
    levels = gpu_levels(ms, pair)
    res = run_levels(levels)
    err = orc.pose_error(res.transform.numpy(), oracle_chain(levels))
    print("640x480 3-level pose error vs oracle", err)
    assert err[0] < 1e-6 and err[1] < 1e-6
    assert res.iterations == 10


@pytest.mark.parametrize("storage", [torch.bfloat16, torch.float16])
def test_1280x720_16bit_descriptors_match_the_oracle(storage):
    pair = synthetic.make_pair(502, 1280, 720, kcam=synthetic.intrinsics(1280, 720),
                               channels=32, descriptor=True)
    ms = MultiscaleRegistration([(1.0, ICPOdometry(3, **KW))])
    levels = gpu_levels(ms, pair, storage=storage)
    assert levels[0][1].src_points.dtype == storage
    res = run_levels(levels)
    err = orc.pose_error(res.transform.numpy(), oracle_chain(levels))
    print(f"1280x720 {storage} C=32 pose error vs oracle", err)
    assert err[0] < 1e-6 and err[1] < 1e-6


def test_1280x720_fused_sums_match_dropin_partials():
    # Bilateral-filtered depth (raw noise at this resolution makes the
    # normals too rough for the 20 degree gate), near the true motion, so
    # that most pixels match.
    P = prepared_pair(44, 1280, 720, channels=3, bilateral=True)
    T = synthetic.random_se3(np.random.default_rng(4), 0.2, 0.002) @ P["T_gt"]
    gw, fw = 10.0, 1.0
    g = to_dev(P, DEV)
    level = Level(P["kcam"], g["src_pts"], g["src_nrm"], g["src_mask"],
                  g["tgt_pts"], g["tgt_nrm"], g["tgt_mask"], g["src_feats"],
                  g["tgt_feats"])
    params = make_params(gw, fw, 0.1, math.radians(20), 2.75, True)
    run = runner_for(DEV)
    run.reserve(level.n, level.height, level.width)
    run.begin(level, torch.from_numpy(T))
    run.linearize(level, params)
    st = run.read_state()
    rt = torch.from_numpy(T).to(DEV, torch.float32)
    k = torch.from_numpy(P["kcam"]).to(DEV)
    cmap = CorrespondenceMap(0.1, math.radians(20))(
        g["src_pts"].view(-1, 3), g["src_nrm"].view(-1, 3), g["src_mask"].view(-1),
        rt, g["tgt_pts"], g["tgt_nrm"], g["tgt_mask"], k)
    n = level.n
    jtj = torch.empty(n, 6, 6, device=DEV)
    jr = torch.empty(n, 6, device=DEV)
    r2 = torch.empty(n, device=DEV)
    cg = ICPJacobian.estimate_geometric(g["src_pts"].view(-1, 3), g["src_mask"].view(-1),
                                        rt, g["tgt_pts"], g["tgt_nrm"], k, cmap,
                                        jtj, jr, r2)
    JtJ = jtj.double().sum(0) * gw * gw
    Jr = jr.double().sum(0) * gw
    cf = ICPJacobian.estimate_feature(g["src_pts"].view(-1, 3),
                                      g["src_feats"].view(3, -1), g["src_mask"].view(-1),
                                      rt, g["tgt_feats"], k, cmap, 2.75, jtj, jr, r2)
    JtJ = (JtJ + jtj.double().sum(0) * fw * fw).cpu().numpy()
    Jr = (Jr - jr.double().sum(0) * fw).cpu().numpy()
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    got_jtj = np.array(st.last_jtj).reshape(6, 6)
    got_jr = np.array(st.last_jr)
    print("1280x720 fused vs drop-in:", rel(got_jtj, JtJ), rel(got_jr, Jr), cg, cf)
    assert st.last_geom_count == cg and st.last_feat_count == cf
    assert cg > 0.5 * n
    assert rel(got_jtj, JtJ) < 1e-5 and rel(got_jr, Jr) < 1e-5


def test_3840x2160_self_registration_is_exact():
    W, H = 3840, 2160
    k = synthetic.intrinsics(W, H).astype(np.float32)
    pair = synthetic.make_pair(9, W, H, kcam=k, channels=1)
    kcam = KCamera(torch.from_numpy(k))
    d = torch.from_numpy(pair["src_depth"]).to(DEV)
    fp = FramePointCloud.from_frame(Frame(FrameInfo(kcam), depth_image=d), device=DEV)
    icp = ICPOdometry(3, geom_weight=1.0, feat_weight=0.0, distance_threshold=0.1,
                      normals_angle_thresh=math.radians(20))
    res = icp.estimate(torch.from_numpy(k), fp.points, fp.normals, fp.mask,
                       target_points=fp.points, target_mask=fp.mask,
                       target_normals=fp.normals)
    valid = int(fp.mask.sum().item())
    assert valid > 0.5 * W * H
    assert res.iterations == 3
    assert torch.equal(res.transform, torch.eye(4, dtype=torch.float64))
    assert round(res.match_ratio * W * H) == valid
    assert res.residual == 0.0
    # The map itself: every valid pixel owns its own position.
    cmap = CorrespondenceMap(0.1, math.radians(20))(
        fp.points.view(-1, 3), fp.normals.view(-1, 3), fp.mask.view(-1),
        torch.eye(4, device=DEV), fp.points, fp.normals, fp.mask,
        torch.from_numpy(k).to(DEV)).view(-1)
    owner = cmap >> 32
    pos = torch.arange(W * H, device=DEV)
    m = fp.mask.view(-1)
    assert torch.equal(owner[m], pos[m])
    assert bool((cmap[~m] == 0x0FFFFFFF7F800000).all())


@pytest.mark.parametrize("W,H", [(161, 97), (8, 6)])
def test_ragged_and_tiny_levels_match_the_oracle(W, H):
    P = prepared_pair(31, W, H, channels=3)
    g = to_dev(P, DEV)
    icp = ICPOdometry(4, **KW)
    res = icp.estimate(torch.from_numpy(P["kcam"]), g["src_pts"], g["src_nrm"],
                       g["src_mask"], source_feats=g["src_feats"],
                       target_points=g["tgt_pts"], target_mask=g["tgt_mask"],
                       target_normals=g["tgt_nrm"], target_feats=g["tgt_feats"])
    ref = orc.icp_estimate(P["kcam"], P["src_pts"], P["src_nrm"], P["src_mask"],
                           P["tgt_pts"], P["tgt_nrm"], P["tgt_mask"], 4,
                           src_feats=P["src_feats"], tgt_feats=P["tgt_feats"], **KW)
    assert res.iterations == ref["iters"]
    if ref["transform"] is None:
        assert res.transform is None
    else:
        err = orc.pose_error(res.transform.numpy(), ref["transform"])
        assert err[0] < 1e-6 and err[1] < 1e-6, err


def test_empty_source_mask_stops_at_the_first_iteration():
    P = prepared_pair(12, 80, 60, channels=3)
    Q = dict(P)
    Q["src_mask"] = np.zeros_like(P["src_mask"])
    g = to_dev(Q, DEV)
    icp = ICPOdometry(3, geom_weight=1.0, feat_weight=0.0, distance_threshold=0.1,
                      normals_angle_thresh=math.radians(20))
    res = icp.estimate(torch.from_numpy(P["kcam"]), g["src_pts"], g["src_nrm"],
                       g["src_mask"], target_points=g["tgt_pts"],
                       target_mask=g["tgt_mask"], target_normals=g["tgt_nrm"])
    ref = orc.icp_estimate(P["kcam"], Q["src_pts"], Q["src_nrm"], Q["src_mask"],
                           Q["tgt_pts"], Q["tgt_nrm"], Q["tgt_mask"], 3,
                           geom_weight=1.0, feat_weight=0.0, distance_threshold=0.1,
                           normals_angle_thresh=math.radians(20))
    # No match: JtJ is exactly zero and the factorisation fails at once (the
    # loop's `break`) on both sides.
    assert ref["iters"] == 0 and ref["transform"] is None
    assert res.iterations == 0 and res.transform is None
