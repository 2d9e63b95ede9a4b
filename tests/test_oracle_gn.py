"""Pins the oracle's Gauss-Newton driver (slamtb/registration/icp.py:152-262)
by convergence to the known motion of noise-free synthetic scenes, FP32 vs
FP64 agreement and the failed-factorisation path."""
import math

import numpy as np

from oracle import oracle as orc
from slamtb import synthetic

from _helpers import prepared_pair


def test_converges_to_the_ground_truth_without_noise():
    P = prepared_pair(3, 160, 120, noise=False)
    res = orc.icp_estimate(P["kcam"], P["src_pts"], P["src_nrm"], P["src_mask"],
                           P["tgt_pts"], P["tgt_nrm"], P["tgt_mask"], 25,
                           geom_weight=1, feat_weight=0, distance_threshold=0.1,
                           normals_angle_thresh=math.radians(30))
    start = orc.pose_error(np.eye(4), P["T_gt"])
    err = orc.pose_error(res["transform"], P["T_gt"])
    assert start[0] > 1e-2 or start[1] > 1e-2
    assert err[0] < 1e-3 and err[1] < 1e-3, err
    assert res["iters"] == 25 and 0 < res["match_ratio"] <= 1
    info = res["information"]
    assert np.allclose(info, info.T) and np.linalg.eigvalsh(info).min() > 0


def test_fp32_and_fp64_pixels_agree():
    P = prepared_pair(4, 160, 120)
    kw = dict(geom_weight=1, feat_weight=0, distance_threshold=0.1,
              normals_angle_thresh=math.radians(20))
    a = orc.icp_estimate(P["kcam"], P["src_pts"], P["src_nrm"], P["src_mask"],
                         P["tgt_pts"], P["tgt_nrm"], P["tgt_mask"], 10, **kw)
    b = orc.icp_estimate(P["kcam"], P["src_pts"], P["src_nrm"], P["src_mask"],
                         P["tgt_pts"], P["tgt_nrm"], P["tgt_mask"], 10,
                         dtype=np.float64, **kw)
    err = orc.pose_error(a["transform"], b["transform"])
    assert err[0] < 1e-4 and err[1] < 1e-4, err


def test_failed_factorisation_keeps_no_transform():
    P = prepared_pair(4, 80, 60)
    far = np.eye(4)
    far[0, 3] = 100.0
    res = orc.icp_estimate(P["kcam"], P["src_pts"], P["src_nrm"], P["src_mask"],
                           P["tgt_pts"], P["tgt_nrm"], P["tgt_mask"], 5,
                           transform=far, geom_weight=1, feat_weight=0)
    assert res["transform"] is None and res["iters"] == 0
    assert math.isinf(res["residual"])


def test_residual_and_weights_follow_the_reference_combination():
    # One iteration with both terms equals the explicit combination of the
    # per-term sums: JtJ = fw^2 F + gw^2 G, Jr = -fw f + gw g.
    P = prepared_pair(6, 80, 60)
    gw, fw = 2.0, 0.3
    res = orc.icp_estimate(P["kcam"], P["src_pts"], P["src_nrm"], P["src_mask"],
                           P["tgt_pts"], P["tgt_nrm"], P["tgt_mask"], 1,
                           src_feats=P["src_feats"], tgt_feats=P["tgt_feats"],
                           geom_weight=gw, feat_weight=fw, distance_threshold=0.2,
                           normals_angle_thresh=1.0, feat_residual_thresh=2.75)
    cmap = orc.corresp_map(P["src_pts"], P["src_nrm"], P["src_mask"], np.eye(4),
                           P["tgt_pts"], P["tgt_nrm"], P["tgt_mask"], P["kcam"],
                           0.2, 1.0)
    gj, gr, g2, gc = orc.geometric(P["src_pts"], P["src_mask"], np.eye(4),
                                   P["tgt_pts"], P["tgt_nrm"], P["kcam"], cmap)
    fj, fr, f2, fc = orc.feature(P["src_pts"], P["src_feats"], P["src_mask"],
                                 np.eye(4), P["tgt_feats"], P["kcam"], cmap, 2.75)
    JtJ = fw * fw * fj.sum(0, dtype=np.float64) + gw * gw * gj.sum(0, dtype=np.float64)
    np.testing.assert_allclose(res["jtj"], JtJ, rtol=1e-12)
    count = max(gc, fc)
    expect = (fw * f2.sum(dtype=np.float64) + gw * g2.sum(dtype=np.float64)) / count
    assert abs(res["residual"] - expect) <= 1e-12 * expect
