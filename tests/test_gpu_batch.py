"""Batched multiscale ICP over pairs of uneven resolution: every pair's
result must match the single-pair fused path (and therefore the oracle) on
the same inputs."""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from slamtb import processing, synthetic
from slamtb.camera import KCamera
from slamtb.frame import Frame, FrameInfo, FramePointCloud
from slamtb.registration import ICPOdometry, MultiscaleRegistration
from slamtb.registration.batch import BatchedMultiscaleICP

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")

KW = dict(geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
          normals_angle_thresh=math.radians(20), feat_residual_thresh=2.75)


def registration():
    return MultiscaleRegistration([(1.0, ICPOdometry(6, **KW)),
                                   (0.5, ICPOdometry(4, **KW)),
                                   (0.5, ICPOdometry(3, **KW))])


def make_pairs(sizes, seed0=40, storage=None, channels=3):
    ms = registration()
    pairs, truth = [], []
    for i, (W, H) in enumerate(sizes):
        k = synthetic.intrinsics(W, H)
        pr = synthetic.make_pair(seed0 + i, W, H, kcam=k, channels=channels,
                                 descriptor=channels > 3)
        kcam = KCamera(torch.from_numpy(k))
        fps = []
        for side in ("src", "tgt"):
            d = torch.from_numpy(pr[f"{side}_depth"]).to(DEV)
            d = processing.bilateral_depth_filter(d, None, None, 6, 0.03, 4.5, 1.0)
            fps.append(FramePointCloud.from_frame(Frame(FrameInfo(kcam), depth_image=d),
                                                  device=DEV))
        feats = [torch.from_numpy(pr["src_feats"]).to(DEV),
                 torch.from_numpy(pr["tgt_feats"]).to(DEV)]
        pairs.append(ms.levels(fps[0], fps[1], feats[0], feats[1], storage=storage))
        truth.append(pr["T_gt"])
    return ms, pairs, truth


def single(ms, levels):
    from slamtb.registration._fused import runner_for
    from slamtb.registration.icp import result_from_state
    run = runner_for(DEV)
    for _, L in levels:
        run.reserve(L.n, L.height, L.width)
    ms.enqueue_levels(run, levels)
    return result_from_state(run.read_state(), levels[-1][1].n, True)


@pytest.mark.parametrize("tile_px", [256, 1024])
def test_batch_matches_single_pair_path(tile_px):
    sizes = [(160, 120), (320, 240), (80, 60), (320, 240), (160, 120), (240, 180)]
    ms, pairs, truth = make_pairs(sizes)
    batch = BatchedMultiscaleICP(ms.estimators, pairs, DEV, tile_px=tile_px)
    results = batch.estimate()
    for i, levels in enumerate(pairs):
        ref = single(ms, levels)
        err = orc.pose_error(results[i].transform.numpy(), ref.transform.numpy())
        print(sizes[i], "batched vs single", err, "vs truth",
              orc.pose_error(results[i].transform.numpy(), truth[i]))
        assert err[0] < 1e-6 and err[1] < 1e-6, (i, err)
        assert results[i].iterations == ref.iterations
        assert abs(results[i].match_ratio - ref.match_ratio) < 1e-3
        info_rel = float(np.linalg.norm(results[i].information - ref.information) /
                         np.linalg.norm(ref.information))
        assert info_rel < 1e-4


def test_batch_graph_replay_is_repeatable():
    sizes = [(160, 120), (80, 60), (320, 240)]
    ms, pairs, _ = make_pairs(sizes, seed0=60)
    batch = BatchedMultiscaleICP(ms.estimators, pairs, DEV)
    first = [r.transform.numpy() for r in batch.estimate()]
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            batch.enqueue()
    torch.cuda.current_stream().wait_stream(side)
    for _ in range(3):
        g.replay()
        torch.cuda.synchronize()
        again = [r.transform.numpy() for r in batch.results()]
        for a, b in zip(first, again):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("storage,channels", [(torch.bfloat16, 3), (None, 16),
                                              (torch.float16, 16)])
def test_batch_storage_and_channel_widths(storage, channels):
    """16-bit storage and wide descriptors (feature batches of four) through
    the batched kernels: equal to the single-pair path on the same inputs."""
    sizes = [(160, 120), (240, 180), (80, 60)]
    ms, pairs, _ = make_pairs(sizes, seed0=80, storage=storage, channels=channels)
    results = BatchedMultiscaleICP(ms.estimators, pairs, DEV).estimate()
    for i, levels in enumerate(pairs):
        ref = single(ms, levels)
        err = orc.pose_error(results[i].transform.numpy(), ref.transform.numpy())
        assert err[0] < 1e-6 and err[1] < 1e-6, (i, err)
        assert results[i].iterations == ref.iterations
