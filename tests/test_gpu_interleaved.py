"""Channel-interleaved features (stb_icp_level.feat_layout =
STB_FEAT_INTERLEAVED: source [N, C], target [H, W, C]) against the
reference's channel-planar layout on the same values.

The interleaved Jacobian reads the same texels into the same arithmetic in
the same order, so the whole device state after a multiscale registration --
pose, last system, residuals, counts -- must be bit-identical to the planar
run's; the planar run is itself checked against the oracle by the other GPU
tests (BF16 and FP16 storage at 1280x720 in test_gpu_full_size.py). Sizes
and channel counts cover a ragged width (the right-edge zero padding and
texels past the last row), a channel count that fills no whole 16-byte
chunk (6) and one that fills four (32), and BF16 and FP16 storage; the
batched kernels are checked the same way.
"""
import ctypes
import math

import numpy as np
import pytest
import torch

from slamtb import _lib, processing, synthetic
from slamtb.camera import KCamera
from slamtb.frame import Frame, FrameInfo, FramePointCloud
from slamtb.registration import ICPOdometry, MultiscaleRegistration
from slamtb.registration._fused import GNRunner
from slamtb.registration.batch import BatchedMultiscaleICP

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
KW = dict(geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
          normals_angle_thresh=math.radians(20), feat_residual_thresh=2.75)


def registration():
    return MultiscaleRegistration([(1.0, ICPOdometry(8, **KW)),
                                   (0.5, ICPOdometry(5, **KW)),
                                   (0.5, ICPOdometry(3, **KW))])


def frames(seed, W, H, channels):
    k = synthetic.intrinsics(W, H)
    pr = synthetic.make_pair(seed, W, H, kcam=k, channels=channels,
                             descriptor=channels > 3)
    kcam = KCamera(torch.from_numpy(k))
    fps = []
    for side in ("src", "tgt"):
        d = torch.from_numpy(pr[f"{side}_depth"]).to(DEV)
        d = processing.bilateral_depth_filter(d, d > 0, None, 6, 0.03, 4.5, 1.0)
        fps.append(FramePointCloud.from_frame(Frame(FrameInfo(kcam), depth_image=d),
                                              device=DEV))
    feats = [torch.from_numpy(pr[f"{s}_feats"]).to(DEV) for s in ("src", "tgt")]
    return fps, feats


def state_bytes(levels):
    run = GNRunner(DEV)
    for _, L in levels:
        run.reserve(L.n, L.height, L.width)
    MultiscaleRegistration.enqueue_levels(run, levels)
    st = run.read_state()
    return st, bytes(st)


@pytest.mark.parametrize("W,H,channels,storage", [
    (161, 97, 6, torch.bfloat16),
    (161, 97, 32, torch.bfloat16),
    (320, 240, 32, torch.float16),
    (160, 120, 2, torch.float16),
])
def test_interleaved_equals_planar_bit_for_bit(W, H, channels, storage):
    # Fresh frames for each layout: a frame keeps the pyramid normals it
    # computed once, so a second pyramid of the same frame objects would be
    # other inputs (frame.py:404-408 of the reference).
    ms = registration()
    fps, feats = frames(11 + channels, W, H, channels)
    planar = ms.levels(fps[0], fps[1], feats[0], feats[1], storage=storage)
    fps, feats = frames(11 + channels, W, H, channels)
    inter = ms.levels(fps[0], fps[1], feats[0], feats[1], storage=storage,
                      feat_layout="interleaved")
    L = inter[-1][1]
    assert L.feat_layout == "interleaved" and L.struct.feat_layout == _lib.STB_FEAT_INTERLEAVED
    assert tuple(L.tgt_feats.shape) == (H, W, channels)
    assert tuple(L.src_feats.shape) == (H * W, channels)
    sp, bp = state_bytes(planar)
    si, bi = state_bytes(inter)
    print((W, H, channels, storage), "matches", sp.last_geom_count,
          sp.last_feat_count, "iterations", sp.iterations)
    assert sp.last_feat_count > 0 and sp.iterations > 0
    assert bi == bp


def test_batched_interleaved_equals_batched_planar():
    ms = registration()
    sizes = [(161, 97), (320, 240), (80, 60)]
    planar, inter = [], []
    for i, (W, H) in enumerate(sizes):
        fps, feats = frames(70 + i, W, H, 8)
        planar.append(ms.levels(fps[0], fps[1], feats[0], feats[1],
                                storage=torch.bfloat16))
        fps, feats = frames(70 + i, W, H, 8)
        inter.append(ms.levels(fps[0], fps[1], feats[0], feats[1],
                               storage=torch.bfloat16, feat_layout="interleaved"))
    rp = BatchedMultiscaleICP(ms.estimators, planar, DEV).estimate()
    ri = BatchedMultiscaleICP(ms.estimators, inter, DEV).estimate()
    for a, b in zip(rp, ri):
        assert np.array_equal(a.transform.numpy(), b.transform.numpy())
        assert a.iterations == b.iterations and a.match_ratio == b.match_ratio
        assert np.array_equal(a.information, b.information)


def test_interleaved_fp32_is_refused():
    ms = registration()
    fps, feats = frames(5, 80, 60, 4)
    levels = ms.levels(fps[0], fps[1], feats[0], feats[1],
                       feat_layout="interleaved")
    run = GNRunner(DEV)
    for _, L in levels:
        run.reserve(L.n, L.height, L.width)
    with pytest.raises(RuntimeError, match="interleaved"):
        MultiscaleRegistration.enqueue_levels(run, levels)
