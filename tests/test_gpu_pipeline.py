"""The tracking front end (SurfelSLAM.step without fusion) on the reference's
sample1 frames: the captured-graph step against the eager step (the same
kernels on the same inputs: poses equal), accuracy against the recorded
camera poses, and the per-frame time of both."""
import math
from pathlib import Path

import pytest
import torch

from slamtb import metrics
from slamtb.camera import RTCamera
from slamtb.data import load_ftb
from slamtb.pipeline import SurfelSLAM

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).parent / "golden" / "ftb"


def run(use_graph, frames):
    slam = SurfelSLAM(device="cuda:0", use_graph=use_graph)
    stats = [slam.step(f) for f in frames]
    return stats


def test_graph_step_equals_eager_step_and_tracks_sample1():
    ds = load_ftb(GOLDEN / "sample1")
    frames = [ds[i] for i in range(len(ds))]
    eager = run(False, frames)
    graph = run(True, frames)
    assert eager[0].result is None and graph[0].result is None
    for e, g in zip(eager[1:], graph[1:]):
        assert g.result.iterations == e.result.iterations
        assert torch.allclose(g.rt_camera.matrix, e.rt_camera.matrix, rtol=0,
                              atol=1e-12)
        assert g.accepted == e.accepted
    gt = metrics.set_start_at_identity({i: f.info.rt_cam for i, f in enumerate(frames)})
    pred = {i: s.rt_camera for i, s in enumerate(graph)}
    rmse = lambda x: x.pow(2).mean().sqrt().item()
    ate = rmse(metrics.absolute_translational_error(gt, pred))
    rte = rmse(metrics.relative_translational_error(gt, pred))
    t_e = sorted(s.ms for s in eager[2:])[len(eager[2:]) // 2]
    t_g = sorted(s.ms for s in graph[2:])[len(graph[2:]) // 2]
    print(f"tracking sample1 0..14: ATE {ate:.5f} m RTE {rte:.5f} m; median ms/frame "
          f"eager {t_e:.3f} graph {t_g:.3f}; accepted "
          f"{sum(s.accepted for s in graph[1:])}/{len(graph) - 1}")
    assert ate < 0.02 and rte < 0.004


def test_tracking_steps_match_the_oracle_chain_per_frame():
    """Every frame of sample1 tracked by the captured step against the oracle
    run on the same inputs, composed the same way (surfel_slam.py:111-129):
    each frame's relative pose is the oracle's 3-level registration (10/10/10
    iterations, geometry 10 : features 1 without features, the bilateral
    filter with the frame's depth scale, carried full-resolution normals)
    of that frame onto the previous one, fed the level arrays the GPU
    preprocessing built (bit-identical to the oracle's own preprocessing),
    to 1e-6 rad / 1e-6 m; the camera pose is the product of the relative
    poses, and the verifier's verdict is the one it gives the oracle's
    residual and match ratio with the GPU's information matrix."""
    import numpy as np
    from oracle import oracle as orc
    from slamtb.frame import Frame, FramePointCloud
    from slamtb.processing import bilateral_depth_filter
    from slamtb.registration import RegistrationVerifier

    from _helpers import level_args, level_kwargs

    ds = load_ftb(GOLDEN / "sample1")
    frames = [ds[i] for i in range(len(ds))]
    slam = SurfelSLAM(device="cuda:0", use_graph=True)
    stats = [slam.step(f) for f in frames]

    def fpcl(f):
        d = torch.as_tensor(f.depth_image).to("cuda:0")
        d = bilateral_depth_filter(d, d > 0, depth_scale=f.info.depth_scale)
        p = FramePointCloud.from_frame(Frame(f.info, d))
        p.normals
        return p

    clouds = [fpcl(f) for f in frames]
    verifier = RegistrationVerifier()
    pose = np.eye(4)
    worst = (0.0, 0.0)
    for i in range(1, len(frames)):
        levels = slam.icp.levels(clouds[i], clouds[i - 1])
        T = np.eye(4)
        ref = None
        for est, L in levels:
            ref = orc.icp_estimate(*level_args(L), est.num_iters, transform=T,
                                   **level_kwargs(est, L))
            T = ref["transform"] if ref["transform"] is not None else np.eye(4)
        rel = stats[i].result.transform.numpy()
        err = orc.pose_error(T, rel)
        worst = (max(worst[0], err[0]), max(worst[1], err[1]))
        assert err[0] < 1e-6 and err[1] < 1e-6, (i, err)
        pose = pose @ T
        cam = stats[i].rt_camera.matrix.numpy()
        cerr = orc.pose_error(pose, cam)
        assert cerr[0] < 1e-5 and cerr[1] < 1e-5, (i, cerr)
        n = levels[-1][1].n
        mirrored = type(stats[i].result)(torch.from_numpy(T), stats[i].result.information,
                                         ref["residual"], ref["match_count"] / n)
        assert verifier(mirrored) == stats[i].accepted, i
    print("worst per-frame relative pose difference against the oracle", worst)
