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
