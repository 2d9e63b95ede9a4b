"""FTB frames, TUM trajectories and the trajectory metrics (CPU).

Known answers: the reference's own metrics test pins ATE/RTE/ARE/RRE of the
freiburg1_desk ground-truth and predicted trajectories
(slamtb/metrics/_test/test_metrics.py:52-88); the fixtures are copies of the
reference's test data (tests/golden/make_ftb_fixtures.py).
"""
import json
import math
from pathlib import Path

import numpy as np
import pytest
import torch

from slamtb import metrics
from slamtb.camera import KCamera, RTCamera, quaternion_to_matrix
from slamtb.data import load_ftb, write_ftb
from slamtb.data.tumrgbd import read_trajectory, write_trajectory
from slamtb.processing import ColorSpace, to_color_feature

GOLDEN = Path(__file__).parent / "golden"


@pytest.fixture(scope="module")
def trajs():
    return (read_trajectory(GOLDEN / "trajectory" / "freiburd1_desk_gt.traj"),
            read_trajectory(GOLDEN / "trajectory" / "freiburd1_desk_pred.traj"))


def test_trajectory_metrics_known_answers(trajs):
    gt, pred = trajs
    rmse = lambda e: e.mean().sqrt().item()
    assert abs(rmse(metrics.absolute_translational_error(gt, pred)) - 0.5349) < 5e-5
    assert abs(rmse(metrics.relative_translational_error(gt, pred)) - 0.1008) < 5e-5
    assert abs(rmse(metrics.absolute_rotational_error(gt, pred)) - 0.497086) < 5e-5
    assert abs(rmse(metrics.relative_rotational_error(gt, pred)) - 0.0992) < 5e-5


def test_read_trajectory_and_quaternions(trajs):
    gt, _ = trajs
    assert len(gt) == 500
    first = gt[next(iter(gt))].matrix
    assert torch.allclose(first, torch.eye(4, dtype=torch.float64), atol=1e-9)
    for cam in list(gt.values())[::50]:
        R = cam.matrix[:3, :3]
        assert torch.allclose(R @ R.T, torch.eye(3, dtype=torch.float64), atol=1e-8)
        assert abs(torch.det(R).item() - 1) < 1e-8
    # 90 degrees about z: (w, x, y, z) = (cos 45, 0, 0, sin 45).
    c = math.sqrt(0.5)
    R = quaternion_to_matrix(c, 0, 0, c)
    assert torch.allclose(R, torch.tensor([[0.0, -1, 0], [1, 0, 0], [0, 0, 1]],
                                          dtype=torch.float64), atol=1e-12)


def test_write_trajectory_round_trip(tmp_path, trajs):
    gt, _ = trajs
    sub = dict(list(gt.items())[:40])
    write_trajectory(tmp_path / "t.traj", sub)
    back = read_trajectory(tmp_path / "t.traj")
    assert list(back.keys()) == list(sub.keys())
    for k in sub:
        assert torch.allclose(back[k].matrix, sub[k].matrix, atol=1e-8)


def test_relative_and_absolute_residual_closed_forms():
    Ti = RTCamera.create_from_pos_rot([1.0, 2.0, 3.0], quaternion_to_matrix(0.9, 0.1, 0.3, 0.2))
    step = RTCamera.create_from_pos_rot([0.1, 0.0, -0.05], quaternion_to_matrix(1, 0, 0.05, 0))
    Tj = Ti @ step
    # A prediction off by a known local motion D at j only.
    D = RTCamera.create_from_pos_rot([0.0, 0.02, 0.0], quaternion_to_matrix(1, 0.01, 0, 0))
    Tj_pred = Tj @ D
    gt = {0: Ti, 1: Tj}
    pred = {0: Ti, 1: Tj_pred}
    rte = metrics.relative_translational_error(gt, pred)
    rre = metrics.relative_rotational_error(gt, pred)
    assert abs(rte.item() - 0.02) < 1e-12
    angle = 2 * math.atan2(0.01, 1)
    assert abs(rre.item() - angle) < 1e-12
    ate = metrics.absolute_translational_error(gt, pred)
    assert ate[0].item() < 1e-12 and abs(ate[1].item() - 0.02) < 1e-12
    ident = metrics.set_start_at_identity(gt)
    assert torch.allclose(ident[0].matrix, torch.eye(4, dtype=torch.float64), atol=1e-12)


def test_load_ftb_sample1():
    ds = load_ftb(GOLDEN / "ftb" / "sample1")
    assert len(ds) == 15
    f = ds[0]
    assert f.depth_image.dtype == np.int32 and f.depth_image.shape == (480, 640)
    assert 0.8 < (f.depth_image > 0).mean() < 0.95 and f.depth_image.max() < 4500
    assert f.rgb_image.dtype == np.uint8 and f.rgb_image.shape == (480, 640, 3)
    assert ds[1].rgb_image is None and ds[14].rgb_image is not None
    info = f.info
    assert info.depth_scale == 0.001 and info.kcam.image_size == [640, 480]
    assert torch.allclose(info.kcam.matrix,
                          torch.tensor([[544.4732666015625, 0, 320], [0, 544.4732666015625, 240],
                                        [0, 0, 1]]))
    assert torch.allclose(info.rt_cam.matrix, torch.eye(4, dtype=torch.float64), atol=1e-6)
    raw = json.load(open(GOLDEN / "ftb" / "sample1" / "frames.json"))["root"][14]
    assert torch.equal(ds[14].info.rt_cam.matrix,
                       torch.tensor(raw["info"]["rt_cam"]["matrix"], dtype=torch.float64))


def test_write_ftb_round_trip(tmp_path):
    ds = load_ftb(GOLDEN / "ftb" / "sample2")
    write_ftb(tmp_path, ds, max_frames=2)
    back = load_ftb(tmp_path)
    assert len(back) == 2
    for i in range(2):
        assert np.array_equal(back[i].depth_image, ds[i].depth_image)
        if ds[i].rgb_image is not None:
            assert np.array_equal(back[i].rgb_image, ds[i].rgb_image)
        assert back[i].info.to_json() == ds[i].info.to_json()


def test_to_color_feature():
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, size=(7, 9, 3), dtype=np.uint8)
    f = to_color_feature(img)
    assert f.shape == (3, 7, 9) and f.dtype == torch.float32
    assert torch.allclose(f[1], torch.from_numpy(img[..., 1]).float() / 255)
    g = to_color_feature(np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255], [10, 20, 30]]],
                                  np.uint8), ColorSpace.GRAY)
    assert g.shape == (1, 1, 4)
    assert torch.equal((g * 255).round().int().flatten(),
                       torch.tensor([76, 150, 29, 18], dtype=torch.int32))
    flat = np.full((6, 6, 3), 77, np.uint8)
    assert torch.equal(to_color_feature(flat, blur=True), to_color_feature(flat))
    delta = np.zeros((7, 7, 3), np.uint8)
    delta[3, 3] = 250
    b = (to_color_feature(delta, blur=True) * 255).round()
    assert b[0, 1:6, 1:6].eq(10).all() and b[0, 0].eq(0).all()
    with pytest.raises(NotImplementedError):
        to_color_feature(img, ColorSpace.LAB)


def test_kcamera_and_rtcamera_json():
    k = KCamera.from_json({"matrix": [[500, 0, 320], [0, 500, 240], [0, 0, 1]],
                           "image_size": [640, 480]})
    assert KCamera.from_json(k.to_json()).matrix.equal(k.matrix)
    cam = RTCamera.create_from_pos_quat(1, 2, 3, 1, 0, 0, 0)
    assert cam.center == (1.0, 2.0, 3.0)
    assert torch.equal(RTCamera.from_json(cam.to_json()).matrix, cam.matrix)
    m = torch.eye(4, dtype=torch.float64)
    m[0, 3] = 5
    assert cam.right_transform(m).center == (6.0, 2.0, 3.0)
    assert cam.transform(m).center == (6.0, 2.0, 3.0)
