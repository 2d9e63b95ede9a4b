"""Host logic of the tracking front end (CPU)."""
import math

import pytest
import torch

from slamtb.camera import RTCamera, quaternion_to_matrix
from slamtb.pipeline.surfel_slam import SurfelSLAM, estimate_confidence_weight


def test_confidence_weight():
    a = RTCamera()
    assert estimate_confidence_weight(a, a) == 1.0
    b = RTCamera.create_from_pos_rot([0.004, 0, 0], torch.eye(3))
    assert abs(estimate_confidence_weight(a, b) - 0.6) < 1e-12
    c = RTCamera.create_from_pos_rot([0, 0, 0], quaternion_to_matrix(
        math.cos(0.001), math.sin(0.001), 0, 0))
    assert abs(estimate_confidence_weight(a, c) - 0.8) < 1e-9
    d = RTCamera.create_from_pos_rot([0.5, 0, 0], torch.eye(3))
    assert estimate_confidence_weight(a, d) == 0.5


def test_frame_to_model_is_not_provided():
    with pytest.raises(NotImplementedError):
        SurfelSLAM(tracking_mode='frame-to-model')
