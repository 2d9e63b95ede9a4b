"""Pose-difference metrics (slamtb/metrics/trajectory.py:16-35)."""
import math

import torch


def translational_difference(diff_matrix):
    return diff_matrix[:3, 3].norm(2).item()


def rotational_difference(diff_matrix):
    """Rotation angle of a pose difference, acos((tr R - 1) / 2) evaluated
    as atan2(|vee(R - R^T)| / 2, (tr R - 1) / 2), which is the same angle and
    stays accurate near zero."""
    R = diff_matrix[:3, :3].double()
    c = (torch.trace(R).item() - 1.0) / 2.0
    s = 0.5 * math.sqrt((R[2, 1] - R[1, 2]).item() ** 2 +
                        (R[0, 2] - R[2, 0]).item() ** 2 +
                        (R[1, 0] - R[0, 1]).item() ** 2)
    return math.atan2(s, c)


def pose_error(T_a, T_b):
    """(rotation rad, translation) of T_a^-1 T_b."""
    D = torch.linalg.inv(torch.as_tensor(T_a, dtype=torch.float64)) @ \
        torch.as_tensor(T_b, dtype=torch.float64)
    return rotational_difference(D), translational_difference(D)
