"""Camera-trajectory metrics (slamtb/metrics/trajectory.py).

Trajectories are dicts from timestamps to ``RTCamera`` (or [4, 4] camera-to-
world matrices). The absolute and relative errors follow Sturm et al., "A
benchmark for the evaluation of RGB-D SLAM systems" (IROS 2012), without
trajectory alignment, as the reference computes them: per pose, the
translation norm or rotation angle of the residual transform (not squared).
The RMSE is ``e.pow(2).mean().sqrt()``; the reference's docstrings and tests
summarise with ``e.mean().sqrt()``, the square root of the mean error.
"""
import math

import torch


def _matrix(cam_or_matrix):
    m = getattr(cam_or_matrix, "matrix", cam_or_matrix)
    return torch.as_tensor(m, dtype=torch.float64)


def set_start_at_identity(trajectory):
    """Every pose left-multiplied by the inverse of the first
    (trajectory.py:9-13)."""
    from slamtb.camera import RTCamera
    base = torch.linalg.inv(_matrix(trajectory[next(iter(trajectory))]))
    return {time: RTCamera(base @ _matrix(cam)) for time, cam in trajectory.items()}


def translational_difference(diff_matrix):
    """Norm of the translation of a [3,4] / [4,4] pose difference."""
    return torch.as_tensor(diff_matrix)[:3, 3].double().norm(2).item()


def rotational_difference(diff_matrix):
    """Rotation angle of a pose difference, acos((tr R - 1) / 2) evaluated
    as atan2(|vee(R - R^T)| / 2, (tr R - 1) / 2), which is the same angle and
    stays accurate near zero."""
    R = torch.as_tensor(diff_matrix)[:3, :3].double()
    c = (torch.trace(R).item() - 1.0) / 2.0
    s = 0.5 * math.sqrt((R[2, 1] - R[1, 2]).item() ** 2 +
                        (R[0, 2] - R[2, 0]).item() ** 2 +
                        (R[1, 0] - R[0, 1]).item() ** 2)
    return math.atan2(s, c)


def get_absolute_residual(matrix_true, matrix_pred):
    """true^-1 @ pred (trajectory.py:45-61)."""
    return torch.linalg.inv(_matrix(matrix_true)) @ _matrix(matrix_pred)


def get_relative_residual(matrix_i_true, matrix_j_true, matrix_i_pred,
                          matrix_j_pred):
    """(i_true^-1 j_true)^-1 (i_pred^-1 j_pred) (trajectory.py:128-157)."""
    inv = torch.linalg.inv
    rel_true = inv(_matrix(matrix_i_true)) @ _matrix(matrix_j_true)
    rel_pred = inv(_matrix(matrix_i_pred)) @ _matrix(matrix_j_pred)
    return inv(rel_true) @ rel_pred


def _absolute(trajectory_true, trajectory_pred, diff):
    return torch.tensor([diff(get_absolute_residual(cam, trajectory_pred[k]))
                         for k, cam in trajectory_true.items()],
                        dtype=torch.float64)


def _relative(trajectory_true, trajectory_pred, diff):
    keys = list(trajectory_true.keys())
    return torch.tensor(
        [diff(get_relative_residual(trajectory_true[i], trajectory_true[j],
                                    trajectory_pred[i], trajectory_pred[j]))
         for i, j in zip(keys[:-1], keys[1:])], dtype=torch.float64)


def absolute_translational_error(trajectory_true, trajectory_pred):
    """Per-pose absolute translational error (trajectory.py:64-93)."""
    return _absolute(trajectory_true, trajectory_pred, translational_difference)


def absolute_rotational_error(trajectory_true, trajectory_pred):
    """Per-pose absolute rotational error in radians (trajectory.py:96-125)."""
    return _absolute(trajectory_true, trajectory_pred, rotational_difference)


def relative_translational_error(trajectory_true, trajectory_pred):
    """Per consecutive pair of the true trajectory's keys
    (trajectory.py:181-204)."""
    return _relative(trajectory_true, trajectory_pred, translational_difference)


def relative_rotational_error(trajectory_true, trajectory_pred):
    """Per consecutive pair, in radians (trajectory.py:207-229)."""
    return _relative(trajectory_true, trajectory_pred, rotational_difference)


def pose_error(T_a, T_b):
    """(rotation rad, translation) of T_a^-1 T_b."""
    D = torch.linalg.inv(torch.as_tensor(T_a, dtype=torch.float64)) @ \
        torch.as_tensor(T_b, dtype=torch.float64)
    return rotational_difference(D), translational_difference(D)
