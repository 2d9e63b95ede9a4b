"""Trajectory metrics (slamtb/metrics/__init__.py). The geometric
reconstruction metrics of the reference (``metrics/geometry.py``: triangle
octree, mesh sampling) belong to surfel fusion and are not provided."""
from .trajectory import (absolute_translational_error, absolute_rotational_error,
                         relative_translational_error, relative_rotational_error,
                         get_absolute_residual, get_relative_residual,
                         translational_difference, rotational_difference,
                         set_start_at_identity, pose_error)
