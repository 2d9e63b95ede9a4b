"""Pipelines around the registration step (slamtb/pipeline/)."""
from .surfel_slam import SurfelSLAM, TrackingStats, default_registration
