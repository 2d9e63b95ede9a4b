"""Correspondence search on frames (slamtb/spatial/__init__.py)."""
from .matching import FPCLMatcherOp, FramePointCloudMatcher, PointCloudMatcher
