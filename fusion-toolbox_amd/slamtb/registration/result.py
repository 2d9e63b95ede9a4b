"""Registration result and verifier (slamtb/registration/result.py)."""
import math

import torch


class RegistrationResult:
    """transform: [4, 4] float64 CPU tensor (None when no iteration was
    accepted); information: [6, 6]; residual: float; match_ratio: float."""

    def __init__(self, transform=None, information=None, residual=math.inf,
                 match_ratio=0, source_size=None, target_size=None):
        self.transform = transform
        self.information = information
        self.residual = residual
        self.match_ratio = match_ratio
        self.source_size = source_size
        self.target_size = target_size

    def __str__(self):
        return (f"RegistrationResult with: transform = {self.transform} "
                f"information = {self.information} residual = {self.residual} "
                f"match_ratio = {self.match_ratio}")

    def __repr__(self):
        return str(self)


class RegistrationVerifier:
    """Heuristic acceptance test of an ICP result (result.py:49-95).

    The reference calls ``torch.eig``, which current torch no longer has;
    this uses ``torch.linalg.eigvals`` and keeps the real parts, which is the
    same first column ``torch.eig`` returned. As in the reference the
    information matrix is inverted in its own dtype.
    """

    def __init__(self, match_ratio_threshold=6e-1,
                 covariance_max_threshold=1e-04, residual_threshhold=1e-4):
        self.match_ratio_threshold = match_ratio_threshold
        self.covariance_max_threshold = covariance_max_threshold
        self.residual_threshhold = residual_threshhold

    def __call__(self, result):
        if result.residual > self.residual_threshhold:
            return False
        if result.information is not None:
            eigvals = torch.linalg.eigvals(result.information.inverse()).real
            if torch.any(eigvals > self.covariance_max_threshold):
                return False
        if result.match_ratio < self.match_ratio_threshold:
            return False
        return True

    def __str__(self):
        return (f"RegistrationVerifier with: match_ratio_threshold = "
                f"{self.match_ratio_threshold} covariance_max_threshold = "
                f"{self.covariance_max_threshold}")

    def __repr__(self):
        return str(self)
