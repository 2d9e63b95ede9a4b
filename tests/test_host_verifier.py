"""Known answers of RegistrationVerifier (slamtb/registration/result.py:
76-88): each of its three rejection rules on hand-built results, and the
accept branch. The tracking front end takes its `accepted` flag from it
(slamtb/pipeline/surfel_slam.py:121-125)."""
import math

import numpy as np
import pytest
import torch

from slamtb.registration.result import RegistrationResult, RegistrationVerifier


def result(residual=1e-5, information=None, match_ratio=0.9):
    if information is None:
        information = torch.eye(6, dtype=torch.float64) * 1e5
    return RegistrationResult(torch.eye(4, dtype=torch.float64), information,
                              residual, match_ratio)


def mixed_information(cov_eigs):
    """An information matrix whose inverse (the covariance) has eigenvalues
    cov_eigs and whose eigenvector of the first one is (1, ..., 1) / sqrt(6):
    every diagonal entry of the covariance is then well below that largest
    eigenvalue, so a rule that read the diagonal would accept it."""
    a = np.eye(6)
    a[:, 0] = 1.0
    q, _ = np.linalg.qr(a)
    cov = q @ np.diag(cov_eigs) @ q.T
    return torch.from_numpy(np.linalg.inv(cov))


def test_accepts_a_good_result():
    assert RegistrationVerifier()(result())


def test_rejects_a_residual_above_the_threshold():
    v = RegistrationVerifier()
    assert not v(result(residual=2e-4))
    assert v(result(residual=1e-4))  # the bound itself passes (strict >)
    assert RegistrationVerifier(residual_threshhold=1e-3)(result(residual=2e-4))


def test_rejects_a_covariance_eigenvalue_above_the_threshold():
    v = RegistrationVerifier()
    cov = [1.5e-4] + [1e-6] * 5
    info = mixed_information(cov)
    covariance = np.linalg.inv(info.numpy())
    assert np.diag(covariance).max() < 1e-4  # the diagonal alone looks fine
    assert not v(result(information=info))
    assert v(result(information=mixed_information([0.9e-4] + [1e-6] * 5)))
    # In float32, as a caller may pass it (inverted in its own dtype).
    assert not v(result(information=info.float()))
    assert RegistrationVerifier(covariance_max_threshold=2e-4)(result(information=info))


def test_rejects_a_low_match_ratio():
    v = RegistrationVerifier()
    assert not v(result(match_ratio=0.59))
    assert v(result(match_ratio=0.6))
    assert RegistrationVerifier(match_ratio_threshold=0.5)(result(match_ratio=0.59))


def test_no_information_skips_the_covariance_rule():
    r = result()
    r.information = None
    assert RegistrationVerifier()(r)
    r.match_ratio = 0.1
    assert not RegistrationVerifier()(r)


def test_failed_registration_is_rejected():
    """A registration that accepted no iteration keeps residual = inf."""
    r = RegistrationResult()
    assert r.residual == math.inf
    assert not RegistrationVerifier()(r)
