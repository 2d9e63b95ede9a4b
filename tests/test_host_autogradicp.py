"""Host-side pieces of AutogradICP that need no GPU: the Huber loss as the
reference defines it (residuals below alpha enter linearly, the rest as
alpha*|r| - alpha, autogradicp.py:61-90) and the scipy closure adapter."""
import numpy as np
import scipy.optimize
import torch

from slamtb.registration.autogradicp import AutogradICP, _ClosureBox, _HuberLoss, _to_4x4


def test_huber_loss_definition():
    r = torch.tensor([0.5, 1.0, 3.0, 5.0, 7.0])
    loss = _HuberLoss(4.5)(r)
    expected = np.mean([0.5, 1.0, 3.0]) + np.mean([4.5 * 5 - 4.5, 4.5 * 7 - 4.5])
    assert abs(float(loss) - expected) < 1e-5
    assert _HuberLoss(4.5)(torch.tensor([])) == 0
    assert abs(float(_HuberLoss(1.0)(torch.tensor([2.0]))) - 1.0) < 1e-7


def test_closure_box_drives_bfgs():
    target = torch.tensor([0.1, -0.2, 0.3, 0.05, -0.05, 0.2], dtype=torch.float64)
    x = torch.zeros(6, dtype=torch.float64, requires_grad=True)

    def closure():
        if x.grad is not None:
            x.grad.zero_()
        return ((x - target) ** 2).sum()

    box = _ClosureBox(closure, x)
    res = scipy.optimize.minimize(box.func, np.zeros(6), method="BFGS", jac=True,
                                  tol=1e-12, options=dict(maxiter=50))
    np.testing.assert_allclose(res.x, target.numpy(), atol=1e-8)
    # A closure that formed no term reports a zero cost and gradient.
    value, grad = _ClosureBox(lambda: 0, x).func(np.zeros(6))
    assert value == 0.0 and not grad.any()


def test_to_4x4_and_defaults():
    m = torch.arange(12, dtype=torch.float32).view(3, 4)
    out = _to_4x4(m)
    assert out.shape == (4, 4) and torch.equal(out[:3], m)
    assert torch.equal(out[3], torch.tensor([0.0, 0, 0, 1]))
    icp = AutogradICP(10)
    assert (icp.learning_rate, icp.geom_weight, icp.feat_weight) == (0.01, 0.5, 0.5)
    assert icp.gradient_descent_iters == 100 and icp.huber_loss_alpha == 4.5
