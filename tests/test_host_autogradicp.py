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


class _Icp:
    def __init__(self, gw, fw, thr, alpha):
        self.geom_weight, self.feat_weight = gw, fw
        self.feat_residual_thresh, self.huber_loss_alpha = thr, alpha


def _rows(seed, n=400, C=3):
    from types import SimpleNamespace
    g = torch.Generator().manual_seed(seed)
    mask = torch.rand(n, generator=g) < 0.6
    pts = torch.randn(n, 3, generator=g, dtype=torch.float64) * mask[:, None]
    nrm = torch.randn(n, 3, generator=g, dtype=torch.float64) * mask[:, None]
    feats = (torch.randn(C, n, generator=g, dtype=torch.float64) * 2 * mask).requires_grad_()
    tsrc = torch.randn(n, 3, generator=g, dtype=torch.float64).requires_grad_()
    sf = torch.randn(C, n, generator=g, dtype=torch.float64)
    return SimpleNamespace(points=pts, normals=nrm, features=feats, mask=mask), tsrc, sf


import pytest  # noqa: E402


@pytest.mark.parametrize("gw,fw,thr,alpha", [(1, 0, 0.5, 4.5), (0.5, 0.5, 2.5, 1.5),
                                             (0, 1, -1, 2.0), (0.5, 0.5, 3.0, 0)])
def test_row_loss_equals_the_compacted_reference_cost(gw, fw, thr, alpha):
    """The masked-sum cost AutogradICP evaluates on uncompacted match rows
    equals the reference's cost on compacted matches (autogradicp.py:183-201),
    in value and in the gradients to the moved points and matched features."""
    from slamtb.registration.autogradicp import _compacted_loss, _row_loss
    icp = _Icp(gw, fw, thr, alpha)
    m, tsrc, sf = _rows(7)
    g1, f1 = _row_loss(tsrc, m, sf, icp)
    l1 = g1 * gw + f1 * fw
    l1.backward()
    clone = lambda t: None if t is None else t.clone()
    grads1 = (clone(tsrc.grad), clone(m.features.grad))
    tsrc.grad = None
    m.features.grad = None
    mk = m.mask
    g2, f2 = _compacted_loss(tsrc, m.points[mk], m.normals[mk], m.features[:, mk],
                             mk, sf, icp, _HuberLoss(alpha))
    l2 = g2 * gw + f2 * fw
    l2.backward()
    v1, v2 = float(l1.detach()), float(l2.detach())
    assert abs(v1 - v2) < 1e-12 * max(1.0, abs(v2))
    for a, b in zip(grads1, (tsrc.grad, m.features.grad)):
        assert (a is None) == (b is None)
        if a is not None:
            torch.testing.assert_close(a, b, rtol=1e-10, atol=1e-14)
            assert a.abs().max() > 0
    if fw > 0 and thr > 0:
        # The threshold really removed residuals in this case.
        d = torch.norm(m.features[:, mk] - sf[:, mk], dim=0)
        assert (d.pow(2) >= thr * thr).any() and (d.pow(2) < thr * thr).any()
