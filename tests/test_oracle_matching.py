"""Pins the oracle's FPCLMatcherOp, ProjectOp and RigidTransformOp
restatements (CPU only).

* Projection: the closed form fx*x/z + cx on hand points, and the backward
  pass against central differences of the forward pass.
* Matcher forward: a brute-force matcher written here from the definition
  (rounded projection, owner = map >> 32, copies at the owned pixel) over a
  hand-made map with random owners and empty cells; features linear in
  (u, v), whose bilinear sample at an interior point is exactly that
  linear function, and the zero padding past the right/bottom edge.
* Matcher backward: with linear features the central difference is the
  exact slope, so d loss/d point is J_project^T (sum_c a_c g_c, sum_c b_c g_c);
  at u < h the lower sample clamps to column 0, which changes the slope to
  f(u + h) - f(0) over 2h.
* Rigid transforms: against float64 matrix products and the inverse-transpose
  from numpy's inverse; the backward pass against finite differences.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc

K = np.array([[120.0, 0.0, 31.5], [0.0, 110.0, 23.5], [0.0, 0.0, 1.0]])


def c_round(x):
    """C's round(): half away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def lift(uv, z, k=K):
    """Camera points whose projection is exactly the given (u, v) in f64."""
    x = (uv[:, 0] - k[0, 2]) * z / k[0, 0]
    y = (uv[:, 1] - k[1, 2]) * z / k[1, 1]
    return np.stack([x, y, z], 1)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_project_closed_form(dtype):
    p = np.array([[0.1, 0.2, 1.0], [-0.3, 0.25, 2.0], [0.0, 0.0, 0.5]], dtype)
    uv = orc.project_forward(p, K)
    k = K.astype(dtype)
    exp_u = k[0, 0] * p[:, 0] / p[:, 2] + k[0, 2]
    exp_v = k[1, 1] * p[:, 1] / p[:, 2] + k[1, 2]
    assert np.array_equal(uv[:, 0], exp_u.astype(dtype))
    assert np.array_equal(uv[:, 1], exp_v.astype(dtype))
    np.testing.assert_allclose(uv[0], [43.5, 45.5], rtol=1e-6)


def test_project_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = np.stack([rng.uniform(-1, 1, 50), rng.uniform(-1, 1, 50),
                  rng.uniform(0.5, 3, 50)], 1)
    dy = rng.normal(size=(50, 2))
    got = orc.project_backward(dy, p, K)
    eps = 1e-6
    num = np.zeros_like(p)
    for j in range(3):
        d = np.zeros(3)
        d[j] = eps
        fp = orc.project_forward(p + d, K)
        fm = orc.project_forward(p - d, K)
        num[:, j] = ((fp - fm) / (2 * eps) * dy).sum(1)
    np.testing.assert_allclose(got, num, rtol=1e-6, atol=1e-6)
    # A transposed Jacobian would put fx/z into the z component.
    assert np.abs(got[:, 2]).max() > 0


def ramp_feats(H, W, a, b, c0):
    v, u = np.mgrid[0:H, 0:W].astype(np.float64)
    return np.stack([ai * u + bi * v + ci for ai, bi, ci in zip(a, b, c0)])


def make_case(rng, H=48, W=64, n=400, empty=0.2):
    uv = np.stack([rng.uniform(-0.49, W - 0.51, n),
                   rng.uniform(-0.49, H - 0.51, n)], 1)
    # A few points project outside the image.
    uv[:8, 0] = rng.choice([-3.0, W + 2.0], 8)
    z = rng.uniform(0.5, 3.0, n)
    src = lift(uv, z)
    owner = rng.integers(0, n, size=(H, W)).astype(np.int64)
    ui, vi = c_round(orc.project_forward(src, K)).astype(np.int64).T
    inside = (ui >= 0) & (ui < W) & (vi >= 0) & (vi < H)
    # Make about half the points own their pixel (later points win).
    for i in np.nonzero(inside)[0]:
        if rng.uniform() < 0.5:
            owner[vi[i], ui[i]] = i
    dist_bits = rng.integers(0, 2**31, size=(H, W)).astype(np.int64)
    cmap = (owner << 32) | dist_bits
    cmap[rng.uniform(size=(H, W)) < empty] = orc.CORRESP_INF
    tgt_pts = rng.normal(size=(H, W, 3))
    tgt_nrm = rng.normal(size=(H, W, 3))
    return src, cmap, tgt_pts, tgt_nrm


def test_fpcl_forward_brute_force():
    rng = np.random.default_rng(11)
    H, W = 48, 64
    src, cmap, tp, tn = make_case(rng, H, W)
    a, b, c0 = [0.5, -1.25, 2.0], [0.75, 0.25, -3.0], [1.0, 0.0, 10.0]
    tf = ramp_feats(H, W, a, b, c0)
    op, on, of, mask = orc.fpcl_forward(tp, tn, tf, src, K, cmap)

    uv = orc.project_forward(src, K)
    ui, vi = c_round(uv).astype(np.int64).T
    exp_mask = np.zeros(len(src), bool)
    for i in range(len(src)):
        if 0 <= ui[i] < W and 0 <= vi[i] < H:
            exp_mask[i] = (cmap[vi[i], ui[i]] >> 32) == i
    assert np.array_equal(mask, exp_mask)
    assert 50 < mask.sum() < len(src)
    m = np.nonzero(mask)[0]
    assert np.array_equal(op[m], tp[vi[m], ui[m]])
    assert np.array_equal(on[m], tn[vi[m], ui[m]])
    # Interior points: the bilinear sample of a linear field is the field.
    interior = m[(uv[m, 0] >= 0) & (uv[m, 0] < W - 1) & (uv[m, 1] >= 0)
                 & (uv[m, 1] < H - 1)]
    assert len(interior) > 40
    for ch in range(3):
        exp = a[ch] * uv[interior, 0] + b[ch] * uv[interior, 1] + c0[ch]
        np.testing.assert_allclose(of[ch, interior], exp, rtol=0, atol=1e-10)
    assert not of[:, ~mask].any()


def test_fpcl_forward_zero_padding_at_the_far_edges():
    H, W = 6, 8
    tf = np.ones((1, H, W))
    uv = np.array([[W - 1 + 0.25, 2.0], [3.0, H - 1 + 0.4], [W - 1 + 0.3, H - 1 + 0.2]])
    src = lift(uv, np.ones(3))
    ui, vi = c_round(uv).astype(np.int64).T
    cmap = np.full((H, W), orc.CORRESP_INF, np.int64)
    cmap[vi, ui] = np.arange(3, dtype=np.int64) << 32
    _, _, of, mask = orc.fpcl_forward(np.zeros((H, W, 3)), np.zeros((H, W, 3)),
                                      tf, src, K, cmap)
    assert mask.all()
    ur, vr = uv[:, 0] % 1, uv[:, 1] % 1
    np.testing.assert_allclose(of[0], [1 - ur[0], 1 - vr[1], (1 - ur[2]) * (1 - vr[2])],
                               atol=1e-12)


def test_fpcl_backward_linear_features_closed_form():
    rng = np.random.default_rng(5)
    H, W, n = 40, 50, 300
    uv = np.stack([rng.uniform(1.0, W - 2.0, n), rng.uniform(1.0, H - 2.0, n)], 1)
    src = lift(uv, rng.uniform(0.5, 2.5, n))
    a, b = np.array([0.5, -2.0, 1.5, 0.25]), np.array([1.0, 0.5, -0.75, 3.0])
    tf = ramp_feats(H, W, a, b, [0.0, 1.0, 2.0, 3.0])
    mask = rng.uniform(size=n) < 0.7
    dl = rng.normal(size=(4, n))
    h = 0.005
    dx = orc.fpcl_backward(tf, src, mask, dl, K, h)
    gu = (a[:, None] * dl).sum(0)
    gv = (b[:, None] * dl).sum(0)
    x, y, z = src.T
    fx, fy = K[0, 0], K[1, 1]
    exp = np.stack([fx / z * gu, fy / z * gv,
                    -x * fx / z**2 * gu - y * fy / z**2 * gv], 1)
    exp[~mask] = 0
    np.testing.assert_allclose(dx, exp, rtol=1e-7, atol=1e-9)
    assert not dx[~mask].any()


def test_fpcl_backward_clamps_the_lower_sample_at_zero():
    H, W = 10, 12
    a, b = 2.0, 3.0
    tf = ramp_feats(H, W, [a], [b], [0.0])
    h = 0.2
    uv = np.array([[0.05, 4.0], [4.0, 0.1]])
    src = lift(uv, np.ones(2))
    dx = orc.fpcl_backward(tf, src, np.ones(2, bool), np.ones((1, 2)), K, h)
    # du at u = 0.05: (f(u + h) - f(0)) / 2h = a (u + h) / 2h; dv = b.
    gu0 = a * (uv[0, 0] + h) / (2 * h)
    gv1 = b * (uv[1, 1] + h) / (2 * h)
    fx, fy = K[0, 0], K[1, 1]
    np.testing.assert_allclose(dx[0, 0], fx * gu0, rtol=1e-9)
    np.testing.assert_allclose(dx[0, 1], fy * b, rtol=1e-9)
    np.testing.assert_allclose(dx[1, 0], fx * a, rtol=1e-9)
    np.testing.assert_allclose(dx[1, 1], fy * gv1, rtol=1e-9)
    assert not math.isclose(gu0, a)


def test_fpcl_backward_float32_follows_float64():
    rng = np.random.default_rng(8)
    H, W, n = 30, 40, 200
    uv = np.stack([rng.uniform(0, W - 1, n), rng.uniform(0, H - 1, n)], 1)
    src = lift(uv, rng.uniform(0.5, 2.5, n))
    tf = rng.uniform(0, 1, size=(3, H, W))
    dl = rng.normal(size=(3, n))
    mask = np.ones(n, bool)
    d64 = orc.fpcl_backward(tf, src, mask, dl, K, 0.05)
    d32 = orc.fpcl_backward(tf.astype(np.float32), src.astype(np.float32), mask,
                            dl.astype(np.float32), K, 0.05)
    np.testing.assert_allclose(d32, d64, rtol=2e-3, atol=2e-2 * np.abs(d64).max())


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rigid_points_and_normals(dtype):
    rng = np.random.default_rng(9)
    M = rng.normal(size=(3, 4))
    p = rng.normal(size=(100, 3)).astype(dtype)
    got = orc.rigid_points(M, p)
    exp = p.astype(np.float64) @ M[:, :3].T + M[:, 3]
    np.testing.assert_allclose(got, exp, rtol=1e-5 if dtype == np.float32 else 1e-13,
                               atol=1e-5 if dtype == np.float32 else 1e-13)
    nrm = orc.rigid_points(M, p, normals=True)
    N = np.linalg.inv(M[:, :3]).T
    np.testing.assert_allclose(nrm, p.astype(np.float64) @ N.T,
                               rtol=1e-4 if dtype == np.float32 else 1e-11,
                               atol=1e-4 if dtype == np.float32 else 1e-11)


def test_rigid_backward_matches_finite_differences():
    rng = np.random.default_rng(10)
    M = rng.normal(size=(3, 4))
    p = rng.normal(size=(50, 3))
    dy = rng.normal(size=(50, 3))
    dx, dM = orc.rigid_backward(M, p, dy)
    loss = lambda MM, pp: float((orc.rigid_points(MM, pp) * dy).sum())
    eps = 1e-6
    for r in range(3):
        for c in range(4):
            E = np.zeros((3, 4))
            E[r, c] = eps
            num = (loss(M + E, p) - loss(M - E, p)) / (2 * eps)
            assert abs(num - dM[r, c]) < 1e-6 * max(1, abs(num))
    for i in (0, 17, 49):
        for j in range(3):
            P = np.zeros_like(p)
            P[i, j] = eps
            num = (loss(M, p + P) - loss(M, p - P)) / (2 * eps)
            assert abs(num - dx[i, j]) < 1e-6 * max(1, abs(num))
