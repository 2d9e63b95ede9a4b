"""Pins the oracle's point-cloud restatements (CPU only): exact k nearest
neighbours against numpy brute force, the nearest-neighbour feature
interpolation against hand-computed weights, its finite-difference
gradient against the analytic derivative of the interpolant, and the voxel
merge against hand layouts (including the reference's index aliasing at the
far x edge)."""
import numpy as np
import pytest

from oracle import oracle as orc


def numpy_knn(t, q, k, bound):
    d2 = ((q[:, None, :] - t[None, :, :]) ** 2).sum(-1)
    D = np.full((len(q), k), np.inf)
    I = np.full((len(q), k), len(t), np.int64)
    for r in range(len(q)):
        keep = np.nonzero(d2[r] < bound * bound)[0]
        order = keep[np.lexsort((keep, d2[r, keep]))][:k]
        D[r, :len(order)] = d2[r, order]
        I[r, :len(order)] = order
    return D, I


@pytest.mark.parametrize("k,bound", [(1, np.inf), (5, 0.1), (8, np.inf), (16, 0.2)])
def test_knn_brute_force(k, bound):
    rng = np.random.default_rng(k)
    t = rng.uniform(0, 1, size=(300, 3))
    t[10] = t[20] = t[30]  # equal distances: lower index first
    q = rng.uniform(0, 1, size=(100, 3))
    q[0] = t[10]
    d, i = orc.knn(t, q, k, bound)
    D, I = numpy_knn(t, q, k, bound)
    assert np.array_equal(i, I)
    assert np.array_equal(d, D)
    assert i[0, 0] == 10 and (k == 1 or i[0, 1] == 20)
    if np.isfinite(bound):
        assert (i == len(t)).any() and np.isinf(d[i == len(t)]).all()


def test_nn_forward_hand_weights():
    feats = np.array([[1.0, 2.0, 4.0, 8.0], [0.0, 1.0, 0.0, 1.0]])
    m = feats.shape[1]
    dist = np.array([[0.5, 0.5, 0.5], [0.0, 1.0, np.inf], [np.inf, np.inf, np.inf]])
    index = np.array([[0, 1, 2], [3, 1, m], [m, m, m]])
    out = orc.nn_forward(dist, index, feats)
    np.testing.assert_allclose(out[:, 0], [(1 + 2 + 4) / 3, 1 / 3], rtol=1e-15)
    w0, w1 = 1.0, 0.5  # 1 / (1 + d)
    np.testing.assert_allclose(out[:, 1], [(8 * w0 + 2 * w1) / 1.5, (1 * w0 + 1 * w1) / 1.5],
                               rtol=1e-15)
    assert not out[:, 2].any()


def test_nn_backward_matches_the_analytic_gradient():
    rng = np.random.default_rng(3)
    t = rng.uniform(-1, 1, size=(40, 3))
    src = rng.uniform(-0.5, 0.5, size=(25, 3))
    feats = rng.normal(size=(3, 40))
    dl = rng.normal(size=(3, 25))
    d, idx = orc.knn(t, src, 6)
    eps = orc.nn_epsilon(t, src, idx)
    h = orc.NN_DELTA_H
    dx = orc.nn_backward(eps, idx, feats, dl, h)
    exp = np.zeros_like(src)
    for r in range(len(src)):
        tk = t[idx[r]]
        diff = src[r] - tk                       # x - t_k
        s = 1 / (1 + (diff ** 2).sum(1))
        S = s.sum()
        ds = -(s ** 2)[:, None] * 2 * diff       # d s_k / d x
        dS = ds.sum(0)
        for c in range(3):
            fk = feats[c, idx[r]]
            dF = ((fk[:, None] * (ds * S - s[:, None] * dS)).sum(0)) / S ** 2
            exp[r] += dF * dl[c, r]
    np.testing.assert_allclose(dx, exp, rtol=1e-5, atol=1e-8)
    # The weights of each offset sum to one.
    np.testing.assert_allclose(eps.sum(-1), 1.0, rtol=1e-14)


def test_voxel_merge_hand_layout():
    vs = 0.5
    pts = np.array([[0.0, 0.0, 0.0], [0.1, 0.2, 0.1], [0.6, 0.1, 0.0],
                    [0.7, 0.3, 0.2], [0.2, 0.9, 0.1], [1.0, 0.1, 0.4]], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (6, 1))
    nrm[1] = [0, 1, 0]
    col = np.array([[10, 20, 30], [11, 20, 30], [100, 0, 0], [101, 0, 0],
                    [5, 5, 5], [7, 7, 7]], np.uint8)
    conf = np.arange(6, dtype=np.float32)
    rad = np.full(6, 0.25, np.float32)
    feats = np.arange(12, dtype=np.float32).reshape(2, 6)
    out = orc.voxel_merge(pts, nrm, col, conf, rad, feats, vs)
    # width = int(1.0 / 0.5) = 2, height = int(0.9 / 0.5) = 1: ids are
    # col + 2 * row + 2 * depth. Point 5 (x = 1.0) has col 2 = width and
    # aliases onto row 1, col 0 -- the voxel of point 4.
    np.testing.assert_allclose(out["points"], [[0.05, 0.1, 0.05], [0.65, 0.2, 0.1],
                                               [0.6, 0.5, 0.25]], rtol=1e-6)
    np.testing.assert_allclose(out["normals"][0], [0, 0.5, 0.5])
    assert out["colors"].tolist() == [[10, 20, 30], [100, 0, 0], [6, 6, 6]]
    np.testing.assert_allclose(out["confidences"], [0.5, 2.5, 4.5])
    assert not out["times"].any()
    np.testing.assert_allclose(out["features"], [[0.5, 2.5, 4.5], [6.5, 8.5, 10.5]])
