"""Pins the oracle's per-pixel Jacobians: point-to-plane
(src/icp_jacobian_geometric.cu:54-85, icp_jacobian_common.hpp:37-61), feature
(src/icp_jacobian_feature.cu:58-109, src/feature_map.hpp:8-85) and the
information matrix (icp_jacobian_geometric.cu:117-158) by closed forms."""
import numpy as np

from oracle import oracle as orc

from test_oracle_correspondence import plane_scene


def test_point_to_plane_closed_form():
    k, p, n, m = plane_scene(z=2.0)
    # Tilted target normals, target points offset along them by delta:
    # residual r = (q - p) . n = delta, J = [n; p x n].
    rng = np.random.default_rng(1)
    tn = rng.normal(size=p.shape).astype(np.float32)
    tn[..., 2] = -np.abs(tn[..., 2]) - 2.0
    tn /= np.linalg.norm(tn, axis=-1, keepdims=True)
    delta = np.float32(0.01)
    tp = (p + delta * tn).astype(np.float32)
    cmap = orc.corresp_map(p, n, m, np.eye(4), p, n, m, k, 0.1, 1.5)
    jtj, jr, r2, cnt = orc.geometric(p, m, np.eye(4), tp, tn, k, cmap)
    assert cnt == m.size
    P = p.reshape(-1, 3).astype(np.float64)
    N = tn.reshape(-1, 3).astype(np.float64)
    J = np.concatenate([N, np.cross(P, N)], 1)
    r = np.sqrt(r2.astype(np.float64))
    np.testing.assert_allclose(r, 0.01, rtol=1e-4)
    np.testing.assert_allclose(jr, J * 0.01, rtol=2e-4, atol=1e-8)
    np.testing.assert_allclose(jtj, J[:, :, None] * J[:, None, :], rtol=1e-5,
                               atol=1e-7)
    # The summed system is symmetric positive semi-definite of rank <= 6.
    S = jtj.sum(0, dtype=np.float64)
    assert np.allclose(S, S.T)
    assert np.linalg.eigvalsh(S).min() > -1e-6 * np.abs(S).max()


def test_only_map_owners_contribute():
    k, p, n, m = plane_scene()
    cmap = orc.corresp_map(p, n, m, np.eye(4), p, n, m, k, 0.1, 1.0)
    cmap[:, ::2] = orc.CORRESP_INF  # half of the pixels have no owner
    jtj, jr, r2, cnt = orc.geometric(p, m, np.eye(4), p, n, k, cmap)
    assert cnt == m.size // 2
    assert np.all(jtj.reshape(12, 16, 36)[:, ::2] == 0)
    assert np.all(jtj.reshape(12, 16, 36)[:, 1::2, 14] == 1.0)  # n_z^2


def ramp_scene(H=12, W=16, a=(0.01, 0.03), b=(0.02, -0.01), c=(0.1, 0.5)):
    k, p, n, m = plane_scene(H, W)
    v, u = np.mgrid[0:H, 0:W].astype(np.float32)
    feats = np.stack([a[i] * u + b[i] * v + c[i] for i in range(len(a))]).astype(np.float32)
    return k, p, n, m, feats


def test_feature_linear_ramp_closed_form():
    k, p, n, m, feats = ramp_scene()
    cmap = orc.corresp_map(p, n, m, np.eye(4), p, n, m, k, 0.1, 1.0)
    src_feats = feats.reshape(2, -1)
    jtj, jr, r2, cnt = orc.feature(p, src_feats, m, np.eye(4), feats, k, cmap, 10.0)
    assert cnt == m.size
    # Interior pixels: bilinear of a linear function is exact, so the
    # residual is 0; the h = 0.05 central difference is exact: du = a, dv = b.
    inner = np.zeros((12, 16), bool)
    inner[1:-1, 1:-1] = True
    inner = inner.ravel()
    assert np.all(np.abs(r2[inner]) < 1e-10)
    P = p.reshape(-1, 3).astype(np.float64)
    fx, fy = float(k[0, 0]), float(k[1, 1])
    a, b = 0.03, -0.01  # the LAST channel's slopes (the reference keeps it)
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    g = np.stack([a * fx / z, b * fy / z, -(a * x * fx + b * y * fy) / z ** 2], 1)
    J = np.concatenate([g, np.cross(P, g)], 1)
    np.testing.assert_allclose(jtj[inner], (J[:, :, None] * J[:, None, :])[inner],
                               rtol=2e-3, atol=5e-6)


def test_feature_last_channel_wins_and_residual_sums_channels():
    k, p, n, m, feats = ramp_scene()
    cmap = orc.corresp_map(p, n, m, np.eye(4), p, n, m, k, 0.1, 1.0)
    src = feats.reshape(2, -1).copy()
    src[0] += 0.2
    src[1] -= 0.1
    jtj, jr, r2, cnt = orc.feature(p, src, m, np.eye(4), feats, k, cmap, 10.0)
    inner = np.zeros((12, 16), bool)
    inner[1:-1, 1:-1] = True
    inner = inner.ravel()
    np.testing.assert_allclose(r2[inner], 0.2 ** 2 + 0.1 ** 2, rtol=1e-5)
    # J r uses the last channel's residual (+0.1) only.
    last_only = orc.feature(p, src[1:], m, np.eye(4), feats[1:], k, cmap, 10.0)
    np.testing.assert_array_equal(jtj, last_only[0])
    np.testing.assert_array_equal(jr, last_only[1])
    # Threshold below sqrt(0.05): every pixel rejected, zeroed, uncounted.
    jtj, jr, r2, cnt = orc.feature(p, src, m, np.eye(4), feats, k, cmap, 0.2)
    assert cnt == 0 and not jtj.any() and not r2.any()


def test_information_matrix_closed_form():
    k, p, n, m = plane_scene()
    cmap = orc.corresp_map(p, n, m, np.eye(4), p, n, m, k, 0.1, 1.0)
    gtg = orc.information(p, m, np.eye(4), p, k, cmap)
    Q = p.reshape(-1, 3).astype(np.float64)
    for i in (0, 37, 191):
        x, y, z = Q[i]
        G = np.array([[0, z, -y, 1, 0, 0], [-z, 0, x, 0, 1, 0], [y, -x, 0, 0, 0, 1]])
        np.testing.assert_allclose(gtg[i], G.T @ G, rtol=1e-6, atol=1e-6)
