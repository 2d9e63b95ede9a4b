"""Pins the oracle's SE3 maps (src/exp_rt.hpp, src/exp_rt_to_matrix.cu,
src/matrix_to_exp_rt.cu) and preprocessing (bilateral, normals, erode,
downsample, feature blur) by closed forms and hand-made inputs."""
import math

import numpy as np
import pytest

from oracle import oracle as orc


def rodrigues(w):
    th = np.linalg.norm(w)
    if th == 0:
        return np.eye(3)
    K = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]) / th
    return np.eye(3) + math.sin(th) * K + (1 - math.cos(th)) * K @ K


def test_exp_rt_axis_rotation_and_raw_translation():
    th = 0.7
    m = orc.exp_rt_to_matrix(np.array([0.1, -0.2, 0.3, 0.0, 0.0, th]))
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0],
                   [0, 0, 1]])
    np.testing.assert_allclose(m[:3, :3], Rz, atol=1e-15)
    np.testing.assert_array_equal(m[:, 3], [0.1, -0.2, 0.3])
    np.testing.assert_array_equal(orc.exp_rt_to_matrix(np.zeros(6))[:, :3], np.eye(3))


def test_exp_rt_is_a_rotation():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(20, 6))
    m = orc.exp_rt_to_matrix(x)
    for i in range(20):
        R = m[i, :, :3]
        np.testing.assert_allclose(R.T @ R, np.eye(3), atol=1e-14)
        assert abs(np.linalg.det(R) - 1) < 1e-14
        np.testing.assert_allclose(R, rodrigues(x[i, 3:]), atol=1e-14)


def test_log_exp_round_trip():
    rng = np.random.default_rng(4)
    w = rng.normal(size=(50, 3))
    w *= (rng.uniform(0.01, 3.0, 50) / np.linalg.norm(w, axis=1))[:, None]
    x = np.concatenate([rng.normal(size=(50, 3)), w], 1)
    back = orc.matrix_to_exp_rt(orc.exp_rt_to_matrix(x))
    np.testing.assert_allclose(back, x, atol=1e-12)
    # The literal matrix of the reference's disabled round-trip test
    # (registration/_test/test_se3.py:40-42), projected onto SO(3).
    A = np.array([[0.7350, 0.3997, -0.5477, -0.2654],
                  [-0.4502, 0.8917, 0.0464, 0.3553],
                  [0.5069, 0.2125, 0.8354, -0.3161]])
    U, _, Vt = np.linalg.svd(A[:, :3])
    A[:, :3] = U @ Vt
    np.testing.assert_allclose(orc.exp_rt_to_matrix(orc.matrix_to_exp_rt(A))[0],
                               A, atol=1e-12)


def test_exp_rt_backward_matches_finite_differences():
    # Restates registration/_test/test_se3.py:26-35 (eps 1e-6, atol 1e-4)
    # with a fixed cotangent and also checks the small-angle branch.
    rng = np.random.default_rng(10)
    for x in [rng.uniform(size=6), np.array([0.1, 0.2, 0.3, 1e-8, -2e-8, 1e-8])]:
        y = orc.exp_rt_to_matrix(x)
        G = rng.normal(size=(3, 4))
        got = orc.exp_rt_backward(G, x, y)[0]
        fd = np.zeros(6)
        for i in range(6):
            e = np.zeros(6)
            e[i] = 1e-6
            fd[i] = ((orc.exp_rt_to_matrix(x + e) - orc.exp_rt_to_matrix(x - e))
                     * G).sum() / 2e-6
        np.testing.assert_allclose(got, fd, atol=1e-4)


def test_bilateral_constant_ramp_and_isolated_pixel():
    H, W = 20, 24
    depth = np.full((H, W), 1.5, np.float32)
    mask = np.ones((H, W), bool)
    out = orc.bilateral(depth, mask, 6, 0.03, 4.5, 1.0)
    np.testing.assert_allclose(out, 1.5, rtol=1e-6)
    # A linear ramp is symmetric about every interior pixel, so the weighted
    # mean is the pixel itself.
    ramp = (1.0 + 0.001 * np.arange(W, dtype=np.float32))[None, :].repeat(H, 0)
    out = orc.bilateral(ramp.astype(np.float32), mask, 6, 0.03, 4.5, 1.0)
    np.testing.assert_allclose(out[3:-3, 3:-3], ramp[3:-3, 3:-3], rtol=2e-6)
    iso = np.zeros((H, W), bool)
    iso[10, 10] = True
    out = orc.bilateral(depth, iso, 6, 0.03, 4.5, 1.0)
    assert out[10, 10] == np.float32(1.5) and np.count_nonzero(out) == 1
    # Millimetre integers, as the reference tests feed it: the float mean is
    # converted back by truncation, so 1499.9999 becomes 1499.
    mm = np.full((H, W), 1500, np.int32)
    out = orc.bilateral(mm, mask, 13, 29.9999880000072, 4.50000000225, 1.0)
    assert set(np.unique(out)) <= {1499, 1500}


def _box_mean(depth, mask, half):
    """Mean of the valid pixels of the (2 half + 1)^2 window around each
    valid pixel (0 where the centre is masked out), in FP64."""
    H, W = depth.shape
    out = np.zeros((H, W))
    for r in range(H):
        for c in range(W):
            if not mask[r, c]:
                continue
            win = (slice(max(r - half, 0), r + half + 1),
                   slice(max(c - half, 0), c + half + 1))
            out[r, c] = depth[win][mask[win]].astype(np.float64).mean()
    return out


@pytest.mark.parametrize("filter_width,half", [(6, 3), (5, 2), (4, 2), (2, 1)])
def test_bilateral_wide_sigmas_give_the_window_box_mean(filter_width, half):
    """With both sigmas huge every weight is 1: the output is the plain mean
    of the valid pixels of the (2 floor(fw / 2) + 1)^2 window
    (src/bilateral_depth_filter.cu:47-54,85), which pins the window and the
    masking of neighbours."""
    rng = np.random.default_rng(filter_width)
    H, W = 14, 17
    depth = rng.uniform(1.0, 3.0, size=(H, W)).astype(np.float32)
    mask = rng.uniform(size=(H, W)) < 0.8
    out = orc.bilateral(depth, mask, filter_width, 1e4, 1e4, 1.0)
    np.testing.assert_allclose(out, _box_mean(depth, mask, half), rtol=1e-6, atol=0)
    assert (out[~mask] == 0).all()


def _two_pixel_weight(space_sqr, ddepth, sigma_space, sigma_color):
    return math.exp(-0.5 * (space_sqr / sigma_space ** 2 +
                            ddepth ** 2 / sigma_color ** 2))


@pytest.mark.parametrize("dx,dy", [(2, 1), (-3, 0), (1, -3)])
def test_bilateral_two_pixel_closed_form(dx, dy):
    """Two valid pixels a distance (dx, dy) apart with depths d0, d1: each
    output is (d_self + w d_other) / (1 + w) with
    w = exp(-0.5 (|dx, dy|^2 / sigma_space^2 + (d1 - d0)^2 / sigma_color^2))
    (src/bilateral_depth_filter.cu:56-68). The spatial and range terms are
    of different sizes here, so exchanging the two sigmas, or dropping one
    term, moves the output by far more than the tolerance."""
    H, W = 12, 12
    r0, c0 = 5, 5
    d0, d1 = 1.0, 1.04
    depth = np.zeros((H, W), np.float32)
    mask = np.zeros((H, W), bool)
    depth[r0, c0], depth[r0 + dy, c0 + dx] = d0, d1
    mask[r0, c0] = mask[r0 + dy, c0 + dx] = True
    ss, sc = 2.0, 0.05
    out = orc.bilateral(depth, mask, 6, sc, ss, 1.0)
    w = _two_pixel_weight(dx * dx + dy * dy, np.float32(d1) - np.float32(d0), ss, sc)
    assert 0.01 < w < 0.9  # both terms matter
    np.testing.assert_allclose(out[r0, c0], (d0 + w * d1) / (1 + w), rtol=2e-6)
    np.testing.assert_allclose(out[r0 + dy, c0 + dx], (d1 + w * d0) / (1 + w), rtol=2e-6)
    # The plausible mistakes give other values.
    for wrong in (_two_pixel_weight(dx * dx + dy * dy, d1 - d0, sc, ss),
                  _two_pixel_weight(dx * dx + dy * dy, 0.0, ss, sc),
                  _two_pixel_weight(0.0, d1 - d0, ss, sc)):
        assert abs((d0 + wrong * d1) / (1 + wrong) - out[r0, c0]) > 1e-4
    # A window of half width 2 (filter_width 4 or 5) leaves out a neighbour
    # three pixels away: each pixel keeps its own depth.
    out4 = orc.bilateral(depth, mask, 4, sc, ss, 1.0)
    if max(abs(dx), abs(dy)) == 3:
        assert out4[r0, c0] == np.float32(d0)
    else:
        np.testing.assert_allclose(out4[r0, c0], out[r0, c0], rtol=1e-7)


def test_bilateral_depth_scale_multiplies_in_and_divides_out():
    """Millimetre input with depth_scale 0.001: the weights see metres
    (sigma_color 0.05 m), the output is back in millimetres and, for an
    integer image, truncated (src/bilateral_depth_filter.cu:38,64,72-73)."""
    H, W = 9, 9
    depth = np.zeros((H, W), np.int32)
    mask = np.zeros((H, W), bool)
    depth[4, 4], depth[4, 6] = 1000, 1040
    mask[4, 4] = mask[4, 6] = True
    out = orc.bilateral(depth, mask, 6, 0.05, 2.0, 0.001)
    w = _two_pixel_weight(4.0, 0.04, 2.0, 0.05)
    assert out[4, 4] == int((1.0 + w * 1.04) / (1 + w) * 1000)
    assert out[4, 6] == int((1.04 + w * 1.0) / (1 + w) * 1000)


def test_bilateral_step_edge_does_not_mix():
    """A depth step far above sigma_color: each side is filtered only with
    its own side (the range weight underflows to zero across the step), so
    the output is the space-weighted mean of the same-side pixels."""
    rng = np.random.default_rng(4)
    H, W = 12, 16
    depth = np.where(np.arange(W)[None, :] < 8, 1.0, 2.0).repeat(H, 0)
    depth = (depth + rng.uniform(-1e-4, 1e-4, size=(H, W))).astype(np.float32)
    mask = np.ones((H, W), bool)
    out = orc.bilateral(depth, mask, 6, 0.03, 4.5, 1.0)
    for r, c in [(5, 7), (5, 8), (0, 7), (11, 8)]:
        num = den = 0.0
        for y in range(max(r - 3, 0), min(r + 4, H)):
            for x in range(max(c - 3, 0), min(c + 4, W)):
                if (x < 8) != (c < 8):
                    continue
                w = _two_pixel_weight((x - c) ** 2 + (y - r) ** 2,
                                      float(depth[y, x]) - float(depth[r, c]), 4.5, 0.03)
                num += w * float(depth[y, x])
                den += w
        np.testing.assert_allclose(out[r, c], num / den, rtol=1e-6)


def test_normals_of_planes_and_isolated_pixels():
    H, W, f = 16, 20, 30.0
    k = np.array([[f, 0, (W - 1) / 2], [0, f, (H - 1) / 2], [0, 0, 1]], np.float32)
    v, u = np.mgrid[0:H, 0:W].astype(np.float64)
    # Plane n . X = d with n facing the camera.
    nrm = np.array([0.2, -0.1, -1.0])
    nrm /= np.linalg.norm(nrm)
    dirs = np.stack([(u - k[0, 2]) / f, (v - k[1, 2]) / f, np.ones_like(u)], -1)
    z = (-2.0) / (dirs @ nrm)
    pts = orc.backproject(z.astype(np.float32), k)
    mask = np.ones((H, W), bool)
    out = orc.normals(pts, mask)
    np.testing.assert_allclose(out[1:-1, 1:-1].reshape(-1, 3),
                               np.tile(nrm, ((H - 2) * (W - 2), 1)), atol=1e-4)
    iso = np.zeros((H, W), bool)
    iso[5, 5] = True
    out = orc.normals(pts, iso)
    np.testing.assert_array_equal(out[5, 5], [0, 0, -1])
    assert np.count_nonzero(out) == 1


def test_normals_of_a_sphere_are_radial():
    H, W, f = 40, 40, 40.0
    k = np.array([[f, 0, 19.5], [0, f, 19.5], [0, 0, 1]], np.float32)
    c, r = np.array([0.0, 0.0, 3.0]), 1.0
    v, u = np.mgrid[0:H, 0:W].astype(np.float64)
    d = np.stack([(u - 19.5) / f, (v - 19.5) / f, np.ones_like(u)], -1)
    a = (d * d).sum(-1)
    b = -2 * d @ c
    disc = b * b - 4 * a * (c @ c - r * r)
    z = np.where(disc > 0, (-b - np.sqrt(np.maximum(disc, 0))) / (2 * a), 0)
    pts = orc.backproject(z.astype(np.float32), k)
    mask = orc.erode(z > 0)
    out = orc.normals(pts, mask)
    inner = orc.erode(orc.erode(mask))
    radial = (pts.astype(np.float64) - c) / r
    cos = (out[inner] * radial[inner]).sum(-1)
    assert cos.min() > 0.99


def test_erode_never_tests_the_pixel_below():
    m = np.ones((5, 5), bool)
    m[3, 2] = False  # below (2, 2): ignored by the reference neighbourhood
    out = orc.erode(m)
    assert out[2, 2]  # survives: the pixel below is not tested
    assert not out[4, 2]  # the pixel above (3,2) is tested
    assert not out[3, 1] and not out[3, 3]  # left/right neighbours tested
    m = np.ones((3, 3), bool)
    assert orc.erode(m).all()  # image borders do not erode


def test_downsample_hand_4x4():
    xyz = np.zeros((4, 4, 3), np.float32)
    xyz[..., 2] = np.array([[1, 2, 5, 5], [3, 4, 5, 9], [1, 1, 7, 7], [1, 1, 7, 8]],
                           np.float32)
    mask = np.ones((4, 4), bool)
    mask[0, 3] = False
    out = orc.downsample_xyz(xyz, mask, 0.5)
    # block (0,0): {1,2,3,4} mean 2.5 -> nearest is 2 (first of the ties 2, 3)
    # block (0,1): valid {5, 5, 9} mean 6.33 -> 5; block (1,0): all 1;
    # block (1,1): {7,7,7,8} mean 7.25 -> 7.
    np.testing.assert_array_equal(out[..., 2], [[2, 5], [1, 7]])
    md = orc.downsample_mask(~mask, 0.5)
    np.testing.assert_array_equal(md, [[False, False], [False, False]])
    mask2 = np.zeros((4, 4), bool)
    mask2[0, 2] = True
    np.testing.assert_array_equal(orc.downsample_mask(mask2, 0.5), [[False, True], [False, False]])
    f = np.arange(32, dtype=np.float32).reshape(2, 4, 4)
    np.testing.assert_array_equal(orc.downsample_features(f, 0.5)[0],
                                  [[2.5, 4.5], [10.5, 12.5]])


def test_gaussian_blur_of_a_delta_is_the_normalised_kernel():
    img = np.zeros((1, 9, 9), np.float32)
    img[0, 4, 4] = 1.0
    out = orc.gaussian_blur3(img, 0.849)
    e = math.exp(-1 / (2 * 0.849 ** 2))
    w = np.array([e, 1, e]) / (1 + 2 * e)
    np.testing.assert_allclose(out[0, 3:6, 3:6], np.outer(w, w), rtol=1e-6)
    assert abs(out.sum() - 1) < 1e-6
    const = np.full((2, 5, 7), 0.3, np.float32)
    np.testing.assert_allclose(orc.gaussian_blur3(const), 0.3, rtol=1e-6)


def test_gaussian_blur_border_is_kornias_reflect_default():
    """feature_pyramid blurs with kornia's GaussianBlur2d((3, 3), (0.849,
    0.849)) (slamtb/processing.py:148), whose published default border is
    'reflect' (mirror without repeating the edge texel). kornia is absent, so
    the oracle is checked against the same operation written with torch:
    reflect padding and a 2-D convolution with the normalised sampled
    Gaussian, on random images including every border texel. Replicate and
    zero padding differ at the borders by far more than the tolerance."""
    import torch
    import torch.nn.functional as F
    rng = np.random.default_rng(12)
    img = rng.uniform(0, 1, size=(3, 17, 23)).astype(np.float32)
    sigma = 0.849
    g = np.exp(-np.arange(-1, 2) ** 2 / (2 * sigma ** 2))
    g /= g.sum()
    k2 = torch.from_numpy(np.outer(g, g))[None, None].repeat(3, 1, 1, 1)
    x = torch.from_numpy(img).double()[None]

    def blur(mode):
        pad = F.pad(x, (1, 1, 1, 1), mode=mode) if mode else F.pad(x, (1, 1, 1, 1))
        return F.conv2d(pad, k2, groups=3)[0].numpy()
    out = orc.gaussian_blur3(img, sigma)
    np.testing.assert_allclose(out, blur("reflect"), rtol=0, atol=2e-7)
    assert np.abs(out - blur("replicate")).max() > 1e-2
    assert np.abs(out - blur(None)).max() > 1e-2
