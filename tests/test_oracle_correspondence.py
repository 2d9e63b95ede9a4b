"""Pins the oracle's correspondence map (src/correspondence_map.cu:38-61,
src/correspondence.hpp:69-89, src/merge_map.hpp:37-50) with closed-form
scenes, gate edge cases, the tie rule and a per-source brute force."""
import math

import numpy as np

from oracle import oracle as orc

INF = orc.CORRESP_INF


def plane_scene(H=12, W=16, z=2.0, f=20.0):
    k = np.array([[f, 0, (W - 1) / 2], [0, f, (H - 1) / 2], [0, 0, 1]], np.float32)
    depth = np.full((H, W), z, np.float32)
    pts = orc.backproject(depth, k)
    nrm = np.zeros_like(pts)
    nrm[..., 2] = -1.0
    mask = np.ones((H, W), bool)
    return k, pts, nrm, mask


def idx_of(m):
    return (m >> 32).astype(np.int64)


def dist_of(m):
    return (m & 0xffffffff).astype(np.uint32).view(np.float32)


def test_identity_maps_every_pixel_to_itself():
    k, p, n, m = plane_scene()
    cmap = orc.corresp_map(p, n, m, np.eye(4), p, n, m, k, 0.1, 0.3)
    assert np.array_equal(idx_of(cmap).ravel(), np.arange(cmap.size))
    assert np.all(dist_of(cmap) == 0.0)


def test_one_pixel_shift_is_a_column_permutation():
    H, W, z, f = 12, 16, 2.0, 20.0
    k, p, n, m = plane_scene(H, W, z, f)
    T = np.eye(4)
    T[0, 3] = z / f  # moves every point one pixel to the right
    cmap = orc.corresp_map(p, n, m, T, p, n, m, k, 0.1, 0.3)
    idx = idx_of(cmap)
    assert np.all(cmap[:, 0] == INF)
    rows, cols = np.mgrid[0:H, 1:W]
    assert np.array_equal(idx[:, 1:], rows * W + cols - 1)
    assert np.all(dist_of(cmap[:, 1:]) < 1e-5)


def test_distance_gate_is_inclusive_at_the_threshold():
    k, p, n, m = plane_scene()
    src = p.copy()
    # Source plane 0.09 m or 0.11 m behind the target plane, alternating rows
    # (squared distance compared against float(0.1^2)).
    off = np.where((np.arange(12) % 2 == 0)[:, None], 0.09, 0.11)
    src[..., 2] += off.astype(np.float32)
    src[..., 0] = (np.arange(16)[None, :] - k[0, 2]) * src[..., 2] / k[0, 0]
    src[..., 1] = (np.arange(12)[:, None] - k[1, 2]) * src[..., 2] / k[1, 1]
    cmap = orc.corresp_map(src, n, m, np.eye(4), p, n, m, k, 0.1, 0.3)
    matched = cmap != INF
    assert np.all(matched[0::2]) and not np.any(matched[1::2])


def test_angle_gate_rejects_at_and_above_the_threshold():
    k, p, n, m = plane_scene()
    thr = math.radians(20)
    sn = n.copy()
    ang = np.where((np.arange(16) % 2 == 0)[None, :], math.radians(10),
                   math.radians(30))
    ang = np.broadcast_to(ang, (12, 16))
    sn[..., 0] = np.sin(ang)
    sn[..., 2] = -np.cos(ang)
    cmap = orc.corresp_map(p, sn, m, np.eye(4), p, n, m, k, 0.1, thr)
    matched = cmap != INF
    assert np.all(matched[:, 0::2]) and not np.any(matched[:, 1::2])
    # Threshold 0 rejects everything (|acos| >= 0); above pi nothing.
    assert np.all(orc.corresp_map(p, n, m, np.eye(4), p, n, m, k, 0.1, 0.0) == INF)
    assert np.all(orc.corresp_map(p, sn, m, np.eye(4), p, n, m, k, 0.1, 3.5) != INF)


def test_masks_and_out_of_view():
    k, p, n, m = plane_scene()
    sm = m.copy()
    sm[:, :8] = False
    tm = m.copy()
    tm[:6] = False
    cmap = orc.corresp_map(p, n, sm, np.eye(4), p, n, tm, k, 0.1, 0.3)
    matched = cmap != INF
    expect = np.zeros_like(m)
    expect[6:, 8:] = True
    assert np.array_equal(matched, expect)
    far = np.eye(4)
    far[0, 3] = 100.0
    assert np.all(orc.corresp_map(p, n, m, far, p, n, m, k, 1e9, 3.5) == INF)
    # A source mask of another length is ignored (correspondence_map.cu:40-41).
    cmap1 = orc.corresp_map(p, n, np.zeros(1, bool), np.eye(4), p, n, m, k, 0.1, 0.3)
    assert np.all(cmap1 != INF)


def test_equal_distances_keep_the_lowest_index():
    k = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1]], np.float32)
    tgt = np.array([[[0.0, 0.0, 1.0]]], np.float32)
    tn = np.array([[[0.0, 0.0, -1.0]]], np.float32)
    src = np.array([[0.01, 0.0, 1.0], [0.0, 0.01, 1.0], [-0.01, 0.0, 1.0]], np.float32)
    sn = np.tile(np.array([0.0, 0.0, -1.0], np.float32), (3, 1))
    cmap = orc.corresp_map(src, sn, np.ones(3, bool), np.eye(4), tgt, tn,
                           np.ones((1, 1), bool), k, 1.0, 1.0)
    assert idx_of(cmap)[0, 0] == 0
    # A strictly closer later source wins.
    src[2, 0] = -0.005
    cmap = orc.corresp_map(src, sn, np.ones(3, bool), np.eye(4), tgt, tn,
                           np.ones((1, 1), bool), k, 1.0, 1.0)
    assert idx_of(cmap)[0, 0] == 2


def test_scatter_min_equals_per_source_brute_force():
    rng = np.random.default_rng(0)
    H, W = 12, 16
    k = np.array([[12, 0, 7.5], [0, 12, 5.5], [0, 0, 1]], np.float32)
    depth = rng.uniform(1.0, 1.3, (H, W)).astype(np.float32)
    pts = orc.backproject(depth, k)
    mask = rng.uniform(size=(H, W)) > 0.1
    nrm = orc.normals(pts, mask)
    T = np.eye(4)
    T[:3, 3] = [0.03, -0.02, 0.05]
    full = orc.corresp_map(pts, nrm, mask, T, pts, nrm, mask, k, 0.2, 1.2)
    # Each source alone gives its own candidate (pixel, key); the merged map
    # must hold, per pixel, the lexicographic min of (distance, index).
    best = {}
    for i in range(H * W):
        one = np.zeros(H * W, bool)
        one[i] = mask.ravel()[i]
        if not one.any():
            continue
        m1 = orc.corresp_map(pts, nrm, one, T, pts, nrm, mask, k, 0.2, 1.2)
        hit = np.flatnonzero(m1.ravel() != INF)
        assert len(hit) <= 1
        if len(hit):
            pos = hit[0]
            d = dist_of(m1.ravel()[pos])
            if pos not in best or (d, i) < best[pos]:
                best[pos] = (d, i)
    assert len(best) > 50
    flat = full.ravel()
    for pos in range(H * W):
        if pos in best:
            assert idx_of(flat[pos]) == best[pos][1]
            assert dist_of(flat[pos]) == best[pos][0]
        else:
            assert flat[pos] == INF
