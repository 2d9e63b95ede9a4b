"""Seeded synthetic RGB-D frame pairs for the ICP step.

A scene is 6 random planes and 4 random spheres, ray-cast through a pinhole
camera into a float32 depth map in metres (valid range [0.5, 4.0], invalid
pixels are 0) with Gaussian depth noise of standard deviation 0.0012 z^2.
The target frame renders the same scene moved by a random rigid transform
``T_gt`` (rotation angle at most ``max_angle_deg`` about a uniform axis,
translation at most ``max_trans`` metres), so registering the source frame
onto the target recovers ``T_gt``: ``p_target = T_gt @ p_source``.

Features are attached to the surfaces: each surface has a random albedo per
channel in [0, 1] plus a smooth sinusoidal texture over its 3D position, so
they move with the scene and have non-zero image gradients.  With
``descriptor=True`` the channels are instead per-surface Gaussian N(0, 1)
vectors with a smooth perturbation, normalised to unit length per pixel.
"""
import math

import numpy as np

DEPTH_MIN = 0.5
DEPTH_MAX = 4.0


def intrinsics(width, height, fx=None):
    """Pinhole matrix with fx = fy = 525 * W / 640 and the centre at the
    middle of the pixel grid ((W - 1) / 2, (H - 1) / 2)."""
    if fx is None:
        fx = 525.0 * width / 640.0
    return np.array([[fx, 0.0, (width - 1) / 2.0],
                     [0.0, fx, (height - 1) / 2.0],
                     [0.0, 0.0, 1.0]], dtype=np.float32)


def random_se3(rng, max_angle_deg=5.0, max_trans=0.05):
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    angle = math.radians(max_angle_deg) * rng.uniform()
    K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]],
                  [-axis[1], axis[0], 0]])
    R = np.eye(3) + math.sin(angle) * K + (1 - math.cos(angle)) * (K @ K)
    tdir = rng.normal(size=3)
    tdir /= np.linalg.norm(tdir)
    T = np.eye(4)
    T[:3, :3] = R
    T[:3, 3] = tdir * max_trans * rng.uniform()
    return T


class Scene:
    """Random planes and spheres in the source camera frame."""

    def __init__(self, rng, channels=3, descriptor=False):
        self.planes = []
        for _ in range(6):
            tilt = math.radians(rng.uniform(0, 50))
            az = rng.uniform(0, 2 * math.pi)
            # Normal facing the camera: mostly -z, tilted by up to 50 deg.
            n = np.array([math.sin(tilt) * math.cos(az),
                          math.sin(tilt) * math.sin(az), -math.cos(tilt)])
            z0 = rng.uniform(2.2, 3.8)
            p0 = np.array([rng.uniform(-0.6, 0.6) * z0,
                           rng.uniform(-0.45, 0.45) * z0, z0])
            self.planes.append((n, p0))
        self.spheres = []
        for _ in range(4):
            z = rng.uniform(1.2, 3.0)
            c = np.array([rng.uniform(-0.45, 0.45) * z,
                          rng.uniform(-0.35, 0.35) * z, z])
            self.spheres.append((c, rng.uniform(0.15, 0.45)))
        nsurf = len(self.planes) + len(self.spheres)
        self.channels = channels
        self.descriptor = descriptor
        if descriptor:
            self.albedo = rng.normal(size=(nsurf, channels))
        else:
            self.albedo = rng.uniform(0, 1, size=(nsurf, channels))
        self.tex_dir = rng.normal(size=(channels, 3))
        self.tex_dir /= np.linalg.norm(self.tex_dir, axis=1, keepdims=True)
        self.tex_freq = rng.uniform(12.0, 25.0, size=channels)
        self.tex_phase = rng.uniform(0, 2 * math.pi, size=channels)

    def render(self, kcam, width, height, T=None, rng=None, noise=True):
        """Ray-cast the scene seen after moving it by ``T``.

        Returns depth [H, W] float32 and features [C, H, W] float32.
        """
        if T is None:
            T = np.eye(4)
        R, t = T[:3, :3], T[:3, 3]
        fx, fy, cx, cy = (float(kcam[0, 0]), float(kcam[1, 1]),
                          float(kcam[0, 2]), float(kcam[1, 2]))
        v, u = np.meshgrid(np.arange(height, dtype=np.float64),
                           np.arange(width, dtype=np.float64), indexing="ij")
        d = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)], -1)
        best = np.full((height, width), np.inf)
        surf = np.full((height, width), -1, dtype=np.int64)
        for i, (n, p0) in enumerate(self.planes):
            n2 = R @ n
            p2 = R @ p0 + t
            den = d @ n2
            with np.errstate(divide="ignore", invalid="ignore"):
                z = np.where(np.abs(den) > 1e-9, (n2 @ p2) / den, np.inf)
            z = np.where(z > 0, z, np.inf)
            upd = z < best
            best = np.where(upd, z, best)
            surf = np.where(upd, i, surf)
        for j, (c, r) in enumerate(self.spheres):
            c2 = R @ c + t
            a = np.einsum("hwk,hwk->hw", d, d)
            b = -2.0 * (d @ c2)
            cc = c2 @ c2 - r * r
            disc = b * b - 4 * a * cc
            with np.errstate(invalid="ignore"):
                z = (-b - np.sqrt(disc)) / (2 * a)
            z = np.where((disc > 0) & (z > 0), z, np.inf)
            upd = z < best
            best = np.where(upd, z, best)
            surf = np.where(upd, len(self.planes) + j, surf)
        valid = np.isfinite(best) & (best >= DEPTH_MIN) & (best <= DEPTH_MAX)
        z = np.where(valid, best, 0.0)
        # Features from the surface point in the scene (source) frame.
        X = d * z[..., None]
        Xw = (X - t) @ R  # R^T (X - t)
        feats = np.zeros((self.channels, height, width))
        sidx = np.where(valid, surf, 0)
        for ch in range(self.channels):
            tex = np.sin(self.tex_freq[ch] * (Xw @ self.tex_dir[ch])
                         + self.tex_phase[ch])
            if self.descriptor:
                f = self.albedo[sidx, ch] + 0.5 * tex
            else:
                f = np.clip(0.7 * self.albedo[sidx, ch] + 0.15 * (tex + 1.0), 0, 1)
            feats[ch] = np.where(valid, f, 0.0)
        if self.descriptor:
            norm = np.sqrt((feats ** 2).sum(0))
            feats = np.where(norm > 0, feats / np.maximum(norm, 1e-12), 0.0)
        if noise:
            rng = rng if rng is not None else np.random.default_rng(0)
            z = z + rng.normal(size=z.shape) * 0.0012 * z * z
            z = np.where(valid, np.clip(z, DEPTH_MIN, DEPTH_MAX), 0.0)
        return z.astype(np.float32), feats.astype(np.float32)


def make_pair(seed, width, height, kcam=None, channels=3, descriptor=False,
              max_angle_deg=5.0, max_trans=0.05, noise=True):
    """A seeded (source, target) frame pair and its ground-truth motion.

    Returns a dict with ``kcam`` [3,3] float32, ``src_depth``/``tgt_depth``
    [H,W] float32, ``src_feats``/``tgt_feats`` [C,H,W] float32 and ``T_gt``
    [4,4] float64 with ``p_target = T_gt @ p_source``.
    """
    rng = np.random.default_rng(seed)
    if kcam is None:
        kcam = intrinsics(width, height)
    scene = Scene(rng, channels=channels, descriptor=descriptor)
    T_gt = random_se3(rng, max_angle_deg, max_trans)
    src_depth, src_feats = scene.render(kcam, width, height, None, rng, noise)
    tgt_depth, tgt_feats = scene.render(kcam, width, height, T_gt, rng, noise)
    return dict(kcam=np.asarray(kcam, np.float32), src_depth=src_depth,
                tgt_depth=tgt_depth, src_feats=src_feats, tgt_feats=tgt_feats,
                T_gt=T_gt, seed=seed, width=width, height=height)
