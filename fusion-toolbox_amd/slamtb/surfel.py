"""Surfel clouds (the ``SurfelCloud`` part of slamtb/surfel.py): points with
normals, colours, radii, confidences, times and optional per-point features,
and the voxel merge (``downsample``) that the multiscale point-cloud
registration uses. The surfel model, its allocator and the OpenGL
rasterisers belong to surfel fusion and are not provided.
"""
import math

import torch

from . import _lib
from .camera import RigidTransform


def compute_surfel_radii(cam_points, normals, kcam):
    """r = min((1/sqrt 2)(z / f) / |n_z|, 2 (1/sqrt 2) z / f) with f the mean
    focal length (surfel.py:20-45, src/surfel.cu:64-85)."""
    k = kcam.matrix
    f = (abs(float(k[0, 0])) + abs(float(k[1, 1]))) * 0.5
    radius = 0.7071067811865475 * (cam_points[:, 2] / f)
    return torch.minimum(radius / normals[:, 2].abs(), 2 * radius)


def compute_confidences(kcam, weight, width, height, device):
    """exp(-(d / d_max)^2 / (2 * 0.6^2)) * weight with d the pixel's distance
    to the principal point (surfel.py:48-62, src/surfel.cu:11-36)."""
    cx, cy = kcam.pixel_center
    max_dist = math.sqrt((width - cx) ** 2 + (height - cy) ** 2)
    v, u = torch.meshgrid(torch.arange(height, dtype=torch.float32, device=device),
                          torch.arange(width, dtype=torch.float32, device=device),
                          indexing="ij")
    d = torch.sqrt((u - cx) ** 2 + (v - cy) ** 2) / max_dist
    return torch.exp(-(d * d) / (2.0 * 0.6 ** 2)) * weight


class SurfelCloud:
    """points [N,3], confidences [N], normals [N,3], radii [N], colors [N,3]
    uint8, times [N] int32, features [F,N] or None (surfel.py:82-112)."""

    def __init__(self, points, confidences, normals, radii, colors, times,
                 features=None, sparse_features=None):
        self.points = points
        self.confidences = confidences
        self.normals = normals
        self.radii = radii
        self.colors = colors
        self.times = times
        self.features = features
        self.sparse_features = sparse_features

    @classmethod
    def from_frame_pcl(cls, frame_pcl, confidences=None, confidence_weight=0.5,
                       time=0, features=None):
        """Surfels of the valid pixels of a FramePointCloud, in camera space
        (surfel.py:114-165)."""
        mask = frame_pcl.mask.reshape(-1)
        points = frame_pcl.points.reshape(-1, 3)[mask]
        normals = frame_pcl.normals.reshape(-1, 3)[mask]
        device = points.device
        if confidences is None:
            confidences = compute_confidences(
                frame_pcl.kcam, confidence_weight, frame_pcl.width,
                frame_pcl.height, device).reshape(-1)[mask]
        radii = compute_surfel_radii(points, normals, frame_pcl.kcam)
        if isinstance(time, int):
            time = torch.full((points.size(0),), time, dtype=torch.int32,
                              device=device)
        colors = (frame_pcl.colors.reshape(-1, 3)[mask].to(device)
                  if frame_pcl.colors is not None else
                  torch.zeros(points.size(0), 3, dtype=torch.uint8, device=device))
        if features is not None:
            features = features.reshape(features.size(0), -1)[:, mask]
        return cls(points, confidences, normals, radii, colors, time, features)

    @classmethod
    def from_frame(cls, frame, confidences=None, confidence_weight=0.5, time=0,
                   features=None, device="cuda:0"):
        from .frame import FramePointCloud
        return cls.from_frame_pcl(FramePointCloud.from_frame(frame, device=device),
                                  confidences, confidence_weight, time, features)

    @property
    def device(self):
        return self.points.device

    @property
    def size(self):
        return self.points.size(0)

    def __len__(self):
        return self.size

    @property
    def has_features(self):
        return self.features is not None

    @property
    def feature_size(self):
        return 0 if self.features is None else self.features.size(0)

    def clone(self):
        c = lambda t: None if t is None else t.clone()
        return SurfelCloud(c(self.points), c(self.confidences), c(self.normals),
                           c(self.radii), c(self.colors), c(self.times),
                           c(self.features))

    def transform(self, matrix):
        """New cloud with points and normals moved by a rigid matrix
        (surfel.py:291-310)."""
        rt = RigidTransform(matrix.to(self.device, self.points.dtype))
        return SurfelCloud(rt @ self.points, self.confidences,
                           rt.transform_normals(self.normals), self.radii,
                           self.colors, self.times, self.features)

    def itransform(self, matrix):
        rt = RigidTransform(matrix.to(self.device, self.points.dtype))
        rt.inplace(self.points)
        rt.inplace_normals(self.normals)

    def to(self, device):
        mv = lambda t: None if t is None else t.to(device)
        return SurfelCloud(mv(self.points), mv(self.confidences), mv(self.normals),
                           mv(self.radii), mv(self.colors), mv(self.times),
                           mv(self.features))

    def __getitem__(self, *args):
        idx = args[0]
        g = lambda t: None if t is None else t[idx]
        return SurfelCloud(g(self.points), g(self.confidences), g(self.normals),
                           g(self.radii), g(self.colors), g(self.times),
                           None if self.features is None else self.features[:, idx])

    def downsample(self, voxel_size):
        """Merge the surfels of each voxel of a grid anchored at the cloud's
        minimum (surfel.py:373-399, SurfelVolume): the mean of each voxel's
        points, normals, colours (truncated), confidences, radii and
        features; times 0, as the reference leaves them. Voxels come out in
        increasing voxel id (the reference's order is its hash map's).
        ``voxel_size <= 0`` returns self. Runs on the HIP device."""
        if voxel_size <= 0.0:
            return self
        what = "SurfelCloud.downsample"
        _lib.require_gpu(self.points, what=what)
        pts = self.points.float().contiguous()
        n = pts.size(0)
        dev = pts.device
        if n == 0:
            return self
        mn = pts.min(0)[0].tolist()
        mx = pts.max(0)[0].tolist()
        vs = torch.tensor(voxel_size, dtype=torch.float32).item()
        f32 = lambda x: torch.tensor(x, dtype=torch.float32)
        # Truncated extents, in float32 as the reference's Eigen::Vector3f.
        dims = [int(((f32(b) - f32(a)) / f32(vs)).item()) for a, b in zip(mn, mx)]
        ids = torch.empty(n, dtype=torch.int32, device=dev)
        with torch.cuda.device(dev):
            _lib.check(_lib.lib().stb_voxel_ids(
                _lib.ptr(pts), n, mn[0], mn[1], mn[2], vs, dims[0], dims[1],
                _lib.ptr(ids), _lib.stream()), what)
        sorted_ids, order = torch.sort(ids, stable=True)
        _, counts = torch.unique_consecutive(sorted_ids, return_counts=True)
        V = counts.numel()
        starts = torch.zeros(V + 1, dtype=torch.int64, device=dev)
        starts[1:] = torch.cumsum(counts, 0)
        C = self.feature_size
        feats = None if C == 0 else self.features.float().contiguous()
        out = SurfelCloud(torch.empty(V, 3, device=dev),
                          torch.empty(V, device=dev),
                          torch.empty(V, 3, device=dev),
                          torch.empty(V, device=dev),
                          torch.empty(V, 3, dtype=torch.uint8, device=dev),
                          torch.empty(V, dtype=torch.int32, device=dev),
                          None if C == 0 else torch.empty(C, V, device=dev))
        c = lambda t, dt=torch.float32: t.to(dev, dt).contiguous()
        srcs = (c(self.normals), c(self.colors, torch.uint8), c(self.confidences),
                c(self.radii))
        with torch.cuda.device(dev):
            _lib.check(_lib.lib().stb_voxel_merge(
                _lib.ptr(pts), _lib.ptr(srcs[0]), _lib.ptr(srcs[1]),
                _lib.ptr(srcs[2]), _lib.ptr(srcs[3]), _lib.ptr(feats), C, n,
                _lib.ptr(order), _lib.ptr(starts), V, _lib.ptr(out.points),
                _lib.ptr(out.normals), _lib.ptr(out.colors),
                _lib.ptr(out.confidences), _lib.ptr(out.radii),
                _lib.ptr(out.times), _lib.ptr(out.features), _lib.stream()), what)
        return out
