"""Pinhole intrinsics (the KCamera part of slamtb/camera.py:19-277)."""
import math

import torch

from ._utils import ensure_torch


class KCamera:
    """Intrinsic pinhole camera: ``matrix`` is a [3, 3] float32 tensor."""

    def __init__(self, matrix, undist_coeff=None, image_size=None):
        self.matrix = ensure_torch(matrix)
        self.undist_coeff = undist_coeff if undist_coeff is not None else []
        self.image_size = image_size

    @classmethod
    def from_params(cls, flen_x, flen_y, center_point, undist_coeff=None,
                    image_size=None):
        """camera.py:78-103: K = translate(c) @ scale(f) in float32."""
        center_x, center_y = center_point
        k_trans = torch.tensor([[1.0, 0.0, float(center_x)],
                                [0.0, 1.0, float(center_y)],
                                [0.0, 0.0, 1.0]])
        k_scale = torch.tensor([[float(flen_x), 0.0, 0.0],
                                [0.0, float(flen_y), 0.0],
                                [0.0, 0.0, 1.0]])
        return cls(k_trans @ k_scale, undist_coeff, image_size)

    @classmethod
    def from_estimation_by_fov(cls, hfov, vfov, img_width, img_height):
        flen_x = img_width * .5 / math.tan(hfov / 2)
        flen_y = img_height * .5 / math.tan(vfov / 2)
        return cls.from_params(flen_x, flen_y, (img_width * .5, img_height * .5))

    @property
    def pixel_center(self):
        return (self.matrix[0, 2].item(), self.matrix[1, 2].item())

    def backproject(self, points):
        """[N, 3] (u, v, z) -> camera-space xyz (camera.py:127-149)."""
        xyz = points.clone()
        fx, fy = self.matrix[0, 0], self.matrix[1, 1]
        cx, cy = self.matrix[0, 2], self.matrix[1, 2]
        z = xyz[:, 2]
        xyz[:, 0] = (xyz[:, 0] - cx) * z / fx
        xyz[:, 1] = (xyz[:, 1] - cy) * z / fy
        return xyz

    def project(self, points):
        """[N, 3] xyz -> [N, 3] (u, v, z) with division by z (camera.py:151-170)."""
        points = (self.matrix.to(points.device, points.dtype)
                  @ points.reshape(-1, 3, 1)).reshape(-1, 3)
        z = points[:, 2]
        points[:, :2] /= z.reshape(-1, 1)
        return points

    def scaled(self, xscale, yscale=None):
        """Intrinsics for a rescaled image (camera.py:172-198): fx, cx scale
        by xscale and fy, cy by yscale, computed in float32."""
        if yscale is None:
            yscale = xscale
        image_size = None
        if self.image_size is not None:
            image_size = (int(self.image_size[0] * xscale),
                          int(self.image_size[1] * yscale))
        m = self.matrix
        return KCamera.from_params(
            (m[0, 0] * xscale).item(), (m[1, 1] * yscale).item(),
            ((m[0, 2] * xscale).item(), (m[1, 2] * yscale).item()),
            self.undist_coeff, image_size)

    def clone(self):
        return KCamera(self.matrix.clone(), list(self.undist_coeff),
                       self.image_size)

    def to(self, device):
        return KCamera(self.matrix.to(device), self.undist_coeff, self.image_size)

    def __str__(self):
        return f"KCamera({self.matrix.tolist()})"
