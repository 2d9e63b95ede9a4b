"""Frame preprocessing for the ICP step (slamtb/processing.py of the reference):
bilateral depth filter, normals, erosion, pyramid downsampling and the
feature pyramid. Every function runs on the HIP device of its inputs,
except the host-side colour conversion of :func:`to_color_feature`.
"""
from enum import Enum

import numpy as np
import torch

from . import _lib
from ._cslamtb import (DownsampleXYZMethod, EstimateNormalsMethod,
                       Processing as _Processing)
from ._utils import depth_image_to_uvz, empty_ensured_size, ensure_torch

FEATURE_BLUR_SIGMA = 0.849


def bilateral_depth_filter(depth, mask=None, out_tensor=None, filter_width=6,
                           sigma_color=29.9999880000072,
                           sigma_space=4.50000000225, depth_scale=1.0):
    """processing.py:17-55 (ElasticFusion defaults; depth in millimetres).
    For depth in metres use sigma_color=0.03."""
    depth = ensure_torch(depth)
    if mask is None:
        mask = depth > 0
    else:
        mask = ensure_torch(mask, dtype=torch.bool)
    out_tensor = empty_ensured_size(out_tensor, depth.size(0), depth.size(1),
                                    dtype=depth.dtype, device=depth.device)
    _Processing.bilateral_depth_filter(depth, mask, out_tensor, filter_width,
                                       sigma_color, sigma_space, depth_scale)
    return out_tensor


class BilateralDepthFilter:
    """Reusable-output form of :func:`bilateral_depth_filter` (processing.py:58-76)."""

    def __init__(self, filter_width=6, sigma_color=29.9999880000072,
                 sigma_space=4.50000000225, depth_scale=1.0):
        self.filter_width = filter_width
        self.sigma_color = sigma_color
        self.sigma_space = sigma_space
        self.depth_scale = depth_scale
        self._out_tensor = None

    def __call__(self, depth, mask):
        self._out_tensor = empty_ensured_size(self._out_tensor, depth.size(0),
                                              depth.size(1), device=depth.device,
                                              dtype=depth.dtype)
        _Processing.bilateral_depth_filter(depth, mask, self._out_tensor,
                                           self.filter_width, self.sigma_color,
                                           self.sigma_space, self.depth_scale)
        return self._out_tensor


def estimate_normals(depth_image, frame_info, mask,
                     method=EstimateNormalsMethod.CentralDifferences,
                     out_tensor=None):
    """processing.py:83-97: normals of the back-projected depth image."""
    image_points = depth_image_to_uvz(ensure_torch(depth_image), frame_info)
    xyz_img = frame_info.kcam.backproject(
        image_points.reshape(-1, 3)).reshape(image_points.shape)
    if out_tensor is None:
        out_tensor = torch.empty(xyz_img.size(0), xyz_img.size(1), 3,
                                 dtype=xyz_img.dtype, device=xyz_img.device)
    _Processing.estimate_normals(xyz_img, ensure_torch(mask, dtype=torch.bool).to(xyz_img.device),
                                 out_tensor, method)
    return out_tensor


def downsample_xyz(xyz_image, mask, scale, dst=None, normalize=False,
                   method=DownsampleXYZMethod.Nearest):
    """processing.py:104-115."""
    dst = empty_ensured_size(dst, int(xyz_image.size(0) * scale),
                             int(xyz_image.size(1) * scale), xyz_image.size(2),
                             dtype=xyz_image.dtype, device=xyz_image.device)
    _Processing.downsample_xyz(xyz_image, mask, scale, dst, normalize, method)
    return dst


def downsample_mask(mask, scale, dst=None):
    """processing.py:118-125."""
    dst = empty_ensured_size(dst, int(mask.size(0) * scale),
                             int(mask.size(1) * scale), dtype=mask.dtype,
                             device=mask.device)
    _Processing.downsample_mask(mask, scale, dst)
    return dst


def downsample_features(feature_image, scale, dst=None):
    """processing.py:128-138: 2x2 box mean per channel."""
    dst = empty_ensured_size(dst, feature_image.size(0),
                             int(feature_image.size(1) * scale),
                             int(feature_image.size(2) * scale),
                             dtype=feature_image.dtype,
                             device=feature_image.device)
    _Processing.downsample_features(feature_image, scale, dst)
    return dst


def gaussian_blur3x3(feature_image, sigma=FEATURE_BLUR_SIGMA):
    """Normalised separable 3x3 Gaussian with reflect borders over
    [C, H, W]; replaces kornia's GaussianBlur2d((3, 3), (s, s)) call of the
    reference feature pyramid (processing.py:148-153)."""
    _lib.require_gpu(feature_image, what="gaussian_blur3x3")
    src = feature_image.contiguous()
    dst = torch.empty_like(src)
    C, H, W = src.shape
    with torch.cuda.device(src.device):
        _lib.check(_lib.lib().stb_gaussian_blur3x3(
            _lib.dtype_code(src), _lib.ptr(src), C, H, W, float(sigma),
            _lib.ptr(dst), _lib.stream()), "gaussian_blur3x3")
    return dst


def feature_pyramid(feature_map, scales):
    """processing.py:141-159: blur then 2x2 mean for each scale < 1,
    returned coarse first."""
    pyramid = []
    for scale in scales:
        if scale < 1.0:
            feature_map = downsample_features(gaussian_blur3x3(feature_map), scale)
        pyramid.append(feature_map)
    pyramid.reverse()
    return pyramid


def erode_mask(in_mask):
    """processing.py:166-171 (keeps the reference neighbourhood, which never
    tests the pixel below)."""
    out_mask = torch.empty(in_mask.size(), dtype=torch.bool,
                           device=in_mask.device)
    _Processing.erode_mask(in_mask, out_mask)
    return out_mask


class ColorSpace(Enum):
    """Colour spaces of :func:`to_color_feature` (processing.py:178-185)."""
    RGB = 0
    GRAY = 1
    INTENSITY = 1
    LAB = 2
    HSV = 3


def _box_blur5(img):
    """5x5 mean with mirrored borders (gfedcb|abcdefgh|gfedcba), rounded
    to the nearest integer as an 8-bit image filter stores it."""
    pad = [(2, 2), (2, 2)] + [(0, 0)] * (img.ndim - 2)
    x = np.pad(img.astype(np.int32), pad, mode="reflect")
    H, W = img.shape[:2]
    acc = np.zeros(img.shape, np.int32)
    for dy in range(5):
        for dx in range(5):
            acc += x[dy:dy + H, dx:dx + W]
    return np.floor(acc / 25.0 + 0.5).astype(np.uint8)


def to_color_feature(rgb_image, color_space=ColorSpace.RGB, blur=False):
    """RGB uint8 [H, W, 3] image -> float32 [C, H, W] features in [0, 1]
    (processing.py:188-206): optional 5x5 box blur, optional conversion to
    gray (0.299 R + 0.587 G + 0.114 B, rounded; C = 1), then /255 in
    channel-first layout. The LAB and HSV conversions of the reference's
    OpenCV call are not provided."""
    img = np.asarray(rgb_image.cpu() if torch.is_tensor(rgb_image) else rgb_image)
    if blur:
        img = _box_blur5(img)
    if color_space == ColorSpace.GRAY:
        y = img[..., 0] * 0.299 + img[..., 1] * 0.587 + img[..., 2] * 0.114
        img = np.floor(y + 0.5).clip(0, 255).astype(np.uint8)
    elif color_space != ColorSpace.RGB:
        raise NotImplementedError(f"to_color_feature: {color_space} is not provided")
    t = torch.from_numpy(np.ascontiguousarray(img)).float().div_(255.0)
    return t.unsqueeze(0) if t.dim() == 2 else t.permute(2, 0, 1).contiguous()


__all__ = ["bilateral_depth_filter", "BilateralDepthFilter", "estimate_normals",
           "downsample_xyz", "downsample_mask", "downsample_features",
           "gaussian_blur3x3", "feature_pyramid", "erode_mask",
           "EstimateNormalsMethod", "DownsampleXYZMethod", "ColorSpace",
           "to_color_feature"]
