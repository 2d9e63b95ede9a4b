"""Frames and frame point clouds (the ICP-facing part of slamtb/frame.py).

``FramePointCloud`` keeps a point cloud on its image grid: (u, v, z) image
points, a validity mask, intrinsics, and lazily computed camera-space points
and normals. Construction, back-projection, normals and the pyramid run on
the HIP device through ``libslamtb_hip.so`` when the frame lives there.
"""
import torch

from . import _lib
from ._cslamtb import EstimateNormalsMethod, Processing as _Processing
from ._utils import depth_image_to_uvz, ensure_torch
from .camera import KCamera
from .processing import downsample_mask, downsample_xyz, erode_mask


class FrameInfo:
    """Camera and depth-scaling description of a frame (frame.py:19-53)."""

    def __init__(self, kcam=None, depth_scale=1.0, depth_bias=0.0,
                 depth_max=4500.0, timestamp=None, rgb_kcam=None, rt_cam=None):
        self.kcam = kcam
        self.depth_scale = depth_scale
        self.depth_bias = depth_bias
        self.depth_max = depth_max
        self.rgb_kcam = rgb_kcam
        self.rt_cam = rt_cam
        self.timestamp = timestamp

    @classmethod
    def from_json(cls, json):
        """FTB JSON form (frame.py:51-76)."""
        from .camera import KCamera, RTCamera
        kcam = json.get('kcam', None)
        rgb_kcam = json.get('rgb_kcam', None)
        rt_cam = json.get('rt_cam', None)
        return cls(KCamera.from_json(kcam) if kcam is not None else None,
                   json.get('depth_scale', 1.0), json.get('depth_bias', 0.0),
                   json.get('depth_max', 4500.0), json.get('timestamp', None),
                   KCamera.from_json(rgb_kcam) if rgb_kcam is not None else None,
                   RTCamera.from_json(rt_cam) if rt_cam is not None else None)

    def to_json(self):
        json = {}
        for name, value in vars(self).items():
            if value is None:
                continue
            if hasattr(value, 'to_json'):
                value = value.to_json()
            json[name] = value
        return json

    def clone(self):
        return FrameInfo(self.kcam.clone() if self.kcam is not None else None,
                         self.depth_scale, self.depth_bias, self.depth_max,
                         self.timestamp, self.rgb_kcam, self.rt_cam)


class Frame:
    """A depth (and optional colour/normal) image with its FrameInfo."""

    def __init__(self, info, depth_image=None, rgb_image=None, seg_image=None,
                 normal_image=None):
        self.info = info
        self.depth_image = depth_image
        self.rgb_image = rgb_image
        self.seg_image = seg_image
        self.normal_image = normal_image

    def clone(self):
        def c(x):
            return None if x is None else x.copy() if hasattr(x, "copy") else x.clone()
        return Frame(self.info.clone(), c(self.depth_image), c(self.rgb_image),
                     c(self.seg_image), c(self.normal_image))


class FramePointCloud:
    """Point cloud embedded in its frame (frame.py:235-546).

    Attributes: ``image_points`` [H, W, 3] (u, v, z) float32, ``mask`` [H, W]
    bool, ``kcam``, ``rt_cam``, lazily ``points`` and ``normals`` [H, W, 3].
    """

    def __init__(self, image_points, mask, kcam, rt_cam=None, points=None,
                 normals=None, colors=None):
        self.image_points = image_points
        self.mask = mask
        self.kcam = kcam
        self.rt_cam = rt_cam
        self._points = points
        self._normals = normals
        self.colors = colors

    @classmethod
    def from_frame(cls, frame, ignore_seg_background=False, device=None):
        """frame.py:272-306: uvz grid, mask = erode(depth > 0)."""
        depth = ensure_torch(frame.depth_image)
        if device is not None:
            depth = depth.to(device)
        image_points = depth_image_to_uvz(depth, frame.info)
        mask = erode_mask(depth > 0)
        if frame.seg_image is not None and ignore_seg_background:
            seg = ensure_torch(frame.seg_image).to(mask.device)
            mask = torch.logical_and(seg > 0, mask)
        colors = None
        if frame.rgb_image is not None:
            colors = ensure_torch(frame.rgb_image).to(mask.device)
        normals = None
        if frame.normal_image is not None:
            normals = ensure_torch(frame.normal_image).to(mask.device)
        return cls(image_points, mask, frame.info.kcam, frame.info.rt_cam,
                   normals=normals, colors=colors)

    @property
    def device(self):
        return self.mask.device

    @property
    def points(self):
        """Camera-space points [H, W, 3] (frame.py:308-322)."""
        if self._points is None:
            if self.kcam is None:
                raise RuntimeError("Frame doesn't have intrinsics camera")
            kc = self.kcam
            ip = self.image_points
            if ip.is_cuda and ip.dtype == torch.float32:
                m = kc.matrix
                pts = torch.empty_like(ip.contiguous())
                depth = ip[:, :, 2].contiguous()
                with torch.cuda.device(ip.device):
                    _lib.check(_lib.lib().stb_depth_to_points(
                        _lib.STB_F32, _lib.ptr(depth), ip.size(0), ip.size(1),
                        float(m[0, 0]), float(m[1, 1]), float(m[0, 2]),
                        float(m[1, 2]), 1.0, 0.0, _lib.ptr(pts),
                        _lib.stream()), "FramePointCloud.points")
                self._points = pts
            else:
                self._points = kc.backproject(
                    ip.reshape(-1, 3)).reshape(ip.shape)
        return self._points

    @property
    def normals(self):
        """Central-difference normals, computed on first use (frame.py:324-342)."""
        if self._normals is None:
            pts = self.points
            self._normals = torch.empty(pts.size(0), pts.size(1), 3,
                                        dtype=pts.dtype, device=pts.device)
            _Processing.estimate_normals(pts, self.mask, self._normals,
                                         EstimateNormalsMethod.CentralDifferences)
        return self._normals

    @normals.setter
    def normals(self, normals):
        self._normals = normals

    @property
    def width(self):
        return self.mask.size(1)

    @property
    def height(self):
        return self.mask.size(0)

    def downsample(self, scale, downsample_xyz_method=0, colored=False):
        """frame.py:388-422: nearest-to-mean xyz, top-left mask, scaled K.
        Normals are carried over only when already computed."""
        points = downsample_xyz(self.points, self.mask, scale,
                                method=downsample_xyz_method)
        normals = None
        if self._normals is not None:
            normals = downsample_xyz(self._normals, self.mask, scale,
                                     normalize=True, method=downsample_xyz_method)
        mask = downsample_mask(self.mask, scale)
        colors = None
        if colored and self.colors is not None:
            colors = torch.nn.functional.interpolate(
                self.colors.permute(2, 0, 1).unsqueeze(0).float(),
                scale_factor=scale, mode="bilinear", align_corners=False)
            colors = colors.squeeze(0).permute(1, 2, 0).byte()
        kcam = self.kcam.scaled(scale)
        return FramePointCloud(None, mask, kcam, rt_cam=self.rt_cam,
                               points=points, normals=normals, colors=colors)

    def pyramid(self, scales, downsample_xyz_method=0, colored=False):
        """frame.py:424-453: cumulative downsampling, returned coarse first."""
        pyramid = []
        curr = self
        for scale in scales:
            if scale < 1.0:
                curr = curr.downsample(scale, downsample_xyz_method, colored)
            pyramid.append(curr)
        pyramid.reverse()
        return pyramid

    def to(self, dst):
        """Device/dtype change of the arrays (dtype applies to points and
        normals only)."""
        def mv(x, fp=False):
            if x is None:
                return None
            if fp and isinstance(dst, torch.dtype):
                return x.to(dst)
            if isinstance(dst, torch.dtype):
                return x
            return x.to(dst)
        kcam = self.kcam
        return FramePointCloud(mv(self.image_points, True), mv(self.mask),
                               kcam, self.rt_cam, mv(self._points, True),
                               mv(self._normals, True), mv(self.colors))

    def clone(self):
        def c(x):
            return None if x is None else x.clone()
        return FramePointCloud(c(self.image_points), c(self.mask), self.kcam,
                               self.rt_cam, c(self._points), c(self._normals),
                               c(self.colors))


__all__ = ["Frame", "FrameInfo", "FramePointCloud", "KCamera"]
