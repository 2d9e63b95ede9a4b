"""Pinhole intrinsics, the differentiable projection and rigid transforms
(slamtb/camera.py:19-403)."""
import math

import torch

from ._cslamtb import ProjectOp as _ProjectOp, RigidTransformOp as _RigidTransformOp
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
    def from_json(cls, json):
        """FTB JSON form (camera.py:45-52)."""
        return cls(torch.tensor(json['matrix'], dtype=torch.float).view(-1, 3),
                   undist_coeff=json.get('undist_coeff', None),
                   image_size=json.get('image_size', None))

    def to_json(self):
        json = {'matrix': self.matrix.tolist()}
        if self.undist_coeff is not None:
            json['undist_coeff'] = self.undist_coeff
        if self.image_size is not None:
            json['image_size'] = list(self.image_size)
        return json

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


class RTCamera:
    """Extrinsic camera: ``matrix`` is the [4, 4] float64 camera-to-world
    transform (camera.py:420-607)."""

    def __init__(self, matrix=None):
        if matrix is not None:
            self.matrix = ensure_torch(matrix, dtype=torch.double)
        else:
            self.matrix = torch.eye(4, dtype=torch.double)

    @classmethod
    def create_from_pos_rot(cls, position, rotation_matrix):
        m = torch.eye(4, dtype=torch.double)
        m[:3, :3] = torch.as_tensor(rotation_matrix, dtype=torch.double)
        m[:3, 3] = torch.as_tensor(position, dtype=torch.double)
        return cls(m)

    @classmethod
    def create_from_pos_quat(cls, x, y, z, qw, qx, qy, qz):
        m = torch.eye(4, dtype=torch.double)
        m[:3, :3] = quaternion_to_matrix(qw, qx, qy, qz)
        m[:3, 3] = torch.tensor([x, y, z], dtype=torch.double)
        return cls(m)

    @classmethod
    def from_json(cls, json):
        return cls(torch.tensor(json['matrix'], dtype=torch.double).view(-1, 4))

    def to_json(self):
        return {'matrix': self.matrix.tolist()}

    @property
    def cam_to_world(self):
        return self.matrix

    @property
    def world_to_cam(self):
        return torch.inverse(self.matrix)

    def right_transform(self, matrix):
        """matrix @ self.matrix (camera.py:525-535)."""
        return RTCamera(torch.as_tensor(matrix).cpu().double() @ self.matrix)

    def transform(self, matrix):
        """self.matrix @ matrix (camera.py:537-547)."""
        return RTCamera(self.matrix @ torch.as_tensor(matrix).cpu().double())

    def __matmul__(self, other):
        return RTCamera(self.matrix @ other.matrix.double())

    def translate(self, tx, ty, tz):
        t = torch.eye(4, dtype=torch.double)
        t[:3, 3] = torch.tensor([tx, ty, tz], dtype=torch.double)
        return RTCamera(self.matrix @ t)

    def difference(self, other):
        """Relative transform world_to_cam(self) @ other (the reference
        multiplies elementwise here, camera.py:571-582, which is not a
        transform; the matrix product is what its docstring describes)."""
        return RigidTransform(self.world_to_cam @ other.matrix)

    @property
    def center(self):
        return (self.matrix[0, 3].item(), self.matrix[1, 3].item(),
                self.matrix[2, 3].item())

    def clone(self):
        return RTCamera(self.matrix.clone())

    def __str__(self):
        return f"RTCamera({self.matrix.tolist()})"

    __repr__ = __str__


def quaternion_to_matrix(qw, qx, qy, qz):
    """[3, 3] float64 rotation of the quaternion w + xi + yj + zk (normalised
    first), R = I + 2w[v]x + 2[v]x^2 with v = (x, y, z)."""
    q = torch.tensor([qw, qx, qy, qz], dtype=torch.double)
    w, x, y, z = (q / q.norm()).tolist()
    return torch.tensor([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]],
        dtype=torch.double)


class Project(torch.autograd.Function):
    """Differentiable pinhole projection (camera.py:280-308): [N, 3] camera
    points and a [3, 3] intrinsic matrix -> [N, 2] image points, both
    passes on the HIP device."""

    @staticmethod
    def forward(ctx, points, intrinsics):
        ctx.save_for_backward(points, intrinsics)
        return _ProjectOp.forward(points, intrinsics)

    @staticmethod
    def backward(ctx, dy_grad):
        points, intrinsics = ctx.saved_tensors
        return _ProjectOp.backward(dy_grad, points, intrinsics), None


class _RigidPoints(torch.autograd.Function):
    """p' = R p + t on the HIP device, differentiable in the matrix and the
    points: one transform launch forward, one transpose-and-reduce launch
    pair backward (FP64 sums in a fixed order)."""

    @staticmethod
    def forward(ctx, matrix, points):
        out = torch.empty_like(points)
        _RigidTransformOp.transform(matrix, points, out)
        ctx.save_for_backward(matrix, points)
        return out

    @staticmethod
    def backward(ctx, dy):
        matrix, points = ctx.saved_tensors
        dx = torch.empty_like(points) if ctx.needs_input_grad[1] else None
        d34 = torch.empty(3, 4, dtype=points.dtype, device=points.device)
        _RigidTransformOp.transform_backward(matrix, points, dy, dx, d34)
        dm = None
        if ctx.needs_input_grad[0]:
            dm = d34 if matrix.size(0) == 3 else torch.cat(
                [d34, torch.zeros(1, 4, dtype=d34.dtype, device=d34.device)])
            dm = dm.to(matrix.dtype)
        return dm, dx


def _on_gpu(*tensors):
    return all(t.is_cuda and t.dtype in (torch.float32, torch.float64)
               for t in tensors)


class RigidTransform:
    """A [4, 4] or [3, 4] rigid matrix applied to [N, 3] points with ``@``
    (camera.py:311-403).

    On HIP tensors ``@`` is one transform kernel (differentiable in matrix
    and points) and ``transform_normals`` of tensors that need no gradient
    is one kernel with the cofactor inverse-transpose of the reference's
    native RigidTransform (include/slamtb/camera.hpp:99-116); a normal
    transform that needs a gradient uses torch's inverse, as the reference
    does. CPU tensors take the reference's torch expressions."""

    def __init__(self, matrix):
        self.matrix = matrix
        self._normal_matrix = None

    def __matmul__(self, points):
        if _on_gpu(points):
            m = self.matrix.to(points.device, points.dtype)
            return _RigidPoints.apply(m, points.reshape(-1, 3).contiguous()).squeeze()
        points = self.matrix[:3, :3].matmul(points.view(-1, 3, 1))
        points += self.matrix[:3, 3].view(3, 1)
        return points.squeeze()

    def transform_normals(self, normals):
        """Normals through the inverse transpose of the rotation block."""
        if (_on_gpu(normals) and not normals.requires_grad
                and not (torch.is_grad_enabled() and self.matrix.requires_grad)):
            n = normals.reshape(-1, 3).contiguous()
            out = torch.empty_like(n)
            _RigidTransformOp.transform_normals(self.matrix.detach(), n, out)
            return out
        return (self.normal_matrix @ normals.view(-1, 3, 1)).view(-1, 3)

    def inplace(self, points):
        points = points.view(-1, 3)
        if _on_gpu(points):
            _RigidTransformOp.transform_inplace(self.matrix, points)
        else:
            points.copy_(self.matrix[:3, :3].matmul(points.view(-1, 3, 1)).view(-1, 3)
                         + self.matrix[:3, 3])
        return points

    def inplace_normals(self, normals):
        normals = normals.view(-1, 3)
        if _on_gpu(normals):
            _RigidTransformOp.transform_normals_inplace(self.matrix, normals)
        else:
            normals.copy_(self.transform_normals(normals))
        return normals

    @property
    def normal_matrix(self):
        if self._normal_matrix is None:
            self._normal_matrix = normal_transform_matrix(self.matrix)
        return self._normal_matrix

    def translation(self):
        return self.matrix[:3, 3]


def normal_transform_matrix(matrix):
    """Transpose of the inverse of the [3, 3] rotation block
    (camera.py:392-403)."""
    return torch.inverse(matrix[:3, :3]).transpose(1, 0)
