"""Drop-in replacement of the reference's native module ``slamtb._cslamtb``
(src/_cslamtb.cpp:22-53) for the ICP hot path, backed by libslamtb_hip.so.

Same class names, static-method names, positional arguments and in-place
output convention as the pybind11 bindings:

* ``ICPJacobian`` -- src/icp_jacobian.cpp:7-13
* ``CorrespondenceMap`` -- src/matching.cpp:12-16
* ``Processing``, ``EstimateNormalsMethod``, ``DownsampleXYZMethod`` --
  src/processing.cpp:7-22
* ``ExpRtToMatrixOp``, ``MatrixToExpRtOp`` -- src/se3.cpp:6-15
* ``FPCLMatcherOp`` -- src/matching.cpp:8-9
* ``ProjectOp`` -- src/camera.cpp:11-12
* ``RigidTransformOp`` -- src/camera.cpp:40-45
* ``NearestNeighborsOp`` -- src/nearest_neighbors.cpp:6-12

Differences, all stricter or documented: tensors must live on a HIP device
(the reference also had a CPU instantiation; this library has no CPU path),
output tensors must be contiguous, and ``kcam``/``rt_cam`` are cast to the
source dtype instead of failing the accessor type check.
"""
import enum
import math

import torch

from . import _lib
from ._lib import check, dtype_code, lib, ptr, require_contiguous_out, require_gpu, stream


class EstimateNormalsMethod(enum.IntEnum):
    CentralDifferences = 0
    Average8 = 1


class DownsampleXYZMethod(enum.IntEnum):
    Nearest = 0


def _in(t, dtype=None):
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    return t.contiguous()


def _mask(t):
    if t.dtype != torch.bool and t.dtype != torch.uint8:
        t = t.to(torch.bool)
    return t.contiguous()


def _rt(rt_cam, dtype):
    rt = _in(rt_cam, dtype)
    if rt.dim() != 2 or rt.size(1) != 4 or rt.size(0) not in (3, 4):
        raise RuntimeError("rt_cam must be a [3,4] or [4,4] matrix")
    return rt


def _kcam(kcam, dtype):
    k = _in(kcam, dtype)
    if k.dim() != 2 or k.size(0) < 2 or k.size(1) != 3:
        raise RuntimeError("kcam must be a [3,3] (or [2,3]) matrix")
    if k.size(0) == 2:
        k = torch.cat([k, torch.tensor([[0, 0, 1]], dtype=k.dtype, device=k.device)])
    return k


class CorrespondenceMap:
    @staticmethod
    def compute_correspondence_map(source_points, source_normals, source_mask,
                                   rt_cam, target_points, target_normals,
                                   target_mask, kcam, corresp_map,
                                   distance_thresh, normal_angle_thesh):
        """include/slamtb/matching.hpp:11-21. ``corresp_map`` [H, W] int64 is
        min-merged in place with ``(index << 32) | float_bits(distance)``."""
        what = "CorrespondenceMap.compute_correspondence_map"
        require_gpu(source_points, source_normals, source_mask, rt_cam,
                    target_points, target_normals, target_mask, kcam,
                    corresp_map, what=what)
        require_contiguous_out(corresp_map, what=what)
        if corresp_map.dtype != torch.int64:
            raise RuntimeError(f"{what}: corresp_map must be int64")
        dt = source_points.dtype
        sp = _in(source_points, dt).view(-1, 3)
        sn = _in(source_normals, dt).view(-1, 3)
        sm = _mask(source_mask).view(-1)
        tp = _in(target_points, dt)
        tn = _in(target_normals, dt)
        tm = _mask(target_mask)
        H, W = tm.shape[0], tm.shape[1]
        if corresp_map.shape[0] != H or corresp_map.shape[1] != W:
            raise RuntimeError(f"{what}: corresp_map must be [{H}, {W}]")
        with torch.cuda.device(sp.device):
            check(lib().stb_compute_correspondence_map(
                dtype_code(sp), ptr(sp), ptr(sn), ptr(sm), sm.numel(),
                sp.size(0), ptr(_rt(rt_cam, dt)), _rt(rt_cam, dt).size(0),
                ptr(tp), ptr(tn), ptr(tm), H, W, ptr(_kcam(kcam, dt)),
                ptr(corresp_map), float(distance_thresh),
                float(normal_angle_thesh), stream()), what)


class ICPJacobian:
    @staticmethod
    def estimate_geometric(src_points, src_mask, rt_cam, tgt_points,
                           tgt_normals, kcam, corresp_map, JtJ_partial,
                           Jr_partial, squared_residual):
        """include/slamtb/icp_jacobian.hpp:72-77; returns the match count."""
        what = "ICPJacobian.estimate_geometric"
        require_gpu(src_points, src_mask, rt_cam, tgt_points, tgt_normals,
                    kcam, corresp_map, JtJ_partial, Jr_partial,
                    squared_residual, what=what)
        require_contiguous_out(JtJ_partial, Jr_partial, squared_residual,
                               what=what)
        dt = src_points.dtype
        sp = _in(src_points, dt).view(-1, 3)
        n = sp.size(0)
        for t, shape in ((JtJ_partial, (n, 6, 6)), (Jr_partial, (n, 6)),
                         (squared_residual, (n,))):
            if tuple(t.shape) != shape or t.dtype != dt:
                raise RuntimeError(f"{what}: expected {shape} {dt} output")
        if tgt_points.dtype != torch.float32 or tgt_normals.dtype != torch.float32:
            raise RuntimeError(f"{what}: target maps must be float32 "
                               "(src/icp_jacobian_geometric.cu:22-23)")
        sm = _mask(src_mask).view(-1)
        cmap = _in(corresp_map)
        H, W = cmap.shape
        count = torch.empty(1, dtype=torch.int32, device=sp.device)
        rt = _rt(rt_cam, dt)
        with torch.cuda.device(sp.device):
            check(lib().stb_icp_estimate_geometric(
                dtype_code(sp), ptr(sp), ptr(sm), n, ptr(rt), rt.size(0),
                ptr(_in(tgt_points)), ptr(_in(tgt_normals)), H, W,
                ptr(_kcam(kcam, dt)), ptr(cmap), ptr(JtJ_partial),
                ptr(Jr_partial), ptr(squared_residual), ptr(count), stream()),
                what)
        return int(count.item())

    @staticmethod
    def get_information_matrix(src_points, src_mask, rt_cam, tgt_points, kcam,
                               corresp_map, GtG_partial):
        """include/slamtb/icp_jacobian.hpp:79-85."""
        what = "ICPJacobian.get_information_matrix"
        require_gpu(src_points, src_mask, rt_cam, tgt_points, kcam,
                    corresp_map, GtG_partial, what=what)
        require_contiguous_out(GtG_partial, what=what)
        dt = src_points.dtype
        sp = _in(src_points, dt).view(-1, 3)
        n = sp.size(0)
        if tuple(GtG_partial.shape) != (n, 6, 6) or GtG_partial.dtype != dt:
            raise RuntimeError(f"{what}: expected ({n}, 6, 6) {dt} output")
        sm = _mask(src_mask).view(-1)
        cmap = _in(corresp_map)
        H, W = cmap.shape
        rt = _rt(rt_cam, dt)
        with torch.cuda.device(sp.device):
            check(lib().stb_icp_get_information_matrix(
                dtype_code(sp), ptr(sp), ptr(sm), n, ptr(rt), rt.size(0),
                ptr(_in(tgt_points, torch.float32)), H, W,
                ptr(_kcam(kcam, dt)), ptr(cmap), ptr(GtG_partial), stream()),
                what)

    @staticmethod
    def estimate_feature(src_points, src_feats, src_mask, rt_cam, tgt_feats,
                         kcam, corresp_map, residual_thresh, JtJ_partial,
                         Jr_partial, squared_residuals):
        """include/slamtb/icp_jacobian.hpp:140-146; returns the match count."""
        what = "ICPJacobian.estimate_feature"
        require_gpu(src_points, src_feats, src_mask, rt_cam, tgt_feats, kcam,
                    corresp_map, JtJ_partial, Jr_partial, squared_residuals,
                    what=what)
        require_contiguous_out(JtJ_partial, Jr_partial, squared_residuals,
                               what=what)
        dt = src_points.dtype
        sp = _in(src_points, dt).view(-1, 3)
        n = sp.size(0)
        tf = _in(tgt_feats, dt)
        C, H, W = tf.shape
        sf = _in(src_feats, dt).view(C, n)
        for t, shape in ((JtJ_partial, (n, 6, 6)), (Jr_partial, (n, 6)),
                         (squared_residuals, (n,))):
            if tuple(t.shape) != shape or t.dtype != dt:
                raise RuntimeError(f"{what}: expected {shape} {dt} output")
        sm = _mask(src_mask).view(-1)
        cmap = _in(corresp_map)
        count = torch.empty(1, dtype=torch.int32, device=sp.device)
        rt = _rt(rt_cam, dt)
        with torch.cuda.device(sp.device):
            check(lib().stb_icp_estimate_feature(
                dtype_code(sp), ptr(sp), ptr(sf), ptr(sm), n, ptr(rt),
                rt.size(0), ptr(tf), C, H, W, ptr(_kcam(kcam, dt)), ptr(cmap),
                float(residual_thresh), ptr(JtJ_partial), ptr(Jr_partial),
                ptr(squared_residuals), ptr(count), stream()), what)
        return int(count.item())

    @staticmethod
    def estimate_feature_so3(*args):
        """include/slamtb/icp_jacobian.hpp:201-207. The reference's only
        caller passes 16 arguments to this 11-parameter binding
        (slamtb/registration/icp.py:71-76), so the path cannot run there
        either; it is not provided."""
        raise NotImplementedError(
            "estimate_feature_so3 is not provided: the reference's SO3 path "
            "is unreachable (wrong call arity in icp.py:71-76)")


class Processing:
    @staticmethod
    def bilateral_depth_filter(input, mask, result, filter_width, sigma_color,
                               sigma_space, depth_scale):
        """src/bilateral_depth_filter.cu:79-110 (positional parameters as the
        binding passes them)."""
        what = "Processing.bilateral_depth_filter"
        require_gpu(input, mask, result, what=what)
        require_contiguous_out(result, what=what)
        inp = _in(input)
        if result.dtype != inp.dtype or result.shape != inp.shape:
            raise RuntimeError(f"{what}: result must match input")
        m = _mask(mask)
        H, W = inp.shape
        with torch.cuda.device(inp.device):
            check(lib().stb_bilateral_depth_filter(
                dtype_code(inp), ptr(inp), ptr(m), ptr(result), H, W,
                int(filter_width), float(sigma_color), float(sigma_space),
                float(depth_scale), stream()), what)
        return result

    @staticmethod
    def estimate_normals(xyz, mask, out, method):
        """src/normals.cu:190-208 (CentralDifferences only)."""
        what = "Processing.estimate_normals"
        require_gpu(xyz, mask, out, what=what)
        require_contiguous_out(out, what=what)
        x = _in(xyz)
        if out.dtype != x.dtype or out.shape != x.shape:
            raise RuntimeError(f"{what}: out must match xyz")
        m = _mask(mask)
        H, W = x.shape[:2]
        with torch.cuda.device(x.device):
            check(lib().stb_estimate_normals(
                dtype_code(x), ptr(x), ptr(m), ptr(out), H, W, int(method),
                stream()), what)

    @staticmethod
    def downsample_xyz(src, mask, scale, dst, normalize, method):
        """src/downsample.cu:75-98."""
        what = "Processing.downsample_xyz"
        require_gpu(src, mask, dst, what=what)
        require_contiguous_out(dst, what=what)
        s = _in(src)
        m = _mask(mask)
        H, W = s.shape[:2]
        with torch.cuda.device(s.device):
            check(lib().stb_downsample_xyz(
                dtype_code(s), ptr(s), ptr(m), H, W, float(scale), ptr(dst),
                dst.size(0), dst.size(1), int(bool(normalize)), int(method),
                stream()), what)

    @staticmethod
    def downsample_mask(mask, scale, dst):
        """src/downsample.cu:119-131."""
        what = "Processing.downsample_mask"
        require_gpu(mask, dst, what=what)
        require_contiguous_out(dst, what=what)
        m = _mask(mask)
        H, W = m.shape
        with torch.cuda.device(m.device):
            check(lib().stb_downsample_mask(
                ptr(m), H, W, float(scale), ptr(dst), dst.size(0), dst.size(1),
                stream()), what)

    @staticmethod
    def downsample_features(feat, scale, dst):
        """src/downsample.cu:171-191."""
        what = "Processing.downsample_features"
        require_gpu(feat, dst, what=what)
        require_contiguous_out(dst, what=what)
        f = _in(feat)
        C, H, W = f.shape
        with torch.cuda.device(f.device):
            check(lib().stb_downsample_features(
                dtype_code(f), ptr(f), C, H, W, float(scale), ptr(dst),
                dst.size(1), dst.size(2), stream()), what)

    @staticmethod
    def erode_mask(in_mask, out_mask):
        """src/erode_mask.cu:42-54."""
        what = "Processing.erode_mask"
        require_gpu(in_mask, out_mask, what=what)
        require_contiguous_out(out_mask, what=what)
        m = _mask(in_mask)
        H, W = m.shape
        with torch.cuda.device(m.device):
            check(lib().stb_erode_mask(ptr(m), ptr(out_mask), H, W, stream()),
                  what)


class ExpRtToMatrixOp:
    @staticmethod
    def forward(exp_rt, matrix):
        """src/exp_rt_to_matrix.cu:41-52: [B,6] -> [B,3,4]."""
        what = "ExpRtToMatrixOp.forward"
        require_gpu(exp_rt, matrix, what=what)
        require_contiguous_out(matrix, what=what)
        x = _in(exp_rt).view(-1, 6)
        with torch.cuda.device(x.device):
            check(lib().stb_exp_rt_to_matrix_forward(
                dtype_code(x), ptr(x), x.size(0), ptr(matrix), stream()), what)

    @staticmethod
    def backward(d_matrix, exp_rt, y_matrix, d_exp_rt):
        """src/exp_rt_to_matrix.cu:104-121."""
        what = "ExpRtToMatrixOp.backward"
        require_gpu(d_matrix, exp_rt, y_matrix, d_exp_rt, what=what)
        require_contiguous_out(d_exp_rt, what=what)
        x = _in(exp_rt).view(-1, 6)
        dt = x.dtype
        with torch.cuda.device(x.device):
            check(lib().stb_exp_rt_to_matrix_backward(
                dtype_code(x), ptr(_in(d_matrix, dt)), ptr(x),
                ptr(_in(y_matrix, dt)), x.size(0), ptr(d_exp_rt), stream()),
                what)


class MatrixToExpRtOp:
    @staticmethod
    def forward(matrix, exp_rt):
        """src/matrix_to_exp_rt.cu:38-50: [B,3,4] -> [B,6]."""
        what = "MatrixToExpRtOp.forward"
        require_gpu(matrix, exp_rt, what=what)
        require_contiguous_out(exp_rt, what=what)
        m = _in(matrix)
        if m.dim() == 2:
            m = m.unsqueeze(0)
        m = m[:, :3, :4].contiguous()
        with torch.cuda.device(m.device):
            check(lib().stb_matrix_to_exp_rt_forward(
                dtype_code(m), ptr(m), m.size(0), ptr(exp_rt), stream()), what)


class FPCLMatcherOp:
    @staticmethod
    def forward(target_points, target_normals, target_features, source_points,
                kcam, merge_map, out_points, out_normals, out_features,
                match_mask):
        """src/fpcl_matcher.cu:85-123. For each source point: the target pixel
        it projects onto (rounded) must hold its own index in ``merge_map``;
        then ``match_mask`` is set and the target point, normal and the
        bilinear feature sample at the unrounded projection are written.
        Unmatched rows of the outputs are written as zeros (the reference
        leaves them unwritten)."""
        what = "FPCLMatcherOp.forward"
        require_gpu(target_points, target_normals, target_features,
                    source_points, kcam, merge_map, out_points, out_normals,
                    out_features, match_mask, what=what)
        require_contiguous_out(out_points, out_normals, out_features,
                               match_mask, what=what)
        dt = target_points.dtype
        tp = _in(target_points, dt)
        tn = _in(target_normals, dt)
        tf = _in(target_features, dt)
        C, H, W = tf.shape
        if tuple(tp.shape[:2]) != (H, W) or tuple(tn.shape[:2]) != (H, W):
            raise RuntimeError(f"{what}: target points/normals must be [{H}, {W}, 3]")
        sp = _in(source_points, dt).view(-1, 3)
        n = sp.size(0)
        cmap = _in(merge_map)
        if cmap.dtype != torch.int64 or tuple(cmap.shape) != (H, W):
            raise RuntimeError(f"{what}: merge_map must be int64 [{H}, {W}]")
        for t, shape in ((out_points, (n, 3)), (out_normals, (n, 3)),
                         (out_features, (C, n))):
            if tuple(t.shape) != shape or t.dtype != dt:
                raise RuntimeError(f"{what}: expected {shape} {dt} output")
        if tuple(match_mask.shape) != (n,) or match_mask.dtype not in (
                torch.bool, torch.uint8):
            raise RuntimeError(f"{what}: match_mask must be bool [{n}]")
        with torch.cuda.device(sp.device):
            check(lib().stb_fpcl_match_forward(
                dtype_code(sp), ptr(tp), ptr(tn), ptr(tf), C, H, W, ptr(sp), n,
                ptr(_kcam(kcam, dt)), ptr(cmap), ptr(out_points),
                ptr(out_normals), ptr(out_features), ptr(match_mask),
                stream()), what)

    @staticmethod
    def backward(target_features, source_points, match_mask, dl_feature, kcam,
                 grad_precision, dx_points):
        """src/fpcl_matcher.cu:187-229: d loss / d source point through the
        central-difference gradient of the bilinear feature sample (step
        ``grad_precision`` pixels) and the projection Jacobian; zero rows
        where ``match_mask`` is false."""
        what = "FPCLMatcherOp.backward"
        require_gpu(target_features, source_points, match_mask, dl_feature,
                    kcam, dx_points, what=what)
        require_contiguous_out(dx_points, what=what)
        dt = target_features.dtype
        tf = _in(target_features, dt)
        C, H, W = tf.shape
        sp = _in(source_points, dt).view(-1, 3)
        n = sp.size(0)
        dl = _in(dl_feature, dt)
        if tuple(dl.shape) != (C, n):
            raise RuntimeError(f"{what}: dl_feature must be [{C}, {n}]")
        mm = _mask(match_mask).view(-1)
        if mm.numel() != n:
            raise RuntimeError(f"{what}: match_mask must have {n} elements")
        if tuple(dx_points.shape) != (n, 3) or dx_points.dtype != dt:
            raise RuntimeError(f"{what}: expected ({n}, 3) {dt} output")
        with torch.cuda.device(sp.device):
            check(lib().stb_fpcl_match_backward(
                dtype_code(sp), ptr(tf), C, H, W, ptr(sp), n, ptr(mm), ptr(dl),
                ptr(_kcam(kcam, dt)), float(grad_precision), ptr(dx_points),
                stream()), what)


class ProjectOp:
    @staticmethod
    def forward(points, intrinsics):
        """src/camera.cu:34-60: [..., 3] points -> new [N, 2] (u, v) with
        u = fx*x/z + cx, v = fy*y/z + cy."""
        what = "ProjectOp.forward"
        require_gpu(points, intrinsics, what=what)
        p = _in(points).view(-1, 3)
        out = torch.empty(p.size(0), 2, dtype=p.dtype, device=p.device)
        with torch.cuda.device(p.device):
            check(lib().stb_project_forward(
                dtype_code(p), ptr(p), p.size(0), ptr(_kcam(intrinsics, p.dtype)),
                ptr(out), stream()), what)
        return out

    @staticmethod
    def backward(dy_grad, points, intrinsics):
        """src/camera.cu:91-117: new [N, 3] gradient through the projection
        Jacobian (include/slamtb/camera.hpp:65-78)."""
        what = "ProjectOp.backward"
        require_gpu(dy_grad, points, intrinsics, what=what)
        p = _in(points).view(-1, 3)
        dy = _in(dy_grad, p.dtype).view(-1, 2)
        if dy.size(0) != p.size(0):
            raise RuntimeError(f"{what}: dy_grad must be [{p.size(0)}, 2]")
        out = torch.empty_like(p)
        with torch.cuda.device(p.device):
            check(lib().stb_project_backward(
                dtype_code(p), ptr(dy), ptr(p), p.size(0),
                ptr(_kcam(intrinsics, p.dtype)), ptr(out), stream()), what)
        return out


def _matrix34(matrix, like):
    m = matrix.to(device=like.device, dtype=like.dtype)
    if m.dim() != 2 or m.size(1) != 4 or m.size(0) not in (3, 4):
        raise RuntimeError("matrix must be [3,4] or [4,4]")
    return m[:3].contiguous()


class RigidTransformOp:
    @staticmethod
    def transform_inplace(matrix, points):
        """src/camera.cu:134-174: points [N,3] <- R p + t in place."""
        RigidTransformOp.transform(matrix, points, points)

    @staticmethod
    def transform_normals_inplace(matrix, normals):
        """src/camera.cu:176-240: normals [N,3] <- inverse-transpose(R) n."""
        RigidTransformOp.transform_normals(matrix, normals, normals)

    @staticmethod
    def transform(matrix, points, out):
        """out [N,3] = R p + t (out may be points). Not a reference binding."""
        what = "RigidTransformOp.transform"
        require_gpu(points, out, what=what)
        require_contiguous_out(points, out, what=what)
        m = _matrix34(matrix, points)
        n = points.numel() // 3
        with torch.cuda.device(points.device):
            check(lib().stb_rigid_transform_points(
                dtype_code(points), ptr(m), ptr(points), n, ptr(out), stream()),
                what)

    @staticmethod
    def transform_normals(matrix, normals, out):
        """out [N,3] = inverse-transpose(R) n. Not a reference binding."""
        what = "RigidTransformOp.transform_normals"
        require_gpu(normals, out, what=what)
        require_contiguous_out(normals, out, what=what)
        m = _matrix34(matrix, normals)
        n = normals.numel() // 3
        with torch.cuda.device(normals.device):
            check(lib().stb_rigid_transform_normals(
                dtype_code(normals), ptr(m), ptr(normals), n, ptr(out), stream()),
                what)

    @staticmethod
    def transform_backward(matrix, points, dy, dx_points, d_matrix):
        """Backward of out = R p + t: dx_points [N,3] = R^T dy (None skips
        it) and d_matrix [3,4] = sum_i dy_i [p_i; 1]^T. Not a reference
        binding (the reference differentiates a torch matmul)."""
        what = "RigidTransform backward"
        require_gpu(points, dy, dx_points, d_matrix, what=what)
        require_contiguous_out(dx_points, d_matrix, what=what)
        p = _in(points).view(-1, 3)
        g = _in(dy, p.dtype).view(-1, 3)
        m = _matrix34(matrix, p)
        ws_bytes = lib().stb_rigid_backward_workspace_bytes()
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=p.device)
        with torch.cuda.device(p.device):
            check(lib().stb_rigid_transform_backward(
                dtype_code(p), ptr(m), ptr(p), p.size(0), ptr(g),
                ptr(dx_points), ptr(d_matrix), ptr(ws), ws_bytes, stream()),
                what)

    @staticmethod
    def rodrigues(rot_matrix, rodrigues):
        """src/camera.cpp:15-37 is CPU-only and fills only element (0, 0) of
        its rotation before taking the axis, so its output does not depend
        on the input's rotation; it is not provided."""
        raise NotImplementedError("RigidTransformOp.rodrigues is not provided")


NN_DELTA_H = 0.0005000000237487257  # float(0.0005f), the reference's step


class NearestNeighborsOp:
    @staticmethod
    def forward(nn_distances, nn_index, features, out_features):
        """src/nearest_neighbors.cu:56-79: out_features [C, N] = neighbour
        features averaged with normalised weights 1 / (1 + d)."""
        what = "NearestNeighborsOp.forward"
        require_gpu(nn_distances, nn_index, features, out_features, what=what)
        require_contiguous_out(out_features, what=what)
        d = _in(nn_distances)
        dt = d.dtype
        idx = _in(nn_index, torch.int64)
        f = _in(features, dt)
        n, k = d.shape
        C, m = f.shape
        if tuple(idx.shape) != (n, k):
            raise RuntimeError(f"{what}: nn_index must be [{n}, {k}]")
        if tuple(out_features.shape) != (C, n) or out_features.dtype != dt:
            raise RuntimeError(f"{what}: out_features must be [{C}, {n}] {dt}")
        with torch.cuda.device(d.device):
            check(lib().stb_nn_features_forward(
                dtype_code(d), ptr(d), ptr(idx), n, k, ptr(f), C, m,
                ptr(out_features), stream()), what)

    @staticmethod
    def compute_epsilon_distances(target_xyz, source_xyz, nn_index,
                                  epsilon_distances):
        """src/nearest_neighbors.cu:145-164: epsilon_distances [N, 6, k]."""
        what = "NearestNeighborsOp.compute_epsilon_distances"
        require_gpu(target_xyz, source_xyz, nn_index, epsilon_distances, what=what)
        require_contiguous_out(epsilon_distances, what=what)
        dt = target_xyz.dtype
        t = _in(target_xyz, dt).view(-1, 3)
        sx = _in(source_xyz, dt).view(-1, 3)
        idx = _in(nn_index, torch.int64)
        n, k = idx.shape
        if tuple(epsilon_distances.shape) != (n, 6, k) or epsilon_distances.dtype != dt:
            raise RuntimeError(f"{what}: epsilon_distances must be [{n}, 6, {k}] {dt}")
        with torch.cuda.device(t.device):
            check(lib().stb_nn_epsilon_distances(
                dtype_code(t), ptr(t), t.size(0), ptr(sx), ptr(idx), n, k,
                NN_DELTA_H, ptr(epsilon_distances), stream()), what)

    @staticmethod
    def backward(epsilon_distances, nn_index, features, dl_features, dl_xyz):
        """src/nearest_neighbors.cu:265-293: dl_xyz [N, 3]."""
        what = "NearestNeighborsOp.backward"
        require_gpu(epsilon_distances, nn_index, features, dl_features, dl_xyz,
                    what=what)
        require_contiguous_out(dl_xyz, what=what)
        e = _in(epsilon_distances)
        dt = e.dtype
        idx = _in(nn_index, torch.int64)
        n, k = idx.shape
        f = _in(features, dt)
        C, m = f.shape
        dl = _in(dl_features, dt)
        if tuple(dl.shape) != (C, n) or tuple(dl_xyz.shape) != (n, 3):
            raise RuntimeError(f"{what}: expected dl_features [{C}, {n}], dl_xyz [{n}, 3]")
        with torch.cuda.device(e.device):
            check(lib().stb_nn_features_backward(
                dtype_code(e), ptr(e), ptr(idx), n, k, ptr(f), C, m, ptr(dl),
                NN_DELTA_H, ptr(dl_xyz), stream()), what)


class KnnGrid:
    """Uniform grid over target points for the bounded neighbour search:
    cells at least ``cell`` wide (one cell of slack on each axis), the
    targets stably sorted by cell, and each cell's start. Built on the
    device; the bounding box is read once on the host."""

    MAX_CELLS = 1 << 24

    def __init__(self, target, cell):
        what = "KnnGrid"
        require_gpu(target, what=what)
        t = _in(target).view(-1, 3)
        self.dtype = t.dtype
        self.m = t.size(0)
        finite = torch.isfinite(t).all(1)
        pts = t[finite] if not bool(finite.all()) else t
        if pts.size(0) == 0:
            lo, hi = [0.0] * 3, [0.0] * 3
        else:
            lo, hi = pts.min(0)[0].tolist(), pts.max(0)[0].tolist()
        cell = float(cell) * (1 + 1e-4)
        dims = [int((b - a) / cell) + 2 for a, b in zip(lo, hi)]
        total = dims[0] * dims[1] * dims[2]
        if total > self.MAX_CELLS:
            cell *= (total / self.MAX_CELLS) ** (1 / 3) * 1.01
            dims = [int((b - a) / cell) + 2 for a, b in zip(lo, hi)]
        self.lo, self.cell, self.dims = lo, cell, dims
        ncell = dims[0] * dims[1] * dims[2]
        cells = torch.empty(self.m, dtype=torch.int32, device=t.device)
        with torch.cuda.device(t.device):
            check(lib().stb_grid_cells(dtype_code(t), ptr(t), self.m, lo[0], lo[1],
                                       lo[2], cell, dims[0], dims[1], dims[2],
                                       ptr(cells), stream()), what)
        sorted_cells, order = torch.sort(cells, stable=True)
        if t.dtype == torch.float32:
            tag = order.to(torch.int32).view(torch.float32)
        else:
            tag = order.to(t.dtype)
        self.entries = torch.cat([t[order], tag[:, None]], 1).contiguous()
        self.start = torch.searchsorted(
            sorted_cells, torch.arange(ncell + 1, dtype=torch.int32,
                                       device=t.device)).to(torch.int64).contiguous()

    def query(self, query, k, distance_upper_bound):
        """Threads take the queries in grid-cell order (a stable sort of
        their cells), so that a wavefront's queries scan the same cells."""
        what = "knn"
        q = _in(query, self.dtype).view(-1, 3)
        n = q.size(0)
        d = torch.empty(n, k, dtype=q.dtype, device=q.device)
        i = torch.empty(n, k, dtype=torch.int64, device=q.device)
        nx, ny, nz = self.dims
        with torch.cuda.device(q.device):
            cells = torch.empty(n, dtype=torch.int32, device=q.device)
            check(lib().stb_grid_cells(dtype_code(q), ptr(q), n, self.lo[0],
                                       self.lo[1], self.lo[2], self.cell, nx, ny,
                                       nz, ptr(cells), stream()), what)
            order = torch.sort(cells, stable=True)[1].contiguous()
            check(lib().stb_knn_grid(
                dtype_code(q), ptr(self.entries), ptr(self.start),
                nx, ny, nz, self.lo[0], self.lo[1], self.lo[2], self.cell, self.m,
                ptr(q), n, ptr(order), int(k), float(distance_upper_bound), ptr(d),
                ptr(i), stream()), what)
        return d, i


def knn(target, query, k, distance_upper_bound=float("inf"), grid=None,
        method="auto"):
    """Exact k nearest neighbours on the device: (squared distances [n, k],
    index [n, k] int64), (inf, m) in missing slots. Not a reference binding:
    it stands in for the host k-d tree query of KDTreeLayer. A finite bound
    searches a uniform grid (``grid``, a ``KnnGrid`` built with a cell at
    least the bound, or one built here); an infinite bound, or
    ``method="brute"``, searches all targets."""
    what = "knn"
    require_gpu(target, query, what=what)
    q = _in(query).view(-1, 3)
    t = _in(target, q.dtype).view(-1, 3)
    n = q.size(0)
    if method not in ("auto", "brute"):
        raise ValueError(f"{what}: method must be 'auto' or 'brute'")
    if method == "auto" and math.isfinite(distance_upper_bound) and t.size(0) > 0:
        if grid is None or grid.cell < distance_upper_bound or grid.dtype != q.dtype:
            grid = KnnGrid(t, distance_upper_bound)
        return grid.query(q, k, distance_upper_bound)
    d = torch.empty(n, k, dtype=q.dtype, device=q.device)
    i = torch.empty(n, k, dtype=torch.int64, device=q.device)
    with torch.cuda.device(q.device):
        check(lib().stb_knn(dtype_code(q), ptr(t), t.size(0), ptr(q), n, int(k),
                            float(distance_upper_bound), ptr(d), ptr(i), stream()),
              what)
    return d, i


__all__ = ["CorrespondenceMap", "ICPJacobian", "Processing",
           "EstimateNormalsMethod", "DownsampleXYZMethod", "ExpRtToMatrixOp",
           "MatrixToExpRtOp", "FPCLMatcherOp", "ProjectOp", "RigidTransformOp",
           "NearestNeighborsOp", "knn", "KnnGrid", "_lib"]
