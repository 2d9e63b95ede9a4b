"""ORACLE — test infrastructure only.

numpy/ctypes front end of ``liboracle.so``, the sequential CPU restatement of
slamtb's ICP Gauss-Newton step (see ``icp_oracle.cpp`` for the per-function
citations of the reference and for the arithmetic contract).  Only the tests,
the smoke check and the CPU-baseline leg of ``bench.py`` import this module,
and only as the checker; the product path never does.

The GN driver here (:func:`icp_estimate`) restates
``slamtb/registration/icp.py:152-262``: per iteration it builds one
correspondence map per active term, fills per-pixel partials, sums them,
combines the terms with the reference's weights (J^T J by w^2, J^T r by w and
by -w for the feature term, ``icp.py:221-234``), solves by Cholesky in FP64 and
left-multiplies the ``exp_rt_to_matrix`` update (``icp.py:244-247``,
``src/exp_rt_to_matrix.cu:22-36``).  A failed factorisation stops the loop
(``icp.py:236-242``).
"""
import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

CORRESP_INF = 0x0FFFFFFF7F800000

_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_i32p = ctypes.POINTER(ctypes.c_int32)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build():
    """Compile liboracle.so in place (needs g++; used on the build host)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = _HERE / "liboracle.so"
        if not path.exists():
            build()
        _LIB = ctypes.CDLL(str(path))
        _LIB.orc_geometric_f32.restype = ctypes.c_int
        _LIB.orc_geometric_f64.restype = ctypes.c_int
        _LIB.orc_feature_f32.restype = ctypes.c_int
        _LIB.orc_feature_f64.restype = ctypes.c_int
        _LIB.orc_num_threads.restype = ctypes.c_int
        _LIB.orc_set_num_threads.argtypes = [ctypes.c_int]
        _LIB.orc_set_num_threads.restype = None
    return _LIB


def _p(a):
    """ctypes pointer of a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"], "oracle inputs must be contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f32"
    if dtype == np.float64:
        return "f64"
    raise TypeError(f"oracle supports float32/float64, got {dtype}")


def _c(a, dtype=None):
    return np.ascontiguousarray(a, dtype=dtype)


def _rt34(rt, dtype):
    rt = np.asarray(rt, dtype=np.float64)
    return _c(rt.reshape(-1, 4)[:3], dtype)


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    """Thread count of the parallel (OpenMP) oracle loops and of torch's CPU
    reductions that the parallel variant uses."""
    import torch
    lib().orc_set_num_threads(int(n))
    torch.set_num_threads(int(n))


# ---------------------------------------------------------------------------
# Per-pixel kernels
# ---------------------------------------------------------------------------

def corresp_map(src_pts, src_nrm, src_mask, rt, tgt_pts, tgt_nrm, tgt_mask,
                kcam, distance_thresh, angle_thresh, init_map=None,
                parallel=False):
    """Correspondence map, reference layout ``(index << 32) | f32 bits``."""
    dt = np.asarray(src_pts).dtype
    sfx = _sfx(dt)
    sp = _c(np.asarray(src_pts).reshape(-1, 3), dt)
    sn = _c(np.asarray(src_nrm).reshape(-1, 3), dt)
    sm = _c(np.asarray(src_mask).reshape(-1), np.uint8)
    H, W = tgt_mask.shape[:2]
    tp = _c(tgt_pts, dt)
    tn = _c(tgt_nrm, dt)
    tm = _c(tgt_mask, np.uint8)
    k = _c(np.asarray(kcam, dtype=np.float64).reshape(3, 3), dt)
    if init_map is None:
        out = np.full((H, W), CORRESP_INF, dtype=np.int64)
    else:
        out = _c(init_map, np.int64).copy()
    getattr(lib(), f"orc_corresp_map_{sfx}")(
        _p(sp), _p(sn), _p(sm), ctypes.c_int64(sm.size),
        ctypes.c_int64(sp.shape[0]), _p(_rt34(rt, dt)), _p(tp), _p(tn), _p(tm),
        ctypes.c_int(H), ctypes.c_int(W), _p(k), _p(out),
        ctypes.c_double(distance_thresh), ctypes.c_double(angle_thresh),
        ctypes.c_int(int(parallel)))
    return out


def geometric(src_pts, src_mask, rt, tgt_pts, tgt_nrm, kcam, cmap,
              parallel=False):
    """Per-pixel point-to-plane partials: (JtJ[N,6,6], Jr[N,6], r2[N], count)."""
    dt = np.asarray(src_pts).dtype
    sfx = _sfx(dt)
    sp = _c(np.asarray(src_pts).reshape(-1, 3), dt)
    sm = _c(np.asarray(src_mask).reshape(-1), np.uint8)
    n = sp.shape[0]
    H, W = cmap.shape
    jtj = np.empty((n, 6, 6), dt)
    jr = np.empty((n, 6), dt)
    r2 = np.empty((n,), dt)
    k = _c(np.asarray(kcam, dtype=np.float64).reshape(3, 3), dt)
    cnt = getattr(lib(), f"orc_geometric_{sfx}")(
        _p(sp), _p(sm), ctypes.c_int64(n), _p(_rt34(rt, dt)),
        _p(_c(tgt_pts, np.float32)), _p(_c(tgt_nrm, np.float32)),
        ctypes.c_int(H), ctypes.c_int(W), _p(k), _p(_c(cmap, np.int64)),
        _p(jtj), _p(jr), _p(r2), ctypes.c_int(int(parallel)))
    return jtj, jr, r2, cnt


def feature(src_pts, src_feats, src_mask, rt, tgt_feats, kcam, cmap,
            residual_thresh, parallel=False):
    """Per-pixel feature partials (last channel's J kept, as the reference)."""
    dt = np.asarray(src_pts).dtype
    sfx = _sfx(dt)
    sp = _c(np.asarray(src_pts).reshape(-1, 3), dt)
    n = sp.shape[0]
    C = tgt_feats.shape[0]
    sf = _c(np.asarray(src_feats).reshape(C, n), dt)
    sm = _c(np.asarray(src_mask).reshape(-1), np.uint8)
    H, W = cmap.shape
    jtj = np.empty((n, 6, 6), dt)
    jr = np.empty((n, 6), dt)
    r2 = np.empty((n,), dt)
    k = _c(np.asarray(kcam, dtype=np.float64).reshape(3, 3), dt)
    cnt = getattr(lib(), f"orc_feature_{sfx}")(
        _p(sp), _p(sf), _p(sm), ctypes.c_int64(n), _p(_rt34(rt, dt)),
        _p(_c(tgt_feats, dt)), ctypes.c_int(C), ctypes.c_int(H),
        ctypes.c_int(W), _p(k), _p(_c(cmap, np.int64)),
        ctypes.c_float(residual_thresh), _p(jtj), _p(jr), _p(r2),
        ctypes.c_int(int(parallel)))
    return jtj, jr, r2, cnt


def information(src_pts, src_mask, rt, tgt_pts, kcam, cmap, parallel=False):
    dt = np.asarray(src_pts).dtype
    sfx = _sfx(dt)
    sp = _c(np.asarray(src_pts).reshape(-1, 3), dt)
    sm = _c(np.asarray(src_mask).reshape(-1), np.uint8)
    n = sp.shape[0]
    H, W = cmap.shape
    gtg = np.empty((n, 6, 6), dt)
    k = _c(np.asarray(kcam, dtype=np.float64).reshape(3, 3), dt)
    getattr(lib(), f"orc_information_{sfx}")(
        _p(sp), _p(sm), ctypes.c_int64(n), _p(_rt34(rt, dt)),
        _p(_c(tgt_pts, np.float32)), ctypes.c_int(H), ctypes.c_int(W), _p(k),
        _p(_c(cmap, np.int64)), _p(gtg), ctypes.c_int(int(parallel)))
    return gtg


# ---------------------------------------------------------------------------
# SE3
# ---------------------------------------------------------------------------

def fpcl_forward(tgt_pts, tgt_nrm, tgt_feats, src_pts, kcam, cmap):
    """FPCLMatcherOp forward: (points [N,3], normals [N,3], features [C,N],
    mask [N] bool). Unmatched rows are zero, as in the HIP kernel (the
    reference leaves them unwritten)."""
    dt = np.asarray(src_pts).dtype
    sfx = _sfx(dt)
    tf = _c(tgt_feats, dt)
    C, H, W = tf.shape
    sp = _c(np.asarray(src_pts).reshape(-1, 3), dt)
    n = sp.shape[0]
    op = np.zeros((n, 3), dt)
    on = np.zeros((n, 3), dt)
    of = np.zeros((C, n), dt)
    mask = np.zeros((n,), np.uint8)
    k = _c(np.asarray(kcam, dtype=np.float64).reshape(3, 3), dt)
    getattr(lib(), f"orc_fpcl_forward_{sfx}")(
        _p(_c(tgt_pts, dt)), _p(_c(tgt_nrm, dt)), _p(tf), ctypes.c_int(C),
        ctypes.c_int(H), ctypes.c_int(W), _p(sp), ctypes.c_int64(n), _p(k),
        _p(_c(cmap, np.int64)), _p(op), _p(on), _p(of), _p(mask))
    return op, on, of, mask.astype(bool)


def fpcl_backward(tgt_feats, src_pts, mask, dl_feats, kcam, grad_precision):
    """FPCLMatcherOp backward: d loss / d source points [N, 3]."""
    dt = np.asarray(src_pts).dtype
    sfx = _sfx(dt)
    tf = _c(tgt_feats, dt)
    C, H, W = tf.shape
    sp = _c(np.asarray(src_pts).reshape(-1, 3), dt)
    n = sp.shape[0]
    dx = np.empty((n, 3), dt)
    k = _c(np.asarray(kcam, dtype=np.float64).reshape(3, 3), dt)
    getattr(lib(), f"orc_fpcl_backward_{sfx}")(
        _p(tf), ctypes.c_int(C), ctypes.c_int(H), ctypes.c_int(W), _p(sp),
        ctypes.c_int64(n), _p(_c(np.asarray(mask).reshape(-1), np.uint8)),
        _p(_c(np.asarray(dl_feats).reshape(C, n), dt)), _p(k),
        ctypes.c_double(grad_precision), _p(dx))
    return dx


def project_forward(points, kcam):
    dt = np.asarray(points).dtype
    p = _c(np.asarray(points).reshape(-1, 3), dt)
    uv = np.empty((p.shape[0], 2), dt)
    k = _c(np.asarray(kcam, dtype=np.float64).reshape(3, 3), dt)
    getattr(lib(), f"orc_project_forward_{_sfx(dt)}")(
        _p(p), ctypes.c_int64(p.shape[0]), _p(k), _p(uv))
    return uv


def project_backward(dy, points, kcam):
    dt = np.asarray(points).dtype
    p = _c(np.asarray(points).reshape(-1, 3), dt)
    dx = np.empty_like(p)
    k = _c(np.asarray(kcam, dtype=np.float64).reshape(3, 3), dt)
    getattr(lib(), f"orc_project_backward_{_sfx(dt)}")(
        _p(_c(np.asarray(dy).reshape(-1, 2), dt)), _p(p),
        ctypes.c_int64(p.shape[0]), _p(k), _p(dx))
    return dx


def rigid_points(matrix, points, normals=False):
    """RigidTransformOp: [N,3] points through [R | t], or normals through
    the cofactor inverse-transpose of R."""
    dt = np.asarray(points).dtype
    p = _c(np.asarray(points).reshape(-1, 3), dt)
    out = np.empty_like(p)
    getattr(lib(), f"orc_rigid_points_{_sfx(dt)}")(
        _p(_rt34(matrix, dt)), _p(p), ctypes.c_int64(p.shape[0]), _p(out),
        ctypes.c_int(int(normals)))
    return out


def rigid_backward(matrix, points, dy):
    """Backward of p' = R p + t: (dx [N,3], dM [3,4])."""
    dt = np.asarray(points).dtype
    p = _c(np.asarray(points).reshape(-1, 3), dt)
    dx = np.empty_like(p)
    dM = np.empty((3, 4), dt)
    getattr(lib(), f"orc_rigid_backward_{_sfx(dt)}")(
        _p(_rt34(matrix, dt)), _p(p), _p(_c(np.asarray(dy).reshape(-1, 3), dt)),
        ctypes.c_int64(p.shape[0]), _p(dx), _p(dM))
    return dx, dM


NN_DELTA_H = float(np.float32(0.0005))  # the reference's float constant


def knn(target, query, k, distance_upper_bound=np.inf):
    """Exact k nearest neighbours: (squared distances [n,k], index [n,k]),
    missing slots (inf, m)."""
    dt = np.asarray(query).dtype
    t = _c(np.asarray(target).reshape(-1, 3), dt)
    q = _c(np.asarray(query).reshape(-1, 3), dt)
    d = np.empty((q.shape[0], k), dt)
    i = np.empty((q.shape[0], k), np.int64)
    getattr(lib(), f"orc_knn_{_sfx(dt)}")(
        _p(t), ctypes.c_int64(t.shape[0]), _p(q), ctypes.c_int64(q.shape[0]),
        ctypes.c_int(k), ctypes.c_double(distance_upper_bound), _p(d), _p(i))
    return d, i


def nn_forward(dist, index, features):
    dt = np.asarray(dist).dtype
    d = _c(dist, dt)
    n, k = d.shape
    f = _c(features, dt)
    C, m = f.shape
    out = np.empty((C, n), dt)
    getattr(lib(), f"orc_nn_forward_{_sfx(dt)}")(
        _p(d), _p(_c(index, np.int64)), ctypes.c_int64(n), ctypes.c_int(k),
        _p(f), ctypes.c_int(C), ctypes.c_int64(m), _p(out))
    return out


def nn_epsilon(target, source, index, h=NN_DELTA_H):
    dt = np.asarray(source).dtype
    t = _c(np.asarray(target).reshape(-1, 3), dt)
    s = _c(np.asarray(source).reshape(-1, 3), dt)
    idx = _c(index, np.int64)
    n, k = idx.shape
    eps = np.zeros((n, 6, k), dt)
    getattr(lib(), f"orc_nn_epsilon_{_sfx(dt)}")(
        _p(t), ctypes.c_int64(t.shape[0]), _p(s), _p(idx), ctypes.c_int64(n),
        ctypes.c_int(k), ctypes.c_double(h), _p(eps))
    return eps


def nn_backward(eps, index, features, dl, h=NN_DELTA_H):
    dt = np.asarray(eps).dtype
    idx = _c(index, np.int64)
    n, k = idx.shape
    f = _c(features, dt)
    C, m = f.shape
    dx = np.empty((n, 3), dt)
    getattr(lib(), f"orc_nn_backward_{_sfx(dt)}")(
        _p(_c(eps, dt)), _p(idx), ctypes.c_int64(n), ctypes.c_int(k), _p(f),
        ctypes.c_int(C), ctypes.c_int64(m), _p(_c(dl, dt)), ctypes.c_double(h),
        _p(dx))
    return dx


def voxel_merge(points, normals, colors, confidences, radii, features, voxel_size):
    """SurfelVolume merge: dict of per-voxel means, voxels in increasing id."""
    n = len(points)
    f = None if features is None else _c(features, np.float32)
    C = 0 if f is None else f.shape[0]
    out = dict(points=np.empty((n, 3), np.float32), normals=np.empty((n, 3), np.float32),
               colors=np.empty((n, 3), np.uint8), confidences=np.empty(n, np.float32),
               radii=np.empty(n, np.float32), times=np.empty(n, np.int32),
               features=np.empty((C, max(n, 1)), np.float32))
    lib().orc_voxel_merge.restype = ctypes.c_int64
    V = lib().orc_voxel_merge(
        _p(_c(points, np.float32)), _p(_c(normals, np.float32)), _p(_c(colors, np.uint8)),
        _p(_c(confidences, np.float32)), _p(_c(radii, np.float32)),
        _p(f) if f is not None else ctypes.c_void_p(0), ctypes.c_int(C),
        ctypes.c_int64(n), ctypes.c_float(voxel_size), _p(out["points"]),
        _p(out["normals"]), _p(out["colors"]), _p(out["confidences"]),
        _p(out["radii"]), _p(out["times"]), _p(out["features"]))
    res = {key: v[:V] for key, v in out.items() if key != "features"}
    res["features"] = out["features"][:, :V]
    return res


def exp_rt_to_matrix(x):
    x = np.asarray(x)
    dt = x.dtype if x.dtype in (np.float32, np.float64) else np.float64
    xb = _c(x.reshape(-1, 6), dt)
    out = np.empty((xb.shape[0], 3, 4), dt)
    getattr(lib(), f"orc_exp_rt_to_matrix_{_sfx(dt)}")(
        _p(xb), _p(out), ctypes.c_int64(xb.shape[0]))
    return out.reshape(x.shape[:-1] + (3, 4))


def exp_rt_backward(d_matrix, x, y):
    xb = _c(np.asarray(x).reshape(-1, 6), np.float64)
    out = np.empty((xb.shape[0], 6), np.float64)
    lib().orc_exp_rt_backward_f64(
        _p(_c(np.asarray(d_matrix).reshape(-1, 3, 4), np.float64)), _p(xb),
        _p(_c(np.asarray(y).reshape(-1, 3, 4), np.float64)), _p(out),
        ctypes.c_int64(xb.shape[0]))
    return out


def matrix_to_exp_rt(m):
    mb = _c(np.asarray(m, dtype=np.float64).reshape(-1, 3, 4)
            if np.asarray(m).shape[-2] == 3 else
            np.asarray(m, dtype=np.float64).reshape(-1, 4, 4)[:, :3], np.float64)
    out = np.empty((mb.shape[0], 6), np.float64)
    lib().orc_matrix_to_exp_rt_f64(_p(mb), _p(out), ctypes.c_int64(mb.shape[0]))
    return out


def exp_rt_4x4(x):
    m = np.eye(4)
    m[:3] = exp_rt_to_matrix(np.asarray(x, dtype=np.float64))
    return m


# ---------------------------------------------------------------------------
# Preprocessing
# ---------------------------------------------------------------------------

def bilateral(depth, mask, filter_width=6, sigma_color=29.9999880000072,
              sigma_space=4.50000000225, depth_scale=1.0):
    depth = np.asarray(depth)
    if depth.dtype == np.float32:
        fn, dt = lib().orc_bilateral_f32, np.float32
    elif depth.dtype == np.int32:
        fn, dt = lib().orc_bilateral_i32, np.int32
    else:
        raise TypeError(depth.dtype)
    d = _c(depth, dt)
    m = _c(mask, np.uint8)
    out = np.empty_like(d)
    H, W = d.shape
    fn(_p(d), _p(m), _p(out), ctypes.c_int(H), ctypes.c_int(W),
       ctypes.c_int(filter_width), ctypes.c_float(sigma_color),
       ctypes.c_float(sigma_space), ctypes.c_float(depth_scale))
    return out


def normals(xyz, mask):
    xyz = np.asarray(xyz)
    dt = xyz.dtype
    x = _c(xyz, dt)
    out = np.empty_like(x)
    H, W = x.shape[:2]
    getattr(lib(), f"orc_normals_{_sfx(dt)}")(
        _p(x), _p(_c(mask, np.uint8)), _p(out), ctypes.c_int(H), ctypes.c_int(W))
    return out


def erode(mask):
    m = _c(mask, np.uint8)
    out = np.empty_like(m)
    H, W = m.shape
    lib().orc_erode(_p(m), _p(out), ctypes.c_int(H), ctypes.c_int(W))
    return out.astype(bool)


def downsample_xyz(xyz, mask, scale):
    x = _c(xyz, np.float32)
    H, W = x.shape[:2]
    dh, dw = int(H * scale), int(W * scale)
    out = np.empty((dh, dw, 3), np.float32)
    lib().orc_downsample_xyz_f32(_p(x), _p(_c(mask, np.uint8)), ctypes.c_int(H),
                                 ctypes.c_int(W), ctypes.c_float(scale), _p(out),
                                 ctypes.c_int(dh), ctypes.c_int(dw))
    return out


def downsample_mask(mask, scale):
    m = _c(mask, np.uint8)
    H, W = m.shape
    dh, dw = int(H * scale), int(W * scale)
    out = np.empty((dh, dw), np.uint8)
    lib().orc_downsample_mask(_p(m), ctypes.c_int(H), ctypes.c_int(W),
                              ctypes.c_float(scale), _p(out), ctypes.c_int(dh),
                              ctypes.c_int(dw))
    return out.astype(bool)


def downsample_features(feats, scale):
    f = _c(feats, np.float32)
    C, H, W = f.shape
    dh, dw = int(H * scale), int(W * scale)
    out = np.empty((C, dh, dw), np.float32)
    lib().orc_downsample_features_f32(_p(f), ctypes.c_int(C), ctypes.c_int(H),
                                      ctypes.c_int(W), ctypes.c_float(scale),
                                      _p(out), ctypes.c_int(dh), ctypes.c_int(dw))
    return out


def gaussian_blur3(feats, sigma=0.849):
    f = _c(feats, np.float32)
    C, H, W = f.shape
    out = np.empty_like(f)
    lib().orc_gaussian_blur3_f32(_p(f), ctypes.c_int(C), ctypes.c_int(H),
                                 ctypes.c_int(W), ctypes.c_double(sigma), _p(out))
    return out


def feature_pyramid(feats, scales):
    """``slamtb/processing.py:141-159``: blur then 2x2 mean per scale < 1."""
    pyr = []
    cur = np.asarray(feats, np.float32)
    for s in scales:
        if s < 1.0:
            cur = downsample_features(gaussian_blur3(cur), s)
        pyr.append(cur)
    pyr.reverse()
    return pyr


def backproject(depth, kcam, depth_scale=1.0, depth_bias=0.0):
    """``slamtb/_utils.py:11-37`` + ``KCamera.backproject`` (camera.py:127-149)."""
    k = np.asarray(kcam, np.float32).reshape(3, 3)
    z = np.asarray(depth).astype(np.float32) * np.float32(depth_scale) + np.float32(depth_bias)
    H, W = z.shape
    v, u = np.meshgrid(np.arange(H, dtype=np.float32),
                       np.arange(W, dtype=np.float32), indexing="ij")
    x = (u - k[0, 2]) * z / k[0, 0]
    y = (v - k[1, 2]) * z / k[1, 1]
    return np.stack([x, y, z], 2).astype(np.float32)


# ---------------------------------------------------------------------------
# Gauss-Newton driver
# ---------------------------------------------------------------------------

def _cholesky_solve(JtJ, Jr):
    try:
        L = np.linalg.cholesky(JtJ)
    except np.linalg.LinAlgError:
        return None
    if not np.all(np.isfinite(L)):
        return None
    y = np.linalg.solve(L, Jr)
    return np.linalg.solve(L.T, y)


def _sum0(a, sum_dtype, parallel):
    """Sum over the pixel axis, widened to FP64 for the solve. The parallel
    (CPU-baseline) variant sums with torch on the CPU in the partials' own
    dtype, as the reference does (icp.py:78-80); the sequential oracle sums
    in ``sum_dtype`` with numpy."""
    if parallel:
        import torch
        return torch.from_numpy(np.ascontiguousarray(a)).sum(0).double().numpy()
    return np.asarray(a.sum(0, dtype=sum_dtype), np.float64)


def icp_estimate(kcam, src_pts, src_nrm, src_mask, tgt_pts, tgt_nrm, tgt_mask,
                 num_iters, src_feats=None, tgt_feats=None, transform=None,
                 geom_weight=1.0, feat_weight=1.0, distance_threshold=0.5,
                 normals_angle_thresh=np.pi / 4, feat_residual_thresh=2.75,
                 dtype=np.float32, sum_dtype=np.float64, parallel=False,
                 with_information=True):
    """One pyramid level of ``ICPOdometry.estimate`` (icp.py:152-262).

    ``dtype`` is the per-pixel arithmetic type; the partials are summed in
    ``sum_dtype`` (FP64 for the pose oracle); with ``parallel`` (the CPU
    baseline: OpenMP pixel loops) they are summed by torch in FP32 like the
    reference. The system is solved in FP64.
    """
    dt = np.dtype(dtype)
    T = np.eye(4) if transform is None else np.asarray(transform, np.float64).copy()
    sp = np.asarray(src_pts, dt).reshape(-1, 3)
    sn = np.asarray(src_nrm, dt).reshape(-1, 3)
    sm = np.asarray(src_mask).reshape(-1)
    tp = np.asarray(tgt_pts, dt)
    tn = np.asarray(tgt_nrm, dt)
    n = sp.shape[0]
    has_feat = (src_feats is not None and tgt_feats is not None and feat_weight > 0)
    has_geom = geom_weight > 0
    if has_feat:
        sf = np.asarray(src_feats, dt).reshape(np.asarray(src_feats).shape[0], -1)
        tf = np.asarray(tgt_feats, dt)
    result = dict(transform=None, jtj=None, residual=np.inf, match_ratio=0.0,
                  information=None, iters=0, match_count=0)
    tp32 = np.asarray(tgt_pts, np.float32)
    tn32 = np.asarray(tgt_nrm, np.float32)
    for _ in range(num_iters):
        JtJ = np.zeros((6, 6))
        Jr = np.zeros(6)
        residual = 0.0
        count = 0
        rt = T[:3].astype(dt)
        if has_feat:
            cmap = corresp_map(sp, sn, sm, rt, tp, tn, tgt_mask, kcam,
                               distance_threshold, normals_angle_thresh,
                               parallel=parallel)
            jtj, jr, r2, cnt = feature(sp, sf, sm, rt, tf, kcam, cmap,
                                       feat_residual_thresh, parallel=parallel)
            JtJ = _sum0(jtj, sum_dtype, parallel) * feat_weight * feat_weight
            Jr = _sum0(jr, sum_dtype, parallel) * -feat_weight
            residual = float(_sum0(r2, sum_dtype, parallel)) * feat_weight
            count = cnt
        if has_geom:
            cmap = corresp_map(sp, sn, sm, rt, tp, tn, tgt_mask, kcam,
                               distance_threshold, normals_angle_thresh,
                               parallel=parallel)
            jtj, jr, r2, cnt = geometric(sp, sm, rt, tp32, tn32, kcam, cmap,
                                         parallel=parallel)
            JtJ = JtJ + _sum0(jtj, sum_dtype, parallel) * geom_weight * geom_weight
            Jr = Jr + _sum0(jr, sum_dtype, parallel) * geom_weight
            residual = residual + float(_sum0(r2, sum_dtype, parallel)) * geom_weight
            count = max(cnt, count)
        x = _cholesky_solve(JtJ, Jr)
        if x is None:
            break
        U = exp_rt_4x4(x)
        T = U @ T
        result.update(transform=T.copy(), jtj=JtJ, residual=residual / count,
                      match_ratio=count / n, match_count=count)
        result["iters"] += 1
    if with_information and has_geom:
        rt = T[:3].astype(dt)
        cmap = corresp_map(sp, sn, sm, rt, tp, tn, tgt_mask, kcam,
                           distance_threshold, normals_angle_thresh,
                           parallel=parallel)
        result["information"] = information(sp, sm, rt, tp32, kcam, cmap,
                                            parallel=parallel).sum(0, dtype=np.float64)
    result["final_transform"] = T
    return result


def pose_error(Ta, Tb):
    """(rotation angle, translation distance) of Ta^-1 Tb
    (``slamtb/metrics/trajectory.py:16-35``)."""
    D = np.linalg.inv(np.asarray(Ta, np.float64)) @ np.asarray(Tb, np.float64)
    R = D[:3, :3]
    c = (np.trace(R) - 1.0) / 2.0
    s = 0.5 * np.linalg.norm([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]])
    # atan2 form of acos((tr R - 1) / 2): same angle, accurate near zero.
    return float(np.arctan2(s, c)), float(np.linalg.norm(D[:3, 3]))
