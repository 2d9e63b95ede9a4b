"""ctypes binding of ``libslamtb_hip.so`` (declared in ``include/slamtb_hip.h``).

The library is the only compute path: every public op in this package runs
on a HIP device through it, and a missing library or a CPU tensor raises
instead of falling back to anything else.
"""
import ctypes
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().parent / "libslamtb_hip.so"

STB_F32, STB_F64, STB_BF16, STB_F16, STB_I32, STB_U8, STB_I64, STB_I16 = range(8)
STB_FEAT_PLANAR, STB_FEAT_INTERLEAVED = 0, 1

_TORCH_DTYPE = {
    torch.float32: STB_F32,
    torch.float64: STB_F64,
    torch.bfloat16: STB_BF16,
    torch.float16: STB_F16,
    torch.int32: STB_I32,
    torch.uint8: STB_U8,
    torch.bool: STB_U8,
    torch.int64: STB_I64,
    torch.int16: STB_I16,
}

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f32 = ctypes.c_float
_f64 = ctypes.c_double
_sz = ctypes.c_size_t


class StbIcpLevel(ctypes.Structure):
    _fields_ = [
        ("src_points", _vp), ("src_normals", _vp), ("src_mask", _vp),
        ("src_feats", _vp), ("tgt_points", _vp), ("tgt_normals", _vp),
        ("tgt_mask", _vp), ("tgt_feats", _vp), ("n_src", _i64),
        ("height", _i32), ("width", _i32), ("channels", _i32),
        ("storage", _i32), ("fx", _f32), ("fy", _f32), ("cx", _f32),
        ("cy", _f32), ("feat_layout", _i32), ("reserved", _i32),
    ]


class StbIcpParams(ctypes.Structure):
    _fields_ = [
        ("geom_weight", _f64), ("feat_weight", _f64),
        ("distance_thresh", _f64), ("normal_angle_thresh", _f64),
        ("feat_residual_thresh", _f32), ("use_geom", _i32),
        ("use_feat", _i32), ("feat_mode", _i32), ("grid_update", _i32),
        ("grid_corresp", _i32),
    ]


class StbIcpState(ctypes.Structure):
    _fields_ = [
        ("transform", _f64 * 16), ("best_transform", _f64 * 16),
        ("best_jtj", _f64 * 36), ("information", _f64 * 36),
        ("best_residual", _f64), ("best_match_count", _f64),
        ("last_jtj", _f64 * 36), ("last_jr", _f64 * 6),
        ("last_geom_r2", _f64), ("last_feat_r2", _f64),
        ("last_geom_count", _f64), ("last_feat_count", _f64),
        ("rt", _f32 * 12), ("nm", _f32 * 9), ("stopped", _i32),
        ("iterations", _i32), ("map_epoch", _i32), ("map_sel", _i32),
    ]


class StbIcpBatchSizes(ctypes.Structure):
    _fields_ = [
        ("num_pairs", _i32), ("total_tiles", _i32), ("tile_px", _i32),
        ("storage", _i32), ("total_map", _i64), ("plan_bytes", _sz),
        ("scratch_bytes", _sz), ("max_channels", _i32), ("feat_layout", _i32),
    ]


_BATCH = ctypes.POINTER(StbIcpBatchSizes)
_PARAMS = ctypes.POINTER(StbIcpParams)

_SIGNATURES = {
    "stb_last_error": (ctypes.c_char_p, []),
    "stb_abi_version": (ctypes.c_int, []),
    "stb_icp_state_bytes": (_sz, []),
    "stb_angle_gate_cos_f32": (_f32, [_f64]),
    "stb_angle_gate_cos_f64": (_f64, [_f64]),
    "stb_compute_correspondence_map": (ctypes.c_int, [
        ctypes.c_int, _vp, _vp, _vp, _i64, _i64, _vp, ctypes.c_int, _vp, _vp,
        _vp, _i32, _i32, _vp, _vp, _f64, _f64, _vp]),
    "stb_icp_estimate_geometric": (ctypes.c_int, [
        ctypes.c_int, _vp, _vp, _i64, _vp, ctypes.c_int, _vp, _vp, _i32, _i32,
        _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "stb_icp_get_information_matrix": (ctypes.c_int, [
        ctypes.c_int, _vp, _vp, _i64, _vp, ctypes.c_int, _vp, _i32, _i32, _vp,
        _vp, _vp, _vp]),
    "stb_icp_estimate_feature": (ctypes.c_int, [
        ctypes.c_int, _vp, _vp, _vp, _i64, _vp, ctypes.c_int, _vp, _i32, _i32,
        _i32, _vp, _vp, _f32, _vp, _vp, _vp, _vp, _vp]),
    "stb_bilateral_depth_filter": (ctypes.c_int, [
        ctypes.c_int, _vp, _vp, _vp, _i32, _i32, _i32, _f32, _f32, _f32, _vp]),
    "stb_estimate_normals": (ctypes.c_int, [
        ctypes.c_int, _vp, _vp, _vp, _i32, _i32, ctypes.c_int, _vp]),
    "stb_downsample_xyz": (ctypes.c_int, [
        ctypes.c_int, _vp, _vp, _i32, _i32, _f32, _vp, _i32, _i32,
        ctypes.c_int, ctypes.c_int, _vp]),
    "stb_downsample_mask": (ctypes.c_int, [
        _vp, _i32, _i32, _f32, _vp, _i32, _i32, _vp]),
    "stb_downsample_features": (ctypes.c_int, [
        ctypes.c_int, _vp, _i32, _i32, _i32, _f32, _vp, _i32, _i32, _vp]),
    "stb_erode_mask": (ctypes.c_int, [_vp, _vp, _i32, _i32, _vp]),
    "stb_gaussian_blur3x3": (ctypes.c_int, [
        ctypes.c_int, _vp, _i32, _i32, _i32, _f64, _vp, _vp]),
    "stb_depth_to_points": (ctypes.c_int, [
        ctypes.c_int, _vp, _i32, _i32, _f32, _f32, _f32, _f32, _f32, _f32,
        _vp, _vp]),
    "stb_exp_rt_to_matrix_forward": (ctypes.c_int, [
        ctypes.c_int, _vp, _i64, _vp, _vp]),
    "stb_exp_rt_to_matrix_backward": (ctypes.c_int, [
        ctypes.c_int, _vp, _vp, _vp, _i64, _vp, _vp]),
    "stb_matrix_to_exp_rt_forward": (ctypes.c_int, [
        ctypes.c_int, _vp, _i64, _vp, _vp]),
    "stb_fpcl_match_forward": (ctypes.c_int, [
        _i32, _vp, _vp, _vp, _i32, _i32, _i32, _vp, _i64, _vp, _vp, _vp, _vp,
        _vp, _vp, _vp]),
    "stb_fpcl_match_backward": (ctypes.c_int, [
        _i32, _vp, _i32, _i32, _i32, _vp, _i64, _vp, _vp, _vp, _f64, _vp, _vp]),
    "stb_project_forward": (ctypes.c_int, [_i32, _vp, _i64, _vp, _vp, _vp]),
    "stb_project_backward": (ctypes.c_int, [_i32, _vp, _vp, _i64, _vp, _vp, _vp]),
    "stb_rigid_transform_points": (ctypes.c_int, [_i32, _vp, _vp, _i64, _vp, _vp]),
    "stb_rigid_transform_normals": (ctypes.c_int, [_i32, _vp, _vp, _i64, _vp, _vp]),
    "stb_rigid_backward_workspace_bytes": (_sz, []),
    "stb_knn": (ctypes.c_int, [_i32, _vp, _i64, _vp, _i64, _i32, _f64, _vp, _vp, _vp]),
    "stb_grid_cells": (ctypes.c_int, [
        _i32, _vp, _i64, _f64, _f64, _f64, _f64, _i32, _i32, _i32, _vp, _vp]),
    "stb_knn_grid": (ctypes.c_int, [
        _i32, _vp, _vp, _i32, _i32, _i32, _f64, _f64, _f64, _f64, _i64, _vp,
        _i64, _vp, _i32, _f64, _vp, _vp, _vp]),
    "stb_knn_grid_set_lane_threshold": (_i64, [_i64]),
    "stb_nn_features_forward": (ctypes.c_int, [
        _i32, _vp, _vp, _i64, _i32, _vp, _i32, _i64, _vp, _vp]),
    "stb_nn_epsilon_distances": (ctypes.c_int, [
        _i32, _vp, _i64, _vp, _vp, _i64, _i32, _f64, _vp, _vp]),
    "stb_nn_features_backward": (ctypes.c_int, [
        _i32, _vp, _vp, _i64, _i32, _vp, _i32, _i64, _vp, _f64, _vp, _vp]),
    "stb_voxel_ids": (ctypes.c_int, [
        _vp, _i64, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
        _i32, _i32, _vp, _vp]),
    "stb_voxel_merge": (ctypes.c_int, [
        _vp, _vp, _vp, _vp, _vp, _vp, _i32, _i64, _vp, _vp, _i64, _vp, _vp, _vp,
        _vp, _vp, _vp, _vp, _vp]),
    "stb_rigid_transform_backward": (ctypes.c_int, [
        _i32, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _sz, _vp]),
    "stb_icp_noalloc_scratch_bytes": (_sz, []),
    "stb_icp_estimate_noalloc": (ctypes.c_int, [
        ctypes.POINTER(StbIcpLevel), ctypes.POINTER(StbIcpParams), _vp, _i32,
        _i32, _vp, _vp]),
    "stb_icp_workspace_bytes": (_sz, [_i64, _i32, _i32]),
    "stb_icp_begin": (ctypes.c_int, [_vp, _sz, _vp, _i32, _i64, _i32, _i32, _vp]),
    "stb_icp_iterate": (ctypes.c_int, [
        ctypes.POINTER(StbIcpLevel), ctypes.POINTER(StbIcpParams), _vp, _sz,
        _i32, _vp]),
    "stb_icp_correspond": (ctypes.c_int, [
        ctypes.POINTER(StbIcpLevel), ctypes.POINTER(StbIcpParams), _vp, _sz,
        _vp]),
    "stb_icp_update": (ctypes.c_int, [
        ctypes.POINTER(StbIcpLevel), ctypes.POINTER(StbIcpParams), _vp, _sz,
        _vp]),
    "stb_icp_jacobians": (ctypes.c_int, [
        ctypes.POINTER(StbIcpLevel), ctypes.POINTER(StbIcpParams), _vp, _sz,
        _vp]),
    "stb_icp_linearize": (ctypes.c_int, [
        ctypes.POINTER(StbIcpLevel), ctypes.POINTER(StbIcpParams), _vp, _sz,
        _vp]),
    "stb_icp_information": (ctypes.c_int, [
        ctypes.POINTER(StbIcpLevel), ctypes.POINTER(StbIcpParams), _vp, _sz,
        _vp]),
    "stb_icp_batch_plan": (ctypes.c_int, [
        _i32, ctypes.POINTER(StbIcpLevel), _i32, _vp, _sz, _BATCH]),
    "stb_icp_batch_begin": (ctypes.c_int, [
        _vp, _BATCH, _vp, _vp, _i32, _vp, _sz, _vp]),
    "stb_icp_batch_iterate": (ctypes.c_int, [
        _vp, _BATCH, _PARAMS, _vp, _vp, _sz, _i32, _vp]),
    "stb_icp_batch_linearize": (ctypes.c_int, [
        _vp, _BATCH, _PARAMS, _vp, _vp, _sz, _vp]),
    "stb_icp_batch_information": (ctypes.c_int, [
        _vp, _BATCH, _PARAMS, _vp, _vp, _sz, _vp]),
    "stb_icp_batch_system64": (ctypes.c_int, [
        _vp, _BATCH, _PARAMS, _vp, _vp, _vp, _sz, _vp]),
    "stb_icp_batch_adjoint": (ctypes.c_int, [
        _vp, _BATCH, _PARAMS, _vp, _vp, _vp, _vp, _vp, _sz, _vp]),
}

EXPORTED = tuple(_SIGNATURES)

_LIB = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` "
                "(or `make -C fusion-toolbox_amd/csrc`). There is no CPU "
                "fallback.")
        handle = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def check(rc, what=""):
    if rc != 0:
        msg = lib().stb_last_error().decode(errors="replace")
        raise RuntimeError(f"{what}: {msg}" if what else msg)


def dtype_code(t):
    try:
        return _TORCH_DTYPE[t.dtype]
    except KeyError as err:
        raise TypeError(f"unsupported dtype {t.dtype}") from err


def ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def stream(device=None):
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def require_gpu(*tensors, what="slamtb"):
    """The reference's STB_CHECK_DEVICE (include/slamtb/error.hpp:43-45):
    every tensor on one device -- which here must be a HIP device."""
    dev = None
    for t in tensors:
        if t is None:
            continue
        if not t.is_cuda:
            raise RuntimeError(
                f"Check failed: {what}: Expected a gpu tensor (this library "
                "runs on HIP devices only)")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise RuntimeError(
                f"Check failed: {what}: Expected all tensors on {dev}, got "
                f"{t.device}")
    return dev


def require_contiguous_out(*tensors, what="slamtb"):
    for t in tensors:
        if t is not None and not t.is_contiguous():
            raise RuntimeError(f"{what}: output tensors must be contiguous")
