"""The C-ABI library loads without a GPU, exports every symbol the header
declares, and the host-side helpers behave (no device compute here)."""
import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest
import torch

from slamtb import _lib

HEADER = Path(__file__).resolve().parents[1] / "include" / "slamtb_hip.h"


def declared_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(stb_\w+)\s*\(", text)))


def test_header_symbols_exported():
    names = declared_functions()
    assert len(names) >= 20
    lib = _lib.lib()
    for name in names:
        assert hasattr(lib, name), name


def test_binding_covers_header():
    assert set(declared_functions()) <= set(_lib.EXPORTED)


def test_state_struct_size_matches_library_layout():
    # The ctypes mirror has the size the library was compiled with:
    # 16+16+36+36+1+1+36+6+4 doubles, 12+9 floats, 4 ints (+4 bytes of tail
    # padding to the 8-byte alignment).
    assert ctypes.sizeof(_lib.StbIcpState) == _lib.lib().stb_icp_state_bytes()
    assert ctypes.sizeof(_lib.StbIcpState) == 152 * 8 + 21 * 4 + 4 * 4 + 4
    assert _lib.StbIcpState.rt.offset == 152 * 8
    assert _lib.StbIcpState.map_sel.offset == 152 * 8 + 21 * 4 + 3 * 4


def test_workspace_size_grows_with_the_target():
    lib = _lib.lib()
    a = lib.stb_icp_workspace_bytes(320 * 240, 240, 320)
    b = lib.stb_icp_workspace_bytes(640 * 480, 480, 640)
    assert b > a >= 2 * 320 * 240 * 8


@pytest.mark.parametrize("thr", [0.0, 0.1, math.radians(20), math.pi / 4, 1.0,
                                 math.pi / 2, 3.0, math.pi, 3.5])
def test_angle_gate_threshold_is_exact(thr):
    """c* splits [-1, 1] exactly where the float acos gate flips: every float
    within 64 ulps of c* agrees with acosf(c) >= thr (float32 libm)."""
    lib = _lib.lib()
    cstar = np.float32(lib.stb_angle_gate_cos_f32(thr))
    thr32 = np.float32(thr)
    libm = ctypes.CDLL("libm.so.6")
    libm.acosf.restype = ctypes.c_float
    libm.acosf.argtypes = [ctypes.c_float]

    def rejects(c):
        a = libm.acosf(float(c))
        return not math.isnan(a) and np.float32(abs(a)) >= thr32

    if cstar < -1:
        assert not rejects(np.float32(-1.0))
        return
    neighbours = [cstar]
    for direction in (np.float32(2), np.float32(-2)):
        c = cstar
        for _ in range(64):
            c = np.nextafter(c, direction, dtype=np.float32)
            neighbours.append(c)
    for c in neighbours:
        if -1.0 <= c <= 1.0:
            assert rejects(c) == (c <= cstar), (c, cstar)


def test_cpu_tensors_are_rejected():
    from slamtb._cslamtb import Processing
    m = torch.ones(4, 4, dtype=torch.bool)
    with pytest.raises(RuntimeError, match="gpu tensor"):
        Processing.erode_mask(m, torch.empty_like(m))


def test_batch_plan_layout_on_the_host():
    """stb_icp_batch_plan is host-only: tile and map tables of uneven pairs."""
    lib = _lib.lib()
    dims = [(120, 160), (240, 320), (60, 80)]
    levels = (_lib.StbIcpLevel * 3)()
    for i, (H, W) in enumerate(dims):
        fake = ctypes.c_void_p(0x1000 * (i + 1))
        levels[i] = _lib.StbIcpLevel(fake, fake, fake, None, fake, fake, fake,
                                     None, H * W, H, W, 0, _lib.STB_F32,
                                     1.0, 1.0, 0.0, 0.0)
    sizes = _lib.StbIcpBatchSizes()
    assert lib.stb_icp_batch_plan(3, levels, 1024, None, 0, ctypes.byref(sizes)) == 0
    tiles = [-(-H * W // 1024) for H, W in dims]
    assert sizes.total_tiles == sum(tiles)
    assert sizes.total_map == sum(H * W for H, W in dims)
    buf = (ctypes.c_uint8 * sizes.plan_bytes)()
    assert lib.stb_icp_batch_plan(3, levels, 1024, buf, sizes.plan_bytes,
                                  ctypes.byref(sizes)) == 0
    raw = bytes(buf)
    assert np.frombuffer(raw[:16], np.int32).tolist() == [3, sum(tiles), 1024, _lib.STB_F32]
    pairs = np.frombuffer(raw[-((sum(tiles) * 4 + 63) // 64 * 64):], np.int32)[:sum(tiles)]
    assert np.array_equal(pairs, np.repeat(np.arange(3), tiles))
    # tile_px must be a multiple of the 256-lane workgroup
    assert lib.stb_icp_batch_plan(3, levels, 100, None, 0, ctypes.byref(sizes)) != 0


def test_knn_size_limits_are_errors():
    """The neighbour searches give each query its own thread and the grid
    search keeps 31-bit indices: sizes past either limit are refused on the
    host with STB_ERR_INVALID before anything reaches the device."""
    lib = _lib.lib()
    fake = ctypes.c_void_p(256)  # never dereferenced: the checks fail first
    i64 = ctypes.cast(fake, ctypes.POINTER(ctypes.c_int64))
    too_many = (1 << 28) + 1
    rc = lib.stb_knn(0, fake, 10, fake, too_many, 8, math.inf, fake, i64, None)
    assert rc == 1 and b"2^28" in lib.stb_last_error()
    rc = lib.stb_knn_grid(0, fake, i64, 2, 2, 2, 0.0, 0.0, 0.0, 0.1, 10, fake,
                          too_many, None, 8, 0.1, fake, i64, None)
    assert rc == 1 and b"2^28" in lib.stb_last_error()
    rc = lib.stb_knn_grid(0, fake, i64, 2, 2, 2, 0.0, 0.0, 0.0, 0.1, 1 << 31,
                          fake, 10, None, 8, 0.1, fake, i64, None)
    assert rc == 1 and b"31-bit" in lib.stb_last_error()


def test_level_struct_layout_and_interleaved_validation():
    """The level descriptor ends with feat_layout (8-byte aligned, 112
    bytes). Interleaved features are refused on the host, before anything
    reaches the device, unless the storage is 16-bit and the channel count
    even and at most 32: by the batch planner, and by the single-pair entry
    points."""
    lib = _lib.lib()
    assert ctypes.sizeof(_lib.StbIcpLevel) == 112
    assert _lib.StbIcpLevel.feat_layout.offset == 104

    fake = ctypes.c_void_p(0x1000)

    def level(storage, channels, layout):
        return _lib.StbIcpLevel(fake, fake, fake, fake, fake, fake, fake, fake,
                                64 * 48, 48, 64, channels, storage, 1.0, 1.0,
                                0.0, 0.0, layout, 0)

    sizes = _lib.StbIcpBatchSizes()
    for storage, channels, ok in [(_lib.STB_BF16, 32, True),
                                  (_lib.STB_F16, 6, True),
                                  (_lib.STB_BF16, 3, False),
                                  (_lib.STB_BF16, 34, False),
                                  (_lib.STB_F32, 8, False)]:
        arr = (_lib.StbIcpLevel * 2)(level(storage, channels, _lib.STB_FEAT_INTERLEAVED),
                                     level(storage, channels, _lib.STB_FEAT_INTERLEAVED))
        rc = lib.stb_icp_batch_plan(2, arr, 1024, None, 0, ctypes.byref(sizes))
        assert (rc == 0) == ok, (storage, channels)
        if ok:
            assert sizes.feat_layout == _lib.STB_FEAT_INTERLEAVED
    mixed = (_lib.StbIcpLevel * 2)(level(_lib.STB_BF16, 8, _lib.STB_FEAT_INTERLEAVED),
                                   level(_lib.STB_BF16, 8, _lib.STB_FEAT_PLANAR))
    assert lib.stb_icp_batch_plan(2, mixed, 1024, None, 0, ctypes.byref(sizes)) != 0
    assert b"feature layout" in lib.stb_last_error()

    params = _lib.StbIcpParams(10.0, 1.0, 0.1, 0.3, 2.75, 1, 1, 0, 0, 0)
    for storage, channels in [(_lib.STB_F32, 8), (_lib.STB_BF16, 7)]:
        lv = level(storage, channels, _lib.STB_FEAT_INTERLEAVED)
        rc = lib.stb_icp_iterate(ctypes.byref(lv), ctypes.byref(params), fake,
                                 1 << 30, 1, None)
        assert rc != 0 and b"interleaved" in lib.stb_last_error()
    lv = level(_lib.STB_BF16, 8, 7)
    rc = lib.stb_icp_iterate(ctypes.byref(lv), ctypes.byref(params), fake,
                             1 << 30, 1, None)
    assert rc != 0 and b"feat_layout" in lib.stb_last_error()
