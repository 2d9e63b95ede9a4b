"""The no-allocation entry point (stb_icp_estimate_noalloc), called through
the C ABI, against the workspace entry points it varies (stb_icp_begin +
stb_icp_iterate + stb_icp_information) on the same inputs: the same kernels
and launch shapes, so the final states must be bit-identical. Two calls in
a row, the second on a smaller level, show that the static scratch is
cleared by every call; a target beyond the scratch is refused."""
import ctypes
import math

import numpy as np
import pytest
import torch

from slamtb import _lib
from slamtb.registration import ICPOdometry
from slamtb.registration._fused import GNRunner, STATE_BYTES

from _helpers import prepared_pair, to_dev

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def level_of(seed, W, H, channels=3):
    g = to_dev(prepared_pair(seed, W, H, channels=channels), DEV)
    est = ICPOdometry(7, geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
                      normals_angle_thresh=math.radians(20))
    L = est.level(g["kcam"], g["src_pts"], g["src_nrm"], g["src_mask"],
                  g["src_feats"], g["tgt_pts"], g["tgt_mask"], g["tgt_nrm"],
                  g["tgt_feats"])
    return est, L


def with_workspace(est, L, T0):
    run = GNRunner(DEV)
    run.reserve(L.n, L.height, L.width)
    est.enqueue(run, L, T0)
    return run.state_tensor().clone()


def noalloc(est, L, T0, out):
    t0 = torch.as_tensor(T0, dtype=torch.float64).to(DEV).contiguous()
    params = est.params(L.has_features)
    lib = _lib.lib()
    _lib.check(lib.stb_icp_estimate_noalloc(
        ctypes.byref(L.struct), ctypes.byref(params), _lib.ptr(t0), est.num_iters,
        1, _lib.ptr(out), _lib.stream()), "stb_icp_estimate_noalloc")
    torch.cuda.synchronize()
    return out.clone()


def test_noalloc_equals_the_workspace_path_twice_in_a_row():
    rng = np.random.default_rng(3)
    T0 = np.eye(4)
    T0[:3, 3] = rng.uniform(-0.01, 0.01, 3)
    out = torch.zeros(STATE_BYTES, dtype=torch.uint8, device=DEV)
    for seed, (W, H) in ((5, (640, 480)), (6, (320, 240))):
        est, L = level_of(seed, W, H)
        ref = with_workspace(est, L, torch.from_numpy(T0))
        got = noalloc(est, L, T0, out)
        st_ref = _lib.StbIcpState.from_buffer_copy(ref.cpu().numpy().tobytes())
        st_got = _lib.StbIcpState.from_buffer_copy(got.cpu().numpy().tobytes())
        print(W, H, "iterations", st_got.iterations, "pose",
              np.array(list(st_got.transform))[:12])
        assert st_ref.iterations == est.num_iters and st_ref.stopped == 0
        # Every field (the struct's 4 bytes of tail padding are never written).
        end = _lib.StbIcpState.map_sel.offset + 4
        assert torch.equal(got[:end], ref[:end]), (W, H)
        assert np.abs(np.array(list(st_got.information))).sum() > 0


def test_noalloc_refuses_a_target_beyond_its_scratch():
    est, L = level_of(7, 160, 120)
    need = _lib.lib().stb_icp_workspace_bytes(L.n, 2000, 2000)
    assert need > _lib.lib().stb_icp_noalloc_scratch_bytes()
    big = _lib.StbIcpLevel.from_buffer_copy(bytes(L.struct))
    big.height, big.width = 2000, 2000
    out = torch.zeros(STATE_BYTES, dtype=torch.uint8, device=DEV)
    params = est.params(True)
    rc = _lib.lib().stb_icp_estimate_noalloc(ctypes.byref(big), ctypes.byref(params),
                                             None, 3, 1, _lib.ptr(out), _lib.stream())
    assert rc == 1
    assert b"scratch" in _lib.lib().stb_last_error()
