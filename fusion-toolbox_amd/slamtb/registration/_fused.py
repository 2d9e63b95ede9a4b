"""Device-resident Gauss-Newton driver over the fused entry points of
libslamtb_hip.so (``stb_icp_begin`` / ``stb_icp_iterate`` /
``stb_icp_information``).

A :class:`GNRunner` owns one workspace per device: the pose state, two
correspondence maps and the per-block reduction partials. Everything is
enqueued on the current stream; the host reads the state once, when a
result is requested. Because all kernel arguments are fixed device
pointers, an enqueued sequence can also be captured in a hipGraph
(``torch.cuda.CUDAGraph``) and replayed.
"""
import ctypes

import torch

from slamtb import _lib

STATE_BYTES = ctypes.sizeof(_lib.StbIcpState)

_STORAGE = {torch.float32: _lib.STB_F32, torch.bfloat16: _lib.STB_BF16,
            torch.float16: _lib.STB_F16}


FEAT_LAYOUTS = {"planar": _lib.STB_FEAT_PLANAR,
                "interleaved": _lib.STB_FEAT_INTERLEAVED}


class Level:
    """One pyramid level of one frame pair, as the C struct plus the tensors
    it points to (kept alive here).

    ``feat_layout="interleaved"`` stores the features channel-innermost
    (source [N, C], target [H, W, C]) instead of the reference's planar
    [C, N] / [C, H, W]: the fused Jacobian then reads a texel's channels in
    16-byte pieces. Inputs are given planar either way; the copy to the
    interleaved layout is made here, once per level."""

    def __init__(self, kcam, src_points, src_normals, src_mask, tgt_points,
                 tgt_normals, tgt_mask, src_feats=None, tgt_feats=None,
                 storage=None, feat_layout="planar"):
        if feat_layout not in FEAT_LAYOUTS:
            raise ValueError(f"feat_layout must be one of {sorted(FEAT_LAYOUTS)}, "
                             f"got {feat_layout!r}")
        _lib.require_gpu(src_points, src_normals, src_mask, tgt_points,
                         tgt_normals, tgt_mask, src_feats, tgt_feats,
                         what="ICPOdometry")
        if storage is None:
            storage = src_points.dtype if src_points.dtype in _STORAGE else torch.float32
        if storage not in _STORAGE:
            raise TypeError(f"map storage must be float32, bfloat16 or float16, got {storage}")
        H, W = tgt_points.size(0), tgt_points.size(1)
        self.src_points = src_points.reshape(-1, 3).to(storage).contiguous()
        n = self.src_points.size(0)
        self.src_normals = src_normals.reshape(-1, 3).to(storage).contiguous()
        mask = src_mask.reshape(-1)
        if mask.numel() != n:
            # The correspondence kernel ignores a mask of another size
            # (src/correspondence_map.cu:40-41); treat it as all-valid.
            mask = torch.ones(n, dtype=torch.bool, device=self.src_points.device)
        self.src_mask = mask.to(torch.bool).contiguous()
        self.tgt_points = tgt_points.to(storage).contiguous()
        self.tgt_normals = tgt_normals.to(storage).contiguous()
        self.tgt_mask = tgt_mask.to(torch.bool).contiguous()
        self.channels = 0
        self.src_feats = self.tgt_feats = None
        self.feat_layout = "planar"
        if src_feats is not None and tgt_feats is not None:
            C = tgt_feats.size(0)
            self.tgt_feats = tgt_feats.to(storage).contiguous()
            self.src_feats = src_feats.reshape(C, -1).to(storage).contiguous()
            if self.src_feats.size(1) != n:
                raise RuntimeError("source features must be [C, N]")
            self.channels = C
            if feat_layout == "interleaved":
                self.tgt_feats = self.tgt_feats.permute(1, 2, 0).contiguous()
                self.src_feats = self.src_feats.t().contiguous()
                self.feat_layout = "interleaved"
        k = kcam.matrix if hasattr(kcam, "matrix") else kcam
        k = torch.as_tensor(k).detach().to("cpu", torch.float32)
        self.fx, self.fy = float(k[0, 0]), float(k[1, 1])
        self.cx, self.cy = float(k[0, 2]), float(k[1, 2])
        self.n = n
        self.height, self.width = H, W
        self.device = self.src_points.device
        self.struct = _lib.StbIcpLevel(
            _lib.ptr(self.src_points), _lib.ptr(self.src_normals),
            _lib.ptr(self.src_mask), _lib.ptr(self.src_feats),
            _lib.ptr(self.tgt_points), _lib.ptr(self.tgt_normals),
            _lib.ptr(self.tgt_mask), _lib.ptr(self.tgt_feats), n, H, W,
            self.channels, _STORAGE[storage], self.fx, self.fy, self.cx,
            self.cy, FEAT_LAYOUTS[self.feat_layout], 0)

    @property
    def has_features(self):
        return self.channels > 0


def make_params(geom_weight, feat_weight, distance_threshold,
                normals_angle_thresh, feat_residual_thresh, use_feat):
    use_geom = geom_weight > 0
    return _lib.StbIcpParams(float(geom_weight), float(feat_weight),
                             float(distance_threshold),
                             float(normals_angle_thresh),
                             float(feat_residual_thresh), int(use_geom),
                             int(bool(use_feat)), 0, 0, 0)


class GNRunner:
    """Workspace and launch helper for the fused GN loop on one device."""

    def __init__(self, device):
        self.device = torch.device(device)
        self.workspace = None
        self.t0 = torch.empty(4, 4, dtype=torch.float64, device=self.device)

    def reserve(self, n_src, height, width):
        need = int(_lib.lib().stb_icp_workspace_bytes(n_src, height, width))
        if self.workspace is None or self.workspace.numel() < need:
            self.workspace = torch.empty(need, dtype=torch.uint8,
                                         device=self.device)
        return self.workspace

    def _ws(self):
        return _lib.ptr(self.workspace), self.workspace.numel()

    def begin(self, level, transform=None, chain=False):
        """Reset the state to ``transform`` (4x4, any device) or identity; with
        ``chain`` start from the previous run's accepted pose (identity if
        none), the pyramid hand-over."""
        self.reserve(level.n, level.height, level.width)
        t0 = None
        if transform is not None and not chain:
            self.t0.copy_(torch.as_tensor(transform).reshape(4, 4))
            t0 = self.t0
        ws, nbytes = self._ws()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().stb_icp_begin(
                ws, nbytes, _lib.ptr(t0), int(bool(chain)), level.n,
                level.height, level.width, _lib.stream()), "stb_icp_begin")

    def iterate(self, level, params, num_iters):
        ws, nbytes = self._ws()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().stb_icp_iterate(
                ctypes.byref(level.struct), ctypes.byref(params), ws, nbytes,
                int(num_iters), _lib.stream()), "stb_icp_iterate")

    def correspond(self, level, params):
        """First half of one iteration (correspondence kernel only)."""
        ws, nbytes = self._ws()
        _lib.check(_lib.lib().stb_icp_correspond(
            ctypes.byref(level.struct), ctypes.byref(params), ws, nbytes,
            _lib.stream()), "stb_icp_correspond")

    def update(self, level, params):
        """Second half of one iteration (Jacobians, reduction, solve)."""
        ws, nbytes = self._ws()
        _lib.check(_lib.lib().stb_icp_update(
            ctypes.byref(level.struct), ctypes.byref(params), ws, nbytes,
            _lib.stream()), "stb_icp_update")

    def linearize(self, level, params):
        ws, nbytes = self._ws()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().stb_icp_linearize(
                ctypes.byref(level.struct), ctypes.byref(params), ws, nbytes,
                _lib.stream()), "stb_icp_linearize")

    def jacobians(self, level, params):
        """The Jacobian launch alone over the last ``linearize`` call's map
        (timing)."""
        ws, nbytes = self._ws()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().stb_icp_jacobians(
                ctypes.byref(level.struct), ctypes.byref(params), ws, nbytes,
                _lib.stream()), "stb_icp_jacobians")

    def information(self, level, params):
        ws, nbytes = self._ws()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().stb_icp_information(
                ctypes.byref(level.struct), ctypes.byref(params), ws, nbytes,
                _lib.stream()), "stb_icp_information")

    def state_tensor(self):
        """The device bytes of the state (no copy)."""
        return self.workspace[:STATE_BYTES]

    def read_state(self):
        """Copy the state to the host (synchronises the stream)."""
        raw = self.state_tensor().cpu().numpy().tobytes()
        return _lib.StbIcpState.from_buffer_copy(raw)


_RUNNERS = {}


def runner_for(device):
    device = torch.device(device)
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    key = (device.type, device.index)
    if key not in _RUNNERS:
        _RUNNERS[key] = GNRunner(device)
    return _RUNNERS[key]


def state_matrix(values, n=4):
    return torch.tensor(list(values), dtype=torch.float64).view(n, n)
