"""Internal helpers (slamtb/_utils.py of the reference)."""
import contextlib
import cProfile

import numpy as np
import torch


def depth_image_to_uvz(depth_image, finfo):
    """[H, W] depth -> [H, W, 3] grid of (u, v, z) with z = depth * scale +
    bias in float32 (slamtb/_utils.py:11-37)."""
    depth_image = depth_image.float() * finfo.depth_scale + finfo.depth_bias
    device = depth_image.device
    dtype = depth_image.dtype
    ys, xs = torch.meshgrid(
        torch.arange(depth_image.size(0), dtype=dtype, device=device),
        torch.arange(depth_image.size(1), dtype=dtype, device=device),
        indexing="ij")
    return torch.stack([xs, ys, depth_image], 2)


def ensure_torch(x, dtype=None):
    """numpy/list -> tensor; tensors pass through (slamtb/_utils.py:40-55)."""
    if isinstance(x, (np.ndarray, list, tuple)):
        x = torch.from_numpy(np.asarray(x))
    if dtype is not None:
        return x.type(dtype)
    return x


def empty_ensured_size(tensor, *sizes, dtype=torch.float, device=None):
    """Reuse ``tensor`` when it already has ``sizes``, dtype and device, else
    allocate (slamtb/_utils.py:58-67)."""
    if (tensor is None or tensor.size() != torch.Size(sizes)
            or tensor.dtype != dtype
            or (device is not None and tensor.device != torch.device(device))):
        tensor = torch.empty(sizes, dtype=dtype, device=device)
    return tensor


@contextlib.contextmanager
def profile(output_file, really=True):
    """cProfile context manager (slamtb/_utils.py:70-81)."""
    if really:
        prof = cProfile.Profile()
        prof.enable()
        yield
        prof.disable()
        prof.dump_stats(str(output_file))
    else:
        yield
