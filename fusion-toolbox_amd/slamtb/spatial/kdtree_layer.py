"""Differentiable nearest-neighbour feature layer
(slamtb/spatial/kdtree_layer.py).

Same protocol as the reference: ``setup(target_xyz)`` registers the target
points and returns the layer's ``apply``; ``query(xyz, k,
distance_upper_bound)`` finds neighbours and keeps them in ``last_query``;
``apply(features, xyz)`` interpolates target features at the query points
from those neighbours, differentiably in ``xyz``.

The reference searched a host k-d tree (pykdtree, approximate with
eps = 0.5, points copied to the host per query); here the search is the
exact brute-force kernel of libslamtb_hip.so on the device, so ``tree`` is
the target point tensor itself.
"""
import math

import torch

from slamtb._cslamtb import KnnGrid, NearestNeighborsOp as _NearestNeighborsOp, knn


class KDTreeLayer(torch.autograd.Function):
    """kdtree_layer.py:12-107."""
    tree = None
    target_xyz = None
    last_query = (None, None)

    _gradcheck = False
    _grid = None
    _key = None

    @staticmethod
    def _target_key(target_xyz):
        """Identity of the target's current contents: its storage, shape,
        dtype and version counter (bumped by every in-place write to the
        tensor or a view of it)."""
        return (target_xyz.data_ptr(), tuple(target_xyz.shape), target_xyz.dtype,
                target_xyz._version)

    @staticmethod
    def _refresh():
        """Re-read the target (and drop the search grid) when it is another
        tensor or has been written in place since the last read."""
        key = KDTreeLayer._target_key(KDTreeLayer.target_xyz)
        if key != KDTreeLayer._key:
            KDTreeLayer._key = key
            KDTreeLayer.tree = KDTreeLayer.target_xyz.detach().reshape(-1, 3).contiguous()
            KDTreeLayer._grid = None

    @staticmethod
    def setup(target_xyz):
        KDTreeLayer.target_xyz = target_xyz
        KDTreeLayer._refresh()
        return KDTreeLayer.apply

    @staticmethod
    def query(xyz, k=1, distance_upper_bound=math.inf):
        """Returns the [N] mask of queries with at least one neighbour;
        ``last_query`` = (squared distances [N, k], index [N, k]). A finite
        bound reuses the target grid built for the first such query while
        the target is unchanged (the reference rebuilt its k-d tree in every
        setup; here the grid is rebuilt when the target is another tensor or
        was written in place)."""
        KDTreeLayer._refresh()
        tree = KDTreeLayer.tree
        q = xyz.detach().reshape(-1, 3)
        grid = KDTreeLayer._grid
        if math.isfinite(distance_upper_bound) and (
                grid is None or grid.cell < distance_upper_bound
                or grid.dtype != q.dtype):
            grid = KDTreeLayer._grid = KnnGrid(tree.to(q.dtype), distance_upper_bound)
        distance, index = knn(tree.to(q.dtype), q, k, distance_upper_bound, grid=grid)
        KDTreeLayer.last_query = (distance, index)
        return index[:, 0] < tree.size(0)

    @staticmethod
    def forward(ctx, features, xyz):
        if KDTreeLayer._gradcheck:
            KDTreeLayer.query(xyz, k=8)
        distances, index = KDTreeLayer.last_query
        ctx.save_for_backward(xyz, index, features)
        out = torch.empty(features.size(0), xyz.size(0), dtype=features.dtype,
                          device=features.device)
        _NearestNeighborsOp.forward(distances.to(features.dtype), index,
                                    features, out)
        return out

    @staticmethod
    def backward(ctx, dl_feat):
        source_xyz, index, features = ctx.saved_tensors
        eps = torch.empty(index.size(0), 6, index.size(1), dtype=features.dtype,
                          device=features.device)
        _NearestNeighborsOp.compute_epsilon_distances(
            KDTreeLayer.target_xyz.detach().to(features.dtype), source_xyz.detach(),
            index, eps)
        dl_xyz = torch.empty(index.size(0), 3, dtype=features.dtype,
                             device=features.device)
        _NearestNeighborsOp.backward(eps, index, features, dl_feat.contiguous(),
                                     dl_xyz)
        return None, dl_xyz
