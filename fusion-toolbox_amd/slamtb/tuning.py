"""Autotuner of the fused Gauss-Newton launch shapes.

The fused kernels take their workgroup counts from ``stb_icp_params``
(``grid_update`` for the Jacobian/reduction kernel, ``grid_corresp`` for the
correspondence kernel) and the batched kernels also take a tile size. The
best values depend on resolution, channel count and storage type, so they
are measured: each candidate runs a few iterations on the real inputs and is
timed with HIP events on the current stream; the fastest is kept in a
per-process table keyed by the level's shape.
"""
import torch

from slamtb.registration._fused import runner_for

# 0: the library default shape (at most 256 Jacobian blocks, of 1024 threads
# above 131,072 pixels); a count: that many blocks of 256 threads.
UPDATE_CANDIDATES = (0, 512, 1024, 2048)
# 0: one pixel per thread up to the library's caps; a count: that cap.
CORRESP_CANDIDATES = (0, 256, 1024, 2048)
TILE_CANDIDATES = (256, 512, 1024, 2048)

_TABLE = {}


def _key(level):
    return (level.height, level.width, level.channels,
            int(level.struct.storage), level.n)


def _time(fn, reps):
    s = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    fn()  # warm
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / reps


def tune_level(level, params, reps=20):
    """Pick grid_update / grid_corresp for one level; returns the params
    with both fields set (the input struct is updated in place)."""
    key = _key(level) + (int(params.use_geom), int(params.use_feat))
    if key in _TABLE:
        params.grid_update, params.grid_corresp = _TABLE[key]
        return params
    run = runner_for(level.device)
    run.begin(level)
    run.iterate(level, params, 2)
    best_u, best_t = 0, float("inf")
    # Whole loops are timed: a loop's correspondence launches also apply
    # the previous iteration's update, so their cost depends on both shapes.
    for g in UPDATE_CANDIDATES:
        params.grid_update = g
        run.begin(level)
        t = _time(lambda: run.iterate(level, params, 5), reps)
        if t < best_t:
            best_u, best_t = g, t
    params.grid_update = best_u
    best_c, best_t = 0, float("inf")
    for g in CORRESP_CANDIDATES:
        params.grid_corresp = g
        run.begin(level)
        t = _time(lambda: run.iterate(level, params, 5), reps)
        if t < best_t:
            best_c, best_t = g, t
    params.grid_corresp = best_c
    _TABLE[key] = (best_u, best_c)
    return params


def tune_batch(estimators, pairs, device, reps=5):
    """Pick the tile size (and grid) of a batched registration by timing the
    whole multi-level loop; returns the tuned BatchedMultiscaleICP."""
    from slamtb.registration.batch import BatchedMultiscaleICP
    best, best_t = None, float("inf")
    for tile in TILE_CANDIDATES:
        for grid in (1024, 4096, 16384, 1 << 20):  # 1 << 20: one block per tile
            b = BatchedMultiscaleICP(estimators, pairs, device, tile_px=tile,
                                     grid_update=grid, grid_corresp=grid)
            t = _time(b.enqueue, reps)
            if t < best_t:
                best, best_t = b, t
    return best, best_t


def table():
    return dict(_TABLE)
