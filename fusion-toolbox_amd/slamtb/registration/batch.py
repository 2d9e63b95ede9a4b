"""Batched multiscale ICP: many frame pairs of uneven resolution registered
together on one GPU (the batch entry points of libslamtb_hip.so).

For every pyramid level, all pairs' source pixels form one flattened tile
list; each Gauss-Newton iteration is one correspondence launch and one
Jacobian launch over that list (one partial per tile), then one launch with
a workgroup per pair that sums the pair's partials, solves its system and
updates its pose on the device. The result equals running
``MultiscaleRegistration.estimate_frame`` on each pair with the same
estimators (same per-pixel arithmetic; only the order of the sums over
pixels differs).
"""
import ctypes

import torch

from slamtb import _lib

from ._fused import Level, STATE_BYTES
from .icp import ICPOdometry, result_from_state

DEFAULT_TILE_PX = 1024


class _PlanLevel:
    def __init__(self, est, levels, tile_px, device):
        self.est = est
        self.levels = levels
        self.params = est.params(levels[0].has_features)
        arr = (_lib.StbIcpLevel * len(levels))(*[L.struct for L in levels])
        self.sizes = _lib.StbIcpBatchSizes()
        lib = _lib.lib()
        _lib.check(lib.stb_icp_batch_plan(len(levels), arr, tile_px, None, 0,
                                          ctypes.byref(self.sizes)),
                   "stb_icp_batch_plan")
        host = (ctypes.c_uint8 * self.sizes.plan_bytes)()
        _lib.check(lib.stb_icp_batch_plan(len(levels), arr, tile_px, host,
                                          self.sizes.plan_bytes,
                                          ctypes.byref(self.sizes)),
                   "stb_icp_batch_plan")
        self.plan = torch.frombuffer(bytearray(host), dtype=torch.uint8).to(device)


class BatchedMultiscaleICP:
    """Registers a list of pairs with the estimators of a
    ``MultiscaleRegistration`` made of ``ICPOdometry`` levels.

    ``pairs`` is a list of per-pair level lists (coarse first), as returned
    by ``MultiscaleRegistration.levels`` for each pair.
    """

    def __init__(self, estimators, pairs, device="cuda:0", tile_px=DEFAULT_TILE_PX,
                 grid_update=0, grid_corresp=0):
        if not all(isinstance(est, ICPOdometry) for _, est in estimators):
            raise TypeError("batched registration needs ICPOdometry estimators")
        self.device = torch.device(device)
        self.num_pairs = len(pairs)
        nlev = len(pairs[0])
        if any(len(p) != nlev for p in pairs):
            raise ValueError("every pair needs the same number of levels")
        self.plans = []
        for li in range(nlev):
            est = pairs[0][li][0]
            plan = _PlanLevel(est, [p[li][1] for p in pairs], tile_px, self.device)
            plan.params.grid_update = grid_update
            plan.params.grid_corresp = grid_corresp
            self.plans.append(plan)
        self.states = torch.zeros(self.num_pairs, STATE_BYTES, dtype=torch.uint8,
                                  device=self.device)
        need = max(p.sizes.scratch_bytes for p in self.plans)
        self.scratch = torch.empty(need, dtype=torch.uint8, device=self.device)
        self.t0 = torch.empty(self.num_pairs, 4, 4, dtype=torch.float64,
                              device=self.device)

    def _args(self, plan):
        return (_lib.ptr(plan.plan), ctypes.byref(plan.sizes))

    def _scratch(self):
        return _lib.ptr(self.scratch), self.scratch.numel()

    def begin_level(self, i, transforms=None):
        """Reset every pair for level ``i`` (coarse first): level 0 starts
        from the staged initial poses when ``transforms`` is given (see
        ``set_initial``), else identity; later levels start from each pair's
        last accepted pose of the level before."""
        scratch, nbytes = self._scratch()
        t0 = _lib.ptr(self.t0) if (i == 0 and transforms is not None) else None
        _lib.check(_lib.lib().stb_icp_batch_begin(
            *self._args(self.plans[i]), _lib.ptr(self.states), t0, int(i > 0),
            scratch, nbytes, _lib.stream()), "stb_icp_batch_begin")

    def iterate_level(self, i, num_iters=None):
        """``num_iters`` Gauss-Newton iterations of every pair at level ``i``
        (default: the level's estimator's count). Iterations enqueued one
        call at a time give the same states as one call for all of them."""
        plan = self.plans[i]
        n = plan.est.num_iters if num_iters is None else int(num_iters)
        scratch, nbytes = self._scratch()
        _lib.check(_lib.lib().stb_icp_batch_iterate(
            *self._args(plan), ctypes.byref(plan.params), _lib.ptr(self.states),
            scratch, nbytes, n, _lib.stream()), "stb_icp_batch_iterate")

    def information_level(self, i):
        """Every pair's information matrix at its current pose, level ``i``."""
        plan = self.plans[i]
        scratch, nbytes = self._scratch()
        _lib.check(_lib.lib().stb_icp_batch_information(
            *self._args(plan), ctypes.byref(plan.params), _lib.ptr(self.states),
            scratch, nbytes, _lib.stream()), "stb_icp_batch_information")

    def enqueue(self, transforms=None):
        """Enqueue every level of every pair (capturable in a hipGraph when
        ``transforms`` is None or already staged with ``set_initial``)."""
        for i, plan in enumerate(self.plans):
            self.begin_level(i, transforms)
            self.iterate_level(i)
            if i == len(self.plans) - 1 and plan.est._geom_step is not None:
                self.information_level(i)

    def read_states(self):
        """Every pair's device state (one host read for the batch)."""
        raw = self.states.cpu().numpy()
        return [_lib.StbIcpState.from_buffer_copy(raw[p].tobytes())
                for p in range(self.num_pairs)]

    def set_initial(self, transforms):
        self.t0.copy_(torch.as_tensor(transforms, dtype=torch.float64).view(-1, 4, 4))

    def results(self):
        """Per-pair RegistrationResult (one host read for the batch)."""
        last = self.plans[-1]
        return [result_from_state(st, last.levels[p].n,
                                  last.est._geom_step is not None)
                for p, st in enumerate(self.read_states())]

    def estimate(self, transforms=None):
        if transforms is not None:
            self.set_initial(transforms)
        self.enqueue(transforms)
        return self.results()
