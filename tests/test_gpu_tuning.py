"""The launch-shape tuner (slamtb/tuning.py).

A launch shape changes only the order of the FP32 lane sums, never a
correspondence or a threshold decision: every candidate shape of a level
must register as the oracle does on the same inputs, the tuner's choices
must come from its candidate lists and be cached per level shape, and a
tuned batched registration must give each pair the single-pair result.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from slamtb import synthetic, tuning
from slamtb.registration._fused import Level, make_params, runner_for

from _helpers import prepared_pair, state_pose, to_dev
from test_gpu_batch import make_pairs, single

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
ANG20 = math.radians(20)
ITERS = 4


def _level_and_oracle(W, H):
    P = prepared_pair(21, W, H, channels=3)
    T0 = synthetic.random_se3(np.random.default_rng(5), 2.0, 0.02)
    g = to_dev(P, DEV)
    level = Level(P["kcam"], g["src_pts"], g["src_nrm"], g["src_mask"],
                  g["tgt_pts"], g["tgt_nrm"], g["tgt_mask"], g["src_feats"],
                  g["tgt_feats"])
    ref = orc.icp_estimate(P["kcam"], P["src_pts"], P["src_nrm"], P["src_mask"],
                           P["tgt_pts"], P["tgt_nrm"], P["tgt_mask"], ITERS,
                           src_feats=P["src_feats"], tgt_feats=P["tgt_feats"],
                           geom_weight=10.0, feat_weight=1.0,
                           distance_threshold=0.1, normals_angle_thresh=ANG20,
                           feat_residual_thresh=2.75, transform=T0)
    return level, T0, ref


def _params():
    return make_params(10.0, 1.0, 0.1, ANG20, 2.75, True)


def _register(level, T0, params):
    run = runner_for(DEV)
    run.begin(level, torch.from_numpy(T0))
    run.iterate(level, params, ITERS)
    st = run.read_state()
    return state_pose(st), st.iterations


# 320x240: one pixel per thread; 640x480: the pipelined default shape
# (grid_update 0) against 256-thread blocks looping over pixels.
@pytest.mark.parametrize("W,H", [(320, 240), (640, 480)])
def test_every_candidate_shape_registers_like_the_oracle(W, H):
    level, T0, ref = _level_and_oracle(W, H)
    for gu in tuning.UPDATE_CANDIDATES:
        for gc in tuning.CORRESP_CANDIDATES:
            params = _params()
            params.grid_update, params.grid_corresp = gu, gc
            T, iters = _register(level, T0, params)
            err = orc.pose_error(T, ref["transform"])
            assert iters == ref["iters"] == ITERS, (gu, gc)
            assert err[0] < 1e-6 and err[1] < 1e-6, (gu, gc, err)


def test_tune_level_picks_candidates_and_caches_them():
    level, T0, ref = _level_and_oracle(320, 240)
    tuning._TABLE.clear()
    tuned = tuning.tune_level(level, _params(), reps=3)
    assert tuned.grid_update in tuning.UPDATE_CANDIDATES
    assert tuned.grid_corresp in tuning.CORRESP_CANDIDATES
    assert len(tuning.table()) == 1
    again = tuning.tune_level(level, _params(), reps=3)
    assert (again.grid_update, again.grid_corresp) == (tuned.grid_update,
                                                       tuned.grid_corresp)
    assert len(tuning.table()) == 1
    T, _ = _register(level, T0, tuned)
    err = orc.pose_error(T, ref["transform"])
    assert err[0] < 1e-6 and err[1] < 1e-6, err


def test_tuned_batch_gives_each_pair_the_single_pair_result():
    sizes = [(160, 120), (320, 240), (80, 60)]
    ms, pairs, _ = make_pairs(sizes, seed0=70)
    batch, t = tuning.tune_batch(ms.estimators, pairs, DEV, reps=2)
    assert t > 0
    assert batch.plans[0].params.grid_update in (1024, 4096, 16384, 1 << 20)
    results = batch.estimate()
    for i, levels in enumerate(pairs):
        ref = single(ms, levels)
        err = orc.pose_error(results[i].transform.numpy(), ref.transform.numpy())
        assert err[0] < 1e-6 and err[1] < 1e-6, (sizes[i], err)
        assert results[i].iterations == ref.iterations
