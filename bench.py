#!/usr/bin/env python3
"""Benchmark of the ICP Gauss-Newton step on one MI355X.

Workload (the 640x480 multiscale geometric + feature configuration):
a seeded synthetic scene (6 planes, 4 spheres) rendered to 640x480 float32
depth with fx = fy = 525, cx = 319.5, cy = 239.5 and noise 0.0012 z^2, a
3-channel float32 feature map (albedo + sinusoidal texture), the target moved
by a random rigid motion (<= 5 deg, <= 5 cm). Preprocessing on the GPU:
bilateral filter (sigma_s 4.5 px, sigma_r 0.03 m, width 6), erosion,
back-projection, central-difference normals, 3-level pyramid (640x480,
320x240, 160x120) with feature blur + 2x2 mean. ICP: geometric weight 1,
feature weight 1 (geometric 10), distance 0.1 m, normal angle 20 deg, feature residual
threshold 2.75, 10/5/3 iterations. The 1:0.1 geometric:feature ratio is given as
geom_weight=10, feat_weight=1, the form of the reference preset
(slamtb/registration/preset.py:10); with geom 1 : feat 0.1 the reference
weighting (JtJ by w^2, Jr by w) over-steps the feature term and diverges.

A step is one Gauss-Newton iteration at 640x480 (correspondence kernel +
fused Jacobian/reduction/solve kernel) on inputs resident in HBM; K steps are
captured in a hipGraph and timed with one replay. Also reported: the full
3-level registration per pair, pose error against the oracle fed the same
inputs (FP32 per-pixel arithmetic, FP64 sums and solve) and against the
ground truth, the dominant kernel's roofline and the CPU baseline (the
oracle's OpenMP port of the reference CPU path on this host's cores).
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "fusion-toolbox_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "ICP GN iters/s @640x480; batched pairs/s vs plain; %HBM peak; pose err"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--pairs", type=int, default=64,
                    help="pairs of the batched workload (0 disables it)")
    ap.add_argument("--no-bf16", action="store_true",
                    help="skip the 1280x720 BF16-storage workload")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="only the 640x480 steps and the per-kernel timing "
                         "(for profiler runs: every launch is the same kernel "
                         "at the same size)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="copy a MIB-MiB device buffer 10 times and exit (the "
                         "reference reading for the HBM byte counters)")
    return ap.parse_args()


def copy_probe(mib):
    """A plain device-to-device copy of known size: what the HBM byte
    counters read for it scales the counters of the ICP kernels."""
    import torch
    n = mib << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / 10
    print(json.dumps({"copy_probe_bytes": n, "read_bytes": n, "write_bytes": n,
                      "gbs": 2 * n / dt / 1e9}), flush=True)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def prepare(seed, dev):
    import numpy as np
    import torch
    from slamtb import synthetic
    from slamtb.camera import KCamera
    from slamtb.frame import Frame, FrameInfo, FramePointCloud
    from slamtb.processing import bilateral_depth_filter
    from slamtb.registration import ICPOdometry, MultiscaleRegistration

    k = synthetic.config_kcam(2)
    pair = synthetic.make_pair(seed, 640, 480, kcam=k, channels=3)
    kcam = KCamera(torch.from_numpy(k))
    fpcls = []
    for side in ("src", "tgt"):
        depth = torch.from_numpy(pair[f"{side}_depth"]).to(dev)
        depth = bilateral_depth_filter(depth, depth > 0, None, 6, 0.03, 4.5, 1.0)
        frame = Frame(FrameInfo(kcam, 1.0, 0.0), depth_image=depth)
        fpcls.append(FramePointCloud.from_frame(frame, device=dev))
    ang = math.radians(20)
    kw = dict(geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
              normals_angle_thresh=ang, feat_residual_thresh=2.75)
    ms = MultiscaleRegistration([(1.0, ICPOdometry(10, **kw)),
                                 (0.5, ICPOdometry(5, **kw)),
                                 (0.5, ICPOdometry(3, **kw))])
    feats = [torch.from_numpy(pair[f"{s}_feats"]).to(dev) for s in ("src", "tgt")]
    levels = ms.levels(fpcls[0], fpcls[1], feats[0], feats[1])
    return pair, ms, levels


def time_iteration_kernels(run, level, params, reps):
    """Average duration of each of the two kernels of an iteration, from
    event pairs recorded around each launch on the stream they run on."""
    import torch
    s = torch.cuda.current_stream()
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)]
          for _ in range(reps)]
    for i in range(reps):
        ev[i][0].record(s)
        run.correspond(level, params)
        ev[i][1].record(s)
        run.update(level, params)
        ev[i][2].record(s)
    torch.cuda.synchronize()
    a = sorted(e[0].elapsed_time(e[1]) for e in ev)
    b = sorted(e[1].elapsed_time(e[2]) for e in ev)
    return (sum(a) / reps, a[reps // 2]), (sum(b) / reps, b[reps // 2])


def oracle_levels(levels, T0=None):
    """Run the oracle GN over the same (GPU-prepared) level inputs."""
    import numpy as np
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    T = np.eye(4) if T0 is None else T0
    res = None
    for est, L in levels:
        k = np.array([[L.fx, 0, L.cx], [0, L.fy, L.cy], [0, 0, 1]], np.float32)
        c = lambda t: t.detach().cpu().numpy()
        res = orc.icp_estimate(
            k, c(L.src_points), c(L.src_normals), c(L.src_mask),
            c(L.tgt_points), c(L.tgt_normals), c(L.tgt_mask), est.num_iters,
            src_feats=c(L.src_feats) if L.has_features else None,
            tgt_feats=c(L.tgt_feats) if L.has_features else None, transform=T,
            geom_weight=est.geom_weight, feat_weight=est.feat_weight,
            distance_threshold=est.distance_threshold,
            normals_angle_thresh=est.normals_angle_thresh,
            feat_residual_thresh=est.feat_residual_thresh)
        T = res["transform"] if res["transform"] is not None else np.eye(4)
    return res


def cpu_baseline(level, est, seconds):
    """Reference-structured CPU path (OpenMP per-pixel loops, mutex scatter,
    materialised partials, FP32 sums, FP64 solve) timed on this host."""
    import numpy as np
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    c = lambda t: t.detach().cpu().numpy()
    k = np.array([[level.fx, 0, level.cx], [0, level.fy, level.cy], [0, 0, 1]],
                 np.float32)
    args = (k, c(level.src_points), c(level.src_normals), c(level.src_mask),
            c(level.tgt_points), c(level.tgt_normals), c(level.tgt_mask))
    kw = dict(src_feats=c(level.src_feats), tgt_feats=c(level.tgt_feats),
              geom_weight=est.geom_weight, feat_weight=est.feat_weight,
              distance_threshold=est.distance_threshold,
              normals_angle_thresh=est.normals_angle_thresh,
              feat_residual_thresh=est.feat_residual_thresh,
              sum_dtype=np.float32, parallel=True, with_information=False)
    orc.icp_estimate(*args, 1, **kw)  # warm-up
    iters, t0 = 0, time.perf_counter()
    T = np.eye(4)
    while True:
        r = orc.icp_estimate(*args, 1, transform=T, **kw)
        if r["transform"] is not None:
            T = r["transform"]
        iters += 1
        el = time.perf_counter() - t0
        if el >= seconds and iters >= 3:
            break
    return iters / el, iters, el, orc.num_threads()


def gpu_pair_levels(ms, pair, dev, storage=None):
    """Bilateral-filtered frame point clouds and pyramid levels of a pair."""
    import torch
    from slamtb.camera import KCamera
    from slamtb.frame import Frame, FrameInfo, FramePointCloud
    from slamtb.processing import bilateral_depth_filter
    kcam = KCamera(torch.from_numpy(pair["kcam"]))
    fps = []
    for side in ("src", "tgt"):
        depth = torch.from_numpy(pair[f"{side}_depth"]).to(dev)
        depth = bilateral_depth_filter(depth, depth > 0, None, 6, 0.03, 4.5, 1.0)
        fps.append(FramePointCloud.from_frame(
            Frame(FrameInfo(kcam, 1.0, 0.0), depth_image=depth), device=dev))
    feats = [torch.from_numpy(pair[f"{s}_feats"]).to(dev) for s in ("src", "tgt")]
    return ms.levels(fps[0], fps[1], feats[0], feats[1], storage=storage)


def capture(fn):
    import torch
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()  # run once uncaptured (warms caches, allocations)
        with torch.cuda.graph(g, stream=side):
            fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    return g


def time_graph(g, reps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def batched_section(seed, dev, npairs, tune):
    """64 pairs drawn from {160x120, 320x240, 640x480, 1280x720}, 3 levels
    each (10/5/3 it, geometric + 3-ch features): one batched hipGraph vs the
    same pairs one after another through the single-pair path (also one
    hipGraph)."""
    import numpy as np
    import torch
    from slamtb import synthetic, metrics
    from slamtb.registration import ICPOdometry, MultiscaleRegistration
    from slamtb.registration.batch import BatchedMultiscaleICP
    from slamtb.registration._fused import GNRunner
    from slamtb.registration.icp import result_from_state
    from slamtb import tuning

    rng = np.random.default_rng(seed + 1000)
    choices = [(160, 120), (320, 240), (640, 480), (1280, 720)]
    sizes = [choices[i] for i in rng.integers(0, 4, npairs)]
    kw = dict(geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
              normals_angle_thresh=math.radians(20), feat_residual_thresh=2.75)
    ms = MultiscaleRegistration([(1.0, ICPOdometry(10, **kw)),
                                 (0.5, ICPOdometry(5, **kw)),
                                 (0.5, ICPOdometry(3, **kw))])
    pairs, truths = [], []
    for i, (W, H) in enumerate(sizes):
        pr = synthetic.make_pair(seed + 100 + i, W, H,
                                 kcam=synthetic.intrinsics(W, H))
        pairs.append(gpu_pair_levels(ms, pr, dev))
        truths.append(pr["T_gt"])
    log(f"batched: {npairs} pairs prepared")
    if tune:
        batch, _ = tuning.tune_batch(ms.estimators, pairs, dev)
    else:
        batch = BatchedMultiscaleICP(ms.estimators, pairs, dev)
    g_batch = capture(batch.enqueue)
    t_batch = time_graph(g_batch, 5)
    res_b = batch.results()
    # Plain: every pair through the single-pair fused path, one after the
    # other (its own runner per pair so its result stays readable).
    runners = []
    for levels in pairs:
        r = GNRunner(dev)
        for _, L in levels:
            r.reserve(L.n, L.height, L.width)
        runners.append(r)

    def plain():
        for r, levels in zip(runners, pairs):
            ms.enqueue_levels(r, levels)
    g_plain = capture(plain)
    t_plain = time_graph(g_plain, 3)
    errs, gt_r, gt_t = [], [], []
    for i, (r, levels) in enumerate(zip(runners, pairs)):
        ref = result_from_state(r.read_state(), levels[-1][1].n, True)
        if res_b[i].transform is None or ref.transform is None:
            errs.append((float("nan"), float("nan")))
            continue
        errs.append(metrics.pose_error(ref.transform, res_b[i].transform))
        e = metrics.pose_error(truths[i], res_b[i].transform)
        gt_r.append(e[0])
        gt_t.append(e[1])
    errs = np.array(errs)
    return {
        "pairs": npairs, "tile_px": batch.plans[0].sizes.tile_px,
        "grid": batch.plans[0].params.grid_update,
        "batched_pairs_per_s": npairs / t_batch,
        "plain_pairs_per_s": npairs / t_plain,
        "speedup_vs_plain": t_plain / t_batch,
        "batched_ms": t_batch * 1e3, "plain_ms": t_plain * 1e3,
        "max_pose_err_vs_single_pair": [float(np.nanmax(errs[:, 0])),
                                         float(np.nanmax(errs[:, 1]))],
        "median_pose_err_vs_gt": [float(np.median(gt_r)), float(np.median(gt_t))],
        "resolutions": {f"{W}x{H}": sizes.count((W, H)) for W, H in choices},
    }


def bf16_section(seed, dev, with_oracle):
    """1280x720, 4 levels (10/5/3/2 it), geometric + 32-channel unit
    descriptors; BF16 storage against FP32 storage of the same scene."""
    import numpy as np
    import torch
    from slamtb import synthetic, metrics
    from slamtb.registration import ICPOdometry, MultiscaleRegistration
    from slamtb.registration._fused import GNRunner
    from slamtb.registration.icp import result_from_state

    kw = dict(geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
              normals_angle_thresh=math.radians(20), feat_residual_thresh=2.75)
    ms = MultiscaleRegistration([(1.0, ICPOdometry(10, **kw)),
                                 (0.5, ICPOdometry(5, **kw)),
                                 (0.5, ICPOdometry(3, **kw)),
                                 (0.5, ICPOdometry(2, **kw))])
    pr = synthetic.make_pair(seed + 500, 1280, 720,
                             kcam=synthetic.intrinsics(1280, 720), channels=32,
                             descriptor=True)
    out = {}
    for name, storage in (("bf16", torch.bfloat16), ("fp32", torch.float32)):
        levels = gpu_pair_levels(ms, pr, dev, storage=storage)
        est, top = levels[-1]
        params = est.params(True)
        run = GNRunner(dev)
        for _, L in levels:
            run.reserve(L.n, L.height, L.width)
        K = 100
        run.begin(top)
        g = capture(lambda: run.iterate(top, params, K))
        run.begin(top)
        t = time_graph(g, 1) / K
        s, C = (2 if storage == torch.bfloat16 else 4), 32
        bytes_it = top.n * (12 * s + 18 + 9 * s + 9 + 2 * C * s)
        g_reg = capture(lambda: ms.enqueue_levels(run, levels))
        t_reg = time_graph(g_reg, 3)
        res = result_from_state(run.read_state(), top.n, True)
        entry = {"iters_per_s": 1.0 / t, "ms_per_iter": t * 1e3,
                 "gbs_all_pixels": bytes_it / t / 1e9,
                 "registration_ms": t_reg * 1e3,
                 "pose_err_vs_gt": list(metrics.pose_error(pr["T_gt"], res.transform))}
        if with_oracle and name == "bf16":
            ref = oracle_levels([(e, L) for e, L in levels])
            entry["pose_err_vs_oracle"] = list(metrics.pose_error(ref["transform"],
                                                                  res.transform))
        out[name] = entry
        out[name + "_transform"] = res.transform
    out["pose_diff_bf16_vs_fp32"] = list(metrics.pose_error(out.pop("fp32_transform"),
                                                            out.pop("bf16_transform")))
    return out


def registration_section(args, ms, levels, run, top, pair):
    """The full 3-level registration of the pair captured in one hipGraph:
    ms per pair, the result and its pose error against the ground truth and
    the oracle fed the same level inputs."""
    import torch
    from slamtb import metrics
    from slamtb.registration.icp import result_from_state
    log("full 3-level registration (graph)")
    reg_graph = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        ms.enqueue_levels(run, levels)  # warm the sequence once uncaptured
        with torch.cuda.graph(reg_graph, stream=side):
            ms.enqueue_levels(run, levels)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    reps_reg = 50
    t0 = time.perf_counter()
    for _ in range(reps_reg):
        reg_graph.replay()
    torch.cuda.synchronize()
    reg_ms = (time.perf_counter() - t0) * 1e3 / reps_reg
    result = result_from_state(run.read_state(), top.n, True)
    T_gpu = result.transform.numpy()
    err_gt = metrics.pose_error(pair["T_gt"], T_gpu)
    err_or = None
    if not args.no_oracle:
        log("oracle registration on the same inputs")
        ref = oracle_levels(levels)
        err_or = metrics.pose_error(ref["transform"], T_gpu)
    return reg_ms, result, list(err_gt), err_or


def main():
    args = parse()
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not "
              "supported (independent frame pairs shard with no exchange; "
              "see the multi-GPU note in DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    from slamtb.registration._fused import runner_for

    dev = torch.device("cuda:0")
    log("preparing 640x480 pair on the GPU")
    pair, ms, levels = prepare(args.seed, dev)
    est, top = levels[-1]
    params = est.params(top.has_features)
    run = runner_for(dev)
    for _, L in levels:
        run.reserve(L.n, L.height, L.width)

    # --- warm-up and timed steps -----------------------------------------
    log(f"warm-up {args.warmup} steps")
    run.begin(top)
    run.iterate(top, params, args.warmup)
    torch.cuda.synchronize()
    K = args.steps
    graph = None
    if not args.no_graph:
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            run.begin(top)
            with torch.cuda.graph(graph, stream=side):
                run.iterate(top, params, K)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
    log(f"timing {K} steps ({'graph' if graph else 'stream'})")
    run.begin(top)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if graph is not None:
        graph.replay()
    else:
        run.iterate(top, params, K)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = run.read_state()
    ms_per_step = (t1 - t0) * 1e3 / K
    value = K / (t1 - t0)
    steps_ok = (st.stopped == 0 and st.iterations == K)

    # --- per-kernel timing for the roofline ----------------------------------
    log("per-kernel event timing")
    run.begin(top)
    run.iterate(top, params, 2)
    reps = max(10, min(200, K)) if args.profile else 200
    (a_avg, a_med), (b_only, b_med) = time_iteration_kernels(run, top, params, reps)
    st2 = run.read_state()
    n = top.n
    s, C = 4, top.channels
    valid = float(top.src_mask.float().mean().item())
    matched = max(st2.last_geom_count, st2.last_feat_count) / n
    bytes_a_all = (12 * s + 18) * n
    bytes_b_all = (9 * s + 9 + 2 * C * s) * n
    # Valid/matched-scaled ("honest") bytes: mask bytes for every pixel,
    # source and gather bytes for valid pixels, Jacobian gathers for matches.
    bytes_a = n * (1 + 8 + valid * (6 * s + 1 + 6 * s + 8))
    bytes_b = n * (1 + valid * (3 * s + 8) + matched * (6 * s + 2 * C * s))
    a_gbs = bytes_a / (a_avg * 1e-3) / 1e9
    b_gbs = bytes_b / (b_only * 1e-3) / 1e9
    if a_avg >= b_only:
        dom = dict(kernel="corresp_fused_kernel", ms=a_avg, gbs=a_gbs,
                   bytes=bytes_a, bytes_all=bytes_a_all)
    else:
        dom = dict(kernel="jacobian_fused_kernel", ms=b_only, gbs=b_gbs,
                   bytes=bytes_b, bytes_all=bytes_b_all)

    if args.profile:
        args.no_oracle = args.no_cpu_baseline = args.no_bf16 = True
        args.pairs = 0
    reg_ms, result, err_gt, err_or = None, None, None, None
    if not args.profile:
        reg_ms, result, err_gt, err_or = registration_section(
            args, ms, levels, run, top, pair)

    batched = None
    if args.pairs > 0:
        log(f"batched workload: {args.pairs} pairs")
        batched = batched_section(args.seed, dev, args.pairs, not args.no_tune)
    bf16 = None
    if not args.no_bf16:
        log("1280x720 BF16 / FP32 storage workload")
        bf16 = bf16_section(args.seed, dev, not args.no_oracle)

    cpu = None
    if not args.no_cpu_baseline:
        log(f"CPU baseline (~{args.cpu_seconds:.0f} s)")
        rate, iters, el, thr = cpu_baseline(top, est, args.cpu_seconds)
        cpu = {"value": rate, "unit": "GN iters/s", "cores": thr,
               "kind": "port",
               "sample": f"{iters} GN iterations at 640x480 (geometric + 3-ch "
                         f"feature, FP32) of the oracle's OpenMP port of the "
                         f"reference CPU path in {el:.1f} s on {thr} threads"}

    out = {
        "metric": METRIC, "value": value, "unit": "GN iters/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "icp-gn-640x480-geom+feat3-fp32",
                   "resolution": "640x480", "feature_channels": C,
                   "levels": "640x480/320x240/160x120 10/5/3 it",
                   "weights": "geom 10 : feat 1 (the 1:0.1 ratio in the reference preset form)", "timing": "hipGraph" if graph else "stream",
                   "steps_all_accepted": steps_ok},
        "roofline": {"bound": "hbm", "achieved": dom["gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": dom["gbs"] / HBM_PEAK_GBS, "traffic": None,
                     "kernel": dom["kernel"], "kernel_ms": dom["ms"],
                     "bytes_per_launch": dom["bytes"],
                     "achieved_all_pixels": dom["bytes_all"] / (dom["ms"] * 1e-3) / 1e9},
        "kernels_ms": {"corresp_fused_kernel": a_avg, "jacobian_fused_kernel": b_only,
                       "corresp_median": a_med, "jacobian_median": b_med},
        "registration_ms_per_pair": reg_ms,
        "pose_err_vs_oracle": err_or, "pose_err_vs_gt": err_gt,
        "match_ratio": None if result is None else result.match_ratio,
        "batched": batched,
        "bf16_1280x720": bf16,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
