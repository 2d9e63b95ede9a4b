#!/usr/bin/env python3
"""Benchmark of the ICP Gauss-Newton step on one MI355X.

Headline workload (640x480 multiscale, geometric + 3-channel features):
a seeded synthetic scene (6 planes, 4 spheres) rendered to 640x480 float32
depth with fx = fy = 525, cx = 319.5, cy = 239.5 and noise 0.0012 z^2, a
3-channel float32 feature map (albedo + sinusoidal texture), the target moved
by a random rigid motion (<= 5 deg, <= 5 cm). Preprocessing on the GPU:
bilateral filter (sigma_s 4.5 px, sigma_r 0.03 m, width 6), erosion,
back-projection, central-difference normals, 3-level pyramid (640x480,
320x240, 160x120) with feature blur + 2x2 mean. ICP: distance 0.1 m, normal
angle 20 deg, feature residual threshold 2.75, 10/5/3 iterations, weights
geometric 10 : feature 1 -- the 1:0.1 ratio in the form of the reference
preset (slamtb/registration/preset.py:10); written as 1 : 0.1 the reference
weighting (JtJ by w^2, Jr by w) over-steps the feature term and diverges.

A step is one Gauss-Newton iteration at 640x480 (correspondence launch,
Jacobian launch, reduction/solve launch) on inputs resident in HBM; K steps
are captured in one hipGraph and timed with one replay between
synchronisations. Also reported:
  * per-launch durations of the correspondence and Jacobian kernels (HIP
    events around a captured graph of back-to-back launches of that kernel
    alone) and the dominant kernel's roofline against 8 TB/s, with the HBM
    bytes that rocprofv3 counters measured for it (profiles/), a
    device-to-device copy ceiling measured in the same process, and the
    whole iteration's fraction of 8 TB/s beside it;
  * the full 3-level registration per pair, its pose error against the
    oracle fed the same level inputs -- FP32 per-pixel arithmetic with FP64
    sums (decision-matched) and all-FP64 -- and against the ground truth;
  * 64 pairs of uneven resolution in one batched hipGraph against the same
    pairs one after another through the single-pair path;
  * 1280x720 with 32-channel descriptors in BF16 and FP16 storage against
    FP32;
  * 5 unrolled differentiable GN steps on 8 pairs at 640x480, forward and
    adjoint backward;
  * the differentiable frame matcher and projection kernels at 640x480 and
    AutogradICP driven by them, with the oracle's matcher beside them;
  * point-cloud registration: surfel clouds of the pair voxel-merged at
    1 cm, the exact neighbour search, the neighbour feature layer and
    AutogradICP.estimate_pcl;
  * the reference's own RGB-D test frames: the Registration notebook's pair
    against the camera poses recorded with the frames, and frame-to-frame
    tracking over 15 frames;
  * config 1 (320x240 geometric, one level, 10 iterations): iterations per
    second, the registration per pair, and its pose, JtJ and Jtr against the
    oracle, free-running and step by step from the GPU's poses;
  * the CPU baseline: the oracle's OpenMP port of the reference CPU path on
    this host, at all threads and at one thread, as medians of 20 runs after
    3 warm-ups: one 640x480 iteration, the 3-level 640x480 registration and
    the config-1 registration.
"""
import argparse
import json
import math
import platform
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "fusion-toolbox_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "ICP GN iters/s @640x480; batched pairs/s vs plain; %HBM peak; pose err"
# The FP32 Jacobian launch of the 640x480 level (pipelined: up to three
# pixels per thread in blocks of 512).
JAC_KERNEL = "jacobian_pipe_kernel"
TRAFFIC_FILE = ROOT / "profiles" / "jacobian_pipe_kernel_640x480_traffic.json"
KW = dict(geom_weight=10.0, feat_weight=1.0, distance_threshold=0.1,
          normals_angle_thresh=math.radians(20), feat_residual_thresh=2.75)


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
    ap.add_argument("--no-pcl", action="store_true",
                    help="skip the point-cloud registration section")
    ap.add_argument("--no-ftb", action="store_true",
                    help="skip the real-frame (FTB sample1) section")
    ap.add_argument("--no-matcher", action="store_true",
                    help="skip the differentiable matcher / AutogradICP section")
    ap.add_argument("--no-diff", action="store_true",
                    help="skip the differentiable 8 x 640x480 workload")
    ap.add_argument("--no-tune", action="store_true",
                    help="batched workload: default tile and grid instead of "
                         "the autotuned ones")
    ap.add_argument("--tune", action="store_true",
                    help="640x480 steps: autotuned grids instead of the "
                         "library defaults")
    ap.add_argument("--profile", action="store_true",
                    help="only the 640x480 steps and the per-kernel timing "
                         "(for profiler runs: every launch is the same kernel "
                         "at the same size)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="copy a MIB-MiB device buffer 10 times and exit (the "
                         "reference reading for the HBM byte counters)")
    return ap.parse_args()


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def copy_rate(mib, reps=10):
    """Device-to-device copy of a known size: (bytes per copy, GB/s read +
    written)."""
    import torch
    n = mib << 20
    a = torch.ones(n, dtype=torch.uint8, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / reps
    return n, 2 * n / dt / 1e9


def copy_probe(mib):
    """What the HBM byte counters read for this copy scales the counters of
    the ICP kernels."""
    n, gbs = copy_rate(mib)
    print(json.dumps({"copy_probe_bytes": n, "read_bytes": n, "write_bytes": n,
                      "gbs": gbs}), flush=True)


def prepare(seed, dev):
    import numpy as np
    from slamtb import synthetic
    from slamtb.registration import ICPOdometry, MultiscaleRegistration

    k = synthetic.intrinsics(640, 480).astype(np.float32)
    pair = synthetic.make_pair(seed, 640, 480, kcam=k, channels=3)
    ms = MultiscaleRegistration([(1.0, ICPOdometry(10, **KW)),
                                 (0.5, ICPOdometry(5, **KW)),
                                 (0.5, ICPOdometry(3, **KW))])
    return pair, ms, gpu_pair_levels(ms, pair, dev)


def gpu_pair_levels(ms, pair, dev, storage=None, feat_layout="planar"):
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
    return ms.levels(fps[0], fps[1], feats[0], feats[1], storage=storage,
                     feat_layout=feat_layout)


def capture(fn):
    """hipGraph of fn() (run once uncaptured first)."""
    import torch
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()
        with torch.cuda.graph(g, stream=side):
            fn()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    return g


def time_graph(g, reps):
    """Mean wall time of one replay (host clock around synchronisations)."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        g.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def launch_ms(fn, reps=100):
    """Average duration of one launch of fn: HIP events on the replaying
    stream around a hipGraph of `reps` back-to-back launches."""
    import torch
    g = capture(lambda: [fn() for _ in range(reps)])
    s = torch.cuda.current_stream()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    g.replay()
    a.record(s)
    g.replay()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / reps


def oracle_levels(levels, T0=None, dtype=None):
    """The oracle GN over the same (GPU-prepared) level inputs: FP32 per-pixel
    arithmetic with FP64 sums (the decision-matched oracle), or everything in
    FP64 with dtype=float64."""
    import numpy as np
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    T = np.eye(4) if T0 is None else T0
    res = None
    for est, L in levels:
        k = np.array([[L.fx, 0, L.cx], [0, L.fy, L.cy], [0, 0, 1]], np.float32)
        c = lambda t: t.detach().float().cpu().numpy()
        res = orc.icp_estimate(
            k, c(L.src_points), c(L.src_normals), L.src_mask.cpu().numpy(),
            c(L.tgt_points), c(L.tgt_normals), L.tgt_mask.cpu().numpy(),
            est.num_iters,
            src_feats=c(L.src_feats) if L.has_features else None,
            tgt_feats=c(L.tgt_feats) if L.has_features else None, transform=T,
            geom_weight=est.geom_weight, feat_weight=est.feat_weight,
            distance_threshold=est.distance_threshold,
            normals_angle_thresh=est.normals_angle_thresh,
            feat_residual_thresh=est.feat_residual_thresh,
            dtype=dtype or np.float32)
        T = res["transform"] if res["transform"] is not None else np.eye(4)
    return res


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _median_time(fn, runs=20, warmup=3):
    for _ in range(warmup):
        fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts)


def _oracle_level_args(L):
    import numpy as np
    c = lambda t: t.detach().float().cpu().numpy()
    k = np.array([[L.fx, 0, L.cx], [0, L.fy, L.cy], [0, 0, 1]], np.float32)
    args = (k, c(L.src_points), c(L.src_normals), L.src_mask.cpu().numpy(),
            c(L.tgt_points), c(L.tgt_normals), L.tgt_mask.cpu().numpy())
    feats = (dict(src_feats=c(L.src_feats), tgt_feats=c(L.tgt_feats))
             if L.has_features else {})
    return args, feats


def _cpu_registration(levels, parallel_kw):
    """The reference CPU path over a pyramid (coarse first): each level's GN
    loop from the previous level's pose."""
    import numpy as np
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    prepared = [(est, _oracle_level_args(L)) for est, L in levels]

    def run():
        T = np.eye(4)
        for est, (args, feats) in prepared:
            r = orc.icp_estimate(*args, est.num_iters, transform=T,
                                 geom_weight=est.geom_weight,
                                 feat_weight=est.feat_weight,
                                 distance_threshold=est.distance_threshold,
                                 normals_angle_thresh=est.normals_angle_thresh,
                                 feat_residual_thresh=est.feat_residual_thresh,
                                 **feats, **parallel_kw)
            if r["transform"] is not None:
                T = r["transform"]
        return T
    return run


def cpu_baseline(levels, config1_levels):
    """The reference-structured CPU path (OpenMP per-pixel loops, a mutex
    scatter into the map, materialised [N,6,6] partials summed by torch on
    the CPU in FP32, FP64 solve) on this host, all threads and one thread:
    the median of 20 runs after 3 warm-ups of one 640x480 GN iteration
    (geometric + 3-channel features), of the whole 3-level 640x480
    registration (10/5/3 iterations) and of the 320x240 geometric
    registration (10 iterations)."""
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    import numpy as np
    kw = dict(sum_dtype=np.float32, parallel=True, with_information=False)
    est, top = levels[-1]
    args, feats = _oracle_level_args(top)

    def one_iter():
        orc.icp_estimate(*args, 1, geom_weight=est.geom_weight,
                         feat_weight=est.feat_weight,
                         distance_threshold=est.distance_threshold,
                         normals_angle_thresh=est.normals_angle_thresh,
                         feat_residual_thresh=est.feat_residual_thresh,
                         **feats, **kw)
    reg2 = _cpu_registration(levels, kw)
    reg1 = _cpu_registration(config1_levels, kw)
    threads = orc.num_threads()
    out = {"threads": threads}
    for name, n in (("all_threads", threads), ("one_thread", 1)):
        orc.set_num_threads(n)
        out[name] = {"iter_640x480_s": _median_time(one_iter),
                     "registration_640x480_3level_s": _median_time(reg2),
                     "registration_320x240_geom_s": _median_time(reg1)}
    orc.set_num_threads(threads)
    return out


def registration_section(args, ms, levels, run, top, pair):
    """The 3-level registration of the pair in one hipGraph: ms per pair and
    its pose error against the ground truth and the oracles."""
    import numpy as np
    from slamtb import metrics
    from slamtb.registration.icp import result_from_state
    log("full 3-level registration (graph)")
    g = capture(lambda: ms.enqueue_levels(run, levels))
    reg_ms = time_graph(g, 50) * 1e3
    result = result_from_state(run.read_state(), top.n, True)
    T_gpu = result.transform.numpy()
    out = {"registration_ms_per_pair": reg_ms,
           "pose_err_vs_gt": list(metrics.pose_error(pair["T_gt"], T_gpu)),
           "match_ratio": result.match_ratio}
    if not args.no_oracle:
        log("oracle registrations on the same inputs (FP32 pixels, then FP64)")
        ref = oracle_levels(levels)
        out["pose_err_vs_oracle"] = list(metrics.pose_error(ref["transform"], T_gpu))
        ref64 = oracle_levels(levels, dtype=np.float64)
        out["pose_err_vs_oracle_fp64"] = list(
            metrics.pose_error(ref64["transform"], T_gpu))
    return out


def config1_section(seed, dev, with_oracle):
    """Config 1: plain geometric point-to-plane ICP at 320x240 (fx = fy =
    262.5, cx = 159.5, cy = 119.5), central-difference normals, one level,
    10 iterations, 0.1 m / 20 deg. GPU iterations per second (hipGraph of
    400 iterations), the whole registration (begin, 10 iterations,
    information) per pair, and the pose, JtJ and Jtr against the oracle run
    on the same inputs (FP32 pixels, FP64 sums), free-running and step by
    step from the GPU's own poses."""
    import numpy as np
    import torch
    from slamtb import metrics, synthetic
    from slamtb.camera import KCamera
    from slamtb.frame import Frame, FrameInfo, FramePointCloud
    from slamtb.registration import ICPOdometry
    from slamtb.registration._fused import GNRunner
    from slamtb.registration.icp import result_from_state

    k = synthetic.intrinsics(320, 240).astype(np.float32)
    pair = synthetic.make_pair(seed + 300, 320, 240, kcam=k)
    kc = KCamera(torch.from_numpy(k))
    fps = [FramePointCloud.from_frame(
        Frame(FrameInfo(kc), depth_image=torch.from_numpy(pair[f"{side}_depth"]).to(dev)),
        device=dev) for side in ("src", "tgt")]
    est = ICPOdometry(10, geom_weight=1.0, feat_weight=0.0, distance_threshold=0.1,
                      normals_angle_thresh=math.radians(20))
    L = est.level(kc, fps[0].points, fps[0].normals, fps[0].mask, None,
                  fps[1].points, fps[1].mask, fps[1].normals, None)
    params = est.params(False)
    run = GNRunner(dev)
    run.reserve(L.n, L.height, L.width)
    K = 400
    g = capture(lambda: (run.begin(L), run.iterate(L, params, K)))
    t_it = time_graph(g, 3) / K
    g_reg = capture(lambda: est.enqueue(run, L))
    t_reg = time_graph(g_reg, 50)
    st = run.read_state()
    res = result_from_state(st, L.n, True)
    out = {"workload": "320x240 geometric point-to-plane, 1 level, 10 iterations",
           "iters_per_s": 1.0 / t_it, "ms_per_iter": t_it * 1e3,
           "registration_ms_per_pair": t_reg * 1e3,
           "bytes_per_iter_all_pixels": L.n * 111,
           "gbs_all_pixels": L.n * 111 / t_it / 1e9,
           "pose_err_vs_gt": list(metrics.pose_error(pair["T_gt"], res.transform)),
           "match_ratio": res.match_ratio}
    if with_oracle:
        sys.path.insert(0, str(ROOT))
        from oracle import oracle as orc
        args, _ = _oracle_level_args(L)
        kw = dict(geom_weight=1.0, feat_weight=0.0, distance_threshold=0.1,
                  normals_angle_thresh=math.radians(20))
        ref = orc.icp_estimate(*args, 10, **kw)
        out["pose_err_vs_oracle"] = list(orc.pose_error(ref["transform"],
                                                        res.transform.numpy()))
        jtj = np.array(list(st.best_jtj)).reshape(6, 6)
        out["jtj_rel_diff_vs_oracle"] = float(np.linalg.norm(jtj - ref["jtj"]) /
                                              np.linalg.norm(ref["jtj"]))
        # Step by step: each iteration from the GPU's pose before it.
        run.begin(L)
        worst, jr_worst = [0.0, 0.0], 0.0
        for _ in range(10):
            T0 = np.array(list(run.read_state().transform)).reshape(4, 4)
            run.iterate(L, params, 1)
            s1 = run.read_state()
            r1 = orc.icp_estimate(*args, 1, transform=T0, with_information=False, **kw)
            e = orc.pose_error(r1["transform"], np.array(list(s1.transform)).reshape(4, 4))
            worst = [max(worst[0], e[0]), max(worst[1], e[1])]
            jr_ref = _oracle_jr(orc, args, T0, kw)
            jr_worst = max(jr_worst, float(np.linalg.norm(np.array(list(s1.last_jr)) - jr_ref) /
                                           np.linalg.norm(jr_ref)))
        out["max_step_err_vs_oracle_from_gpu_pose"] = worst
        out["max_jtr_rel_diff_vs_oracle"] = jr_worst
    return out, [(est, L)]


def _oracle_jr(orc, args, T, kw):
    """The oracle's combined Jr of one iteration at pose T (geometric term)."""
    import numpy as np
    k, sp, sn, sm, tp, tn, tm = args
    rt = T[:3].astype(np.float32)
    cmap = orc.corresp_map(sp, sn, sm, rt, tp, tn, tm, k, kw["distance_threshold"],
                           kw["normals_angle_thresh"])
    _, jr, _, _ = orc.geometric(sp, sm, rt, tp, tn, k, cmap)
    return jr.sum(0, dtype=np.float64) * kw["geom_weight"]


def batched_section(seed, dev, npairs, tune, with_oracle=True):
    """npairs pairs drawn from {160x120, 320x240, 640x480, 1280x720}, 3 levels
    each (10/5/3 it, geometric + 3-ch features): one batched hipGraph against
    the same pairs one after another through the single-pair path (also one
    hipGraph)."""
    import numpy as np
    from slamtb import synthetic, metrics
    from slamtb.registration import ICPOdometry, MultiscaleRegistration
    from slamtb.registration.batch import BatchedMultiscaleICP
    from slamtb.registration._fused import GNRunner
    from slamtb.registration.icp import result_from_state
    from slamtb import tuning

    rng = np.random.default_rng(seed + 1000)
    choices = [(160, 120), (320, 240), (640, 480), (1280, 720)]
    sizes = [choices[i] for i in rng.integers(0, 4, npairs)]
    ms = MultiscaleRegistration([(1.0, ICPOdometry(10, **KW)),
                                 (0.5, ICPOdometry(5, **KW)),
                                 (0.5, ICPOdometry(3, **KW))])
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
    px_iters = sum(L.n * est.num_iters for levels in pairs for est, L in levels)
    # Algorithmic bytes of the whole batch (every level and iteration of
    # every pair): 12s+18 (correspondence) + 9s+9+2Cs (Jacobians) per
    # source pixel and iteration counting all pixels, and the honest count
    # (mask and map words for every pixel, source data and gathers for
    # valid pixels, Jacobian gathers for matched pixels; each pair's final
    # match ratio stands for its matched fraction at every level).
    s_b, C = 4, 3
    b_all = b_honest = 0.0
    for i, levels in enumerate(pairs):
        mr = res_b[i].match_ratio
        for est, L in levels:
            n, it = L.n, est.num_iters
            valid = float(L.src_mask.float().mean().item())
            b_all += n * it * (12 * s_b + 18 + 9 * s_b + 9 + 2 * C * s_b)
            b_honest += n * it * (1 + 8 + valid * (6 * s_b + 1 + 6 * s_b + 8) +
                                  1 + valid * (3 * s_b + 8) +
                                  min(mr, valid) * (6 * s_b + 2 * C * s_b))
    out = {
        "pairs": npairs, "tile_px": batch.plans[0].sizes.tile_px,
        "grid": batch.plans[0].params.grid_update,
        "batched_pairs_per_s": npairs / t_batch,
        "plain_pairs_per_s": npairs / t_plain,
        "speedup_vs_plain": t_plain / t_batch,
        "batched_ms": t_batch * 1e3, "plain_ms": t_plain * 1e3,
        "batched_gpx_iters_per_s": px_iters / t_batch / 1e9,
        "roofline": {"bound": "hbm", "unit": "GB/s", "peak": HBM_PEAK_GBS,
                     "bytes_all_pixels": b_all, "bytes_honest": b_honest,
                     "achieved_all_pixels": b_all / t_batch / 1e9,
                     "achieved": b_honest / t_batch / 1e9,
                     "frac": b_honest / t_batch / 1e9 / HBM_PEAK_GBS,
                     "frac_all_pixels": b_all / t_batch / 1e9 / HBM_PEAK_GBS,
                     "plain_frac_all_pixels": b_all / t_plain / 1e9 / HBM_PEAK_GBS},
        "free_run_max_pose_diff_vs_single_pair": [float(np.nanmax(errs[:, 0])),
                                                  float(np.nanmax(errs[:, 1]))],
        "median_pose_err_vs_gt": [float(np.median(gt_r)), float(np.median(gt_t))],
        "resolutions": {f"{W}x{H}": sizes.count((W, H)) for W, H in choices},
    }
    if with_oracle:
        out["pose_err_vs_decision_matched_oracle"] = batched_oracle_check(batch, pairs)
    return out


def batched_oracle_check(batch, pairs):
    """Each pair's last Gauss-Newton iteration against the oracle run from
    the same pose (the same FP32 per-pixel decisions, FP64 sums and solve):
    every level but the last, and the last level but its last iteration,
    run as in the timed graph; the states are read; the last iteration runs;
    the oracle takes that one iteration from the read poses. Per pair the
    (rotation, translation) difference; the maximum over the pairs."""
    import numpy as np
    sys.path.insert(0, str(ROOT))
    from oracle import oracle as orc
    last = len(batch.plans) - 1
    for li in range(last + 1):
        batch.begin_level(li)
        n = batch.plans[li].est.num_iters
        batch.iterate_level(li, n - 1 if li == last else n)
    before = batch.read_states()
    batch.iterate_level(last, 1)
    after = batch.read_states()
    pose = lambda st: np.array(list(st.transform), np.float64).reshape(4, 4)
    per_pair = []
    for p, levels in enumerate(pairs):
        est, L = levels[last]
        k = np.array([[L.fx, 0, L.cx], [0, L.fy, L.cy], [0, 0, 1]], np.float32)
        c = lambda t: t.detach().float().cpu().numpy()
        ref = orc.icp_estimate(
            k, c(L.src_points), c(L.src_normals), L.src_mask.cpu().numpy(),
            c(L.tgt_points), c(L.tgt_normals), L.tgt_mask.cpu().numpy(), 1,
            src_feats=c(L.src_feats), tgt_feats=c(L.tgt_feats),
            transform=pose(before[p]), geom_weight=est.geom_weight,
            feat_weight=est.feat_weight, distance_threshold=est.distance_threshold,
            normals_angle_thresh=est.normals_angle_thresh,
            feat_residual_thresh=est.feat_residual_thresh, with_information=False)
        per_pair.append(orc.pose_error(ref["transform"], pose(after[p])))
    per_pair = np.array(per_pair)
    return {"max": [float(per_pair[:, 0].max()), float(per_pair[:, 1].max())],
            "per_pair": per_pair.tolist(),
            "check": "last iteration of the finest level of every pair, from "
                     "the GPU's pose before it"}


def bf16_section(seed, dev, with_oracle):
    """1280x720, 4 levels (10/5/3/2 it), geometric + 32-channel unit
    descriptors; BF16 and FP16 storage against FP32 storage of the same
    scene, the 16-bit ones also with channel-interleaved features (the
    same results bit for bit): iterations per second at the top level, its
    correspondence and Jacobian launches alone, the whole registration, and
    pose errors."""
    import numpy as np
    import torch
    from slamtb import synthetic, metrics
    from slamtb.registration import ICPOdometry, MultiscaleRegistration
    from slamtb.registration._fused import GNRunner
    from slamtb.registration.icp import result_from_state

    ms = MultiscaleRegistration([(1.0, ICPOdometry(10, **KW)),
                                 (0.5, ICPOdometry(5, **KW)),
                                 (0.5, ICPOdometry(3, **KW)),
                                 (0.5, ICPOdometry(2, **KW))])
    pr = synthetic.make_pair(seed + 500, 1280, 720,
                             kcam=synthetic.intrinsics(1280, 720), channels=32,
                             descriptor=True)
    out, poses = {}, {}
    for name, storage, layout in (
            ("bf16", torch.bfloat16, "planar"),
            ("bf16_interleaved", torch.bfloat16, "interleaved"),
            ("fp16", torch.float16, "planar"),
            ("fp16_interleaved", torch.float16, "interleaved"),
            ("fp32", torch.float32, "planar")):
        levels = gpu_pair_levels(ms, pr, dev, storage=storage, feat_layout=layout)
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
        s_b, C = (4 if storage == torch.float32 else 2), 32
        bytes_it = top.n * (12 * s_b + 18 + 9 * s_b + 9 + 2 * C * s_b)
        run.begin(top)
        run.iterate(top, params, 2)
        corr = launch_ms(lambda: run.correspond(top, params), 50)
        run.linearize(top, params)
        jac = launch_ms(lambda: run.jacobians(top, params), 50)
        matched = run.read_state().last_geom_count
        g_reg = capture(lambda: ms.enqueue_levels(run, levels))
        t_reg = time_graph(g_reg, 3)
        res = result_from_state(run.read_state(), top.n, True)
        entry = {"iters_per_s": 1.0 / t, "ms_per_iter": t * 1e3,
                 "gbs_all_pixels": bytes_it / t / 1e9,
                 "launch_ms": {"correspondence": corr, "jacobian": jac},
                 "matched_top_level": matched,
                 "registration_ms": t_reg * 1e3,
                 "pose_err_vs_gt": list(metrics.pose_error(pr["T_gt"], res.transform))}
        if with_oracle and layout == "planar" and name != "fp32":
            ref = oracle_levels([(e, L) for e, L in levels])
            entry["pose_err_vs_oracle"] = list(metrics.pose_error(ref["transform"],
                                                                  res.transform))
        out[name] = entry
        poses[name] = res.transform
    for name in ("bf16", "fp16"):
        out[f"pose_diff_{name}_vs_fp32"] = list(metrics.pose_error(poses["fp32"],
                                                                   poses[name]))
        out[f"{name}_interleaved_equals_planar"] = bool(np.array_equal(
            np.asarray(poses[name]), np.asarray(poses[f"{name}_interleaved"])))
    best16 = max(out["bf16"]["iters_per_s"], out["bf16_interleaved"]["iters_per_s"])
    out["bf16_best_over_fp32"] = best16 / out["fp32"]["iters_per_s"]
    return out


def diff_section(seed, dev, batch=8, steps=5):
    """5 unrolled GN steps on `batch` pairs at 640x480 (bilateral-filtered
    depth, geometric + 3-channel features), forward and the adjoint backward
    of the squared geodesic pose loss to the ground truth, with respect to
    the source depth, target points and normals, both feature maps and the
    initial pose."""
    import numpy as np
    import torch
    from slamtb import synthetic
    from slamtb.camera import KCamera
    from slamtb.frame import Frame, FrameInfo, FramePointCloud
    from slamtb.registration.diff import DifferentiableICP, pose_loss

    k = synthetic.intrinsics(640, 480).astype(np.float32)
    kc = KCamera(torch.from_numpy(k))
    pairs = [synthetic.make_pair(seed + 700 + i, 640, 480, kcam=k, channels=3)
             for i in range(batch)]

    from slamtb.processing import bilateral_depth_filter

    def filtered(depth):
        d = torch.from_numpy(depth).to(dev)
        return bilateral_depth_filter(d, d > 0, None, 6, 0.03, 4.5, 1.0)

    def fpc(depth):
        return FramePointCloud.from_frame(
            Frame(FrameInfo(kc), depth_image=filtered(depth)), device=dev)
    src = [fpc(p["src_depth"]) for p in pairs]
    tgt = [fpc(p["tgt_depth"]) for p in pairs]
    st = lambda xs: torch.stack(xs).contiguous()
    depth = st([filtered(p["src_depth"]) for p in pairs])
    src_n, src_m = st([f.normals for f in src]), st([f.mask for f in src])
    tgt_p, tgt_n = st([f.points for f in tgt]), st([f.normals for f in tgt])
    tgt_m = st([f.mask for f in tgt])
    sf = st([torch.from_numpy(p["src_feats"]) for p in pairs]).to(dev)
    tf = st([torch.from_numpy(p["tgt_feats"]) for p in pairs]).to(dev)
    T_gt = torch.from_numpy(np.stack([p["T_gt"] for p in pairs])).to(dev)
    leaves = [depth, tgt_p, tgt_n, sf, tf]
    for t in leaves:
        t.requires_grad_()
    x0 = torch.zeros(batch, 6, dtype=torch.float64, device=dev, requires_grad=True)
    icp = DifferentiableICP(steps, **KW)

    def forward():
        return icp(k, depth, src_n, src_m, tgt_p, tgt_n, tgt_m, x0,
                   src_feats=sf, tgt_feats=tf)

    def step():
        for t in leaves + [x0]:
            t.grad = None
        pose_loss(forward(), T_gt).sum().backward()

    step()
    torch.cuda.synchronize()
    fwd, tot = [], []
    for _ in range(5):
        t0 = time.perf_counter()
        with torch.no_grad():
            forward()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        fwd.append(t1 - t0)
        tot.append(t2 - t1)
    ms_b = statistics.median(tot) * 1e3
    return {"batch": batch, "steps": steps, "resolution": "640x480",
            "forward_ms": statistics.median(fwd) * 1e3,
            "forward_backward_ms": ms_b, "pairs_per_s": batch / ms_b * 1e3,
            "grad_finite": bool(all(torch.isfinite(t.grad).all().item()
                                    for t in leaves + [x0]))}


def matcher_section(pair, top, dev, with_cpu):
    """The differentiable frame matcher at 640x480 (every valid source pixel,
    moved by the true motion, matched against the target frame with its
    3-channel features) and AutogradICP driven by it: per-launch durations
    of the matcher and projection kernels, the whole autograd forward +
    backward, and the oracle's single-thread matcher on the same inputs."""
    import numpy as np
    import torch
    from slamtb._cslamtb import FPCLMatcherOp as Op, ProjectOp
    from slamtb.camera import KCamera
    from slamtb.registration.autogradicp import AutogradICP
    from slamtb.registration.correspondence_map import CorrespondenceMap
    from slamtb.spatial.matching import FPCLMatcherOp, FramePointCloudMatcher

    k = torch.tensor([[top.fx, 0, top.cx], [0, top.fy, top.cy], [0, 0, 1]],
                     dtype=torch.float32)
    kd = k.to(dev)
    T = torch.from_numpy(pair["T_gt"]).to(dev, torch.float32)
    m = top.src_mask.view(-1).bool()
    sp = (top.src_points.view(-1, 3)[m].float() @ T[:3, :3].T + T[:3, 3]).contiguous()
    sn = (top.src_normals.view(-1, 3)[m].float() @ T[:3, :3].T).contiguous()
    tp, tn = top.tgt_points.float(), top.tgt_normals.float()
    tf = top.tgt_feats.float().contiguous()
    C, H, W = tf.shape
    n = sp.size(0)
    ones = torch.ones(1, dtype=torch.bool, device=dev)
    eye = torch.eye(4, device=dev)
    cm = CorrespondenceMap(0.1, math.radians(20))
    cmap = cm(sp, sn, ones, eye, tp, tn, top.tgt_mask, kd)
    op = torch.empty(n, 3, device=dev)
    on = torch.empty(n, 3, device=dev)
    of = torch.empty(C, n, device=dev)
    mask = torch.empty(n, dtype=torch.bool, device=dev)
    dl = torch.randn(C, n, device=dev)
    dx = torch.empty(n, 3, device=dev)
    fwd = lambda: Op.forward(tp, tn, tf, sp, kd, cmap, op, on, of, mask)
    bwd = lambda: Op.backward(tf, sp, mask, dl, kd, 0.005, dx)
    fwd()
    matched = float(mask.float().mean().item())
    t_map = launch_ms(lambda: cm(sp, sn, ones, eye, tp, tn, top.tgt_mask, kd), 50)
    t_fwd = launch_ms(fwd, 100)
    t_bwd = launch_ms(bwd, 100)
    dy = torch.randn(n, 2, device=dev)
    t_pf = launch_ms(lambda: ProjectOp.forward(sp, kd), 100)
    t_pb = launch_ms(lambda: ProjectOp.backward(dy, sp, kd), 100)
    # Algorithmic bytes: forward reads the point (12) and its map word (8);
    # a matched point gathers the target point and normal (24) and one texel
    # per channel (neighbouring taps are shared between neighbouring points),
    # and every point writes its rows (24 + 4C) and mask byte. Backward reads
    # the mask, and for matched points the point (12), incoming gradient (4C)
    # and one texel per channel; it writes 12.
    b_fwd = n * (12 + 8 + 1 + 24 + 4 * C) + n * matched * (24 + 4 * C)
    b_bwd = n * (1 + 12) + n * matched * (12 + 8 * C)
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9

    kc = KCamera(k)
    matcher = FramePointCloudMatcher(tp, tn, top.tgt_mask, tf, kc, 0.1,
                                     math.radians(20))
    src = sp.clone().requires_grad_(True)

    def autograd_step():
        src.grad = None
        pts, nrm, feats, msk = matcher.find_correspondences(src, sn)
        loss = ((pts - src[msk]) * nrm).sum(1).pow(2).mean() + feats.pow(2).mean()
        loss.backward()
    autograd_step()
    torch.cuda.synchronize()
    walls = []
    for _ in range(20):
        t0 = time.perf_counter()
        autograd_step()
        torch.cuda.synchronize()
        walls.append((time.perf_counter() - t0) * 1e3)

    # AutogradICP from the identity, counting cost evaluations.
    calls = [0]

    class Counting(FramePointCloudMatcher):
        def match_rows(self, *a):
            calls[0] += 1
            return super().match_rows(*a)
    sm = top.src_mask.bool()
    src_pts = top.src_points.float()[sm].view(-1, 3)
    src_nrm = top.src_normals.float()[sm].view(-1, 3)
    src_f = top.src_feats.float()[:, sm].contiguous()
    icp = AutogradICP(20, learning_rate=0.1, geom_weight=0.5, feat_weight=0.5,
                      distance_threshold=0.1, normals_angle_thresh=math.radians(20),
                      feat_residual_thresh=2.75, gradient_descent_iters=20)
    cnt = Counting(tp, tn, top.tgt_mask, tf, kc, 0.1, math.radians(20))
    AutogradICP(2, learning_rate=0.1, gradient_descent_iters=2)._estimate(
        src_pts, src_nrm, src_f, cnt)  # warm-up (first-call costs)
    calls[0] = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = icp._estimate(src_pts, src_nrm, src_f, cnt)
    el = time.perf_counter() - t0
    from slamtb import metrics
    out = {"points": n, "matched_frac": matched, "channels": C,
           "launch_ms": {"correspondence_map": t_map, "fpcl_forward_kernel": t_fwd,
                         "fpcl_backward_kernel": t_bwd,
                         "project_forward_kernel": t_pf,
                         "project_backward_kernel": t_pb},
           "fpcl_forward_gbs": gbs(b_fwd, t_fwd), "fpcl_backward_gbs": gbs(b_bwd, t_bwd),
           "fpcl_forward_bytes": b_fwd, "fpcl_backward_bytes": b_bwd,
           "autograd_match_forward_backward_ms": statistics.median(walls),
           "autogradicp": {"evaluations": calls[0], "seconds": el,
                           "ms_per_evaluation": el * 1e3 / max(1, calls[0]),
                           "pose_err_vs_truth": [float(e) for e in metrics.pose_error(
                               res.transform.double().numpy(), pair["T_gt"])]}}
    if with_cpu:
        sys.path.insert(0, str(ROOT))
        from oracle import oracle as orc
        c = lambda x: x.detach().cpu().numpy()
        a = [c(tp), c(tn), c(tf), c(sp), c(kd), c(cmap)]
        t0 = time.perf_counter()
        r = orc.fpcl_forward(*a)
        t1 = time.perf_counter()
        orc.fpcl_backward(a[2], a[3], r[3], c(dl), a[4], 0.005)
        t2 = time.perf_counter()
        out["cpu_oracle_1thread_ms"] = {"fpcl_forward": (t1 - t0) * 1e3,
                                        "fpcl_backward": (t2 - t1) * 1e3}
    return out


def ftb_section(dev):
    """The reference's own RGB-D frames (FTB copies under tests/golden/ftb):
    the Registration notebook's pair, sample1 0 -> 14 with RGB features
    (10/15/20 iterations, weights 1 : 1), and depth-only frame-to-frame
    tracking over frames 0..14; accuracy against the recorded camera poses
    and the time per registration (preprocessing on the GPU included,
    PNG decoding excluded)."""
    import torch
    from slamtb import metrics
    from slamtb.camera import RTCamera
    from slamtb.data import load_ftb
    from slamtb.frame import Frame, FramePointCloud
    from slamtb.processing import bilateral_depth_filter, to_color_feature
    from slamtb.registration import ICPOdometry, MultiscaleRegistration

    ds = load_ftb(ROOT / "tests" / "golden" / "ftb" / "sample1")
    frames = [ds[i] for i in range(len(ds))]
    depth = [torch.from_numpy(f.depth_image).to(dev) for f in frames]

    def fpcl(i):
        d = bilateral_depth_filter(depth[i])
        return FramePointCloud.from_frame(Frame(frames[i].info, d), device=dev)

    kw = dict(geom_weight=1.0, feat_weight=1.0, distance_threshold=0.5,
              normals_angle_thresh=math.pi / 4, feat_residual_thresh=2.75)
    ms = MultiscaleRegistration([(1.0, ICPOdometry(10, **kw)),
                                 (0.5, ICPOdometry(15, **kw)),
                                 (0.5, ICPOdometry(20, **kw))])
    sf = to_color_feature(frames[14].rgb_image).to(dev)
    tf = to_color_feature(frames[0].rgb_image).to(dev)
    pair = lambda: ms.estimate_frame(fpcl(14), fpcl(0), source_feats=sf,
                                     target_feats=tf)
    res = pair()
    walls = []
    for _ in range(10):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = pair()
        walls.append(time.perf_counter() - t0)
    gt = {0: frames[0].info.rt_cam, 1: frames[14].info.rt_cam}
    pred = {0: RTCamera(), 1: RTCamera(res.transform)}
    out = {"notebook_pair_0_14": {
        "rte_m": metrics.relative_translational_error(gt, pred).item(),
        "rre_rad": metrics.relative_rotational_error(gt, pred).item(),
        "notebook_rte_m": 0.011427460885717639,
        "notebook_rre_rad": 0.015468830242753029,
        "ms_per_registration": statistics.median(walls) * 1e3}}

    kw = dict(geom_weight=1.0, feat_weight=0.0, distance_threshold=0.1,
              normals_angle_thresh=math.pi / 4)
    ms = MultiscaleRegistration([(1.0, ICPOdometry(10, **kw)),
                                 (0.5, ICPOdometry(10, **kw)),
                                 (0.5, ICPOdometry(10, **kw))])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prev = fpcl(0)
    pose = torch.eye(4, dtype=torch.float64)
    gt, pred = {0: frames[0].info.rt_cam}, {0: RTCamera(pose)}
    for i in range(1, len(frames)):
        cur = fpcl(i)
        pose = pose @ ms.estimate_frame(cur, prev).transform
        gt[i], pred[i] = frames[i].info.rt_cam, RTCamera(pose)
        prev = cur
    el = time.perf_counter() - t0
    gt = metrics.set_start_at_identity(gt)
    rmse = lambda e: e.pow(2).mean().sqrt().item()
    out["tracking_0_14_depth"] = {
        "ate_rmse_m": rmse(metrics.absolute_translational_error(gt, pred)),
        "rte_rmse_m": rmse(metrics.relative_translational_error(gt, pred)),
        "rre_rmse_rad": rmse(metrics.relative_rotational_error(gt, pred)),
        "ms_per_frame": el * 1e3 / (len(frames) - 1)}

    # The tracking front end (SurfelSLAM.step without fusion): the default
    # GN preset, eager and as one captured graph per step.
    from slamtb.pipeline import SurfelSLAM
    for mode in ("eager", "graph"):
        slam = SurfelSLAM(device=dev, use_graph=mode == "graph")
        stats = [slam.step(f) for f in frames]
        stats += [slam.step(f) for f in frames[1:]]  # second pass: warm
        times = sorted(st.ms for st in stats[len(frames):])
        gtt = metrics.set_start_at_identity(
            {i: f.info.rt_cam for i, f in enumerate(frames)})
        pr = {i: st.rt_camera for i, st in enumerate(stats[:len(frames)])}
        out[f"surfel_slam_tracking_{mode}"] = {
            "ms_per_frame_median": times[len(times) // 2],
            "ate_rmse_m": rmse(metrics.absolute_translational_error(gtt, pr)),
            "accepted": sum(st.accepted for st in stats[1:len(frames)])}
    return out


def pcl_section(pair, dev, with_cpu):
    """Point-cloud registration at 640x480: both frames of the bench pair as
    surfel clouds, voxel-merged at 1 cm, then the exact 8-nearest-neighbour
    search within 0.1 m (uniform grid, and brute force beside it), the
    nearest-neighbour feature layer forward and
    backward, and AutogradICP.estimate_pcl; the oracle's neighbour search
    on a 2,000-query sample beside it."""
    import numpy as np
    import torch
    from slamtb import metrics
    from slamtb._cslamtb import NearestNeighborsOp, knn
    from slamtb.camera import KCamera
    from slamtb.frame import Frame, FrameInfo, FramePointCloud
    from slamtb.registration.autogradicp import AutogradICP
    from slamtb.surfel import SurfelCloud

    kc = KCamera(torch.from_numpy(pair["kcam"]))
    clouds = []
    for side in ("src", "tgt"):
        d = torch.from_numpy(pair[f"{side}_depth"]).to(dev)
        fp = FramePointCloud.from_frame(Frame(FrameInfo(kc), depth_image=d), device=dev)
        clouds.append(SurfelCloud.from_frame_pcl(
            fp, features=torch.from_numpy(pair[f"{side}_feats"]).to(dev)))
    src_full, tgt_full = clouds
    tgt_full.downsample(0.01)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        tgt_full.downsample(0.01)
    torch.cuda.synchronize()
    t_ds = (time.perf_counter() - t0) * 1e3 / 5  # host-synchronising: wall clock
    src, tgt = src_full.downsample(0.01), tgt_full.downsample(0.01)
    n, m = src.size, tgt.size
    from slamtb._cslamtb import KnnGrid
    KnnGrid(tgt.points, 0.1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    grid = KnnGrid(tgt.points, 0.1)
    torch.cuda.synchronize()
    t_build = (time.perf_counter() - t0) * 1e3  # host-synchronising: wall clock
    t_knn = launch_ms(lambda: grid.query(src.points, 8, 0.1), 20)
    t_brute = launch_ms(lambda: knn(tgt.points, src.points, 8, 0.1, method="brute"), 3)
    dist, index = grid.query(src.points, 8, 0.1)
    C = tgt.feature_size
    out = torch.empty(C, n, device=dev)
    eps = torch.empty(n, 6, 8, device=dev)
    dl = torch.randn(C, n, device=dev)
    dx = torch.empty(n, 3, device=dev)
    t_fwd = launch_ms(lambda: NearestNeighborsOp.forward(dist, index, tgt.features, out), 50)
    t_eps = launch_ms(lambda: NearestNeighborsOp.compute_epsilon_distances(
        tgt.points, src.points, index, eps), 50)
    t_bwd = launch_ms(lambda: NearestNeighborsOp.backward(eps, index, tgt.features,
                                                           dl, dx), 50)
    icp = AutogradICP(20, learning_rate=0.1, geom_weight=1.0, feat_weight=0.0,
                      distance_threshold=0.1, gradient_descent_iters=20)
    icp.estimate_pcl(src, tgt)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = icp.estimate_pcl(src, tgt)
    el = time.perf_counter() - t0
    out_d = {"surfels_full": [src_full.size, tgt_full.size],
             "surfels_1cm": [n, m],
             "launch_ms": {"voxel_downsample_1cm": t_ds, "knn_grid_build": t_build,
                           "knn_k8_0.1m_grid": t_knn, "knn_k8_0.1m_brute": t_brute,
                           "nn_features_forward": t_fwd,
                           "nn_epsilon_distances": t_eps,
                           "nn_features_backward": t_bwd},
             "knn_brute_pairs_per_s": n * m / (t_brute * 1e-3),
             "estimate_pcl_s": el,
             "estimate_pcl_pose_err_vs_truth": [float(e) for e in metrics.pose_error(
                 res.transform.double().numpy(), pair["T_gt"])]}
    if with_cpu:
        sys.path.insert(0, str(ROOT))
        from oracle import oracle as orc
        q = src.points[:2000].cpu().numpy()
        tg = tgt.points.cpu().numpy()
        t0 = time.perf_counter()
        orc.knn(tg, q, 8, 0.1)
        out_d["cpu_oracle_knn_pairs_per_s"] = len(q) * m / (time.perf_counter() - t0)
    return out_d


def main():
    args = parse()
    if args.copy_probe > 0:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not "
              "supported (independent frame pairs shard with no exchange "
              "between GPUs: one process per GPU, each with its own pairs)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from slamtb.registration._fused import runner_for
    if args.profile:
        args.no_oracle = args.no_cpu_baseline = args.no_bf16 = True
        args.no_diff = args.no_matcher = args.no_ftb = args.no_pcl = True
        args.pairs = 0

    dev = torch.device("cuda:0")
    log("preparing 640x480 pair on the GPU")
    pair, ms, levels = prepare(args.seed, dev)
    est, top = levels[-1]
    params = est.params(top.has_features)
    if args.tune:
        from slamtb import tuning
        params = tuning.tune_level(top, params)
        log(f"tuned grids: update {params.grid_update} corresp {params.grid_corresp}")
    run = runner_for(dev)
    for _, L in levels:
        run.reserve(L.n, L.height, L.width)

    # --- warm-up and the timed steps ---------------------------------------
    log(f"warm-up {args.warmup} steps")
    run.begin(top)
    run.iterate(top, params, args.warmup)
    torch.cuda.synchronize()
    K = args.steps
    graph = None if args.no_graph else capture(
        lambda: (run.begin(top), run.iterate(top, params, K)))
    log(f"timing {K} steps ({'graph' if graph else 'stream'})")

    def timed():
        if graph is not None:
            graph.replay()
        else:
            run.begin(top)
            run.iterate(top, params, K)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    timed()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    st = run.read_state()
    value = K / (t1 - t0)
    ms_per_step = (t1 - t0) * 1e3 / K
    steps_ok = (st.stopped == 0 and st.iterations == K)
    reruns = []
    for _ in range(5):
        torch.cuda.synchronize()
        a = time.perf_counter()
        timed()
        torch.cuda.synchronize()
        reruns.append((time.perf_counter() - a) * 1e3 / K)

    # --- per-launch durations and the roofline -------------------------------
    log("per-launch event timing")
    run.begin(top)
    run.iterate(top, params, 2)
    reps = 100
    corr_ms = launch_ms(lambda: run.correspond(top, params), reps)
    run.linearize(top, params)  # the map the Jacobian launches read
    jac_ms = launch_ms(lambda: run.jacobians(top, params), reps)
    run.begin(top)
    run.iterate(top, params, 2)
    st2 = run.read_state()
    n = top.n
    s, C = 4, top.channels
    valid = float(top.src_mask.float().mean().item())
    matched = max(st2.last_geom_count, st2.last_feat_count) / n
    # Algorithmic bytes per launch: all-pixel figures (12s+18, 9s+9+2Cs B/px)
    # and the honest ones: mask bytes for every pixel, source and gather
    # bytes for valid pixels, Jacobian gathers for matched pixels.
    bytes_a_all = (12 * s + 18) * n
    bytes_b_all = (9 * s + 9 + 2 * C * s) * n
    bytes_a = n * (1 + 8 + valid * (6 * s + 1 + 6 * s + 8))
    bytes_b = n * (1 + valid * (3 * s + 8) + matched * (6 * s + 2 * C * s))
    if corr_ms >= jac_ms:
        dom = dict(kernel="corresp_fused_kernel", ms=corr_ms, bytes=bytes_a,
                   bytes_all=bytes_a_all)
    else:
        dom = dict(kernel=JAC_KERNEL, ms=jac_ms, bytes=bytes_b,
                   bytes_all=bytes_b_all)
    achieved = dom["bytes"] / (dom["ms"] * 1e-3) / 1e9
    traffic = None
    if TRAFFIC_FILE.exists() and dom["kernel"] == JAC_KERNEL:
        traffic = json.loads(TRAFFIC_FILE.read_text()).get("hbm_bytes_per_launch")
    _, copy_gbs = copy_rate(256)

    out = {
        "metric": METRIC, "value": value, "unit": "GN iters/s", "n_gpus": 1,
        "steps": K, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": "icp-gn-640x480-geom+feat3-fp32",
                   "resolution": "640x480", "feature_channels": C,
                   "levels": "640x480/320x240/160x120 10/5/3 it",
                   "weights": "geom 10 : feat 1 (the 1:0.1 ratio in the reference preset form)",
                   "timing": "hipGraph" if graph else "stream",
                   "grid_update": params.grid_update,
                   "grid_corresp": params.grid_corresp,
                   "steps_all_accepted": steps_ok},
        "ms_per_step_median_of_5": statistics.median(reruns),
        "roofline": {"bound": "hbm", "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": dom["kernel"], "kernel_ms": dom["ms"],
                     "bytes_per_launch": dom["bytes"],
                     "bytes_per_launch_all_pixels": dom["bytes_all"],
                     "achieved_all_pixels": dom["bytes_all"] / (dom["ms"] * 1e-3) / 1e9,
                     "copy_ceiling_gbs": copy_gbs,
                     # The whole loop iteration (both launches and the gap
                     # between them) over the same algorithmic bytes.
                     "iteration_bytes": bytes_a + bytes_b,
                     "iteration_bytes_all_pixels": bytes_a_all + bytes_b_all,
                     "iteration_achieved": (bytes_a + bytes_b) / (ms_per_step * 1e-3) / 1e9,
                     "iteration_frac": (bytes_a + bytes_b) / (ms_per_step * 1e-3) / 1e9
                                       / HBM_PEAK_GBS,
                     "iteration_frac_all_pixels": (bytes_a_all + bytes_b_all)
                                                  / (ms_per_step * 1e-3) / 1e9
                                                  / HBM_PEAK_GBS},
        # A loop iteration is a correspondence launch that first applies the
        # previous update (corresp_solve_kernel) and the Jacobian launch; the
        # plain correspondence kernel runs only in a loop's first iteration.
        "launch_ms": {"corresp_fused_kernel": corr_ms,
                      JAC_KERNEL: jac_ms,
                      "corresp_solve_kernel_and_gaps": statistics.median(reruns) - jac_ms},
    }
    c1_levels = None
    if not args.profile:
        out.update(registration_section(args, ms, levels, run, top, pair))
        log("config 1: 320x240 geometric, 10 iterations")
        out["config1_320x240"], c1_levels = config1_section(args.seed, dev,
                                                            not args.no_oracle)
    if args.pairs > 0:
        log(f"batched workload: {args.pairs} pairs")
        out["batched"] = batched_section(args.seed, dev, args.pairs, not args.no_tune,
                                         not args.no_oracle)
    if not args.no_bf16:
        log("1280x720 BF16 / FP32 storage workload")
        out["bf16_1280x720"] = bf16_section(args.seed, dev, not args.no_oracle)
    if not args.no_ftb:
        log("reference RGB-D frames (FTB sample1)")
        out["ftb_sample1"] = ftb_section(dev)
    if not args.no_pcl:
        log("point-cloud registration at 640x480")
        out["pcl_640x480"] = pcl_section(pair, dev, not args.no_oracle)
    if not args.no_matcher:
        log("differentiable matcher and AutogradICP at 640x480")
        out["matcher_640x480"] = matcher_section(pair, top, dev, not args.no_oracle)
    if not args.no_diff:
        log("differentiable GN: 8 pairs at 640x480, 5 steps, forward + backward")
        out["diff_8x640x480"] = diff_section(args.seed, dev)
    if not args.no_cpu_baseline and c1_levels is not None:
        log("CPU baseline (~40 s)")
        cb = cpu_baseline(levels, c1_levels)
        best = max(("all_threads", "one_thread"),
                   key=lambda k: 1.0 / cb[k]["iter_640x480_s"])
        out["cpu_baseline"] = {
            "value": 1.0 / cb["all_threads"]["iter_640x480_s"], "unit": "GN iters/s",
            "cores": cb["threads"], "kind": "port",
            "sample": f"median of 20 runs after 3 warm-ups of one 640x480 GN "
                      f"iteration (geometric + 3-ch feature, FP32) of the "
                      f"oracle's OpenMP port of the reference CPU path on "
                      f"{cb['threads']} threads of {cpu_model()}; also the "
                      f"3-level 640x480 and the 320x240 geometric registrations "
                      f"and one thread",
            "cpu_model": cpu_model(), "faster_thread_setting": best,
            "all_threads": cb["all_threads"], "one_thread": cb["one_thread"]}
        if "config1_320x240" in out:
            c1 = out["config1_320x240"]
            c1["cpu_registration_s"] = {k: cb[k]["registration_320x240_geom_s"]
                                        for k in ("all_threads", "one_thread")}
            c1["speedup_vs_cpu_best"] = (min(c1["cpu_registration_s"].values()) /
                                         (c1["registration_ms_per_pair"] * 1e-3))
        out["registration_speedup_vs_cpu_best"] = (
            min(cb[k]["registration_640x480_3level_s"]
                for k in ("all_threads", "one_thread")) /
            (out["registration_ms_per_pair"] * 1e-3)) if "registration_ms_per_pair" in out else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
