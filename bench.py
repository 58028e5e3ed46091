"""Benchmark: frames/s of the dense-reconstruction pipeline on one MI355X.

Workload (C1): 300 synthetic 640x480 RGB-D frames of a 4x4x3 m room with 8
seeded primitives along a 1.5 m orbit, 5 mm voxels, 15 mm truncation, 2^20
buckets, 3-level ICP+RGB tracking with {10, 5, 3} Gauss-Newton iterations.
A step is one frame of System::process_images: image pyramid -> tracking ->
block allocation + TSDF integration -> raycast -> vertex/normal re-pyramid.
All frames are resident in HBM before the timed region starts.

Prints one JSON line.  `roofline` describes the TSDF integration kernel
(achieved algorithmic bytes / its average HIP-event duration, measured inside
the timed region on the system's stream); `cpu_baseline` is the CPU oracle's
FP32 instantiation (the same method, OpenMP over the host cores) on a bounded
sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_frames(sc, n):
    d = np.zeros((n, sc.height, sc.width), np.uint16)
    c = np.zeros((n, sc.height, sc.width, 3), np.uint8)
    for i in range(n):
        d[i], c[i] = sc.frame(i)
    return d, c


def cpu_baseline(sc, params, K, frames_d, frames_c, n_frames):
    from oracle import oracle as O

    threads = O.max_threads()
    op = O.Params(params.num_buckets, params.num_entries, params.num_blocks, params.max_triangles,
                  params.zmin_raycast, params.zmax_raycast, params.zmin_update, params.zmax_update,
                  params.voxel_size, params.truncation)
    oK = O.Intr(K.width, K.height, K.fx, K.fy, K.cx, K.cy, K.invfx, K.invfy)
    from libfusion_amd import configs
    s = O.System(op, oK, configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale, np.float32)
    t0 = time.perf_counter()
    for i in range(n_frames):
        s.process(frames_d[i], frames_c[i])
    dt = time.perf_counter() - t0
    # the same oracle on one core, over the first frames only
    n1 = max(1, min(n_frames, 6))
    O.set_threads(1)
    s1 = O.System(op, oK, configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale, np.float32)
    t1 = time.perf_counter()
    for i in range(n1):
        s1.process(frames_d[i], frames_c[i])
    dt1 = time.perf_counter() - t1
    O.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": n_frames / dt, "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": "C1 frames 0-%d (640x480, 5 mm), FP32 oracle System, %d OpenMP threads" % (n_frames - 1, threads),
            "one_core": {"value": round(n1 / dt1, 3), "sample": "C1 frames 0-%d, 1 thread" % (n1 - 1)},
            "nproc": os.cpu_count(), "cpu_model": model}


def batch_throughput(n, warmup, steps, single_fps, precision=0, separate=True, repeats=3):
    """Aggregate frames/s of n independent C1 systems stepped together: system i
    is its own seeded scene and orbit (seed 100 + i) with its own map.  Two
    ways: lf_batch (pyramid passes and Gauss-Newton steps of all systems in one
    launch each, each map fused on its own stream, one batched z-range +
    raycast launch per stage for all maps), and (separate) n lf_system objects
    on their own streams with no shared launches."""
    import torch

    from libfusion_amd import capi, configs
    from libfusion_amd.synth import Scene

    frames, cfgs = [], []
    for i in range(n):
        sc = Scene(1, 100 + i)
        d, c = make_frames(sc, warmup + steps)
        frames.append((torch.from_numpy(d).cuda(), torch.from_numpy(c).cuda()))
        K = capi.intrinsics(sc.width, sc.height, sc.fx, sc.fy, sc.cx, sc.cy)
        _, _, params = configs.c1()
        params.precision = precision
        cfgs.append((K, params, sc.depth_scale))
    torch.cuda.synchronize()

    def timed(step, sync):
        for f in range(warmup):
            step(f)
        sync()
        t0 = time.perf_counter()
        for f in range(warmup, warmup + steps):
            step(f)
        sync()
        return time.perf_counter() - t0

    # the median of `repeats` runs, each a fresh batch over the same frames
    runs = []
    for _ in range(repeats):
        batch = capi.Batch([capi.system_config(K, p, configs.C1_LEVELS, configs.C1_ITERATIONS, ds)
                            for K, p, ds in cfgs])
        runs.append(timed(lambda f: batch.process([d[f] for d, _ in frames], [c[f] for _, c in frames]),
                          batch.synchronize))
        batch.close()
        del batch
        torch.cuda.empty_cache()
    dt_b = sorted(runs)[len(runs) // 2]
    v_b = n * steps / dt_b
    name = "plain FP32" if precision == 0 else "mixed-precision"
    out = {"systems": n, "value": round(v_b, 1), "unit": "frames/s (all systems)",
           "ms_per_step": round(1000.0 * dt_b / steps, 4), "vs_single_system": round(v_b / single_fps, 3),
           "steps": steps, "warmup": warmup, "runs": [round(n * steps / t, 1) for t in runs],
           "statistic": "median of %d runs" % repeats,
           "workload": "%d independent %s C1 systems (seeds 100-%d: own scene, orbit and map), one frame "
                       "of each per step" % (n, name, 99 + n),
           "launch": "lf_batch: one descriptor table per step; pyramid passes, Gauss-Newton steps, block allocation, "
                     "visible lists, integration, z-range and raycast of all systems in one launch each"}
    if not separate:
        return out
    systems = [capi.System(K, p, configs.C1_LEVELS, configs.C1_ITERATIONS, ds) for K, p, ds in cfgs]

    def step_systems(f):
        for s, (d, c) in zip(systems, frames):
            s.process(d[f], c[f])

    def sync_systems():
        for s in systems:
            s.synchronize()

    dt_s = timed(step_systems, sync_systems)
    for s in systems:
        s.close()
    v_s = n * steps / dt_s
    out["separate_systems"] = {"value": round(v_s, 1), "ms_per_step": round(1000.0 * dt_s / steps, 4),
                               "launch": "separate lf_system objects on their own streams, no shared launches"}
    return out


def timed_system(K, params, dev_d, dev_c, warmup, n_total, depth_scale):
    """frames/s of one System over frames warmup..n_total-1 (after warmup frames)."""
    import torch

    from libfusion_amd import capi, configs

    s = capi.System(K, params, configs.C1_LEVELS, configs.C1_ITERATIONS, depth_scale)
    for i in range(warmup):
        s.process(dev_d[i], dev_c[i])
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(warmup, n_total):
        s.process(dev_d[i], dev_c[i])
    s.synchronize()
    dt = time.perf_counter() - t0
    pose = s.pose()
    s.close()
    return (n_total - warmup) / dt, dt, pose


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=290)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cpu-frames", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stages", action="store_true", help="skip the instrumented stage-breakdown pass")
    ap.add_argument("--batch", type=int, default=16,
                    help="also time this many independent systems stepped together (0: skip)")
    ap.add_argument("--batch-steps", type=int, default=50)
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision passes")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device (a plain copy kernel of known size, "
                         "to scale rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the path runs one process per GPU with no "
              "cross-GPU exchange in this build; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    from libfusion_amd import capi, configs

    capi.load()
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    sc, K, params = configs.c1()
    n_total = args.warmup + args.steps
    if n_total > sc.num_frames:
        raise SystemExit("warmup + steps must be <= %d frames" % sc.num_frames)
    t = time.perf_counter()
    fd, fc = make_frames(sc, n_total)
    log("generated %d frames in %.1fs" % (n_total, time.perf_counter() - t))
    dev_d = torch.from_numpy(fd).cuda()
    dev_c = torch.from_numpy(fc).cuda()
    torch.cuda.synchronize()

    sysm = capi.System(K, params, configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale)
    for i in range(args.warmup):
        sysm.process(dev_d[i], dev_c[i])
    sysm.synchronize()
    c0 = sysm.counters()
    # only the integrate kernel is bracketed in the timed region (its average
    # launch duration is the roofline's denominator); each event record costs
    # stream time, so the stage breakdown comes from a separate pass below
    sysm.set_timing(capi.TIMING_INTEGRATE)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.warmup, n_total):
        sysm.process(dev_d[i], dev_c[i])
    sysm.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    c1 = sysm.counters()
    pose = sysm.pose()
    fps_evented = args.steps / elapsed
    log("timed %d frames in %.3fs: %.1f frames/s" % (args.steps, elapsed, fps_evented))

    n_ig, ms_ig = sysm.read_timing(0)
    integrate_avg_ms = ms_ig / max(n_ig, 1)
    # stage breakdown: a second system over the same frames with every stage bracketed
    stage_names = {0: "integrate_kernel", 1: "raycast_kernel", 2: "tracking", 4: "frame"}
    stages = {}
    if not args.no_stages:
        del sysm
        sysm = capi.System(K, params, configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale)
        for i in range(args.warmup):
            sysm.process(dev_d[i], dev_c[i])
        sysm.set_timing(capi.TIMING_ALL)
        for i in range(args.warmup, n_total):
            sysm.process(dev_d[i], dev_c[i])
        sysm.synchronize()
        for sid, name in stage_names.items():
            n, ms = sysm.read_timing(sid)
            stages[name] = {"launches": n, "avg_ms": ms / max(n, 1)}
    # two more runs of the same frames without the integrate events; the
    # reported frame rate is the median of the three
    run_fps = [fps_evented]
    for _ in range(2):
        del sysm
        torch.cuda.empty_cache()
        f_r, _, _ = timed_system(K, params, dev_d, dev_c, args.warmup, n_total, sc.depth_scale)
        run_fps.append(f_r)
        sysm = None
    fps = sorted(run_fps)[1]
    elapsed = args.steps / fps
    log("runs %s frames/s, median %.1f" % (", ".join("%.1f" % f for f in run_fps), fps))
    vis_sum = (int(c1[capi.CNT_VIS_SUM_LO]) | (int(c1[capi.CNT_VIS_SUM_HI]) << 32)) - \
              (int(c0[capi.CNT_VIS_SUM_LO]) | (int(c0[capi.CNT_VIS_SUM_HI]) << 32))
    updates = int(c1[capi.CNT_UPDATES]) - int(c0[capi.CNT_UPDATES])
    n_vis = vis_sum / max(updates, 1)
    P = K.width * K.height
    # TSDF integration kernel, algorithmic bytes per launch: every visible block's
    # 512 voxels read and written (sdf int16 + weight f32 + rgb 3 x u8 = 9 B),
    # its list entry (coordinate and pool index, 16 B), plus one depth (4 B),
    # normal x/z (8 B) and colour (3 B) gather per pixel.
    bytes_per_launch = n_vis * (2 * 512 * 9 + 16) + P * 15
    avg_s = integrate_avg_ms * 1e-3
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    traffic = None
    tf = os.path.join(ROOT, "profiles", "integrate_c1_traffic.json")
    if os.path.exists(tf):  # HBM bytes per launch measured with rocprofv3 --pmc (see that file)
        traffic = json.load(open(tf)).get("hbm_bytes_per_launch")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "k_integrate<true,true,true,false> (FP32, packed samples)", "algorithmic_bytes_per_launch": int(bytes_per_launch),
                "avg_launch_ms": round(integrate_avg_ms, 5), "visible_blocks_avg": round(n_vis, 1)}

    # raycast: compulsory bytes (vertex map out, sdf + weight of the distinct
    # blocks the march reads, their hash entries, z-range tiles), with the
    # distinct-block count measured by the oracle's instrumented march on C1
    stage_rooflines = {}
    rf = os.path.join(ROOT, "profiles", "raycast_c1_touched.json")
    if "raycast_kernel" in stages and os.path.exists(rf):
        rb = json.load(open(rf))["algorithmic_bytes_per_raycast"]
        rs = stages["raycast_kernel"]["avg_ms"] * 1e-3
        ra = rb / rs / 1e9
        stage_rooflines["raycast"] = {"bound": "latency (dependent sample chain)", "achieved": round(ra, 1),
                                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ra / HBM_PEAK_GBS, 4),
                                      "algorithmic_bytes_per_launch": rb,
                                      "avg_launch_ms": round(stages["raycast_kernel"]["avg_ms"], 5)}

    batched = None
    del sysm
    torch.cuda.empty_cache()
    if args.batch > 0:
        batched = batch_throughput(args.batch, args.warmup, args.batch_steps, fps)
        log("batched: %d systems, %.1f frames/s aggregate" % (args.batch, batched["value"]))
    # mixed precision (FP16 TSDF + 8-bit weight codes, FP16 maps): the same
    # frames through one system, then batched; frames/s against this run's
    # plain FP32 single system
    mixed = None
    if not args.no_mixed:
        pm = configs.c1()[2]
        pm.precision = capi.PRECISION_MIXED
        m_fps, m_dt, m_pose = timed_system(K, pm, dev_d, dev_c, args.warmup, n_total, sc.depth_scale)
        log("mixed precision: %.1f frames/s" % m_fps)
        mixed = {"value": round(m_fps, 2), "unit": "frames/s", "ms_per_step": round(1000.0 * m_dt / args.steps, 4),
                 "vs_plain_fp32": round(m_fps / fps, 3),
                 "final_pose_t": [round(float(x), 7) for x in m_pose[:3, 3]],
                 "storage": "FP16 TSDF + 8-bit weight codes + 3 x u8 colour (6 B/voxel); FP16 normal / intensity / "
                            "gradient maps (vertex maps FP32); FP32 pairwise normal-equation sums; FP64 solve and SE(3)"}
        if args.batch > 0:
            torch.cuda.empty_cache()
            mb = batch_throughput(args.batch, args.warmup, args.batch_steps, fps, capi.PRECISION_MIXED,
                                  separate=False)
            mb["vs_plain_fp32_single_system"] = round(mb["value"] / fps, 3)
            mixed["batched"] = mb
            log("mixed batched: %d systems, %.1f frames/s aggregate (%.2fx plain FP32)"
                % (args.batch, mb["value"], mb["value"] / fps))

    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline(sc, params, K, fd, fc, args.cpu_frames)
        log("cpu baseline %.2f frames/s in %.1fs" % (cpu["value"], time.perf_counter() - t))
    out = {
        "metric": "frames/s at 640x480, 5 mm voxels: mixed-prec+batched vs plain FP32",
        "value": round(fps, 2),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded analytic room, Perlin albedo, depth noise + 2% dropouts)",
        "config": {"workload": "C1 plain FP32 GPU path: 300-frame 640x480 orbit, room + 8 primitives, 5 mm voxels, "
                               "15 mm truncation, 2^20 buckets, 3-level ICP+RGB {10,5,3}; frames %d-%d timed"
                               % (args.warmup, n_total - 1),
                   "width": K.width, "height": K.height, "voxel_mm": 5, "buckets": params.num_buckets,
                   "pool_blocks": params.num_blocks, "precision": "plain FP32"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stage_rooflines": stage_rooflines,
        "stages_ms": {k: round(v["avg_ms"], 4) for k, v in stages.items()},
        "runs": [round(f, 2) for f in run_fps],
        "statistic": "median of 3 runs over the same frames (the first with HIP events around the integration kernel)",
        "batched": batched,
        "mixed_precision": mixed,
        "mixed_batched_vs_plain_fp32": (mixed["batched"]["vs_plain_fp32_single_system"]
                                        if mixed and mixed.get("batched") else None),
        "allocated_blocks": int(min(c1[capi.CNT_ALLOCATED], params.num_blocks)),
        "overflow": int(c1[capi.CNT_OVERFLOW]),
        "final_pose_t": [round(float(x), 7) for x in pose[:3, 3]],
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
