"""Benchmark: frames/s at 640x480, 5 mm voxels, mixed precision + batched
against the plain FP32 path, on one MI355X.

Workload: 16 independent synthetic RGB-D streams at 640x480 (seeds 100-115:
each its own 4x4x3 m room with 8 seeded primitives, Perlin albedo, depth noise
and 2% dropouts, and its own 1.5 m orbit), 5 mm voxels, 15 mm truncation,
2^20 buckets, 3-level ICP + RGB tracking with {10, 5, 3} Gauss-Newton
iterations.  A step is one frame of every stream through System::process_images
(upload -> image pyramid -> tracking -> block allocation + TSDF integration ->
raycast -> vertex/normal re-pyramid).  All frames are resident in HBM before
the timed region starts.

value       mixed precision (FP16 TSDF + 8-bit weights, FP16 normal / intensity /
            gradient maps, FP64 sums and solve) with the 16 streams stepped
            together by lf_batch: aggregate frames/s over exactly --steps timed
            steps after --warmup steps, median of 3 runs.
vs_plain_fp32  value / the plain FP32 GPU path (the reference's method, one
            System per stream) over the same streams and the same frames.

Also printed: the integration kernel's roofline (algorithmic bytes / its
average HIP-event duration on the batch stream, inside the first timed run),
the raycast's, a stage breakdown, the same step with every frame copied from
pinned host memory (PCIe-inclusive), the uneven 16-stream batched
configuration, and the CPU oracle (the same method in FP32, OpenMP over the
host cores) on a bounded sample of the same streams.
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
METRIC = "frames/s at 640x480, 5 mm voxels: mixed-prec+batched vs plain FP32"
SEED0 = 100  # streams are seeds SEED0 .. SEED0 + n - 1


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_frames(sc, n):
    d = np.zeros((n, sc.height, sc.width), np.uint16)
    c = np.zeros((n, sc.height, sc.width, 3), np.uint8)
    for i in range(n):
        d[i], c[i] = sc.frame(i)
    return d, c


def cpu_baseline(params, n_frames, seconds, seconds_one):
    """The FP32 oracle System (the same method, OpenMP over the host cores) on
    the first frames of stream SEED0: frames until n_frames or `seconds` of
    work, then the same on one core until `seconds_one`.  It generates its own
    frames, so the sample never depends on the GPU legs' frame counts."""
    from libfusion_amd import configs
    from libfusion_amd.synth import Scene
    from oracle import oracle as O

    sc = Scene(1, SEED0)
    K = _intrinsics(sc)
    op = O.Params(params.num_buckets, params.num_entries, params.num_blocks, params.max_triangles,
                  params.zmin_raycast, params.zmax_raycast, params.zmin_update, params.zmax_update,
                  params.voxel_size, params.truncation)
    oK = O.Intr(K.width, K.height, K.fx, K.fy, K.cx, K.cy, K.invfx, K.invfy)

    def run(limit_frames, limit_s):
        s = O.System(op, oK, configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale, np.float32)
        done, busy = 0, 0.0
        while done < limit_frames and busy < limit_s:
            d, c = sc.frame(done)  # generation is outside the timing
            t0 = time.perf_counter()
            s.process(d, c)
            busy += time.perf_counter() - t0
            done += 1
        return done, busy

    threads = O.max_threads()
    n, dt = run(n_frames, seconds)
    O.set_threads(1)
    n1, dt1 = run(n_frames, seconds_one)
    O.set_threads(threads)
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return {"value": round(n / dt, 3), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": "stream %d frames 0-%d (640x480, 5 mm, C1 tracking schedule), FP32 oracle System, "
                      "%d OpenMP threads, %.1f s of work" % (SEED0, n - 1, threads, dt),
            "one_core": {"value": round(n1 / dt1, 3), "sample": "stream %d frames 0-%d, 1 thread, %.1f s"
                                                                 % (SEED0, n1 - 1, dt1)},
            "nproc": os.cpu_count(), "cpu_model": model}


def _intrinsics(sc):
    from libfusion_amd import capi
    return capi.intrinsics(sc.width, sc.height, sc.fx, sc.fy, sc.cx, sc.cy)


def c1_params(precision):
    from libfusion_amd import configs
    p = configs.c1()[2]
    p.precision = precision
    return p


def median(xs):
    return sorted(xs)[len(xs) // 2]


def vis_sum(c):
    from libfusion_amd import capi
    return int(c[capi.CNT_VIS_SUM_LO]) | (int(c[capi.CNT_VIS_SUM_HI]) << 32)


class Streams:
    """The frames of n streams, device-resident (and pinned host copies on demand)."""

    def __init__(self, n, frames, config=1):
        import torch

        from libfusion_amd.synth import Scene
        t = time.perf_counter()
        self.scenes = [Scene(config, SEED0 + i) for i in range(n)]
        self.host = [make_frames(sc, frames) for sc in self.scenes]
        self.dev = [(torch.from_numpy(d).cuda(), torch.from_numpy(c).cuda()) for d, c in self.host]
        torch.cuda.synchronize()
        self.n, self.frames = n, frames
        self.pinned = None
        log("generated %d x %d frames in %.1fs" % (n, frames, time.perf_counter() - t))

    def pin(self):
        import torch
        if self.pinned is None:
            self.pinned = [(torch.from_numpy(d).pin_memory(), torch.from_numpy(c).pin_memory())
                           for d, c in self.host]
        return self.pinned


def batch_runs(streams, cfgs, warmup, steps, runs, timing_first=0, host=False, graphs=False, options=None):
    """Aggregate frames/s of the streams stepped together by lf_batch: `runs`
    fresh batches over the same frames, each W untimed steps then exactly K
    timed steps bracketed by synchronisation.  The first run records the
    stage events of timing_first (HIP events on the batch stream).  graphs:
    steady-state steps replayed from HIP graphs (lf_batch_set_graphs).  options:
    {BATCH_OPT_*: value} set on each batch."""
    import torch

    from libfusion_amd import capi
    n = streams.n
    src = streams.pin() if host else streams.dev
    out = []
    for r in range(runs):
        b = capi.Batch(cfgs)
        for o, v in (options or {}).items():
            b.set_option(o, v)
        if graphs:
            b.set_graphs(True)
        step = (lambda f: b.process_host([d[f] for d, _ in src], [c[f] for _, c in src])) if host else \
            (lambda f: b.process([d[f] for d, _ in src], [c[f] for _, c in src]))
        for f in range(warmup):
            step(f)
        b.synchronize()
        c0 = [b.counters(i) for i in range(n)]
        if r == 0 and timing_first:
            b.set_timing(timing_first)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for f in range(warmup, warmup + steps):
            step(f)
        b.synchronize()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        c1 = [b.counters(i) for i in range(n)]
        rec = {"s": dt, "fps": n * steps / dt,
               "vis_blocks_per_step": sum(vis_sum(y) - vis_sum(x) for x, y in zip(c0, c1)) / steps,
               "overflow": int(max(int(c[capi.CNT_OVERFLOW]) for c in c1)),
               "tracked": [b.tracking(i)[0] for i in range(n)]}
        if r == 0 and timing_first:
            rec["timing"] = {s: b.read_timing(s) for s in range(5) if timing_first & (1 << s)}
        out.append(rec)
        b.close()
        del b
        torch.cuda.empty_cache()
    return out


def single_runs(streams, params, warmup, steps, runs, timing=0, host=False):
    """The same streams and frames, each stream alone through one System (the
    plain path: one stream at a time): frames / summed wall time of the timed
    frames.  timing: System stage mask recorded in the first run."""
    import torch

    from libfusion_amd import capi, configs
    src = streams.pin() if host else streams.dev
    out = []
    for r in range(runs):
        total, ev, vsum = 0.0, {}, 0
        for i, sc in enumerate(streams.scenes):
            d, c = src[i]
            s = capi.System(_intrinsics(sc), params, configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale)
            proc = s.process_host if host else s.process
            for f in range(warmup):
                proc(d[f], c[f])
            s.synchronize()
            c0 = s.counters()
            if r == 0 and timing:
                s.set_timing(timing)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for f in range(warmup, warmup + steps):
                proc(d[f], c[f])
            s.synchronize()
            total += time.perf_counter() - t0
            vsum += vis_sum(s.counters()) - vis_sum(c0)
            if r == 0 and timing:
                for sid in range(5):
                    if timing & (1 << sid):
                        k, ms = s.read_timing(sid)
                        a = ev.setdefault(sid, [0, 0.0])
                        a[0] += k
                        a[1] += ms
            s.close()
        rec = {"s": total, "fps": streams.n * steps / total, "vis_blocks_per_frame": vsum / (streams.n * steps)}
        if ev:
            rec["timing"] = {k: tuple(v) for k, v in ev.items()}
        out.append(rec)
        torch.cuda.empty_cache()
    return out


def integrate_bytes(vis_blocks, pixels, voxel_bytes):
    """Algorithmic bytes of TSDF integration: every visible block's 512 voxels
    read and written, its list entry (coordinate + pool index, 16 B), and one
    depth (4 B) + normal x / z (8 B) + colour (3 B) gather per pixel."""
    return vis_blocks * (2 * 512 * voxel_bytes + 16) + pixels * 15


def roofline(bytes_per_launch, launches, total_ms, kernel, extra=None):
    avg_s = total_ms * 1e-3 / max(launches, 1)
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
           "algorithmic_bytes_per_launch": int(bytes_per_launch), "avg_launch_ms": round(total_ms / max(launches, 1), 5),
           "launches": launches}
    if extra:
        out.update(extra)
    return out


def profile_value(name, key):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get(key)
    return None


def tuned_leg(streams, cfgs, warmup, steps, runs):
    """lf_batch_autotune on the headline batch (one settling and two timed
    steps per candidate, two rounds, over the streams' frames forth and back,
    so consecutive steps stay neighbouring frames), then the headline
    measurement again: batches of this shape now start with the tuned launch
    options."""
    from libfusion_amd import capi
    b = capi.Batch(cfgs)
    b.autotune(2, 2)
    f = 0
    period = max(1, 2 * streams.frames - 2)
    while b.tuning():
        k = f % period
        k = k if k < streams.frames else period - k
        b.process([d[k] for d, _ in streams.dev], [c[k] for _, c in streams.dev])
        f += 1
    b.synchronize()
    names = {capi.BATCH_OPT_INTEGRATE_WORKGROUPS: "integrate_workgroups_per_map",
             capi.BATCH_OPT_TRACK_WORKGROUPS: "tracking_workgroups_per_system",
             capi.BATCH_OPT_ZRANGE_WORKGROUPS: "zrange_workgroups_per_map",
             capi.BATCH_OPT_UPLOAD_OVERLAP: "upload_overlap"}
    chosen = {name: b.option(o) for o, name in names.items()}
    b.close()
    rs = batch_runs(streams, cfgs, warmup, steps, runs)
    fps = median([r["fps"] for r in rs])
    return {"value": round(fps, 1), "runs": [round(r["fps"], 1) for r in rs], "tuning_steps": f,
            "options": chosen, "note": "0 = the built-in rule (2048 integration workgroups per map, 512 / n "
                                      "tracking workgroups per system, 64 z-range workgroups per map at 16 maps)"}


def c3_leg(warmup, steps, runs):
    """The uneven batched configuration: 16 mixed-precision streams at 160x120,
    320x240, 640x480 and 848x480 (seeds 100-115, resolution by seed), voxels of
    4 / 5 / 8 / 10 mm (truncation 3 voxels), 2^18-2^20 buckets, one lf_batch."""
    from libfusion_amd import capi
    streams = Streams(16, warmup + steps, config=3)
    cfgs, shapes = [], []
    for i, sc in enumerate(streams.scenes):
        vs = (0.004, 0.005, 0.008, 0.010)[i // 4]
        p = capi.map_params(2 ** (18 + i % 3), 1 << 18, vs, 3 * vs, precision=capi.PRECISION_MIXED)
        cfgs.append(capi.system_config(_intrinsics(sc), p, 3, (10, 5, 3), sc.depth_scale))
        shapes.append("%dx%d/%gmm" % (sc.width, sc.height, vs * 1e3))
    rs = batch_runs(streams, cfgs, warmup, steps, runs)
    fps = median([r["fps"] for r in rs])
    pixels = sum(sc.width * sc.height for sc in streams.scenes)
    return {"value": round(fps, 1), "unit": "frames/s (all streams)", "runs": [round(r["fps"], 1) for r in rs],
            "ms_per_step": round(1000.0 * 16 / fps, 4), "steps": steps, "warmup": warmup,
            "megapixels_per_s": round(fps * pixels / 16 / 1e6, 1), "streams": shapes,
            "workload": "16 mixed-precision streams of uneven resolution and voxel size in one lf_batch"}


def se3_exp(xi):
    """4x4 exponential of a twist [translation part; rotation vector]."""
    v, w = np.asarray(xi[:3], np.float64), np.asarray(xi[3:], np.float64)
    th = float(np.linalg.norm(w))
    Wx = np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])
    if th < 1e-12:
        A, B, C = 1.0, 0.5, 1.0 / 6.0
    else:
        A, B, C = np.sin(th) / th, (1.0 - np.cos(th)) / th ** 2, (th - np.sin(th)) / th ** 3
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + A * Wx + B * Wx @ Wx
    T[:3, 3] = (np.eye(3) + B * Wx + C * Wx @ Wx) @ v
    return T


def drifted_walk(sc, n, seed=7, metres=4e-4, degrees=0.02):
    """The walk's poses as a tracker with drift reports them: the true pose
    composed with a seeded random walk of the given per-frame steps."""
    rng = np.random.default_rng(seed)
    E = np.eye(4)
    out = []
    for f in range(n):
        out.append(sc.pose(f) @ E)
        w = rng.normal(size=3)
        w *= np.deg2rad(degrees) / np.linalg.norm(w)
        t = rng.normal(size=3)
        t *= metres / np.linalg.norm(t)
        E = E @ se3_exp(np.concatenate([t, w]))
    return out


def c4_leg(threads=16):
    """The large-volume workload: the 8x8x3 m multi-room (seed 4, 40
    primitives), its 1000-frame 640x480 random walk fused in mixed precision
    (4 mm voxels, 12 mm truncation, 2^22 buckets) at drifting poses (the
    reference method's tracking loses this walk: its 1.5 cm steps exceed the
    1 cm ICP gate), a keyframe every 50 frames; then the 20 keyframes re-fused
    at their true poses in one batched call, a 1280x960 raycast (visibility,
    z-range tiles and ray march) and full marching-cubes extraction.  Each
    stage is timed on the GPU alone (frames rendered beforehand on the host)."""
    import concurrent.futures
    import torch

    from libfusion_amd import capi
    from libfusion_amd.synth import Scene
    sc = Scene(4, 4)
    K = _intrinsics(sc)
    W, H = 2 * K.width, 2 * K.height
    K2 = capi.intrinsics(W, H, 2 * K.fx, 2 * K.fy, 2 * K.cx + 0.5, 2 * K.cy + 0.5)
    p = capi.map_params(2 ** 22, 1 << 19, 0.004, 0.012, precision=capi.PRECISION_MIXED)
    n = sc.num_frames
    poses = drifted_walk(sc, n)
    m = capi.Map(p, W, H)
    img = capi.Image(K, 1)
    keys = list(range(0, n, 50))
    kimg = {}
    fuse_s = 0.0
    pool = concurrent.futures.ThreadPoolExecutor(threads)
    chunk = 50
    for c0 in range(0, n, chunk):
        frames = list(pool.map(sc.frame, range(c0, min(n, c0 + chunk))))
        dev = [(torch.from_numpy(d).cuda(), torch.from_numpy(c).cuda()) for d, c in frames]
        pose_dev = [torch.from_numpy(capi.pose24(poses[f])).cuda() for f in range(c0, c0 + len(frames))]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i, (d, c) in enumerate(dev):
            img.upload(d, c, sc.depth_scale)
            m.update_image(img, pose_dev[i])
        torch.cuda.synchronize()
        fuse_s += time.perf_counter() - t0
        for i, (d, c) in enumerate(dev):
            if c0 + i in keys:
                ki = capi.Image(K, 1)
                ki.upload(d, c, sc.depth_scale)
                kimg[c0 + i] = ki
    pool.shutdown()
    torch.cuda.synchronize()
    blocks = int(m.counters()[capi.CNT_ALLOCATED])
    old = [torch.from_numpy(capi.pose24(poses[k])).cuda() for k in keys]
    new = [torch.from_numpy(capi.pose24(sc.pose(k))).cuda() for k in keys]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    m.refuse_keyframes([kimg[k] for k in keys], old, new)
    torch.cuda.synchronize()
    refuse_ms = 1e3 * (time.perf_counter() - t0)
    vm = torch.empty((H, W, 4), device="cuda")
    rgb = torch.empty((H, W, 3), dtype=torch.uint8, device="cuda")
    pose = torch.from_numpy(capi.pose24(sc.pose(n - 1))).cuda()
    ray = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        m.raycast_check_visibility(vm, pose, K2, rgb_out=rgb)
        torch.cuda.synchronize()
        ray.append(1e3 * (time.perf_counter() - t0))
    hits = int((~torch.isnan(vm[..., 0])).sum())
    cap = 48_000_000  # the walk's full mesh is about 34M triangles
    vtx = torch.empty((cap, 3, 3), device="cuda")
    nt = ctypes_uint()
    lib = capi.load()
    mesh = []
    for _ in range(2):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        capi.check(lib.lf_create_mesh(ctypes_ref(m.params), ctypes_ref(m.storage), capi.ptr(vtx), capi.ptr(None),
                                      capi.ptr(None), cap, ctypes_ref(nt), capi.stream_ptr()))
        torch.cuda.synchronize()
        mesh.append(1e3 * (time.perf_counter() - t0))
    for ki in kimg.values():
        ki.close()
    img.close()
    out = {"workload": "multi-room 8x8x3 m (seed 4), 1000-frame 640x480 random walk, mixed precision, 4 mm voxels, "
                       "12 mm truncation, 2^22 buckets; 20 keyframes re-fused; 1280x960 raycast; full mesh",
           "frames": n, "fuse_ms_per_frame": round(1e3 * fuse_s / n, 4), "blocks": blocks,
           "refuse_20_keyframes_ms": round(refuse_ms, 3), "raycast_1280x960_ms": round(min(ray), 3),
           "raycast_hits": hits, "mesh_ms": round(min(mesh), 3), "triangles": int(nt.value),
           "poses": "true walk composed with a seeded drift of 0.4 mm / 0.02 deg per frame (keyframes re-fused "
                    "to the true poses)"}
    del m, vm, rgb, vtx
    torch.cuda.empty_cache()
    return out


def ctypes_uint():
    import ctypes
    return ctypes.c_uint(0)


def ctypes_ref(x):
    import ctypes
    return ctypes.byref(x)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50, help="timed steps (one frame of every stream each)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--streams", type=int, default=16)
    ap.add_argument("--runs", type=int, default=3, help="runs of the timed steps (median reported)")
    ap.add_argument("--cpu-frames", type=int, default=60, help="at most this many frames in the CPU sample")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="at most this much CPU work, all cores")
    ap.add_argument("--cpu-seconds-one", type=float, default=6.0, help="at most this much CPU work, one core")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the headline and its plain FP32 comparison (no breakdown, PCIe or C3 legs)")
    ap.add_argument("--upload-overlap", type=int, default=1, choices=(0, 1),
                    help="0: the headline batch uploads each step's frames on its own stream ahead of the "
                         "tracking (no kernel runs beside the integration; for kernel traces)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 256 MiB device-to-device (a plain copy kernel of known size, "
                         "to scale rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build's path runs on one GPU (independent streams "
              "would shard one process per GPU with no exchange); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0 or args.streams < 1:
        raise SystemExit("need --steps >= 1, --warmup >= 0, --streams >= 1")

    import torch

    from libfusion_amd import capi

    capi.load()
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    W, K, n = args.warmup, args.steps, args.streams
    if W + K > 300:
        raise SystemExit("warmup + steps must be <= 300 frames (the orbit's length)")
    streams = Streams(n, W + K)
    P = streams.scenes[0].width * streams.scenes[0].height
    mixed, plain = capi.PRECISION_MIXED, capi.PRECISION_FP32

    def cfgs(precision):
        from libfusion_amd import configs
        return [capi.system_config(_intrinsics(sc), c1_params(precision), configs.C1_LEVELS, configs.C1_ITERATIONS,
                                   sc.depth_scale) for sc in streams.scenes]

    # headline: mixed precision, batched
    kmask = (1 << capi.BATCH_INTEGRATE) | (1 << capi.BATCH_RAYCAST)
    head = batch_runs(streams, cfgs(mixed), W, K, args.runs, timing_first=kmask,
                      options={capi.BATCH_OPT_UPLOAD_OVERLAP: args.upload_overlap})
    fps = median([r["fps"] for r in head])
    log("mixed batched: runs %s frames/s, median %.1f" % (", ".join("%.1f" % r["fps"] for r in head), fps))
    # the comparison: plain FP32, one System per stream, same frames
    base = single_runs(streams, c1_params(plain), W, K, args.runs, timing=capi.TIMING_INTEGRATE)
    fps_plain = median([r["fps"] for r in base])
    ratio = fps / fps_plain
    log("plain FP32 single systems: runs %s frames/s, median %.1f; ratio %.3f"
        % (", ".join("%.1f" % r["fps"] for r in base), fps_plain, ratio))

    # roofline of the batched integration kernel (HIP events in the first timed run)
    t = head[0]["timing"]
    ig_n, ig_ms = t[capi.BATCH_INTEGRATE]
    ig_bytes = integrate_bytes(head[0]["vis_blocks_per_step"], n * P, 6)
    rl = roofline(ig_bytes, ig_n, ig_ms, "k_integrate_batch<true> (mixed: FP16 sdf + a 32-bit word of 8-bit weight "
                  "code and colour per voxel, %d maps per launch)" % n,
                  {"visible_blocks_per_launch": round(head[0]["vis_blocks_per_step"], 1),
                   "basis": "per visible block 2*512*6 + 16 B (mixed voxel: FP16 sdf + 32-bit weight code / colour word), "
                           "per pixel 15 B"})
    rl["traffic"] = profile_value("integrate_batch_mixed_traffic.json", "hbm_bytes_per_launch")
    stage_rooflines = {}
    rc_n, rc_ms = t[capi.BATCH_RAYCAST]
    touched = profile_value("raycast_batch_touched.json", "blocks_touched_mean_per_map")
    if touched:
        # compulsory: the vertex map out (16 B) and the hit voxel's colour
        # word read (4 B: raycast_with_colour checks it) per pixel, the FP16
        # sdf plane of the distinct blocks the march reads and their 20 B
        # entries, the z-range tiles
        rb = n * (P * (16 + 4) + touched * (512 * 2 + 20) + 2 * (P // 64) * 4)
        stage_rooflines["raycast"] = roofline(rb, rc_n, rc_ms, "k_raycast_batch<true, true> (%d maps per launch)" % n,
                                              {"bound": "VALU issue (dependent sample chain)",
                                               "blocks_touched_per_map": touched})
        stage_rooflines["raycast"]["traffic"] = profile_value("raycast_batch_mixed_traffic.json",
                                                              "hbm_bytes_per_launch")
    else:
        stage_rooflines["raycast"] = {"avg_launch_ms": round(rc_ms / max(rc_n, 1), 5), "launches": rc_n}
    pt = base[0]["timing"]
    plain_rl = roofline(integrate_bytes(base[0]["vis_blocks_per_frame"], P, 9), pt[0][0], pt[0][1],
                        "k_integrate<true, true, true, false> (FP32, one map)")

    extras = {}
    if not args.no_extras:
        # stage breakdown of the batched step (every stage bracketed: slower)
        st = batch_runs(streams, cfgs(mixed), W, K, 1, timing_first=capi.BATCH_TIMING_ALL)[0]["timing"]
        names = {capi.BATCH_INTEGRATE: "integrate_kernel", capi.BATCH_RAYCAST: "raycast_kernel",
                 capi.BATCH_TRACKING: "tracking", capi.BATCH_FUSION: "fusion", capi.BATCH_STEP: "batch_stream_step"}
        extras["stages_ms"] = {names[s]: round(ms / max(k, 1), 4) for s, (k, ms) in st.items()}
        pb = batch_runs(streams, cfgs(plain), W, K, 1)[0]
        ms = single_runs(streams, c1_params(mixed), W, K, 1)[0]
        gr = batch_runs(streams, cfgs(mixed), W, K, args.runs, graphs=True)
        fps_graph = median([r["fps"] for r in gr])
        extras["breakdown"] = {
            "plain_fp32_batched": {"value": round(pb["fps"], 1), "vs_plain_fp32": round(pb["fps"] / fps_plain, 3)},
            "mixed_single_systems": {"value": round(ms["fps"], 1), "vs_plain_fp32": round(ms["fps"] / fps_plain, 3)},
            "mixed_batched_graph_replay": {"value": round(fps_graph, 1), "runs": [round(r["fps"], 1) for r in gr],
                                           "vs_plain_fp32": round(fps_graph / fps_plain, 3),
                                           "note": "the headline's steps replayed from HIP graphs "
                                                   "(lf_batch_set_graphs); the headline runs them eagerly"}}
        log("breakdown: plain batched %.1f, mixed single %.1f, mixed batched with graphs %.1f"
            % (pb["fps"], ms["fps"], fps_graph))
        # the same headline step with every frame copied from pinned host memory
        ph = batch_runs(streams, cfgs(mixed), W, K, 1, host=True)[0]
        extras["pcie_inclusive"] = {
            "value": round(ph["fps"], 1), "unit": "frames/s (all streams)",
            "h2d_bytes_per_step": int(n * P * 5),
            "note": "lf_batch_process_host from pinned host buffers: each step's 16 depth + RGB frames are copied "
                    "on the upload stream, overlapping the previous step's fusion and raycast"}
        log("pcie-inclusive mixed batched: %.1f" % ph["fps"])
        # last on this shape: batches of it start with the tuned options from here on
        extras["autotuned"] = tuned_leg(streams, cfgs(mixed), W, K, args.runs)
        extras["autotuned"]["vs_plain_fp32"] = round(extras["autotuned"]["value"] / fps_plain, 3)
        log("autotuned mixed batched: %s" % extras["autotuned"])
        del streams
        torch.cuda.empty_cache()
        extras["c3_uneven_streams"] = c3_leg(W, K, args.runs)
        log("C3 uneven batched: %.1f frames/s" % extras["c3_uneven_streams"]["value"])
        t0 = time.perf_counter()
        extras["c4_large_volume"] = c4_leg()
        log("C4 large volume: %s (%.1fs)" % (extras["c4_large_volume"], time.perf_counter() - t0))

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(c1_params(plain), args.cpu_frames, args.cpu_seconds, args.cpu_seconds_one)
        log("cpu baseline %.2f frames/s (%d threads), %.2f on one core, in %.1fs"
            % (cpu["value"], cpu["cores"], cpu["one_core"]["value"], time.perf_counter() - t0))

    out = {
        "metric": METRIC,
        "value": round(fps, 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(1000.0 * n / fps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "vs_baseline_note": "no published number exists for this metric; its comparison, mixed+batched over the "
                            "plain FP32 GPU path on the same streams and frames, is vs_plain_fp32",
        "vs_plain_fp32": round(ratio, 3),
        "dtype": "mixed: fp16 storage (TSDF, normal / intensity / gradient maps), u8 weights, fp32 arithmetic, "
                 "fp64 normal-equation sums and solve",
        "data": "synthetic (seeded analytic rooms with 8 primitives, Perlin albedo, depth noise + 2% dropouts)",
        "config": {"workload": "%d independent 640x480 RGB-D streams (seeds %d-%d, own orbit and map each), 5 mm "
                               "voxels, 15 mm truncation, 2^20 buckets, 3-level ICP+RGB {10,5,3}; one step = one "
                               "frame of every stream; frames %d-%d timed" % (n, SEED0, SEED0 + n - 1, W, W + K - 1),
                   "streams": n, "width": 640, "height": 480, "voxel_mm": 5, "buckets": 2 ** 20,
                   "pool_blocks": c1_params(mixed).num_blocks, "precision": "mixed", "launch": "lf_batch", "upload_overlap": args.upload_overlap},
        "roofline": rl,
        "stage_rooflines": stage_rooflines,
        "cpu_baseline": cpu,
        "runs": [round(r["fps"], 1) for r in head],
        "statistic": "median of %d runs over the same frames (the first with HIP events around the integration "
                     "and ray-marching kernels)" % args.runs,
        "plain_fp32": {"value": round(fps_plain, 1), "unit": "frames/s", "runs": [round(r["fps"], 1) for r in base],
                       "ms_per_frame": round(1000.0 / fps_plain, 4), "roofline": plain_rl,
                       "workload": "the same %d streams and frames, each through one plain FP32 System (int16 TSDF, "
                                   "float32 weights and maps, correctly rounded arithmetic)" % n},
        "tracked_streams": int(sum(head[0]["tracked"])),
        "overflow": head[0]["overflow"],
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
