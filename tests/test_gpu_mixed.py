"""GPU: the mixed-precision pipeline (FP16 TSDF with 8-bit weight codes, FP16
normal / intensity / gradient maps, FP32 vertex maps, FP32 pairwise sums,
FP64 solve).

Its results are stated against the FP64 oracle, not bit for bit:
  * per-frame poses within 0.5 mm and 0.05 deg of the FP64 oracle's, tracking
    success flags equal;
  * the fused TSDF within MIXED_TSDF_RMS (0.05) truncation units RMS of the
    FP64 oracle's over the voxels both observed (the FP32 path's own error
    is printed beside it);
and the batched and file paths of a mixed map behave as the plain ones do.
"""
import os

import numpy as np
import pytest
import torch

from libfusion_amd import capi
from libfusion_amd.synth import Scene
from oracle import oracle as O
from tests.helpers import pose_err

pytestmark = pytest.mark.gpu

MIXED = capi.PRECISION_MIXED
# TSDF RMS bar (truncation units).  The mixed error is dominated by the 8-bit
# weight codes: a stored weight is within 2^-5 (3.1%) of the exact running sum,
# which moves each fused mean by that fraction of the spread of its
# observations (up to ~1 truncation unit of depth noise at 3 m).  Measured:
# 0.020 at 320x240 / 10 mm and 0.037 at 640x480 / 5 mm.
MIXED_TSDF_RMS = 0.05


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def half_res(sc):
    return capi.intrinsics(sc.width // 2, sc.height // 2, sc.fx / 2, sc.fy / 2, (sc.cx + 0.5) / 2 - 0.5,
                           (sc.cy + 0.5) / 2 - 0.5)


def o_params(p):
    return O.Params(p.num_buckets, p.num_entries, p.num_blocks, p.max_triangles, p.zmin_raycast, p.zmax_raycast,
                    p.zmin_update, p.zmax_update, p.voxel_size, p.truncation)


def o_K(K):
    return O.Intr(K.width, K.height, K.fx, K.fy, K.cx, K.cy, K.invfx, K.invfy)


def frames_of(sc, K, n):
    out = []
    for f in range(n):
        d16, rgb = sc.render(sc.pose(f), K.width, K.height, K.fx, K.fy, K.cx, K.cy, noise=True, frame_id=f)
        out.append((d16, rgb, dev(d16), dev(rgb)))
    torch.cuda.synchronize()
    return out


def tsdf_rms(gpu_map, oracle_map):
    """RMS of the TSDF difference (truncation units) over the voxels both maps
    observed, matched by block coordinate; and the number of such voxels."""
    gpos, gt, _ = gpu_map
    opos, osdf, ow, _ = oracle_map
    index = {tuple(p): i for i, p in enumerate(opos.tolist())}
    pairs = [(i, index[tuple(p)]) for i, p in enumerate(gpos.tolist()) if tuple(p) in index]
    gi = np.array([a for a, _ in pairs])
    oi = np.array([b for _, b in pairs])
    g = gt[gi]
    o = osdf[oi].astype(np.float64) / 32767.0
    both = ~np.isnan(g) & (ow[oi] > 0)
    d = g[both].astype(np.float64) - o[both]
    return float(np.sqrt(np.mean(d * d))), int(both.sum())


def test_mixed_sequence_vs_fp64_oracle():
    """Half-resolution 12-frame orbit in mixed precision: every frame's pose
    within 0.5 mm / 0.05 deg of the FP64 oracle, and the final TSDF within
    MIXED_TSDF_RMS truncation units RMS of the oracle's."""
    sc = Scene(1, 1)
    K = half_res(sc)
    p32 = capi.map_params(2 ** 18, 1 << 17, 0.01, 0.03)
    p16 = capi.map_params(2 ** 18, 1 << 17, 0.01, 0.03, precision=MIXED)
    mixed = capi.System(K, p16, 3, (10, 5, 3), sc.depth_scale)
    plain = capi.System(K, p32, 3, (10, 5, 3), sc.depth_scale)
    osys = O.System(o_params(p32), o_K(K), 3, (10, 5, 3), sc.depth_scale, np.float64)
    worst = (0.0, 0.0)
    for f, (d16, rgb, t_d, t_rgb) in enumerate(frames_of(sc, K, 12)):
        mixed.process(t_d, t_rgb)
        plain.process(t_d, t_rgb)
        mixed.synchronize()
        plain.synchronize()
        osys.process(d16, rgb)
        dt, dr = pose_err(mixed.pose(), osys.pose())
        assert mixed.tracking()[0] == osys.success(), f
        assert dt <= 5e-4 and dr <= 0.05, (f, dt, dr)
        worst = (max(worst[0], dt), max(worst[1], dr))
    om = osys.map().voxels()
    rms16, n16 = tsdf_rms(capi.read_tsdf(*mixed.map_storage()), om)
    rms32, n32 = tsdf_rms(capi.read_tsdf(*plain.map_storage()), om)
    print("mixed vs FP64: worst pose %.2e m %.4f deg; TSDF RMS %.4f over %d voxels (FP32 path: %.5f over %d)"
          % (worst[0], worst[1], rms16, n16, rms32, n32))
    assert n16 > 100000 and n16 > 0.95 * n32
    assert rms16 <= MIXED_TSDF_RMS


def test_mixed_batch_equals_mixed_systems_alone():
    """Two mixed-precision systems in one lf_batch (batched FP16 pyramids,
    tracking and raycast launches; each system's reductions on half a lone
    system's grid) end within FP32 summation-order noise of the same systems
    run alone: poses within 0.02 mm / 0.002 deg, counters within 0.5%."""
    scenes = [Scene(1, 1), Scene(1, 7)]
    K = half_res(scenes[0])
    p = capi.map_params(2 ** 18, 1 << 17, 0.01, 0.03, precision=MIXED)
    frames = [frames_of(sc, K, 6) for sc in scenes]
    alone = []
    for sc, fs in zip(scenes, frames):
        s = capi.System(K, p, 3, (10, 5, 3), sc.depth_scale)
        for _, _, d, c in fs:
            s.process(d, c)
        s.synchronize()
        alone.append((s.pose(), s.tracking(), s.counters()))
        s.close()
    batch = capi.Batch([capi.system_config(K, p, 3, (10, 5, 3), sc.depth_scale) for sc in scenes])
    for f in range(6):
        batch.process([fs[f][2] for fs in frames], [fs[f][3] for fs in frames])
    batch.synchronize()
    for i, (pose, trk, cnt) in enumerate(alone):
        dt, dr = pose_err(batch.pose(i), pose)
        assert dt <= 2e-5 and dr <= 2e-3, (i, dt, dr)
        assert batch.tracking(i)[0] == trk[0], i
        c = batch.counters(i)
        assert c[capi.CNT_UPDATES] == cnt[capi.CNT_UPDATES], i
        for k in (capi.CNT_ALLOCATED, capi.CNT_RENDER_BLOCKS):
            assert abs(int(c[k]) - int(cnt[k])) <= 0.005 * cnt[k], (i, k, c[k], cnt[k])
    batch.close()


def test_batch_refuses_mixed_and_plain_systems_together():
    sc = Scene(1, 1)
    K = half_res(sc)
    cfgs = [capi.system_config(K, capi.map_params(2 ** 16, 1 << 12, 0.01, 0.03, precision=pr), 3, (10, 5, 3),
                               sc.depth_scale) for pr in (capi.PRECISION_FP32, MIXED)]
    with pytest.raises(capi.LfError):
        capi.Batch(cfgs)


def test_mixed_map_file_round_trip_and_mesh(tmp_path):
    """A mixed map written in the reference's file layout and read back into a
    mixed map keeps every block, every weight code and every TSDF value to
    within the int16 step of the file plus the FP16 rounding; its mesh has the
    triangles of the original."""
    sc = Scene(1, 1)
    K = half_res(sc)
    p = capi.map_params(2 ** 18, 1 << 17, 0.01, 0.03, precision=MIXED)
    s = capi.System(K, p, 3, (10, 5, 3), sc.depth_scale)
    for _, _, d, c in frames_of(sc, K, 4):
        s.process(d, c)
    s.synchronize()
    pos0, t0, w0 = capi.read_tsdf(*s.map_storage())
    path = os.path.join(str(tmp_path), "mixed.map")
    capi.check(capi.load().lf_system_write_map(s.h, path.encode()))
    m = capi.Map(p, K.width, K.height)
    m.read(path)
    torch.cuda.synchronize()
    pos1, t1, w1 = capi.read_tsdf(m.params, m.storage)
    assert np.array_equal(pos0, pos1)
    assert np.array_equal(w0, w1)
    seen = ~np.isnan(t0)
    assert np.array_equal(seen, ~np.isnan(t1))
    tol = 1.0 / 32767 + (np.abs(t0[seen]) + 1.0 / 32767) * 2.0 ** -11 + 2.0 ** -24
    assert np.all(np.abs(t1[seen] - t0[seen]) <= tol)
    cap = 2000000
    v0 = torch.empty((cap, 9), dtype=torch.float32, device="cuda")
    n_a = s.mesh(v0, cap=cap)
    _, _, n_b = m.mesh(mode=0, max_triangles=cap)
    assert 10000 < n_a < cap and abs(n_a - n_b) <= 0.01 * n_a
    s.close()


def test_mixed_full_resolution_sequence_vs_fp64_oracle():
    """The metric's configuration (640x480, 5 mm voxels, 15 mm truncation,
    3-level ICP + RGB {10, 5, 3}) for 20 frames in mixed precision: every
    frame's pose within 0.5 mm / 0.05 deg of the FP64 oracle, tracking
    success equal, TSDF RMS within MIXED_TSDF_RMS truncation units."""
    from libfusion_amd import configs
    sc, K, _ = configs.c1()
    p16 = capi.map_params(2 ** 20, 1 << 17, 0.005, 0.015, precision=MIXED)
    p32 = capi.map_params(2 ** 20, 1 << 17, 0.005, 0.015)
    mixed = capi.System(K, p16, configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale)
    plain = capi.System(K, p32, configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale)
    osys = O.System(o_params(p32), o_K(K), configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale, np.float64)
    worst = (0.0, 0.0)
    for f in range(20):
        d16, rgb = sc.frame(f)
        t_d, t_rgb = dev(d16), dev(rgb)
        mixed.process(t_d, t_rgb)
        plain.process(t_d, t_rgb)
        mixed.synchronize()
        plain.synchronize()
        osys.process(d16, rgb)
        dt, dr = pose_err(mixed.pose(), osys.pose())
        assert mixed.tracking()[0] == osys.success(), f
        assert dt <= 5e-4 and dr <= 0.05, (f, dt, dr)
        worst = (max(worst[0], dt), max(worst[1], dr))
    om = osys.map().voxels()
    rms, n = tsdf_rms(capi.read_tsdf(*mixed.map_storage()), om)
    rms32, n32 = tsdf_rms(capi.read_tsdf(*plain.map_storage()), om)
    print("mixed vs FP64 at 640x480 / 5 mm, 20 frames: worst pose %.2e m %.4f deg; TSDF RMS %.4f over %d voxels "
          "(FP32 path: %.5f over %d)" % (worst[0], worst[1], rms, n, rms32, n32))
    assert n > 1000000 and rms <= MIXED_TSDF_RMS
    mixed.close()
    plain.close()


def test_c3_batched_uneven_streams_vs_fp64_oracle():
    """The batched configuration: 16 independent mixed-precision streams
    (seeds 100-115) at 160x120, 320x240, 640x480 and 848x480, voxels of 4, 5,
    8 and 10 mm (truncation 3 voxels), 2^18-2^20 buckets, stepped together in
    one lf_batch for 6 frames.  Every stream's tracking verdict equals the FP64
    oracle's.  Every tracked stream with voxels of 5 mm or more is within
    0.5 mm / 0.05 deg of the oracle's pose.  At 4 mm voxels these small rooms
    are ill-conditioned: the FP32 method itself (the oracle's FP32
    instantiation, and this library's FP32 path) lands 0.4-0.55 mm from FP64,
    and the mixed weights' 3% quantization about doubles that, so those
    streams are held to 2 mm / 0.2 deg.  Errors are printed beside those of
    the same stream run alone."""
    frames = 6
    runs = []
    for i in range(16):
        sc = Scene(3, 100 + i)
        vs = (0.004, 0.005, 0.008, 0.010)[i // 4]
        K = capi.intrinsics(sc.width, sc.height, sc.fx, sc.fy, sc.cx, sc.cy)
        p = capi.map_params(2 ** (18 + i % 3), 1 << 17, vs, 3 * vs, precision=MIXED)
        host = [sc.frame(f) for f in range(frames)]
        runs.append((sc, K, p, host, [tuple(dev(a) for a in fr) for fr in host]))
    torch.cuda.synchronize()
    assert sorted({(r[1].width, r[1].height) for r in runs}) == [(160, 120), (320, 240), (640, 480), (848, 480)]
    alone = []
    for sc, K, p, _, fs in runs:
        s = capi.System(K, p, 3, (10, 5, 3), sc.depth_scale)
        for d, c in fs:
            s.process(d, c)
        s.synchronize()
        alone.append(s.pose())
        s.close()
    batch = capi.Batch([capi.system_config(K, p, 3, (10, 5, 3), sc.depth_scale) for sc, K, p, _, _ in runs])
    for f in range(frames):
        batch.process([r[4][f][0] for r in runs], [r[4][f][1] for r in runs])
    batch.synchronize()
    tracked = 0
    for i, (sc, K, p, host, _) in enumerate(runs):
        osys = O.System(o_params(p), o_K(K), 3, (10, 5, 3), sc.depth_scale, np.float64)
        for d16, rgb in host:
            osys.process(d16, rgb)
        ok = batch.tracking(i)[0]
        assert ok == osys.success(), i
        eb = pose_err(batch.pose(i), osys.pose())
        ea = pose_err(alone[i], osys.pose())
        print("C3 stream %d %dx%d %g mm: tracked %s; vs FP64 batch %.2e m %.4f deg, alone %.2e m %.4f deg"
              % (i, K.width, K.height, p.voxel_size * 1e3, ok, eb[0], eb[1], ea[0], ea[1]))
        c = batch.counters(i)
        assert c[capi.CNT_UPDATES] == frames and c[capi.CNT_OVERFLOW] == 0, i
        if ok:
            tracked += 1
            bar = (5e-4, 0.05) if p.voxel_size >= 0.005 else (2e-3, 0.2)
            assert eb[0] <= bar[0] and eb[1] <= bar[1], (i, eb)
    assert tracked >= 12
    batch.close()
