"""CPU tests: the oracle functions that the GPU is compared with bit for bit,
each pinned by something other than itself: an independently vectorised FP64
numpy restatement, a brute-force evaluation, a library routine, or a closed
form chosen so that a dropped term, a flipped sign or a transposed operand
fails.

  rgb_step (m_estimator.cu:313-391)          numpy FP64 rows; linear-ramp translation (b = A xi)
  block visibility (depth_fusion.cu:15-47)   brute-force corner projection
  z-range tiles (ray_tracing.cu:63-126,187-210)  brute-force per-tile min / max
  render shading (cuda_imgproc.cu:17-48)     numpy FP64 Phong with a scipy rotation for the reflection
  warp (cuda_imgproc.cu:155-185)             scipy.ndimage.map_coordinates (order 1)
"""
import numpy as np
import scipy.ndimage
from scipy.spatial.transform import Rotation

from libfusion_amd.synth import Scene
from oracle import oracle as O


def params(buckets=2 ** 16, blocks=1 << 15, vs=0.01, trunc=0.03, zmin=0.3, zmax=5.0):
    return O.Params(buckets, buckets + buckets // 4, blocks, 2_000_000, zmin, zmax, zmin, zmax, vs, trunc)


def small_pose(rng, t=0.2, deg=10.0):
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec(np.deg2rad(deg) * rng.normal(size=3) / np.sqrt(3)).as_matrix()
    T[:3, 3] = rng.normal(size=3) * t
    return T


# ------------------------------------------------------------------ rgb_step

def bilinear(img, u, v):
    """interp2 of m_estimator.cu:339-344 as scipy's order-1 spline (exact for
    in-range coordinates; the caller keeps u, v >= 2 and < size - 2)."""
    return scipy.ndimage.map_coordinates(img.astype(np.float64), [v, u], order=1, mode="nearest")


def numpy_rgb_rows(cI, lI, lv, dx, dy, stddev, P, K):
    """FP64 rows of RgbReduction2::compute_jacobian (m_estimator.cu:313-391),
    vectorised over pixels: (accepted rows [n, 7])."""
    H, W = lI.shape
    v = lv.reshape(-1, 4).astype(np.float64)
    il = lI.reshape(-1).astype(np.float64)
    ok = ~(v[:, 3] < 0) & ~np.isnan(v[:, 0]) & np.isfinite(il)
    p = v[:, :3] @ P[:, :3].T + P[:, 3]
    with np.errstate(invalid="ignore", divide="ignore"):
        u0 = p[:, 0] / p[:, 2] * K.fx + K.cx
        v0 = p[:, 1] / p[:, 2] * K.fy + K.cy
        ok &= (u0 >= 2) & (u0 < W - 2) & (v0 >= 2) & (v0 < H - 2)
    uu, vv = np.where(ok, u0, 2.0), np.where(ok, v0, 2.0)
    ic, gx, gy = bilinear(cI, uu, vv), bilinear(dx, uu, vv), bilinear(dy, uu, vv)
    ok &= ((gx > 2) | (gy > 2)) & np.isfinite(ic) & np.isfinite(gx) & np.isfinite(gy)
    p, ic, il, gx, gy = p[ok], ic[ok], il[ok], gx[ok], gy[ok]
    z = p[:, 2]
    left = np.stack([gx * K.fx / z, gy * K.fy / z, np.zeros_like(z)], 1)
    left[:, 2] = -(left[:, 0] * p[:, 0] + left[:, 1] * p[:, 1]) / z
    r = ic - il
    if stddev > 1e-4:
        r = r / stddev
    th = 1.345 * stddev
    w = np.ones_like(r)
    if stddev > 1e-5:
        big = np.abs(r) > th
        w[big] = np.sqrt(th / np.abs(r[big]))
    return np.concatenate([w[:, None] * left, w[:, None] * np.cross(p, left), (w * -r)[:, None]], 1)


def frame_level(sc, frame, K):
    d16, rgb = sc.render(sc.pose(frame), K.width, K.height, K.fx, K.fy, K.cx, K.cy, noise=True, frame_id=frame)
    vm = O.back_project(d16.astype(np.float64) * sc.depth_scale, K, np.float64)
    I = O.grey(rgb, np.float64)
    dx, dy = O.derivative(I, np.float64)
    return vm, I, dx, dy


def test_rgb_step_equals_numpy_fp64_rows():
    """The FP64 oracle's photometric sums equal J^T J / J^T r of the numpy rows
    on a real frame pair, with and without the Huber weighting."""
    sc = Scene(1, 1)
    K = O.make_intr(160, 120, sc.fx / 4, sc.fy / 4, (sc.cx + 0.5) / 4 - 0.5, (sc.cy + 0.5) / 4 - 0.5)
    lv, lI, _, _ = frame_level(sc, 0, K)
    _, cI, cdx, cdy = frame_level(sc, 1, K)
    P = (np.linalg.inv(sc.pose(1)) @ sc.pose(0))[:3]
    for stddev in (0.0, 7.5):
        s = O.rgb_step(cI, lI, lv, cdx, cdy, stddev, P.ravel(), K, np.float64)
        rows = numpy_rgb_rows(cI, lI, lv, cdx, cdy, stddev, P, K)
        assert len(rows) > 1000 and s[28] == len(rows), (stddev, s[28], len(rows))
        M = rows.T @ rows
        A, b, res = O.unpack29(s)
        assert np.allclose(A, M[:6, :6], rtol=1e-9, atol=1e-9 * np.abs(M).max())
        assert np.allclose(b, M[:6, 6], rtol=1e-9, atol=1e-9 * np.abs(M).max())
        assert np.isclose(res[0], M[6, 6], rtol=1e-9)
        if stddev:  # the Huber weights took effect on some rows
            plain = numpy_rgb_rows(cI, lI, lv, cdx, cdy, 1e-6, P, K)
            assert not np.allclose(np.abs(rows[:, 6]), np.abs(plain[:, 6]) / stddev)


def test_rgb_step_linear_ramp_translation_closed_form():
    """A fronto-parallel plane at depth D textured with the intensity ramp
    I = a x + b y + c, seen by a current camera translated by t_x: the current
    image is the ramp shifted by s = f_x t_x / D pixels, bilinear sampling and
    central differences of a ramp are exact, and the photometric rows are
    exactly linear in the motion: every accepted row satisfies row[6] =
    row[0:6] . xi with xi = (-t_x, 0, 0, 0, 0, 0).  So b = A xi and
    sum r^2 = xi^T A xi (a flipped residual sign, a swapped f_x / f_y or a
    wrong Jacobian term fails), and every interior pixel is accepted."""
    W, H, D, a, bgrad, c, tx = 96, 72, 2.0, 3.0, 0.75, 20.0, 0.004
    K = O.make_intr(W, H, 80.0, 70.0, 47.5, 35.5)
    vm = O.back_project(np.full((H, W), D), K, np.float64)
    x, y = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    s = K.fx * tx / D
    lI = a * x + bgrad * y + c
    cI = a * (x + s) + bgrad * y + c
    dx, dy = O.derivative(cI, np.float64)
    P = np.eye(4)[:3]
    out = O.rgb_step(cI, lI, vm, dx, dy, 0.0, P.ravel(), K, np.float64)
    A, b, res = O.unpack29(out)
    xi = np.array([-tx, 0, 0, 0, 0, 0])
    assert out[28] == (W - 4) * (H - 4)  # 2 <= u < W - 2, 2 <= v < H - 2
    assert np.allclose(b, A @ xi, rtol=1e-9, atol=1e-12 * np.abs(A).max())
    assert np.isclose(res[0], xi @ A @ xi, rtol=1e-9)
    # the photometric error itself: sum r^2 = N (a s)^2
    assert np.isclose(res[0], out[28] * (a * s) ** 2, rtol=1e-9)


# ------------------------------------------------------ visibility, z-range

def corner_projections(blocks, Tcw, K, vs):
    """(u, v, z) of the 8 corners of every block (FP64): [n, 8] each."""
    off = np.array([(c & 1, (c >> 1) & 1, (c >> 2) & 1) for c in range(8)], np.float64)
    w = (blocks[:, None, :].astype(np.float64) + off[None]) * (8 * vs)
    p = w @ Tcw[:3, :3].T + Tcw[:3, 3]
    with np.errstate(invalid="ignore", divide="ignore"):
        u = K.fx * p[..., 0] / p[..., 2] + K.cx
        v = K.fy * p[..., 1] / p[..., 2] + K.cy
    return u, v, p[..., 2]


def scene_blocks(rng, n=4000):
    """Blocks scattered around and in front of a camera, some behind it."""
    return np.unique(rng.integers([-40, -30, -10], [40, 30, 90], size=(n, 3)), axis=0).astype(np.int32)


def test_block_visibility_equals_brute_force_corner_projection():
    """count_visible_entry on an FP64 map: a block is listed iff one of its 8
    corners projects inside the image (0 <= u <= cols - 1, 0 <= v <= rows - 1)
    with z in [zmin, zmax] (depth_fusion.cu:27-47)."""
    rng = np.random.default_rng(21)
    vs, cols, rows = 0.01, 64, 48
    K = O.make_intr(cols, rows, 50.0, 55.0, 31.5, 23.5)
    p = params(vs=vs, zmin=0.3, zmax=3.0)
    for trial in range(3):
        blocks = scene_blocks(rng)
        m = O.Map(p, np.float64)
        m.set_blocks(blocks, np.zeros((len(blocks), 512), np.int16), np.ones((len(blocks), 512)))
        T = small_pose(rng, t=0.1, deg=8)
        n = m.select_visible(T, K, cols, rows)
        got = {tuple(b) for b in m.visible().tolist()}
        u, v, z = corner_projections(blocks, np.linalg.inv(T), K, p.voxel_size)
        with np.errstate(invalid="ignore"):
            inside = (u >= 0) & (v >= 0) & (u <= cols - 1) & (v <= rows - 1) & (z >= p.zmin_update) & \
                     (z <= p.zmax_update)
        expect = {tuple(b) for b in blocks[inside.any(1)].tolist()}
        assert n == len(got) == len(expect) and got == expect, trial
        assert 50 < len(expect) < len(blocks) - 50


def brute_zrange(blocks, Tcw, K, vs, tx, ty, zmin, zmax):
    """fill_rendering_blocks by brute force: every tile takes the min near z
    and max far z of the blocks whose projected rectangle (floor / ceil of the
    corners on the 1/8 grid, int16-saturated, clipped to the grid) covers it."""
    u, v, z = corner_projections(blocks, Tcw, K, vs)
    with np.errstate(invalid="ignore"):
        fu, cu, fv, cv = np.floor(u / 8), np.ceil(u / 8), np.floor(v / 8), np.ceil(v / 8)
    sat = lambda a: np.clip(a, -32768, 32767)  # noqa: E731
    ulx = np.minimum(tx, sat(np.nanmin(np.where(np.isnan(fu), np.inf, fu), 1)))
    uly = np.minimum(ty, sat(np.nanmin(np.where(np.isnan(fv), np.inf, fv), 1)))
    lrx = np.maximum(-1, sat(np.nanmax(np.where(np.isnan(cu), -np.inf, cu), 1)))
    lry = np.maximum(-1, sat(np.nanmax(np.where(np.isnan(cv), -np.inf, cv), 1)))
    ulx, uly = np.maximum(ulx, 0), np.maximum(uly, 0)
    lrx, lry = np.minimum(lrx, tx - 1), np.minimum(lry, ty - 1)
    near = np.maximum(np.minimum(zmax, z.min(1)), zmin)
    far = np.maximum(zmin, z.max(1))
    ok = (ulx <= lrx) & (uly <= lry)
    X, Y = np.meshgrid(np.arange(tx), np.arange(ty))
    zx, zy = np.full((ty, tx), 100.0), np.zeros((ty, tx))
    for i in np.flatnonzero(ok):
        cover = (X >= ulx[i]) & (X <= lrx[i]) & (Y >= uly[i]) & (Y <= lry[i])
        zx[cover] = np.minimum(zx[cover], near[i])
        zy[cover] = np.maximum(zy[cover], far[i])
    return zx, zy, int(ok.sum())


def test_zrange_tiles_equal_brute_force_rectangles():
    """The oracle's z-range tiles (FP64) over its visible list equal the brute
    force over the same blocks: every tile's near / far z, and the untouched
    tiles' initial 100 / 0."""
    rng = np.random.default_rng(22)
    vs, cols, rows = 0.01, 96, 64
    K = O.make_intr(cols, rows, 70.0, 70.0, 47.5, 31.5)
    p = params(vs=vs, zmin=0.4, zmax=2.5)
    for trial in range(3):
        blocks = scene_blocks(rng, 3000)
        m = O.Map(p, np.float64)
        m.set_blocks(blocks, np.zeros((len(blocks), 512), np.int16), np.ones((len(blocks), 512)))
        T = small_pose(rng, t=0.1, deg=8)
        m.select_visible(T, K, cols, rows)
        vis = m.visible()
        zx, zy = m.zrange(T, K, cols, rows)
        ex, ey, made = brute_zrange(vis, np.linalg.inv(T), K, p.voxel_size, cols // 8, rows // 8, p.zmin_raycast,
                                    p.zmax_raycast)
        assert made > 20
        assert np.array_equal(zx == 100.0, ex == 100.0) and np.array_equal(zy == 0.0, ey == 0.0), trial
        assert np.allclose(zx, ex, rtol=1e-12) and np.allclose(zy, ey, rtol=1e-12), trial
        assert (zx < 100).mean() > 0.3


# --------------------------------------------------------------- rendering

def numpy_phong(P, N, pixel, light=(5.0, 5.0, 5.0)):
    """renderPoint (cuda_imgproc.cu:17-48) in FP64: ambient 0.3 A D, diffuse
    0.5 D max(0, n.l), specular 0.2 S max(0, r.v)^20 with r the light
    direction rotated by 180 degrees about the normal (its mirror image), then
    saturated and truncated to uint8; NaN vertices are (4, 2, 2)."""
    L = np.asarray(light) - P
    L /= np.linalg.norm(L, axis=1, keepdims=True)
    V = -P / np.linalg.norm(P, axis=1, keepdims=True)
    Nn = N / np.linalg.norm(N, axis=1, keepdims=True)
    R = Rotation.from_rotvec(np.pi * Nn).apply(L)
    nl = np.sum(N * L, 1)
    rv = np.sum(R * V, 1)
    A, D, S = pixel[:, 0], pixel[:, 1], pixel[:, 2]
    I = A * 0.3 * D + 0.5 * D * np.maximum(0.0, nl) + 0.2 * S * np.maximum(0.0, rv) ** 20
    g = np.floor(np.clip(I, 0.0, 1.0) * 255.0)
    return g


def sphere_maps(W=80, H=60, f=70.0, centre=(0.1, -0.05, 1.6), r=0.45):
    """Camera-frame vertices and outward unit normals of a sphere (NaN where a ray misses)."""
    x, y = np.meshgrid((np.arange(W) - (W - 1) / 2) / f, (np.arange(H) - (H - 1) / 2) / f)
    d = np.stack([x, y, np.ones_like(x)], -1)
    d /= np.linalg.norm(d, axis=-1, keepdims=True)
    c = np.asarray(centre)
    b = d @ c
    disc = b * b - (c @ c - r * r)
    t = b - np.sqrt(np.where(disc > 0, disc, np.nan))
    P = d * t[..., None]
    N = (P - c) / r
    vm = np.concatenate([P, np.ones((H, W, 1))], -1).astype(np.float32)
    nm = np.concatenate([N, np.ones((H, W, 1))], -1).astype(np.float32)
    return vm, nm


def test_render_shading_equals_numpy_phong():
    """renderScene / renderSceneTextured of the oracle within one level of the
    FP64 Phong model everywhere (the float and pow roundings), exact on at
    least 99% of the pixels, with the NaN colour where a ray missed."""
    vm, nm = sphere_maps()
    miss = np.isnan(vm[..., 0])
    assert 0.2 < miss.mean() < 0.8
    rng = np.random.default_rng(23)
    rgb = rng.integers(0, 256, size=vm.shape[:2] + (3,), dtype=np.uint8)
    for tex in (None, rgb):
        out = O.render_scene(vm, nm, tex)
        assert np.all(out[..., 3] == 255)
        assert np.all(out[miss][:, :3] == [4, 2, 2])
        hit = ~miss
        pix = np.ones((hit.sum(), 3)) if tex is None else tex[hit].astype(np.float32) / np.float32(255)
        expect = numpy_phong(vm[hit][:, :3].astype(np.float64), nm[hit][:, :3].astype(np.float64),
                             pix.astype(np.float64))
        got = out[hit][:, 0].astype(np.float64)
        assert np.all(out[hit][:, 0] == out[hit][:, 1]) and np.all(out[hit][:, 0] == out[hit][:, 2])
        assert np.max(np.abs(got - expect)) <= 1 and np.mean(got == expect) >= 0.99
        assert len(np.unique(got)) > 50  # the shading spans many levels


def test_warp_equals_scipy_bilinear():
    """warp_image: dst(x) = src bilinearly sampled where the destination's
    vertex projects under the pose, inside a one-pixel margin, 0 elsewhere;
    scipy's order-1 map_coordinates in FP64 then truncated agrees within one
    level (float rounding of the weights) and exactly on 99% of pixels."""
    rng = np.random.default_rng(24)
    W, H = 64, 48
    K = O.make_intr(W, H, 60.0, 60.0, 31.5, 23.5)
    src = scipy.ndimage.gaussian_filter(rng.uniform(0, 255, (H, W, 3)), (2, 2, 0)).astype(np.uint8)
    vm = O.back_project(rng.uniform(1.0, 2.0, (H, W)).astype(np.float32), K)
    T = small_pose(rng, t=0.03, deg=2)
    P = T[:3].astype(np.float32)
    out = O.warp_image(src, vm, P.ravel(), K)
    q = vm[..., :3].astype(np.float64) @ P[:, :3].T.astype(np.float64) + P[:, 3].astype(np.float64)
    u = K.fx * q[..., 0] / q[..., 2] + K.cx
    v = K.fy * q[..., 1] / q[..., 2] + K.cy
    inside = (u >= 1) & (v >= 1) & (u < W - 1) & (v < H - 1)
    assert 0.5 < inside.mean() < 1.0
    for ch in range(3):
        e = np.floor(scipy.ndimage.map_coordinates(src[..., ch].astype(np.float64), [v[inside], u[inside]],
                                                   order=1) + 1e-9)
        g = out[..., ch][inside].astype(np.float64)
        assert np.max(np.abs(g - e)) <= 1 and np.mean(g == e) >= 0.99
        assert np.all(out[..., ch][~inside] == 0)


# ------------------------------------------------------- colour update (1/d^2)

def test_colour_update_plane_closed_form():
    """update_map_with_colour_kernel (depth_fusion.cu:288-354) on a
    fronto-parallel plane at depth D: sdf = min(1, (D - z) / trunc) and weight
    1/D^2 (no normal term), colour copied, one observation; a second identical
    view doubles the weight and keeps sdf and colour."""
    D, trunc, vs, W, H = 1.5, 0.03, 0.01, 32, 24
    K = O.make_intr(W, H, 30.0, 30.0, (W - 1) / 2, (H - 1) / 2)
    depth = np.full((H, W), D, np.float32)
    rgb = np.full((H, W, 3), (90, 120, 150), np.uint8)
    m = O.Map(params(trunc=trunc), np.float32)
    m.update_weighted(depth, None, rgb, np.eye(4), K)
    blocks, sdf, w, col = m.voxels()
    z = (blocks[:, 2:3] * 8 + (np.arange(512) // 64)[None, :]) * vs
    x = (blocks[:, 0:1] * 8 + (np.arange(512) % 8)[None, :]) * vs
    y = (blocks[:, 1:2] * 8 + ((np.arange(512) // 8) % 8)[None, :]) * vs
    u = np.floor(30.0 * x / z + K.cx + 0.5)
    v = np.floor(30.0 * y / z + K.cy + 0.5)
    inside = (z > 0) & (u >= 0) & (u <= W - 1) & (v >= 0) & (v <= H - 1) & (D - z >= -trunc)
    assert inside.sum() > 1000
    assert np.max(np.abs(sdf[inside] / 32767.0 - np.minimum(1.0, (D - z[inside]) / trunc))) < 2.0 / 32767 + 1e-6
    assert np.allclose(w[inside], 1.0 / D ** 2, rtol=1e-6)
    assert np.all(w[~inside] == 0)
    assert all(np.all(col[:, k][inside] == c) for k, c in enumerate((90, 120, 150)))
    m.update_weighted(depth, None, rgb, np.eye(4), K)
    _, sdf2, w2, col2 = m.voxels()
    assert np.allclose(w2[inside], 2.0 / D ** 2, rtol=1e-6)
    assert np.max(np.abs(sdf2[inside].astype(int) - sdf[inside].astype(int))) <= 1
    assert np.array_equal(col2, col)
