"""ctypes bindings of the CPU oracle (liblf_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the reported CPU baseline.
The product path never loads it.

Every function exists in a float ("f", the method as the GPU runs it) and a
double ("d", the FP64 accuracy oracle) instantiation; pick with dtype.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class Params(ctypes.Structure):
    _fields_ = [("num_buckets", ctypes.c_int), ("num_entries", ctypes.c_int),
                ("num_blocks", ctypes.c_int), ("max_triangles", ctypes.c_int),
                ("zmin_raycast", ctypes.c_float), ("zmax_raycast", ctypes.c_float),
                ("zmin_update", ctypes.c_float), ("zmax_update", ctypes.c_float),
                ("voxel_size", ctypes.c_float), ("truncation", ctypes.c_float)]


class Intr(ctypes.Structure):
    _fields_ = [("width", ctypes.c_int), ("height", ctypes.c_int),
                ("fx", ctypes.c_float), ("fy", ctypes.c_float), ("cx", ctypes.c_float),
                ("cy", ctypes.c_float), ("invfx", ctypes.c_float), ("invfy", ctypes.c_float)]


def make_intr(w, h, fx, fy, cx, cy):
    f32 = np.float32
    return Intr(int(w), int(h), float(f32(fx)), float(f32(fy)), float(f32(cx)), float(f32(cy)),
                float(f32(1.0) / f32(fx)), float(f32(1.0) / f32(fy)))


def intr_pyr_down(K):
    s = np.float32(0.5)
    return make_intr(int(s * np.float32(K.width)), int(s * np.float32(K.height)),
                     s * np.float32(K.fx), s * np.float32(K.fy), s * np.float32(K.cx), s * np.float32(K.cy))


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liblf_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(path)
        P, I, F, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double
        for s, R in (("f", F), ("d", D)):
            def sig(name, res, args):
                fn = getattr(L, name % s)
                fn.restype = res
                fn.argtypes = args
            sig("orc_map_create_%s", P, [P])
            sig("orc_map_free_%s", None, [P])
            sig("orc_map_reset_%s", None, [P])
            sig("orc_map_num_blocks_%s", I, [P])
            sig("orc_map_num_visible_%s", I, [P])
            sig("orc_map_overflow_%s", ctypes.c_uint, [P])
            sig("orc_map_blocks_%s", None, [P, P])
            sig("orc_map_visible_%s", None, [P, P])
            sig("orc_map_voxels_%s", None, [P, P, P, P])
            sig("orc_map_set_blocks_%s", I, [P, I, P, P, P, P])
            sig("orc_allocate_%s", None, [P, P, I, I, P, P])
            sig("orc_update_weighted_%s", None, [P, P, P, P, I, I, P, P])
            sig("orc_raycast_%s", None, [P, P, P, I, I, P, P, I, I])
            sig("orc_ray_blocks_touched_%s", I, [P])
            sig("orc_ray_stats_%s", None, [P, P])
            sig("orc_ray_px_%s", None, [P, P])
            sig("orc_select_visible_%s", I, [P, P, P, I, I])
            sig("orc_track_maps_%s", I, [P, P, P, P, P, I, P, P, I, P])
            sig("orc_zrange_%s", None, [P, P, P, I, I, P, P])
            sig("orc_mesh_%s", ctypes.c_long, [P, I, P, P, P, ctypes.c_long])
            sig("orc_bilateral_%s", None, [P, P, I, I])
            sig("orc_resize_half_%s", None, [P, P, I, I, I])
            sig("orc_pyr_down_%s", None, [P, P, I, I])
            sig("orc_grey_%s", None, [P, P, I, I])
            sig("orc_derivative_%s", None, [P, P, P, I, I])
            sig("orc_back_project_%s", None, [P, P, I, I, P])
            sig("orc_nmap_%s", None, [P, P, I, I])
            sig("orc_icp_reduce_%s", None, [P, P, P, P, I, I, P, P, P])
            sig("orc_rgb_step_%s", None, [P, P, P, P, P, I, I, R, P, P, P])
            sig("orc_icp_rows_%s", None, [P, P, P, P, I, I, P, P, P])
            sig("orc_rgb_rows_%s", None, [P, P, P, P, P, I, I, R, P, P, P])
            sig("orc_system_create_%s", P, [P, P, I, P, D])
            sig("orc_system_free_%s", None, [P])
            sig("orc_system_process_%s", None, [P, P, P])
            sig("orc_system_pose_%s", None, [P, P])
            sig("orc_system_success_%s", I, [P])
            sig("orc_system_map_%s", P, [P])
            sig("orc_system_ref_maps_%s", None, [P, I, P, P])
        L.orc_compute_hash.restype = I
        L.orc_compute_hash.argtypes = [I, I, I, I]
        L.orc_set_threads.argtypes = [I]
        L.orc_max_threads.restype = I
        L.orc_se3_exp.argtypes = [P, P]
        L.orc_se3_log.argtypes = [P, P]
        L.orc_ldlt6.argtypes = [P, P, P]
        L.orc_render_scene.argtypes = [P, P, P, I, I, P]
        L.orc_warp_image.argtypes = [P, P, I, I, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _sfx(dtype):
    return "f" if np.dtype(dtype) == np.float32 else "d"


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def compute_hash(x, y, z, n):
    return int(lib().orc_compute_hash(int(x), int(y), int(z), int(n)))


def se3_exp(xi):
    T = np.zeros((4, 4))
    lib().orc_se3_exp(_p(_c(xi, np.float64)), _p(T))
    return T


def se3_log(T):
    xi = np.zeros(6)
    lib().orc_se3_log(_p(_c(T, np.float64)), _p(xi))
    return xi


def ldlt6(A, b):
    x = np.zeros(6)
    lib().orc_ldlt6(_p(_c(A, np.float64)), _p(_c(b, np.float64)), _p(x))
    return x


def pose24(T, dtype=np.float32):
    """Device-style pose block: T_world_camera 3x4 then its inverse, in dtype."""
    T = np.asarray(T, np.float64)
    Ti = np.eye(4)
    Ti[:3, :3] = T[:3, :3].T
    Ti[:3, 3] = -T[:3, :3].T @ T[:3, 3]
    return np.concatenate([T[:3, :].ravel(), Ti[:3, :].ravel()]).astype(dtype)


class Map:
    def __init__(self, params, dtype=np.float32):
        self.dtype = np.dtype(dtype)
        self.s = _sfx(dtype)
        self.params = params
        self.h = getattr(lib(), "orc_map_create_" + self.s)(ctypes.byref(params))

    def __del__(self):
        if getattr(self, "_borrowed", False):
            return
        try:
            getattr(lib(), "orc_map_free_" + self.s)(self.h)
        except Exception:
            pass

    def _f(self, name):
        return getattr(lib(), name + "_" + self.s)

    def reset(self):
        self._f("orc_map_reset")(self.h)

    def num_blocks(self):
        return int(self._f("orc_map_num_blocks")(self.h))

    def overflow(self):
        return int(self._f("orc_map_overflow")(self.h))

    def blocks(self):
        n = self.num_blocks()
        out = np.zeros((n, 3), np.int32)
        self._f("orc_map_blocks")(self.h, _p(out))
        return out

    def visible(self):
        n = int(self._f("orc_map_num_visible")(self.h))
        out = np.zeros((n, 3), np.int32)
        self._f("orc_map_visible")(self.h, _p(out))
        return out

    def voxels(self):
        """(blocks[n,3], sdf int16[n,512], weight[n,512], rgb uint8[n,3,512]) in allocation order."""
        n = self.num_blocks()
        sdf = np.zeros((n, 512), np.int16)
        w = np.zeros((n, 512), self.dtype)
        rgb = np.zeros((n, 3, 512), np.uint8)
        self._f("orc_map_voxels")(self.h, _p(sdf), _p(w), _p(rgb))
        return self.blocks(), sdf, w, rgb

    def set_blocks(self, xyz, sdf, w, rgb=None):
        xyz = _c(xyz, np.int32)
        r = self._f("orc_map_set_blocks")(self.h, len(xyz), _p(xyz), _p(_c(sdf, np.int16)),
                                          _p(_c(w, self.dtype)), _p(None if rgb is None else _c(rgb, np.uint8)))
        if r != 0:
            raise RuntimeError("oracle pool full")

    def allocate(self, depth, T, K):
        d = _c(depth, self.dtype)
        self._f("orc_allocate")(self.h, _p(d), d.shape[1], d.shape[0], _p(pose24(T, self.dtype)), ctypes.byref(K))

    def update_weighted(self, depth, nmap, rgb, T, K):
        d = _c(depth, self.dtype)
        self._f("orc_update_weighted")(self.h, _p(d), _p(None if nmap is None else _c(nmap, self.dtype)),
                                       _p(_c(rgb, np.uint8)), d.shape[1], d.shape[0],
                                       _p(pose24(T, self.dtype)), ctypes.byref(K))

    def raycast(self, T, K, cols, rows, check_colour=True, want_rgb=False, instrument=False):
        vmap = np.zeros((rows, cols, 4), self.dtype)
        rgb = np.zeros((rows, cols, 3), np.uint8) if want_rgb else None
        self._f("orc_raycast")(self.h, _p(vmap), _p(rgb), cols, rows, _p(pose24(T, self.dtype)),
                               ctypes.byref(K), 1 if check_colour else 0, 1 if instrument else 0)
        return (vmap, rgb) if want_rgb else vmap

    def select_visible(self, T, K, cols, rows):
        """count_visible_entry: the visible list becomes the frustum-visible blocks at T."""
        return int(self._f("orc_select_visible")(self.h, _p(pose24(T, self.dtype)), ctypes.byref(K), cols, rows))

    def ray_stats(self):
        """Counters of the last instrumented raycast: samples, samples in unallocated
        blocks, trilinear reads, block changes along rays, rays started, hits."""
        out = np.zeros(6, np.int64)
        self._f("orc_ray_stats")(self.h, _p(out))
        return dict(zip(("samples", "samples_unallocated", "trilinear", "block_changes", "rays", "hits"),
                        out.tolist()))

    def ray_pixels(self, cols, rows):
        """Per-pixel counters of the last instrumented raycast: [rows, cols, 3] =
        samples, samples in allocated blocks, trilinear reads."""
        out = np.zeros((rows, cols, 3), np.int32)
        self._f("orc_ray_px")(self.h, _p(out))
        return out

    def ray_blocks_touched(self):
        return int(self._f("orc_ray_blocks_touched")(self.h))

    def zrange(self, T, K, cols, rows):
        zx = np.zeros((rows // 8, cols // 8), self.dtype)
        zy = np.zeros((rows // 8, cols // 8), self.dtype)
        self._f("orc_zrange")(self.h, _p(zx), _p(zy), cols, rows, _p(pose24(T, self.dtype)), ctypes.byref(K))
        return zx, zy

    def mesh(self, mode=1, max_tri=None):
        """mode 0 vertices, 1 + normals, 2 + colours.  Returns (tri[n,3,3], nrm or col or None, count)."""
        cnt = int(self._f("orc_mesh")(self.h, mode, None, None, None, 0))
        m = cnt if max_tri is None else min(cnt, max_tri)
        tri = np.zeros((m, 3, 3), np.float32)
        nrm = np.zeros((m, 3, 3), np.float32) if mode == 1 else None
        col = np.zeros((m, 3, 3), np.uint8) if mode == 2 else None
        self._f("orc_mesh")(self.h, mode, _p(tri), _p(nrm), _p(col), m)
        return tri, (nrm if mode == 1 else col), cnt


def _img(name, dtype, *arrays_and_args):
    return getattr(lib(), name + "_" + _sfx(dtype))


def bilateral(depth, dtype=np.float32):
    d = _c(depth, dtype)
    out = np.zeros_like(d)
    _img("orc_bilateral", dtype)(_p(d), _p(out), d.shape[1], d.shape[0])
    return out


def resize_half(img, dtype=np.float32):
    a = _c(img, dtype)
    h, w = a.shape[:2]
    ch = 1 if a.ndim == 2 else a.shape[2]
    out = np.zeros((h // 2, w // 2) + (() if a.ndim == 2 else (ch,)), dtype)
    _img("orc_resize_half", dtype)(_p(a), _p(out), w, h, ch)
    return out


def pyr_down(img, dtype=np.float32):
    a = _c(img, dtype)
    h, w = a.shape
    out = np.zeros(((h + 1) // 2, (w + 1) // 2), dtype)
    _img("orc_pyr_down", dtype)(_p(a), _p(out), w, h)
    return out


def grey(rgb, dtype=np.float32):
    a = _c(rgb, np.uint8)
    out = np.zeros(a.shape[:2], dtype)
    _img("orc_grey", dtype)(_p(a), _p(out), a.shape[1], a.shape[0])
    return out


def derivative(img, dtype=np.float32):
    a = _c(img, dtype)
    dx = np.zeros_like(a)
    dy = np.zeros_like(a)
    _img("orc_derivative", dtype)(_p(a), _p(dx), _p(dy), a.shape[1], a.shape[0])
    return dx, dy


def back_project(depth, K, dtype=np.float32):
    d = _c(depth, dtype)
    out = np.zeros(d.shape + (4,), dtype)
    _img("orc_back_project", dtype)(_p(d), _p(out), d.shape[1], d.shape[0], ctypes.byref(K))
    return out


def nmap(vmap, dtype=np.float32):
    v = _c(vmap, dtype)
    out = np.zeros_like(v)
    _img("orc_nmap", dtype)(_p(v), _p(out), v.shape[1], v.shape[0])
    return out


def render_scene(vmap, nmap, rgb=None):
    """renderScene / renderSceneTextured: RGBA8 Phong shading, light at (5, 5, 5)."""
    v, n = _c(vmap, np.float32), _c(nmap, np.float32)
    out = np.zeros(v.shape[:2] + (4,), np.uint8)
    c = None if rgb is None else _c(rgb, np.uint8)
    lib().orc_render_scene(_p(v), _p(n), _p(c) if c is not None else None, v.shape[1], v.shape[0], _p(out))
    return out


def warp_image(src, vmap_dst, pose12, K):
    """warp_image: src (RGB8) sampled where vmap_dst projects under pose12 (float 3x4)."""
    s, v = _c(src, np.uint8), _c(vmap_dst, np.float32)
    P = _c(pose12, np.float32)
    out = np.zeros_like(s)
    lib().orc_warp_image(_p(s), _p(v), v.shape[1], v.shape[0], _p(P), ctypes.byref(K), _p(out))
    return out


def icp_reduce(curr_v, curr_n, last_v, last_n, pose12, K, dtype=np.float32):
    out = np.zeros(29, dtype)
    cv = _c(curr_v, dtype)
    _img("orc_icp_reduce", dtype)(_p(cv), _p(_c(curr_n, dtype)), _p(_c(last_v, dtype)), _p(_c(last_n, dtype)),
                                  cv.shape[1], cv.shape[0], _p(_c(pose12, dtype)), ctypes.byref(K), _p(out))
    return out


def rgb_step(curr_I, last_I, last_v, dx, dy, stddev, pose12, K, dtype=np.float32):
    out = np.zeros(29, dtype)
    ci = _c(curr_I, dtype)
    _img("orc_rgb_step", dtype)(_p(ci), _p(_c(last_I, dtype)), _p(_c(last_v, dtype)), _p(_c(dx, dtype)),
                                _p(_c(dy, dtype)), ci.shape[1], ci.shape[0], float(stddev),
                                _p(_c(pose12, dtype)), ctypes.byref(K), _p(out))
    return out


def icp_rows(curr_v, curr_n, last_v, last_n, pose12, K, dtype=np.float32):
    """Every reference pixel's ICP row: [H*W, 8] = the 7 row values (0 when
    rejected) and 1 where the pixel was accepted."""
    cv = _c(curr_v, dtype)
    out = np.zeros((cv.shape[0] * cv.shape[1], 8), dtype)
    _img("orc_icp_rows", dtype)(_p(cv), _p(_c(curr_n, dtype)), _p(_c(last_v, dtype)), _p(_c(last_n, dtype)),
                                cv.shape[1], cv.shape[0], _p(_c(pose12, dtype)), ctypes.byref(K), _p(out))
    return out


def rgb_rows(curr_I, last_I, last_v, dx, dy, stddev, pose12, K, dtype=np.float32):
    """Every reference pixel's photometric row, laid out as icp_rows'."""
    ci = _c(curr_I, dtype)
    out = np.zeros((ci.shape[0] * ci.shape[1], 8), dtype)
    _img("orc_rgb_rows", dtype)(_p(ci), _p(_c(last_I, dtype)), _p(_c(last_v, dtype)), _p(_c(dx, dtype)),
                                _p(_c(dy, dtype)), ci.shape[1], ci.shape[0], float(stddev),
                                _p(_c(pose12, dtype)), ctypes.byref(K), _p(out))
    return out


def unpack29(s):
    """create_jtjjtr<6,7> plus residual {sum r^2, count}."""
    A = np.zeros((6, 6))
    b = np.zeros(6)
    k = 0
    for i in range(6):
        for j in range(i, 7):
            v = float(s[k])
            k += 1
            if j == 6:
                b[i] = v
            else:
                A[i, j] = A[j, i] = v
    return A, b, np.array([s[27], s[28]], np.float64)


def track_maps(ref_vmap, ref_rgb, cur_vmap, cur_rgb, K, iters, init=None, context_mode=True, dtype=np.float64):
    """compute_transform on pyramids built from level-0 vertex maps and images
    (set_source_vmap / set_source_image).  Returns (success, update 4x4)."""
    it = np.ascontiguousarray(iters, np.int32)
    T0 = np.eye(4) if init is None else np.asarray(init, np.float64)
    ini = np.ascontiguousarray(T0[:3, :], np.float64)
    out = np.zeros(12, np.float64)
    ok = getattr(lib(), "orc_track_maps_" + _sfx(dtype))(
        _p(_c(ref_vmap, dtype)), _p(_c(ref_rgb, np.uint8)), _p(_c(cur_vmap, dtype)), _p(_c(cur_rgb, np.uint8)),
        ctypes.byref(K), len(it), _p(it), _p(ini), 1 if context_mode else 0, _p(out))
    U = np.eye(4)
    U[:3, :] = out.reshape(3, 4)
    return bool(ok), U


class System:
    def __init__(self, params, K, num_pyr, iters, depth_scale, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.s = _sfx(dtype)
        it = np.array(list(iters) + [0] * (8 - len(iters)), np.int32)
        self.K = K
        self.h = getattr(lib(), "orc_system_create_" + self.s)(ctypes.byref(params), ctypes.byref(K), int(num_pyr),
                                                              _p(it), float(depth_scale))
        self._params = params

    def __del__(self):
        try:
            getattr(lib(), "orc_system_free_" + self.s)(self.h)
        except Exception:
            pass

    def process(self, depth_u16, rgb_u8):
        getattr(lib(), "orc_system_process_" + self.s)(self.h, _p(_c(depth_u16, np.uint16)), _p(_c(rgb_u8, np.uint8)))

    def pose(self):
        T = np.zeros((4, 4))
        getattr(lib(), "orc_system_pose_" + self.s)(self.h, _p(T))
        return T

    def success(self):
        return bool(getattr(lib(), "orc_system_success_" + self.s)(self.h))

    def ref_maps(self, level, w, h):
        v = np.zeros((h, w, 4), self.dtype)
        n = np.zeros((h, w, 4), self.dtype)
        getattr(lib(), "orc_system_ref_maps_" + self.s)(self.h, int(level), _p(v), _p(n))
        return v, n

    def map(self):
        """The system's map, borrowed (not freed with the returned object)."""
        m = Map.__new__(Map)
        m.dtype, m.s, m.params, m._borrowed = self.dtype, self.s, self._params, True
        m.h = getattr(lib(), "orc_system_map_" + self.s)(self.h)
        return m
