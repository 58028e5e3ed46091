"""Seeded synthetic RGB-D scenes (host library liblf_synth.so).

Configurations (scene presets):
  0  single 320x240 frame: 4x4x3 m room + 3 spheres, fx=fy=262.5
  1  300-frame orbit (r 1.5 m, height 1.2 m, 1 cm / 0.5 deg jitter) at 640x480,
     room + 8 spheres/boxes, fx=fy=525
  2  same scene and trajectory as 1 (precision variant)
  3  one stream of the batched set: room of 3-6 m sides, resolution by seed % 4
  4  8x8x3 m multi-room, 40 primitives, 1000-frame random walk
Depth is uint16 at depth_scale metres per unit (1/5000), RGB is uint8 x3.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liblf_synth.so")
        if not os.path.exists(path):
            raise RuntimeError("liblf_synth.so is not built; run libfusion_amd.build.build_all()")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        lib.synth_scene_create.restype = P
        lib.synth_scene_create.argtypes = [ctypes.c_int, ctypes.c_ulonglong]
        lib.synth_scene_free.argtypes = [P]
        lib.synth_num_frames.argtypes = [P]
        lib.synth_intrinsics.argtypes = [P, P, P]
        lib.synth_depth_scale.argtypes = [P]
        lib.synth_depth_scale.restype = ctypes.c_double
        lib.synth_pose.argtypes = [P, ctypes.c_int, P]
        lib.synth_render.argtypes = [P, P, P, P, ctypes.c_int, ctypes.c_int, P, P, P]
        lib.synth_render_frame.argtypes = [P, ctypes.c_int, P, P]
        lib.synth_sdf.argtypes = [P, P, ctypes.c_int, P]
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class Scene:
    def __init__(self, config, seed):
        self._lib = _load()
        self.h = self._lib.synth_scene_create(int(config), int(seed))
        self.config, self.seed = config, seed
        wh = np.zeros(2, np.int32)
        f4 = np.zeros(4, np.float64)
        self._lib.synth_intrinsics(self.h, _ptr(wh), _ptr(f4))
        self.width, self.height = int(wh[0]), int(wh[1])
        self.fx, self.fy, self.cx, self.cy = (float(v) for v in f4)
        self.num_frames = int(self._lib.synth_num_frames(self.h))
        self.depth_scale = float(self._lib.synth_depth_scale(self.h))

    def __del__(self):
        try:
            self._lib.synth_scene_free(self.h)
        except Exception:
            pass

    def pose(self, frame):
        T = np.zeros((4, 4), np.float64)
        self._lib.synth_pose(self.h, int(frame), _ptr(T))
        return T

    def render(self, T, width=None, height=None, fx=None, fy=None, cx=None, cy=None,
               noise=True, frame_id=0, want_true=False):
        w = width or self.width
        h = height or self.height
        f4 = np.array([fx or self.fx, fy or self.fy, self.cx if cx is None else cx,
                       self.cy if cy is None else cy], np.float64)
        wh = np.array([w, h], np.int32)
        T = np.ascontiguousarray(T, np.float64)
        depth = np.zeros((h, w), np.uint16)
        rgb = np.zeros((h, w, 3), np.uint8)
        dtrue = np.zeros((h, w), np.float32) if want_true else None
        self._lib.synth_render(self.h, _ptr(T), _ptr(wh), _ptr(f4), 1 if noise else 0, int(frame_id),
                               _ptr(depth), _ptr(rgb), _ptr(dtrue))
        if want_true:
            return depth, rgb, dtrue
        return depth, rgb

    def frame(self, i):
        depth = np.zeros((self.height, self.width), np.uint16)
        rgb = np.zeros((self.height, self.width, 3), np.uint8)
        self._lib.synth_render_frame(self.h, int(i), _ptr(depth), _ptr(rgb))
        return depth, rgb

    def sdf(self, pts):
        pts = np.ascontiguousarray(pts, np.float64).reshape(-1, 3)
        out = np.zeros(len(pts), np.float64)
        self._lib.synth_sdf(self.h, _ptr(pts), len(pts), _ptr(out))
        return out
