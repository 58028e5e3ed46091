"""Workload definitions: scenes, sensors and volumes of the benchmark configurations.

C0  single 320x240 frame, 10 mm voxels, 30 mm truncation, 2^17 buckets
C1  300-frame 640x480 orbit, 5 mm voxels, 15 mm truncation, 2^20 buckets,
    3-level pyramid with {10, 5, 3} Gauss-Newton iterations (finest first)
"""
from . import capi
from .synth import Scene

SEEDS = {0: 0, 1: 1, 2: 1}


def c0():
    sc = Scene(0, SEEDS[0])
    K = capi.intrinsics(sc.width, sc.height, sc.fx, sc.fy, sc.cx, sc.cy)
    p = capi.map_params(2 ** 17, 2 ** 16, 0.01, 0.03)
    return sc, K, p


def c1(num_blocks=1 << 19):
    sc = Scene(1, SEEDS[1])
    K = capi.intrinsics(sc.width, sc.height, sc.fx, sc.fy, sc.cx, sc.cy)
    p = capi.map_params(2 ** 20, num_blocks, 0.005, 0.015)
    return sc, K, p


C1_ITERATIONS = (10, 5, 3)
C1_LEVELS = 3
