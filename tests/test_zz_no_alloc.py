"""GPU: lf_icp_reduce_noalloc (the fixed-scratch variant of lf_icp_reduce for
callers that cannot allocate) returns exactly lf_icp_reduce's sums, on two
calls in a row, the second on a smaller image: the same kernels and grid run
over the library's fixed device scratch instead of the caller's workspace."""
import ctypes

import numpy as np
import pytest
import torch

from libfusion_amd import capi
from libfusion_amd.synth import Scene
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def frame_maps(sc, frame, K):
    d16, _ = sc.render(sc.pose(frame), K.width, K.height, K.fx, K.fy, K.cx, K.cy, noise=True, frame_id=frame)
    depth = d16.astype(np.float32) * np.float32(sc.depth_scale)
    oK = O.Intr(K.width, K.height, K.fx, K.fy, K.cx, K.cy, K.invfx, K.invfy)
    vm = O.back_project(O.bilateral(depth), oK)
    return vm, O.nmap(vm)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_icp_reduce_noalloc_equals_icp_reduce():
    lib = capi.load()
    sc = Scene(1, 1)
    full = capi.intrinsics(sc.width, sc.height, sc.fx, sc.fy, sc.cx, sc.cy)
    half = capi.intrinsics(sc.width // 2, sc.height // 2, sc.fx / 2, sc.fy / 2, (sc.cx + 0.5) / 2 - 0.5,
                           (sc.cy + 0.5) / 2 - 0.5)
    Trel = np.linalg.inv(sc.pose(1)) @ sc.pose(0)
    pose = dev(Trel[:3, :].astype(np.float32).ravel())
    scratch = torch.zeros(int(lib.lf_reduce_scratch_floats()), device="cuda")
    st = capi.stream_ptr()
    P = capi.ptr
    for K in (full, half):  # the second call is the smaller image
        (cv, cn), (lv, ln) = frame_maps(sc, 1, K), frame_maps(sc, 0, K)
        t = [dev(a) for a in (cv, cn, lv, ln)]
        assert all(x.shape[:2] == (K.height, K.width) for x in t)
        ref = torch.full((29,), float("nan"), device="cuda")
        got = torch.full((29,), float("nan"), device="cuda")
        capi.check(lib.lf_icp_reduce(*[P(x) for x in t], K.width, K.height, P(pose), ctypes.byref(K), P(scratch),
                                     P(ref), st))
        capi.check(lib.lf_icp_reduce_noalloc(*[P(x) for x in t], K.width, K.height, P(pose), ctypes.byref(K),
                                             P(got), st))
        torch.cuda.synchronize()
        r, g = ref.cpu().numpy(), got.cpu().numpy()
        assert r[28] > 1000, r[28]  # correspondences found
        assert np.array_equal(r.view(np.uint32), g.view(np.uint32)), (K.width, r, g)
    # a null output is refused before any device work
    assert lib.lf_icp_reduce_noalloc(*[P(x) for x in t], half.width, half.height, P(pose), ctypes.byref(half),
                                     None, st) == capi.LF_ERR_ARG
