"""Write tests/golden/c0.npz: the oracle's (FP32) result on the C0 workload.

C0: one 320x240 synthetic view (seed 0) of a 4x4x3 m room with 3 spheres, fused
into a 10 mm / 30 mm-truncation voxel hash with 2^17 buckets, then raycast at
the same pose.  Stored: the sorted allocated and visible block sets, the TSDF /
weight of a fixed sample of voxels, the raycast vertex map on a pixel grid and
the marching-cubes triangle count.  The reference ships no test vectors, so
these fixtures pin the build's own oracle (regression) and the GPU path.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from libfusion_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402


def c0_oracle():
    sc, K, p = configs.c0()
    T = sc.pose(0)
    oK = O.Intr(K.width, K.height, K.fx, K.fy, K.cx, K.cy, K.invfx, K.invfy)
    op = O.Params(p.num_buckets, p.num_entries, p.num_blocks, p.max_triangles, p.zmin_raycast, p.zmax_raycast,
                  p.zmin_update, p.zmax_update, p.voxel_size, p.truncation)
    d16, rgb = sc.frame(0)
    depth = d16.astype(np.float32) * np.float32(sc.depth_scale)
    nmap = O.nmap(O.back_project(O.bilateral(depth), oK))
    m = O.Map(op, np.float32)
    m.update_weighted(depth, nmap, rgb, T, oK)
    vm = m.raycast(T, oK, K.width, K.height)
    return sc, K, p, T, depth, nmap, rgb, m, vm


def fixture():
    sc, K, p, T, depth, nmap, rgb, m, vm = c0_oracle()
    blocks, sdf, w, _ = m.voxels()
    order = np.lexsort((blocks[:, 2], blocks[:, 1], blocks[:, 0]))
    rng = np.random.default_rng(1234)
    pick = np.sort(rng.choice(len(order), size=min(256, len(order)), replace=False))
    sel = order[pick]
    vis = m.visible()
    vis = vis[np.lexsort((vis[:, 2], vis[:, 1], vis[:, 0]))]
    _, _, ntri = m.mesh(mode=0, max_tri=0)
    return {
        "blocks": blocks[order].astype(np.int32),
        "visible": vis.astype(np.int32),
        "sample_blocks": blocks[sel].astype(np.int32),
        "sample_sdf": sdf[sel],
        "sample_weight": w[sel],
        "vmap_grid": vm[::4, ::4].astype(np.float32),
        "triangles": np.array([ntri], np.int64),
    }


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "c0.npz"), **fixture())
    print("wrote", os.path.join(HERE, "c0.npz"))
