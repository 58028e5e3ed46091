"""Distinct voxel blocks and hash entries sampled by the C1 raycast (test
infrastructure: the raycast's compulsory bytes for its roofline basis).

Runs the FP32 oracle System over the C1 orbit and, at the sampled frames,
repeats the frame's raycast with the oracle's instrumented march, counting the
distinct blocks read.  The raycast's algorithmic bytes per frame are then
    P * 16 (vertex map) + N_blocks * 512 * 6 (sdf + weight) + N_blocks * 20
    (hash entries) + 2 * tiles * 4 (z-range tiles).
Writes profiles/raycast_c1_touched.json.

    python oracle/raycast_touched.py [frames ...]
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path[0] = ROOT  # the package, not this directory

from libfusion_amd import configs  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    frames = [int(a) for a in sys.argv[1:]] or [50, 150, 250]
    sc, K, p = configs.c1()
    op = O.Params(p.num_buckets, p.num_entries, p.num_blocks, p.max_triangles, p.zmin_raycast, p.zmax_raycast,
                  p.zmin_update, p.zmax_update, p.voxel_size, p.truncation)
    oK = O.make_intr(K.width, K.height, K.fx, K.fy, K.cx, K.cy)
    s = O.System(op, oK, configs.C1_LEVELS, configs.C1_ITERATIONS, sc.depth_scale, np.float32)
    samples = []
    for i in range(max(frames) + 1):
        d, c = sc.frame(i)
        s.process(d, c)
        if i in frames:
            m = s.map()
            m.raycast(s.pose(), oK, K.width, K.height, instrument=True)
            st = m.ray_stats()
            samples.append({"frame": i, "blocks_touched": m.ray_blocks_touched(), "rays": st["rays"],
                            "samples": st["samples"], "hits": st["hits"]})
            print(samples[-1], flush=True)
    P = K.width * K.height
    tiles = (K.width // 8) * (K.height // 8)
    nb = float(np.mean([x["blocks_touched"] for x in samples]))
    out = {"workload": "C1 640x480, 5 mm voxels: raycast at the tracked pose after the frame's update",
           "frames": samples, "blocks_touched_mean": nb,
           "algorithmic_bytes_per_raycast": int(P * 16 + nb * 512 * 6 + nb * 20 + 2 * tiles * 4),
           "command": "python oracle/raycast_touched.py " + " ".join(map(str, frames))}
    path = os.path.join(ROOT, "profiles", "raycast_c1_touched.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, out["algorithmic_bytes_per_raycast"])


if __name__ == "__main__":
    main()
