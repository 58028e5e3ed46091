"""Shared helpers of the parity tests (test infrastructure)."""
import numpy as np


def rot_err_deg(Ra, Rb):
    R = Ra.T @ Rb
    return float(np.degrees(np.arccos(np.clip((np.trace(R) - 1.0) / 2.0, -1.0, 1.0))))


def pose_err(Ta, Tb):
    """(translation error m, rotation error deg) between two 4x4 poses."""
    return float(np.linalg.norm(Ta[:3, 3] - Tb[:3, 3])), rot_err_deg(Ta[:3, :3], Tb[:3, :3])


def sort_blocks(xyz):
    xyz = np.asarray(xyz)
    order = np.lexsort((xyz[:, 2], xyz[:, 1], xyz[:, 0]))
    return order, xyz[order]


def canonical_triangles(tri, extra=None, decimals=5):
    """Sort triangles (n,3,3) by their rounded vertices (then their extra per-vertex
    data) so order-free outputs compare; coincident degenerate triangles from
    neighbouring cubes are ordered by the extra data."""
    tri = np.asarray(tri, np.float64)
    key = np.round(tri.reshape(len(tri), 9), decimals)
    if extra is not None:
        key = np.concatenate([key, np.nan_to_num(np.asarray(extra, np.float64).reshape(len(tri), -1))], axis=1)
    order = np.lexsort(key.T[::-1])
    return tri[order], (None if extra is None else np.asarray(extra)[order])


def vmap_compare(a, b):
    """Fraction of pixels whose validity differs, and max |delta| over pixels valid in both."""
    va = ~np.isnan(a[..., 0])
    vb = ~np.isnan(b[..., 0])
    both = va & vb
    mism = float(np.mean(va != vb))
    d = float(np.max(np.abs(a[both][:, :3] - b[both][:, :3]))) if both.any() else 0.0
    return mism, d, int(both.sum())


def keyframe_pair(sc, a, b, perturb=(0.01, 0.5), seed=0):
    """Half-resolution inputs of re-tracking frame b (source) against frame a
    (reference): bilateral back-projected vertex maps and images, the true
    update (T_ref^-1 T_cur) and an initial estimate of T_cur^-1 T_ref perturbed
    by perturb = (metres, degrees)."""
    from oracle import oracle as O
    from libfusion_amd import capi
    K = capi.intrinsics(sc.width // 2, sc.height // 2, sc.fx / 2, sc.fy / 2, (sc.cx + 0.5) / 2 - 0.5,
                        (sc.cy + 0.5) / 2 - 0.5)
    oK = O.make_intr(K.width, K.height, K.fx, K.fy, K.cx, K.cy)
    out = []
    for f in (a, b):
        d16, rgb = sc.render(sc.pose(f), K.width, K.height, K.fx, K.fy, K.cx, K.cy, noise=True, frame_id=f)
        depth = d16.astype(np.float32) * np.float32(sc.depth_scale)
        out.append((O.back_project(O.bilateral(depth), oK), rgb))
    truth = np.linalg.inv(sc.pose(a)) @ sc.pose(b)
    rng = np.random.default_rng(seed)
    w = rng.normal(size=3)
    w *= np.deg2rad(perturb[1]) / np.linalg.norm(w)
    t = rng.normal(size=3)
    t *= perturb[0] / np.linalg.norm(t)
    P = np.eye(4)
    P[:3, :3] = O.se3_exp(np.concatenate([np.zeros(3), w]))[:3, :3]
    P[:3, 3] = t
    init = P @ np.linalg.inv(truth)
    return K, oK, out[0], out[1], truth, init


def triangles_in_hash_order(tri):
    """Triangles (n,3,3) sorted by a 64-bit mix of their 9 float32 bit patterns
    (-0.0 taken as 0.0): two meshes hold bit-identical triangle multisets iff
    the results are equal (a hash tie between different triangles can only
    make equal multisets compare unequal). Linear-time stand-in for
    canonical_triangles on meshes of millions of triangles."""
    t = np.ascontiguousarray(np.asarray(tri, np.float32) + np.float32(0.0)).reshape(len(tri), 9)
    u = t.view(np.uint32).astype(np.uint64)
    h = np.full(len(t), 0x9E3779B97F4A7C15, np.uint64)
    for k in range(9):
        h = (h ^ u[:, k]) * np.uint64(0xBF58476D1CE4E5B9)
        h ^= h >> np.uint64(31)
    return t[np.argsort(h, kind="stable")]
