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
