"""Topological charge of spin textures (drop-in for magex.texture.topological_charge), on the GPU.

Q = (1/4 pi) sum over triangles of the Berg-Luscher solid angle 2 atan2(s1.(s2 x s3), 1 + s1.s2 +
s2.s3 + s3.s1) (magex/texture.py:30-63).  The layer is triangulated once on the host (periodic Delaunay
over a 3 x 3 tiling, triangles whose centroid lies in the home supercell, each ordered
counter-clockwise) and cached; the sum over triangles runs in libmxllg.so.
"""
from __future__ import annotations

import numpy as np

PLANES = {"xy": (0, 1, 2), "xz": (0, 2, 1), "yz": (1, 2, 0)}


def layer_indices(system, plane="xy", layer=0, tol=0.05):
    """Indices of the spins in one atomic layer normal to the plane (magex/texture.py:14-20)."""
    a, b, n = PLANES[plane]
    z = system.positions[:, n]
    levels = np.unique(np.round(z / tol) * tol)
    lev = levels[layer]
    return np.where(np.abs(z - lev) < tol * 1.01)[0], (a, b, n), len(levels)


def triangles(system, plane="xy", layer=0):
    """(ntri, 3) int32 site indices of the layer's periodic triangulation, counter-clockwise."""
    from scipy.spatial import Delaunay
    idx, (a, b, n), _ = layer_indices(system, plane, layer)
    if len(idx) < 3:
        return np.zeros((0, 3), np.int32)
    P = system.positions[idx][:, [a, b]]
    V3 = np.array(system.L)[:, None] * system.uc.cell
    order = np.argsort(np.abs(V3[:, n]))
    V = V3[order[:2]][:, [a, b]]
    per = [bool(system.pbc[k]) for k in order[:2]]
    shifts = [(i, j) for i in ((-1, 0, 1) if per[0] else (0,)) for j in ((-1, 0, 1) if per[1] else (0,))]
    Pt = np.concatenate([P + i * V[0] + j * V[1] for i, j in shifts])
    try:
        tri = Delaunay(Pt).simplices
    except Exception:
        return np.zeros((0, 3), np.int32)
    frac = np.linalg.solve(V.T, Pt[tri].mean(1).T).T
    if len(shifts) > 1:
        tri = tri[np.all((frac >= -1e-9) & (frac < 1 - 1e-9), axis=1)]
    p1, p2, p3 = Pt[tri[:, 0]], Pt[tri[:, 1]], Pt[tri[:, 2]]
    cw = (p2[:, 0] - p1[:, 0]) * (p3[:, 1] - p1[:, 1]) - (p2[:, 1] - p1[:, 1]) * (p3[:, 0] - p1[:, 0]) < 0
    tri[cw] = tri[cw][:, [0, 2, 1]]
    return np.ascontiguousarray(idx[tri % len(idx)], np.int32)


def topological_charge(system, e, plane="xy", layer=0):
    """Skyrmion number Q of layer `layer` normal to `plane` (magex/texture.py:30-63)."""
    import torch
    from .device import as_device
    key = ("tri", plane, layer, system._version)
    cache = system.__dict__.setdefault("_tri_cache", {})
    tri = cache.get(key)
    if tri is None:
        cache.clear()
        tri = cache[key] = torch.as_tensor(triangles(system, plane, layer), device="cuda")
    if tri.shape[0] == 0:
        return 0.0
    fp32 = str(getattr(e, "dtype", "")) in ("float32", "torch.float32")
    dm = system.device_model("fp32" if fp32 else "fp64")
    hand = 1.0 if plane in ("xy", "yz") else -1.0
    return dm.topological_charge(as_device(e, dm.dtype), tri, hand)
