"""Independent geometric references used to pin the oracle (tests only).

siddon_line_integral: the textbook exact ray-voxel intersection (Siddon 1985):
sort the parametric crossings of every grid plane, weight each voxel by its
chord length.  Float64, no relation to the marcher's slab arithmetic.
"""
import numpy as np


def siddon_line_integral(img, center, size, p1, p2):
    """Exact integral of piecewise-constant img (nx,ny,nz) along p1->p2 (mm)."""
    img = np.asarray(img, np.float64)
    shape = np.array(img.shape)
    center = np.asarray(center, np.float64)
    size = np.asarray(size, np.float64)
    p1 = np.asarray(p1, np.float64)
    p2 = np.asarray(p2, np.float64)
    d = size / shape
    lo = center - size / 2
    D = p2 - p1
    alphas = [np.array([0.0, 1.0])]
    for ax in range(3):
        if D[ax] != 0:
            planes = lo[ax] + d[ax] * np.arange(shape[ax] + 1)
            a = (planes - p1[ax]) / D[ax]
            alphas.append(a[(a > 0) & (a < 1)])
    al = np.unique(np.concatenate(alphas))
    mids = 0.5 * (al[:-1] + al[1:])
    pts = p1[None, :] + mids[:, None] * D[None, :]
    idx = np.floor((pts - lo) / d).astype(np.int64)
    inside = np.all((idx >= 0) & (idx < shape), axis=1)
    seg = np.diff(al) * np.linalg.norm(D)
    idx = idx[inside]
    return float(np.sum(img[idx[:, 0], idx[:, 1], idx[:, 2]] * seg[inside]))


def chord_in_box(center, size, p1, p2):
    """Length (mm) of the part of segment p1->p2 inside the box."""
    center = np.asarray(center, np.float64)
    size = np.asarray(size, np.float64)
    p1 = np.asarray(p1, np.float64)
    p2 = np.asarray(p2, np.float64)
    lo, hi = center - size / 2, center + size / 2
    D = p2 - p1
    t0, t1 = 0.0, 1.0
    for ax in range(3):
        if D[ax] == 0:
            if not (lo[ax] <= p1[ax] <= hi[ax]):
                return 0.0
            continue
        a, b = (lo[ax] - p1[ax]) / D[ax], (hi[ax] - p1[ax]) / D[ax]
        t0, t1 = max(t0, min(a, b)), min(t1, max(a, b))
    return max(0.0, t1 - t0) * float(np.linalg.norm(D))


def random_cylinder_lors(n, radius, axial, seed, tof_sigma=None, box_half=None):
    """Uniform endpoint pairs on a cylinder, float32 (n,6|7)."""
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0, 2 * np.pi, (n, 2))
    z = rng.uniform(-axial / 2, axial / 2, (n, 2))
    l = np.stack([radius * np.cos(phi[:, 0]), radius * np.sin(phi[:, 0]), z[:, 0],
                  radius * np.cos(phi[:, 1]), radius * np.sin(phi[:, 1]), z[:, 1]], 1)
    if tof_sigma is not None:
        # tof = d(p2) - d(p1) for an emission point uniform on the chord, plus noise
        t = rng.uniform(0.2, 0.8, n)
        L = np.linalg.norm(l[:, 3:6] - l[:, 0:3], axis=1)
        tof = L * (1 - 2 * t) + rng.normal(0, 2 * tof_sigma, n)
        l = np.concatenate([l, tof[:, None]], 1)
    return l.astype(np.float32)
