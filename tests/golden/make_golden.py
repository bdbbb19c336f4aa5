"""Writes tests/golden/siddon_cases.npz: special-case LOR sets and the oracle's
term digests and projections for them.

    python tests/golden/make_golden.py

The reference ships no golden vectors for this path (its only hot-path tests,
tests/ops/test_project_mixin.py:21 and test_backproject_mixin.py:23-24, assert
an output length and BP(0) == 0 against a fixture that is not in the repo), so
these fixtures come from the pinned CPU oracle (oracle/srf_oracle.c).  Each case
targets one entry of the quirks register:
  ties |xd| == |yd| (Q3), transaxial < 10 voxels (Q5), |zd| > transaxial (Q2),
  endpoints inside the image (Q4), steep rays (Q6), odd nx (Q8), opposite
  slopes (Q22), axis-aligned and plane-grazing rays, random cylinder LORs.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
sys.path.insert(0, os.path.join(HERE, ".."))

import oracle  # noqa: E402
from siddon_ref import random_cylinder_lors  # noqa: E402

GRIDS = {
    # name: (shape, center, size)
    "even": ((32, 28, 16), [0.0, 0.0, 0.0], [64.0, 56.0, 32.0]),
    "odd": ((31, 27, 15), [1.25, -0.75, 0.5], [62.0, 54.0, 30.0]),
    "aniso_z": ((24, 24, 20), [0.0, 0.0, 0.0], [48.0, 48.0, 30.0]),   # dz = 0.75 voxel units
}
TOF = (3.66, 649.2)


def special_lors(rng):
    L = []
    # ties |xd| == |yd| (exactly representable)
    for off in (0.0, 0.25, -3.5, 7.75):
        L.append([-60.0, -60.0 + off, 0.5, 60.0, 60.0 + off, -0.5])
        L.append([60.0, -60.0 + off, 2.5, -60.0, 60.0 + off, -1.5])
    # transaxial just under / over 10 voxels (dx 2 mm -> 20 mm)
    L.append([-9.9, 0.0, -40.0, 9.9, 0.0, 40.0])
    L.append([-10.1, 0.0, -40.0, 10.1, 0.0, 40.0])
    # |zd| > transaxial
    L.append([-15.0, 3.0, -30.0, 15.0, 3.0, 30.0])
    # endpoints inside the image, both directions
    L.append([-20.3, 1.1, 0.7, 17.9, -4.2, 2.1])
    L.append([17.9, -4.2, 2.1, -20.3, 1.1, 0.7])
    L.append([2.2, -25.5, -1.0, -3.3, 21.7, 3.0])
    # axis aligned through centres and on voxel planes
    L.append([-100.0, 1.0, 1.0, 100.0, 1.0, 1.0])
    L.append([-100.0, 0.0, 0.0, 100.0, 0.0, 0.0])
    L.append([3.0, -100.0, -4.0, 3.0, 100.0, -4.0])
    L.append([0.0, 0.0, -100.0, 30.0, 0.0, 100.0])
    # steep in z, both slope signs
    L.append([-40.0, 5.0, -60.0, 40.0, -5.0, 60.0])
    L.append([40.0, 5.0, -60.0, -40.0, 25.0, 60.0])
    # grazing the box faces
    L.append([-100.0, 28.0, 0.0, 100.0, 28.0, 0.0])
    L.append([-100.0, -28.0, 16.0, 100.0, -27.0, 16.0])
    L = np.array(L, np.float32)
    tof = rng.normal(0, 60, len(L)).astype(np.float32)
    return np.concatenate([L, tof[:, None]], 1)


def build_cases(seed=2024):
    rng = np.random.default_rng(seed)
    lors = np.concatenate([special_lors(rng),
                           random_cylinder_lors(1000, 90.0, 60.0, seed, tof_sigma=25.48)], 0)
    lors[:, 6] = np.where(np.isfinite(lors[:, 6]), lors[:, 6], 0)
    out = {"lors": lors}
    for gname, (shape, center, size) in GRIDS.items():
        img = rng.uniform(0.1, 2.0, shape).astype(np.float32)
        vals = rng.uniform(0.5, 1.5, len(lors)).astype(np.float32)
        out[f"{gname}_image"] = img
        out[f"{gname}_values"] = vals
        for tof in (False, True):
            t = TOF if tof else None
            l = lors if tof else np.ascontiguousarray(lors[:, :6])
            tag = f"{gname}_{'tof' if tof else 'nontof'}"
            for var, vname in ((0, "fp"), (1, "bp")):
                h, c = oracle.digest(l, shape, center, size, tof=t, variant=var)
                out[f"{tag}_{vname}_hash"] = h
                out[f"{tag}_{vname}_count"] = c
            out[f"{tag}_proj"] = oracle.project(img, l, center, size, tof=t, f64=True)
            out[f"{tag}_bp"] = oracle.backproject(vals, l, shape, center, size, tof=t, variant=1).astype(np.float32)
            out[f"{tag}_adj"] = oracle.backproject(vals, l, shape, center, size, tof=t, variant=0).astype(np.float32)
    return out


def main():
    oracle.build()
    out = build_cases()
    path = os.path.join(HERE, "siddon_cases.npz")
    np.savez_compressed(path, **out)
    print(path, sum(v.nbytes for v in out.values()), "bytes raw")


if __name__ == "__main__":
    main()
