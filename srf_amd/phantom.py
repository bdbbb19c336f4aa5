"""Seeded synthetic phantoms and list-mode events (torch; CPU or GPU).

Emission model (shared by all workloads):
  * events are distributed over voxels in proportion to activity (inverse-CDF
    sampling of the activity image, so N is exact);
  * the emission point is uniform inside its voxel;
  * the photon pair direction is isotropic; both ends are intersected with the
    detector cylinder x^2 + y^2 = R^2;
  * events with an end outside the axial length are rejected and redrawn;
  * TOF value tof = |e - p2| - |e - p1| + N(0, (2 sigma)^2): a path difference
    in mm, the sign convention of srfnef/cc/siddon_tof/siddon_tof.cu.cc:44
    (ratio1 = (1 - tof/L)/2 is the emission point's fraction from p1), and
    sigma = FWHM / 2.3548 the position resolution, tof_sigma2 = sigma^2.
"""
import math

import numpy as np
import torch


def voxel_centers(shape, center, size, device="cpu"):
    """Voxel-centre coordinates (mm) of an (nx, ny, nz) grid, each (n,) float64."""
    out = []
    for n, c, s in zip(shape, center, size):
        d = s / n
        out.append(torch.arange(n, dtype=torch.float64, device=device).mul(d).add(c - s / 2 + d / 2))
    return out


def ellipsoid_mask(shape, center, size, c, axes, device="cpu"):
    x, y, z = voxel_centers(shape, center, size, device)
    X = ((x - c[0]) / axes[0]) ** 2
    Y = ((y - c[1]) / axes[1]) ** 2
    Z = ((z - c[2]) / axes[2]) ** 2
    return (X[:, None, None] + Y[None, :, None] + Z[None, None, :]) <= 1.0


def cylinder_spheres_phantom(shape=(64, 64, 32), voxel=3.4, device="cpu"):
    """Water cylinder r=100 mm (activity 1) with two hot spheres r=10 mm
    (activity 4), centred grid of `voxel` mm.  Returns (activity, center, size)."""
    size = [shape[0] * voxel, shape[1] * voxel, shape[2] * voxel]
    center = [0.0, 0.0, 0.0]
    x, y, z = voxel_centers(shape, center, size, device)
    r2 = x[:, None, None] ** 2 + y[None, :, None] ** 2 + 0 * z[None, None, :]
    act = (r2 <= 100.0 ** 2).to(torch.float32)
    for sc in ((40.0, 0.0, 0.0), (-30.0, 35.0, 10.0)):
        act[ellipsoid_mask(shape, center, size, sc, (10.0, 10.0, 10.0), device)] = 4.0
    return act, center, size


def random_ellipsoid_phantom(shape, voxel, seed, body_axes=(180.0, 120.0, 120.0), n_range=(1, 8),
                             act_range=(0.5, 5.0), size_range=(10.0, 60.0), mu_water=0.0096, device="cpu"):
    """Water ellipsoid body (activity 1, mu_water) holding 1-8 random ellipsoids
    of activity U(0.5, 5) placed fully inside it.  Returns (activity, mu, center, size)."""
    rng = np.random.default_rng(seed)
    size = [shape[0] * voxel, shape[1] * voxel, shape[2] * voxel]
    center = [0.0, 0.0, 0.0]
    ba = [min(b, s / 2 - voxel) for b, s in zip(body_axes, size)]
    body = ellipsoid_mask(shape, center, size, (0, 0, 0), ba, device)
    act = body.to(torch.float32)
    mu = body.to(torch.float32) * mu_water
    for _ in range(int(rng.integers(n_range[0], n_range[1] + 1))):
        axes = rng.uniform(size_range[0], size_range[1], 3) / 2
        for _try in range(100):
            c = rng.uniform(-1, 1, 3) * (np.asarray(ba) - axes)
            # fully inside the body: the ellipsoid's bounding box corners inside
            if np.sum(((np.abs(c) + axes) / np.asarray(ba)) ** 2) <= 1.0:
                break
        act[ellipsoid_mask(shape, center, size, tuple(c), tuple(axes), device)] = float(
            rng.uniform(act_range[0], act_range[1]))
    return act, mu, center, size


MU_WATER = 0.0096
MU_BONE = 0.0172


def scatter_phantom(shape, voxel, seed, n_range=(2, 10), device="cpu"):
    """Emission and mu-map pair for the scatter surrogate: a water ellipsoid
    body of random size and offset (activity 1, mu water) holding 2-10
    ellipsoids, each hot tissue (activity U(0.5, 5), mu water), bone (activity
    U(0.1, 1), mu 0.0172/mm) or air (activity 0, mu 0), placed fully inside
    the body.  Returns (activity, mu, center, size), (nx, ny, nz) float32."""
    rng = np.random.default_rng(seed)
    size = [shape[0] * voxel, shape[1] * voxel, shape[2] * voxel]
    center = [0.0, 0.0, 0.0]
    lim = [s / 2 - voxel for s in size]
    ba = [min(rng.uniform(130.0, 200.0), lim[0]), min(rng.uniform(90.0, 150.0), lim[1]),
          min(rng.uniform(80.0, 130.0), lim[2])]
    bc = [rng.uniform(-1, 1) * min(20.0, lim[0] - ba[0]), rng.uniform(-1, 1) * min(20.0, lim[1] - ba[1]), 0.0]
    body = ellipsoid_mask(shape, center, size, bc, ba, device)
    act = body.to(torch.float32)
    mu = body.to(torch.float32) * MU_WATER
    for _ in range(int(rng.integers(n_range[0], n_range[1] + 1))):
        kind = rng.choice(3, p=[0.6, 0.25, 0.15])
        axes = rng.uniform(10.0, 60.0, 3) / 2 if kind != 2 else rng.uniform(30.0, 100.0, 3) / 2
        axes = np.minimum(axes, 0.8 * np.asarray(ba))
        for _try in range(100):
            c = np.asarray(bc) + rng.uniform(-1, 1, 3) * (np.asarray(ba) - axes)
            if np.sum(((np.abs(c - np.asarray(bc)) + axes) / np.asarray(ba)) ** 2) <= 1.0:
                break
        m = ellipsoid_mask(shape, center, size, tuple(c), tuple(axes), device)
        if kind == 0:
            act[m] = float(rng.uniform(0.5, 5.0))
            mu[m] = MU_WATER
        elif kind == 1:
            act[m] = float(rng.uniform(0.1, 1.0))
            mu[m] = MU_BONE
        else:
            act[m] = 0.0
            mu[m] = 0.0
    return act, mu, center, size


def sample_events(activity, center, size, n, radius, axial_length, seed, tof_fwhm_mm=None, device="cpu",
                  chunk=1 << 23):
    """n list-mode events (n, 6|7) float32 drawn from `activity` (nx,ny,nz).
    tof_fwhm_mm: None for non-TOF rows, else the FWHM of the position resolution."""
    dev = torch.device(device)
    gen = torch.Generator(device=dev).manual_seed(int(seed))
    act = activity.to(device=dev, dtype=torch.float64).reshape(-1)
    cdf = torch.cumsum(act, 0)
    total = float(cdf[-1])
    nx, ny, nz = activity.shape
    d = torch.tensor([size[0] / nx, size[1] / ny, size[2] / nz], dtype=torch.float64, device=dev)
    lo = torch.tensor([center[0] - size[0] / 2, center[1] - size[1] / 2, center[2] - size[2] / 2],
                      dtype=torch.float64, device=dev)
    sigma = None if tof_fwhm_mm is None else tof_fwhm_mm / 2.3548200450309493
    cols = 6 if sigma is None else 7
    out = torch.empty((n, cols), dtype=torch.float32, device=dev)
    filled = 0
    while filled < n:
        m = min(chunk, n - filled)
        m_draw = int(m * 1.5) + 1024
        u = torch.rand(m_draw, generator=gen, dtype=torch.float64, device=dev) * total
        idx = torch.searchsorted(cdf, u).clamp_(max=act.numel() - 1)
        ix = idx // (ny * nz)
        iy = (idx // nz) % ny
        iz = idx % nz
        jit = torch.rand((m_draw, 3), generator=gen, dtype=torch.float64, device=dev)
        e = lo + (torch.stack([ix, iy, iz], 1).to(torch.float64) + jit) * d
        v = torch.randn((m_draw, 3), generator=gen, dtype=torch.float64, device=dev)
        v = v / v.norm(dim=1, keepdim=True)
        a = v[:, 0] ** 2 + v[:, 1] ** 2
        b = 2 * (e[:, 0] * v[:, 0] + e[:, 1] * v[:, 1])
        c = e[:, 0] ** 2 + e[:, 1] ** 2 - radius ** 2
        disc = torch.sqrt(torch.clamp(b * b - 4 * a * c, min=0.0))
        a_safe = torch.where(a > 1e-12, a, torch.ones_like(a))
        t1 = (-b - disc) / (2 * a_safe)
        t2 = (-b + disc) / (2 * a_safe)
        p1 = e + t1[:, None] * v
        p2 = e + t2[:, None] * v
        ok = (a > 1e-12) & (p1[:, 2].abs() <= axial_length / 2) & (p2[:, 2].abs() <= axial_length / 2)
        sel = torch.nonzero(ok).squeeze(1)[:m]
        k = sel.numel()
        rows = torch.cat([p1[sel], p2[sel]], 1)
        if sigma is not None:
            d1 = (e[sel] - p1[sel]).norm(dim=1)
            d2 = (e[sel] - p2[sel]).norm(dim=1)
            noise = torch.randn(k, generator=gen, dtype=torch.float64, device=dev) * (2 * sigma)
            rows = torch.cat([rows, (d2 - d1 + noise)[:, None]], 1)
        out[filled:filled + k] = rows.to(torch.float32)
        filled += k
    return out


def tof_sigma2_from_fwhm(fwhm_mm):
    s = fwhm_mm / 2.3548200450309493
    return s * s


def smooth_lognormal_images(batch, shape, sigma_vox, seed, device="cpu"):
    """`batch` images exp(N(0,1)) smoothed by a separable Gaussian of sigma_vox voxels."""
    gen = torch.Generator(device=device).manual_seed(int(seed))
    x = torch.exp(torch.randn((batch,) + tuple(shape), generator=gen, device=device, dtype=torch.float32))
    r = int(math.ceil(3 * sigma_vox))
    t = torch.arange(-r, r + 1, dtype=torch.float32, device=device)
    k = torch.exp(-0.5 * (t / sigma_vox) ** 2)
    k = k / k.sum()
    for dim in (1, 2, 3):
        x = x.movedim(dim, -1)
        sh = x.shape
        y = torch.nn.functional.conv1d(x.reshape(-1, 1, sh[-1]), k.view(1, 1, -1), padding=r)
        x = y.reshape(sh).movedim(-1, dim)
    return x.contiguous()


__all__ = ["voxel_centers", "ellipsoid_mask", "cylinder_spheres_phantom", "random_ellipsoid_phantom",
           "sample_events", "tof_sigma2_from_fwhm", "smooth_lognormal_images"]
