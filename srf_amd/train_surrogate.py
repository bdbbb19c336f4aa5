"""Train the scatter surrogate CNN on (emission, mu) -> SSS sinogram pairs that
the GPU SSS kernel generates, evaluate it on held-out phantoms, save weights.

    python -m srf_amd.train_surrogate --train 2000 --test 200 --epochs 40 \\
        --out srf_amd/weights/scatter_unet_128x128x64.npz

Workload: emission and mu-map 128 x 128 x 64 with 4 mm voxels
(phantom.scatter_phantom: 2-10 ellipsoids, water 0.0096/mm, bone 0.0172/mm,
air 0), scatter points on an 8 mm grid, 400-650 keV window, coarse sinogram 32
rings x 64 angles x 64 radial bins (SssGeometry defaults).  Training seeds are
0..train-1; held-out seeds start at HELD_OUT_SEED0.
"""
import argparse
import os
import time

import numpy as np
import torch

from . import phantom
from .sss import SingleScatterSimulation, SssGeometry
from .surrogate import ScatterSurrogate, SinogramFeatures, UNet, relative_l2, train

SHAPE = (128, 128, 64)
VOXEL = 4.0
HELD_OUT_SEED0 = 1_000_000


def _log(*a):
    print(*a, flush=True)


def workload(device, coarse_spacing=None):
    g = SssGeometry()
    size = [SHAPE[0] * VOXEL, SHAPE[1] * VOXEL, SHAPE[2] * VOXEL]
    center = [0.0, 0.0, 0.0]
    sss = SingleScatterSimulation(g, SHAPE, center, size)
    feat = SinogramFeatures(g, SHAPE, center, size, device, coarse_spacing=coarse_spacing)
    return g, center, size, sss, feat


def varied_case(seed, device, axial_fov=300.0):
    """An (emission, mu) pair like the ones met inside a batched OSEM loop:
    a random grid (in-plane {96,128,192,256}, 64-128 slices, isotropic 2-4 mm
    voxels), the OSEM workload's random ellipsoid phantom or a scatter
    phantom, and an emission that looks like an early OSEM estimate (voxel
    noise of random strength, usually no activity beyond the scanner's axial
    field of view).  Returns (emission, mu, shape, center, size)."""
    rng = np.random.default_rng(seed)
    nxy = int(rng.choice((96, 128, 192, 256)))
    nz = int(rng.integers(64, 129))
    vox = float(rng.uniform(2.0, 4.0))
    shape = (nxy, nxy, nz)
    if rng.uniform() < 0.5:
        act, mu, center, size = phantom.random_ellipsoid_phantom(shape, vox, int(seed), device=device)
    else:
        act, mu, center, size = phantom.scatter_phantom(shape, vox, int(seed), device=device)
    gen = torch.Generator(device=device).manual_seed(int(seed))
    sigma = float(rng.uniform(0.0, 0.8))
    act = torch.clamp(act * (1.0 + sigma * torch.randn(act.shape, generator=gen, device=device)), min=0.0)
    if rng.uniform() < 0.7:
        z = (torch.arange(nz, device=device, dtype=torch.float32) + 0.5) * vox - size[2] / 2
        act = act * (z.abs() <= axial_fov / 2).to(act.dtype)[None, None, :]
    return act, mu, shape, center, size


def osem_cases(seed, device, n_iter=3):
    """(emission, mu, shape, center, size) pairs as met inside a batched OSEM
    loop: a random grid as in varied_case, the OSEM workload's random
    ellipsoid phantom or a scatter phantom, TOF events (log-uniform 1e5 to
    4e6, FWHM 60 mm) with attenuation-corrected values, and the OSEM
    estimates after iterations 1..n_iter (8 subsets) restricted to the body
    (mu > 0), as pipeline.ScatterEstimator sees them."""
    import math
    from .batch import BatchOsem, Scan
    from .corrections import attenuation_factors
    from .data import Umap
    from .functions import uniform_cylinder_lors
    from . import ops
    from .pipeline import AXIAL, FWHM, RADIUS, TOF_BIN, ZERO_GUARD, body_estimate
    rng = np.random.default_rng(seed)
    nxy = int(rng.choice((96, 128, 192, 256)))
    nz = int(rng.integers(64, 129))
    vox = float(rng.uniform(2.0, 4.0))
    shape = (nxy, nxy, nz)
    if rng.uniform() < 0.5:
        act, mu, center, size = phantom.random_ellipsoid_phantom(shape, vox, int(seed), device=device)
    else:
        act, mu, center, size = phantom.scatter_phantom(shape, vox, int(seed), device=device)
    n = int(round(math.exp(rng.uniform(math.log(1e5), math.log(4e6)))))
    lors = phantom.sample_events(act, center, size, n, RADIUS, AXIAL, int(seed) + 7, tof_fwhm_mm=FWHM, device=device)
    inv_att = (1.0 / attenuation_factors(Umap(mu, center, size), lors, physical=True)).contiguous()
    n_e = 500_000
    el = uniform_cylinder_lors(n_e, RADIUS, AXIAL, int(seed) + 11, device)
    emap = ops.from_kernel_layout(ops.backproject(el, None, list(shape), list(center) + list(size), False, 0.0, 0.0,
                                                  0)) / n_e
    b = BatchOsem([Scan(lors, shape, center, size, emap, values=inv_att, seed=int(seed))], 8, TOF_BIN,
                  phantom.tof_sigma2_from_fwhm(FWHM), True, zero_guard=ZERO_GUARD)
    out = []
    for _ in range(n_iter):
        b.iteration()
        out.append((body_estimate(b.image(0), mu, vox), mu, shape, center, size))
    b.close()
    return out


def make_osem_pairs(n_pairs, seed0, sss_geometry, device, log=_log, n_iter=3, coarse_spacing=None):
    """n_pairs in-loop pairs (n_iter per simulated scan): features, cell-
    normalised SSS sinograms and emission means, as make_pairs returns."""
    from .surrogate import SinogramFeatures as SF
    V = sss_geometry.n_rings * sss_geometry.n_angles * sss_geometry.n_radial
    X = torch.empty((n_pairs, V, 8), dtype=torch.int16, device=device)
    S = torch.empty((n_pairs, V), dtype=torch.float32, device=device)
    M = torch.empty(n_pairs, dtype=torch.float32, device=device)
    i, k, t0 = 0, 0, time.time()
    while i < n_pairs:
        for em, mu, shape, center, size in osem_cases(seed0 + k, device, n_iter):
            if i >= n_pairs:
                break
            s_i = SingleScatterSimulation(sss_geometry, shape, center, size, cell_normalized=True)
            f_i = SF(sss_geometry, shape, center, size, device, coarse_spacing=coarse_spacing)
            S[i] = s_i(em, mu).reshape(-1)
            x, m = f_i(em, mu)
            X[i] = x
            M[i] = m[0]
            i += 1
            if i % 200 == 0 or i == n_pairs:
                torch.cuda.synchronize()
                log(f"in-loop pairs {i}/{n_pairs} {time.time() - t0:.0f}s")
        k += 1
    return X, S, M


def make_pairs(seeds, sss, feat, device, log=_log, varied=None, coarse_spacing=None):
    """Features X (n, V, 8) bf16 bits, SSS sinograms S (n, V), emission means
    M (n,).  varied: None (all on the surrogate grid) or a per-seed flag list
    selecting varied_case pairs, whose SSS is cell-volume normalised (equal
    to the plain sum on the surrogate grid)."""
    n, V = len(seeds), feat.n
    X = torch.empty((n, V, 8), dtype=torch.int16, device=device)
    S = torch.empty((n, V), dtype=torch.float32, device=device)
    M = torch.empty(n, dtype=torch.float32, device=device)
    g = sss.g
    t0 = time.time()
    for i, seed in enumerate(seeds):
        if varied is not None and varied[i]:
            act, mu, shape, center, size = varied_case(int(seed), device)
            s_i = SingleScatterSimulation(g, shape, center, size, cell_normalized=True)
            f_i = SinogramFeatures(g, shape, center, size, device, coarse_spacing=coarse_spacing)
            S[i] = s_i(act, mu).reshape(-1)
            x, m = f_i(act, mu)
        else:
            act, mu, _, _ = phantom.scatter_phantom(SHAPE, VOXEL, int(seed), device=device)
            S[i] = sss(act, mu).reshape(-1)
            x, m = feat(act, mu)
        X[i] = x
        M[i] = m[0]
        if (i + 1) % 200 == 0 or i + 1 == n:
            torch.cuda.synchronize()
            log(f"pairs {i + 1}/{n} {time.time() - t0:.0f}s")
    return X, S, M


def evaluate(net, X, S, M, kappa, dims):
    errs = []
    for i in range(X.shape[0]):
        out = net.forward(X[i], dims)
        errs.append(relative_l2(out * (M[i] / kappa), S[i]))
    return np.asarray(errs)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000)
    ap.add_argument("--test", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--time-limit", type=float, default=None, help="seconds of training before stopping")
    ap.add_argument("--varied", type=float, default=0.0,
                    help="fraction of training pairs drawn by varied_case (random grids, noisy emission)")
    ap.add_argument("--test-varied", type=int, default=0, help="held-out varied_case pairs")
    ap.add_argument("--in-loop", action="store_true",
                    help="train on OSEM estimates of simulated scans (osem_cases) instead of phantoms")
    ap.add_argument("--test-in-loop", type=int, default=0, help="held-out in-loop pairs")
    ap.add_argument("--coarse-spacing", type=float, default=0.0,
                    help="> 0: hybrid surrogate, an SSS on this scatter-point spacing (mm) as a fifth input")
    ap.add_argument("--in-loop-fraction", type=float, default=1.0,
                    help="with --in-loop: fraction of training pairs from OSEM estimates, the rest phantoms")
    ap.add_argument("--out", default=os.path.join(os.path.dirname(__file__), "weights",
                                                  "scatter_unet_128x128x64.npz"))
    a = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    cs = a.coarse_spacing or None
    g, center, size, sss, feat = workload(dev, cs)
    dims = feat.dims
    _log(f"generating {a.train} training and {a.test} held-out pairs")
    rng = np.random.default_rng(a.seed + 17)
    flags = list(rng.uniform(size=a.train) < a.varied) if a.varied > 0 else None
    if a.in_loop:
        n_loop = int(round(a.train * a.in_loop_fraction))
        Xtr, Str, Mtr = make_osem_pairs(n_loop, 0, g, dev, coarse_spacing=cs)
        if n_loop < a.train:
            X2, S2, M2 = make_pairs(range(a.train - n_loop), sss, feat, dev, varied=flags, coarse_spacing=cs)
            Xtr, Str, Mtr = torch.cat([Xtr, X2]), torch.cat([Str, S2]), torch.cat([Mtr, M2])
            del X2, S2, M2
    else:
        Xtr, Str, Mtr = make_pairs(range(a.train), sss, feat, dev, varied=flags, coarse_spacing=cs)
    Xte, Ste, Mte = make_pairs(range(HELD_OUT_SEED0, HELD_OUT_SEED0 + a.test), sss, feat, dev, coarse_spacing=cs)
    Xtl = None
    if a.test_in_loop:
        Xtl, Stl, Mtl = make_osem_pairs(a.test_in_loop, HELD_OUT_SEED0, g, dev, coarse_spacing=cs)
    Xtv = None
    if a.test_varied:
        vs = range(HELD_OUT_SEED0 + 500_000, HELD_OUT_SEED0 + 500_000 + a.test_varied)
        Xtv, Stv, Mtv = make_pairs(vs, sss, feat, dev, varied=[True] * a.test_varied)
    # kappa puts the normalised targets at unit mean
    kappa = float(1.0 / (Str.mean(1) / Mtr).mean())
    Ttr = (Str / Mtr[:, None] * kappa).contiguous()
    net = UNet(in_ch=5 if cs else 4, seed=a.seed, device=dev)
    t0 = time.time()
    hist = train(net, Xtr, Ttr, dims, a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=_log,
                 time_limit=a.time_limit, relative=bool(cs))
    t_train = time.time() - t0
    err_tr = evaluate(net, Xtr[:200], Str[:200], Mtr[:200], kappa, dims)
    err = evaluate(net, Xte, Ste, Mte, kappa, dims)
    _log(f"train rel L2 (first 200): mean {err_tr.mean():.4f}")
    _log(f"held-out rel L2: mean {err.mean():.4f} median {np.median(err):.4f} p95 {np.percentile(err, 95):.4f} "
         f"max {err.max():.4f} frac<=2% {(err <= 0.02).mean():.3f}")
    extra = {}
    if Xtv is not None:
        ev = evaluate(net, Xtv, Stv, Mtv, kappa, dims)
        _log(f"held-out varied rel L2: mean {ev.mean():.4f} median {np.median(ev):.4f} "
             f"p95 {np.percentile(ev, 95):.4f} max {ev.max():.4f} frac<=2% {(ev <= 0.02).mean():.3f}")
        extra = dict(varied_test=len(ev), varied_mean=float(ev.mean()), varied_p95=float(np.percentile(ev, 95)),
                     varied_max=float(ev.max()), varied_frac_le_2pct=float((ev <= 0.02).mean()))
    if Xtl is not None:
        el = evaluate(net, Xtl, Stl, Mtl, kappa, dims)
        _log(f"held-out in-loop rel L2: mean {el.mean():.4f} median {np.median(el):.4f} "
             f"p95 {np.percentile(el, 95):.4f} max {el.max():.4f} frac<=2% {(el <= 0.02).mean():.3f}")
        extra.update(in_loop_test=len(el), in_loop_mean=float(el.mean()), in_loop_p95=float(np.percentile(el, 95)),
                     in_loop_max=float(el.max()), in_loop_frac_le_2pct=float((el <= 0.02).mean()))
    # timing: surrogate (features + network) vs SSS on one held-out phantom
    act, mu, _, _ = phantom.scatter_phantom(SHAPE, VOXEL, HELD_OUT_SEED0, device=dev)
    sur = ScatterSurrogate(net, kappa, g, SHAPE, center, size, check_rings=(), coarse_spacing=cs)
    for _ in range(2):
        sur.predict(act, mu)
        sss(act, mu)
    torch.cuda.synchronize()
    t = time.time()
    for _ in range(10):
        sur.predict(act, mu)
    torch.cuda.synchronize()
    t_sur = (time.time() - t) / 10
    t = time.time()
    for _ in range(3):
        sss(act, mu)
    torch.cuda.synchronize()
    t_sss = (time.time() - t) / 3
    _log(f"time per estimate: surrogate {1e3 * t_sur:.2f} ms, SSS {1e3 * t_sss:.1f} ms, ratio {t_sss / t_sur:.1f}x")
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    net.save(a.out, kappa=kappa, shape=list(SHAPE), voxel=VOXEL, train=a.train, test=a.test,
             epochs=len(hist), batch=a.batch, lr=a.lr, train_seconds=round(t_train, 1),
             final_loss=float(hist[-1]), heldout_mean=float(err.mean()), heldout_max=float(err.max()),
             heldout_p95=float(np.percentile(err, 95)), varied_train_fraction=a.varied, in_loop=bool(a.in_loop),
             in_loop_fraction=a.in_loop_fraction if a.in_loop else 0.0, coarse_spacing=cs or 0.0, **extra)
    _log(f"saved {a.out}")


if __name__ == "__main__":
    main()
