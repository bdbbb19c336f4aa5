"""Benchmark: TOF list-mode OSEM on one MI355X.

Workload (the headline configuration): scanner radius 424 mm, axial length
300 mm; image 256x256x128, 2 mm voxels; seeded random-ellipsoid phantom in a
water body; 5e7 TOF events (400 ps = 60 mm FWHM); 10 subsets; attenuation from
a mu-map of 0.0096/mm water.  A step is one full OSEM iteration: 10
sub-iterations, each = TOF projection + fused ratio/backprojection + EM update
over 5e6 events, so one step processes every event once forward and once back.

value = N_events * steps / t (events/s, fwd+back), t = wall time of the timed
steps between two device synchronisations.  Setup (event generation,
attenuation correction, sensitivity image) is outside the timed region.

Also reported: the roofline of the dominant kernel (algorithmic bytes per
launch / mean launch time, HIP events on the launch stream) and a CPU baseline
(the C oracle on the host cores over a bounded sample of the same events).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TOF OSEM events/s (fwd+back); scatter-est time, surrogate vs SSS at ≤2% err"
HBM_PEAK = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_threads():
    """CPU threads for the baselines: every core of this process's affinity,
    capped by OMP_NUM_THREADS where the environment sets it (the GPU host's
    CPU share for one GPU's job)."""
    n = len(os.sched_getaffinity(0))
    env = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(env))) if env.isdigit() and int(env) > 0 else n


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--events", type=float, default=5e7)
    ap.add_argument("--subsets", type=int, default=10)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--cpu-sample", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--event-kernels", action="store_true",
                    help="one-event-per-lane kernels instead of the LDS-tiled plans")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a plain device copy of this many bytes (HBM counter calibration)")
    ap.add_argument("--no-scatter", action="store_true", help="skip the surrogate-vs-SSS scatter timing")
    ap.add_argument("--scatter-phantoms", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched many-scan reconstruction")
    ap.add_argument("--batch-scans", type=int, default=32)
    ap.add_argument("--batch-seed", type=int, default=5)
    ap.add_argument("--batch-only", action="store_true", help="only the batched many-scan reconstruction")
    ap.add_argument("--no-emap", action="store_true", help="skip the crystal-pair sensitivity image")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the MLEM (64x64x32, 1e6 LORs) and differentiable-projector (8 x 128x128x64, "
                         "4e6 LORs) GPU / CPU timings")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the live HBM traffic measurement (two rocprofv3 counter passes of one step)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="(run under rocprofv3 by the traffic measurement) a 1 GiB copy, then one OSEM step")
    ap.add_argument("--tof-window", type=float, default=22.0,
                    help="plan TOF window (terms below 2^-w of the event's TOF peak are not binned; 0 = all terms)")
    return ap.parse_args()


def copy_probe(nbytes):
    """One plain device-to-device copy of `nbytes`: its FETCH_SIZE / WRITE_SIZE
    counters calibrate the per-dispatch HBM traffic of the other kernels."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


TRAFFIC_COPY = 1 << 30
TRAFFIC_KERNELS = {"project": "k_fp_tiled", "backproject": "k_bp_tiled"}


def _counter_means(csv_path, counter):
    """(mean per-dispatch value of `counter` for each kernel name, the value
    of the first device-copy dispatch) from a rocprofv3 counter CSV."""
    import csv
    acc, first_copy = {}, None
    with open(csv_path) as f:
        rows = [r for r in csv.DictReader(f) if r.get("Counter_Name") == counter]
    rows.sort(key=lambda r: int(r.get("Dispatch_Id") or 0))
    for r in rows:
        name = r["Kernel_Name"]
        acc.setdefault(name, []).append(float(r["Counter_Value"]))
        if first_copy is None and "copyBuffer" in name:
            first_copy = float(r["Counter_Value"])
    return {k: sum(v) / len(v) for k, v in acc.items()}, first_copy


def measure_traffic(args):
    """HBM bytes per launch of the tiled kernels, measured now: two child runs
    of this script (a 1 GiB device copy, then one OSEM step of the same
    workload) under rocprofv3, one with FETCH_SIZE, one with WRITE_SIZE.  The
    copy calibrates each counter (bytes copied / counter bytes).  None when
    rocprofv3 is absent or a pass fails."""
    import glob
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which("rocprofv3")
    if not prof:
        return None
    child = [sys.executable, os.path.join(ROOT, "bench.py"), "--traffic-probe", "--events", str(args.events),
             "--subsets", str(args.subsets), "--seed", str(args.seed), "--tof-window", str(args.tof_window)]
    res = {}
    with tempfile.TemporaryDirectory() as td:
        for counter in ("FETCH_SIZE", "WRITE_SIZE"):
            d = os.path.join(td, counter)
            cmd = [prof, "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--"] + child
            log(f"traffic: {counter} pass")
            try:
                subprocess.run(cmd, check=True, timeout=300, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                               env=dict(os.environ, TMPDIR=td))
            except (subprocess.SubprocessError, OSError) as e:
                log(f"traffic: {counter} pass failed ({e})")
                return None
            files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
            if not files:
                return None
            means, copy = _counter_means(files[0], counter)
            if not copy:
                return None
            # the probe's copy (the process's first) moves TRAFFIC_COPY bytes each way
            factor = TRAFFIC_COPY / (copy * 1024.0)   # bytes moved per counted byte
            for part, kname in TRAFFIC_KERNELS.items():
                vals = [v for k, v in means.items() if kname in k]
                if vals:
                    res.setdefault(part, {})[counter] = vals[0] * 1024.0 * factor
            res.setdefault("factor", {})[counter] = factor
    out = {part: d["FETCH_SIZE"] + d["WRITE_SIZE"] for part, d in res.items()
           if part != "factor" and "FETCH_SIZE" in d and "WRITE_SIZE" in d}
    out["parts"] = {k: v for k, v in res.items() if k != "factor"}
    out["calibration_factor"] = res.get("factor")
    return out


def main():
    args = parse()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.traffic_probe:
        copy_probe(TRAFFIC_COPY)
        args.steps, args.warmup = 1, 0
        args.no_cpu = args.no_scatter = args.no_batch = args.no_emap = args.no_configs = args.no_traffic = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); multi-GPU runs (events sharded "
              "across GPUs, one plan per share) are not implemented", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    if args.batch_only:
        print(json.dumps({"batch": batch_bench(args)}), flush=True)
        return
    from srf_amd import ops, phantom
    from srf_amd.corrections import attenuation_factors
    from srf_amd.data import TofConfig, Umap
    from srf_amd.functions import OsemEngine, uniform_cylinder_lors

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    n = int(args.events)
    shape, voxel = (256, 256, 128), 2.0
    radius, axial = 424.0, 300.0
    fwhm = 60.0
    tof_bin, tof_sigma2 = 3.66, phantom.tof_sigma2_from_fwhm(fwhm)

    t0 = time.time()
    act, mu, center, size = phantom.random_ellipsoid_phantom(shape, voxel, args.seed, device=dev)
    lors = phantom.sample_events(act, center, size, n, radius, axial, args.seed + 1, tof_fwhm_mm=fwhm,
                                 device=dev)
    log(f"events generated: {n} in {time.time() - t0:.1f}s")
    geom = list(center) + list(size)
    # attenuation correction of the measured values (AttenuationCorrect semantics)
    y = 1.0 / attenuation_factors(Umap(mu, center, size), lors)
    # sensitivity image: backprojection of ones over n uniform cylinder LORs
    emap_k = torch.zeros((shape[2], shape[1], shape[0]), device=dev)
    n_e, chunk = n, 1 << 23
    for s in range(0, n_e, chunk):
        el = uniform_cylinder_lors(min(chunk, n_e - s), radius, axial, 1000 + s, dev)
        emap_k += ops.backproject(el, None, list(shape), geom, False, 0.0, 0.0, 0)
    emap_k /= float(n_e)
    emap = ops.from_kernel_layout(emap_k)
    torch.cuda.synchronize()
    log(f"setup done in {time.time() - t0:.1f}s")

    perm = np.random.default_rng(args.seed).permutation(n)
    tp = time.time()
    eng = OsemEngine(lors, y, emap, shape, geom, args.subsets, perm,
                     tof=TofConfig(tof_bin, tof_sigma2),
                     ratio_kind=1, eps=1e-5, zero_guard=1e8, tiled=not args.event_kernels,
                     tof_window=0.0 if args.event_kernels else args.tof_window)
    torch.cuda.synchronize()
    plan_s = time.time() - tp
    plan_info = eng.plans[0].info() if eng.plans else None
    log(f"engine ready in {plan_s:.1f}s; plan of subset 0: {plan_info}")
    del lors, y
    # term counts T_e per event: the reference's (the event kernels' digest) and,
    # with a TOF window, the terms the plans hold (the plan's own digest)
    n_sub_ev = eng.n_per_sub
    l0, _ = eng.subset(0)
    _, cnt_fp = ops.lor_digest(l0, list(shape), geom, True, tof_bin, tof_sigma2, 1)
    _, cnt_bp = ops.lor_digest(l0, list(shape), geom, True, tof_bin, tof_sigma2, 0)
    te_ref_fp = float(cnt_fp.double().mean())
    te_ref_bp = float(cnt_bp.double().mean())
    te_fp, te_bp = te_ref_fp, te_ref_bp
    if eng.plans:
        _, c1 = eng.plans[0].digest(1)
        _, c0 = eng.plans[0].digest(0)
        te_fp, te_bp = float(c1.double().mean()), float(c0.double().mean())
    del cnt_fp, cnt_bp
    log(f"mean terms per event: fp {te_fp:.1f} bp {te_bp:.1f} (reference {te_ref_fp:.1f} / {te_ref_bp:.1f})")

    # per-kernel timing with HIP events on the launch stream (torch's current stream)
    stream = torch.cuda.current_stream()
    ev = []

    def sub_iter(lors_s, y_s, plan, timed):
        # timed: events around each half; on the plan path also around the tiled kernels alone
        if timed:
            a, b, c, fa, fb, ba, bb = (torch.cuda.Event(enable_timing=True) for _ in range(7))
            if plan is not None:
                plan.time_next(fa, fb)
            a.record(stream)
        if plan is not None:
            q = plan.project_ratio(eng.x, y_s, 1, 1e-5)
        else:
            p = ops.project(lors_s, eng.x, eng.shape, eng.geom, True, tof_bin, tof_sigma2)
        if timed:
            b.record(stream)
            if plan is not None:
                plan.time_next(ba, bb)
        if plan is not None:
            # backprojection with its tile sums fused into the EM update
            plan.backproject_update(q, eng.x, eng.emap_k, 1e8)
        else:
            bp = ops.backproject_ratio(lors_s, y_s, p, 1, 1e-5, eng.shape, eng.geom, True, tof_bin, tof_sigma2, 0)
        if timed:
            c.record(stream)
            ev.append((a, b, c, (fa, fb, ba, bb) if plan is not None else None))
        if plan is None:
            ops.em_update(eng.x, bp, eng.emap_k, 1e8)

    def step(timed):
        for k in range(eng.n_sub):
            l_s, y_s = eng.subset(k)
            sub_iter(l_s, y_s, eng.plans[k] if eng.plans else None, timed)

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    log("warmup done")
    t_start = time.perf_counter()
    for i in range(args.steps):
        step(True)
        log(f"step {i + 1}/{args.steps} issued")
    torch.cuda.synchronize()
    t = time.perf_counter() - t_start
    n_used = eng.n_per_sub * eng.n_sub
    n_sub = eng.n_sub
    value = n_used * args.steps / t
    # halves: projection + per-event reduction, backprojection + tile sums / EM update
    fp_half_ms = float(np.mean([a.elapsed_time(b) for a, b, _, _ in ev]))
    bp_half_ms = float(np.mean([b.elapsed_time(c) for _, b, c, _ in ev]))
    if ev and ev[0][3] is not None:
        # the tiled kernels alone, from events the library records around their launches
        fp_ms = float(np.mean([k[0].elapsed_time(k[1]) for _, _, _, k in ev]))
        bp_ms = float(np.mean([k[2].elapsed_time(k[3]) for _, _, _, k in ev]))
    else:
        fp_ms, bp_ms = fp_half_ms, bp_half_ms
    nvox = shape[0] * shape[1] * shape[2]
    # algorithmic bytes per launch: LOR (24) + tof (4) + event values + 4 per term of the reference's
    # algorithm (T_e of the reference term set, as the oracle counts it), + the image once.  Event
    # values: on the plan path the projector reads y and writes q (8) and the backprojector reads q
    # (4); on the event-per-lane path the projector writes p (4) and the backprojector reads y and p
    # (8, ratio fused into its load).  With a TOF window the plans hold fewer terms (te_fp, te_bp):
    # the same rate over the terms actually computed is reported beside it.
    fp_bytes = n_sub_ev * (24 + 4 + (4 if args.event_kernels else 8) + 4 * te_ref_fp) + 4 * nvox
    bp_bytes = n_sub_ev * (24 + 4 + (8 if args.event_kernels else 4) + 4 * te_ref_bp) + 4 * nvox
    fp_bytes_held = n_sub_ev * (24 + 4 + (4 if args.event_kernels else 8) + 4 * te_fp) + 4 * nvox
    bp_bytes_held = n_sub_ev * (24 + 4 + (8 if args.event_kernels else 4) + 4 * te_bp) + 4 * nvox
    fp_bw = fp_bytes / (fp_ms * 1e-3)
    bp_bw = bp_bytes / (bp_ms * 1e-3)
    dom = ("backproject", bp_ms, bp_bw, bp_bytes) if bp_ms >= fp_ms else ("project", fp_ms, fp_bw, fp_bytes)
    roof = {"bound": "hbm", "achieved": dom[2] / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": dom[2] / HBM_PEAK, "traffic": None, "kernel": dom[0], "mean_launch_ms": dom[1],
            "algorithmic_bytes_per_launch": dom[3],
            "other": {"project_ms": fp_ms, "backproject_ms": bp_ms,
                      # whole halves: + per-event reduction / + tile sums and EM update
                      "project_half_ms": fp_half_ms, "backproject_half_ms": bp_half_ms,
                      "frac_over_halves": {"project": fp_bytes / (fp_half_ms * 1e-3) / HBM_PEAK,
                                           "backproject": bp_bytes / (bp_half_ms * 1e-3) / HBM_PEAK},
                      "project_GBps": fp_bw / 1e9,
                      "backproject_GBps": bp_bw / 1e9, "terms_per_event_fp": te_fp, "terms_per_event_bp": te_bp,
                      "terms_per_event_reference_fp": te_ref_fp, "terms_per_event_reference_bp": te_ref_bp,
                      "frac_over_terms_computed": {
                          "project": fp_bytes_held / (fp_ms * 1e-3) / HBM_PEAK,
                          "backproject": bp_bytes_held / (bp_ms * 1e-3) / HBM_PEAK}}}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(eng, shape, center, size, tof_bin, tof_sigma2, args.cpu_sample)
    for p in eng.plans or []:
        p.close()
    del eng
    torch.cuda.empty_cache()
    if args.traffic_probe:
        return
    # traffic: measured now by two counter passes of one step; the committed
    # profile's figure only when no pass can run
    tr = None if (args.no_traffic or args.event_kernels) else measure_traffic(args)
    if tr is not None and dom[0] in tr:
        roof["traffic"] = tr[dom[0]]
        roof["traffic_source"] = "rocprofv3 FETCH_SIZE + WRITE_SIZE passes of this run, copy-calibrated"
        roof["traffic_parts"] = tr
    else:
        tf = os.path.join(ROOT, "profiles", "traffic_tof_osem_256x256x128.json")
        if os.path.exists(tf) and not args.event_kernels:
            try:
                roof["traffic"] = json.load(open(tf)).get(dom[0])
                roof["traffic_source"] = "profiles/traffic_tof_osem_256x256x128.json (no live counter pass)"
            except (OSError, ValueError):
                pass
    scatter = None if args.no_scatter else scatter_bench(args)
    batch = None if args.no_batch else batch_bench(args)
    emap = None if args.no_emap else emap_bench(args)
    configs = None if args.no_configs else config_bench(args)
    # the headline workload end to end over 3 iterations, binning included
    e2e = {"iterations": 3, "binning_s": plan_s, "osem_s": 3 * t / args.steps,
           "events_per_s_with_binning": 3 * n_used / (plan_s + 3 * t / args.steps)}

    out = {"metric": METRIC, "value": value, "unit": "events/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": t / args.steps * 1e3, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "TOF list-mode OSEM, 256x256x128 x 2 mm, 5e7 events, 10 subsets, "
                                  "R 424 mm, axial 300 mm, TOF FWHM 60 mm (400 ps), water mu-map",
                      "events": n_used, "subsets": n_sub, "image": list(shape), "voxel_mm": voxel,
                      "tof_bin_mm": tof_bin, "tof_sigma2_mm2": tof_sigma2, "seed": args.seed,
                      "kernels": "event-per-lane" if args.event_kernels else "lds-tiled plans",
                      "tof_window": 0.0 if args.event_kernels else args.tof_window,
                      "setup_binning_s": plan_s, "plan_subset0": plan_info},
           "roofline": roof, "cpu_baseline": cpu, "end_to_end_3_iterations": e2e, "scatter": scatter,
           "batch": batch, "emap": emap, "configs": configs}
    print(json.dumps(out), flush=True)


WEIGHTS = os.path.join(ROOT, "srf_amd", "weights", "scatter_unet_128x128x64.npz")
# the model trained on OSEM estimates (python -m srf_amd.train_surrogate --in-loop), used inside the loop
WEIGHTS_IN_LOOP = os.path.join(ROOT, "srf_amd", "weights", "scatter_unet_in_loop.npz")


def _timed(fn):
    import torch
    torch.cuda.synchronize()
    t = time.perf_counter()
    r = fn()
    torch.cuda.synchronize()
    return r, time.perf_counter() - t


def emap_bench(args):
    """Sensitivity image over every ordered crystal pair of the mCT geometry
    (48 blocks x 4 rings of 13 x 13 crystals, 1.05e9 pairs) on its 200^3 x
    2.05 mm image: GPU time (first call, which allocates the plan, and a second
    one), and the CPU oracle's backprojection of a bounded sample of the same
    pairs (--no-cpu skips it)."""
    import numpy as np
    import torch
    from srf_amd import scanner as S
    from srf_amd.data import ImageConfig
    from srf_amd.functions import crystal_pair_emap
    from srf_amd.scanner import crystal_positions
    cfg = ImageConfig([200, 200, 200], [0.0, 0.0, 0.0], [410.0, 410.0, 410.0])
    st = {}
    e1, t1 = _timed(lambda: crystal_pair_emap(S.scanner_mct, cfg, stats=st))
    s1 = float(e1.double().sum())
    del e1
    e2, t2 = _timed(lambda: crystal_pair_emap(S.scanner_mct, cfg))
    s2 = float(e2.double().sum())
    del e2
    torch.cuda.empty_cache()
    log(f"emap mCT: {st['pairs']} pairs in {t1:.2f} s / {t2:.2f} s ({st['pairs'] / t2:.3e} pairs/s)")
    out = {"workload": "mCT (48 blocks x 4 rings, 13x13 crystals), all ordered crystal pairs, 200^3 x 2.05 mm",
           "pairs": st["pairs"], "segments": st["segments"], "first_s": t1, "s": t2, "pairs_per_s": st["pairs"] / t2,
           "sum_equal_between_runs": s1 == s2}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        pos = crystal_positions(S.scanner_mct)
        n = pos.shape[0]
        m = 1 << 20   # the first 2^20 pairs: crystals 0-32 against every crystal
        k = np.arange(m)
        lors = np.hstack((pos[k // n], pos[k % n])).astype(np.float32)
        threads = host_threads()
        t = time.perf_counter()
        oracle.backproject(np.ones(m, np.float32), lors, (200, 200, 200), [0.0, 0.0, 0.0], [410.0, 410.0, 410.0],
                           f64=False, threads=threads)
        dt = time.perf_counter() - t
        out["cpu_oracle"] = {"pairs_per_s": m / dt, "cores": threads, "sample": f"first {m} pairs, {dt:.1f} s"}
        log(f"emap CPU oracle: {m / dt:.3e} pairs/s on {threads} threads")
    return out


def scatter_bench(args):
    """Scatter estimate per (emission, mu) pair on the surrogate workload
    (128x128x64, 4 mm voxels, 8 mm scatter grid, 32x64x64 sinogram): SSS kernel
    vs the CNN surrogate on held-out phantoms, relative L2 of the surrogate
    against SSS; plus the surrogate with its one-ring SSS check."""
    import numpy as np
    import torch
    from srf_amd import phantom, train_surrogate as ts
    from srf_amd.surrogate import ScatterSurrogate, UNet, relative_l2
    dev = torch.device("cuda", 0)
    net, meta = UNet.load(WEIGHTS, device=dev)
    g, center, size, sss, feat = ts.workload(dev)
    sur = ScatterSurrogate(net, meta["kappa"], g, ts.SHAPE, center, size, check_rings=())
    chk = ScatterSurrogate(net, meta["kappa"], g, ts.SHAPE, center, size, check_rings=(16,))
    errs, t_sss, t_sur, t_chk, fallbacks = [], [], [], [], 0
    for k in range(args.scatter_phantoms + 1):
        act, mu, _, _ = phantom.scatter_phantom(ts.SHAPE, ts.VOXEL, ts.HELD_OUT_SEED0 + k, device=dev)
        ref, a = _timed(lambda: sss(act, mu))
        pred, b = _timed(lambda: sur.predict(act, mu))
        (est, info), c = _timed(lambda: chk.estimate(act, mu))
        if k == 0:
            continue          # first call: allocations and code-object loading
        errs.append(relative_l2(pred, ref))
        t_sss.append(a)
        t_sur.append(b)
        t_chk.append(c)
        fallbacks += int(info["fallback"])
    log(f"scatter: SSS {1e3 * np.median(t_sss):.1f} ms, surrogate {1e3 * np.median(t_sur):.2f} ms, "
        f"rel L2 mean {np.mean(errs):.4f} max {np.max(errs):.4f}")
    return {"workload": "emission + mu 128x128x64 x 4 mm, 8 mm scatter grid, 400-650 keV, sinogram 32 rings x "
                        "64 angles x 64 radial, 4096 virtual detectors",
            "phantoms": len(errs), "sss_ms": 1e3 * float(np.median(t_sss)),
            "surrogate_ms": 1e3 * float(np.median(t_sur)),
            "surrogate_with_check_ms": 1e3 * float(np.median(t_chk)),
            "speedup": float(np.median(t_sss) / np.median(t_sur)),
            "speedup_with_check": float(np.median(t_sss) / np.median(t_chk)),
            "rel_l2_mean": float(np.mean(errs)), "rel_l2_max": float(np.max(errs)),
            "check_fallbacks": fallbacks, "heldout_200_from_training": {
                k: meta[k] for k in ("heldout_mean", "heldout_p95", "heldout_max") if k in meta}}


def batch_bench(args):
    """Batched reconstructions: 32 scans (events log-uniform 1e5-2e7, in-plane
    96-256, 64-128 slices, 2-4 mm voxels), TOF OSEM 8 subsets x 4 iterations with
    scatter re-estimated before iterations 2-4.  Plain method: per-scan launches
    + SSS.  Batched method: one BatchOsem (autotuned schedule, HIP graph) + the
    CNN surrogate with its SSS check.  The batched images are compared with
    per-scan runs using the same surrogate scatter."""
    import tempfile

    import numpy as np
    import torch
    from srf_amd import phantom
    from srf_amd.autotune import Autotuner
    from srf_amd.pipeline import (FWHM, TOF_BIN, TOF_WINDOW, ScatterEstimator, make_scans, reconstruct_batched,
                                  reconstruct_plain)
    from srf_amd.surrogate import UNet
    dev = torch.device("cuda", 0)
    t0 = time.time()
    scans = make_scans(n_scans=args.batch_scans, seed=args.batch_seed, device=dev)
    n_ev = sum(int(s.lors.shape[0]) for s in scans)
    log(f"batch workload: {len(scans)} scans, {n_ev} events, setup {time.time() - t0:.1f}s")
    wpath = WEIGHTS_IN_LOOP if os.path.exists(WEIGHTS_IN_LOOP) else WEIGHTS
    net, meta = UNet.load(wpath, device=dev)
    kappa = meta["kappa"]
    cache = os.environ.get("SRF_AUTOTUNE_CACHE", os.path.join(tempfile.gettempdir(), "srf_amd_autotune.json"))
    tuner = Autotuner(cache, reps=3)
    s2 = phantom.tof_sigma2_from_fwhm(FWHM)
    (cfg, t_tune) = _timed(lambda: tuner.tune(scans, 8, TOF_BIN, s2, True, tof_window=TOF_WINDOW))
    log(f"autotuned {cfg} in {t_tune:.1f}s")
    cs = meta.get("coarse_spacing") or None

    def estimator():
        return ScatterEstimator(net=net, kappa=kappa, coarse_spacing=cs)

    imgs_plain, t_plain = reconstruct_plain(scans, 8, 4, estimator(), scatter="sss", item_records=cfg.item_records,
                                            log=log)
    torch.cuda.empty_cache()
    # the batched method with the same SSS scatter: its images against the
    # per-scan ones (the 1e-4 bar), so that the surrogate's share is apart
    imgs_bs, t_bs, _ = reconstruct_batched(scans, 8, 4, estimator(), config=cfg, graph=True, log=log, scatter="sss")
    eq_sss = [float((a - b).norm() / b.norm()) for a, b in zip(imgs_bs, imgs_plain)]
    bitwise_sss = all(bool(torch.equal(a, b)) for a, b in zip(imgs_bs, imgs_plain))
    del imgs_bs
    torch.cuda.empty_cache()
    est_b = estimator()
    imgs_b, t_b, info = reconstruct_batched(scans, 8, 4, est_b, config=cfg, graph=True, log=log)
    torch.cuda.empty_cache()
    imgs_p, t_p = reconstruct_plain(scans, 8, 4, estimator(), scatter="surrogate", item_records=cfg.item_records,
                                    log=log)
    eq = [float((a - b).norm() / b.norm()) for a, b in zip(imgs_b, imgs_p)]
    bitwise = all(bool(torch.equal(a, b)) for a, b in zip(imgs_b, imgs_p))
    vs_sss = [float((a - b).norm() / b.norm()) for a, b in zip(imgs_b, imgs_plain)]
    checks = [e for _, _, e in est_b.checks]
    log("surrogate checks (scan/estimate: rel L2 on the check ring): " +
        ", ".join(f"{i}/{k}:{e:.3f}" for i, k, e in est_b.checks))
    return {"workload": "32 scans, events log-uniform 1e5-2e7, in-plane {96,128,192,256}, 64-128 slices, "
                        "2-4 mm voxels, random ellipsoid phantoms, TOF FWHM 60 mm, OSEM 8 subsets x 4 "
                        "iterations, scatter before iterations 2-4",
            "scans": len(scans), "events": n_ev,
            "plain_per_scan_sss_s": t_plain, "batched_graph_surrogate_s": t_b,
            "speedup": t_plain / t_b, "per_scan_surrogate_s": t_p,
            "batched_graph_sss_s": t_bs, "speedup_same_scatter": t_plain / t_bs,
            "image_rel_l2_batched_sss_vs_per_scan_sss_max": max(eq_sss),
            "image_bitwise_equal_sss": bitwise_sss,
            "image_rel_l2_batched_vs_per_scan_max": max(eq), "image_bitwise_equal": bitwise,
            "image_rel_l2_surrogate_vs_sss_scatter_mean": float(np.mean(vs_sss)),
            "image_rel_l2_surrogate_vs_sss_scatter_max": float(np.max(vs_sss)),
            "checks": len(checks), "check_err_median": float(np.median(checks)) if checks else None,
            "check_err_p95": float(np.percentile(checks, 95)) if checks else None,
            "check_err_max": max(checks) if checks else None,
            "estimates_over_tolerance": sum(1 for e in checks if not e <= 0.02),
            "fallback_scans": sum(1 for v in est_b.fallback.values() if v),
            "autotuned": cfg.__dict__, "autotune_s": t_tune, "batch_info": info,
            "surrogate_weights": os.path.basename(wpath), "tof_window": TOF_WINDOW}


def config_bench(args):
    """The other two configurations, GPU and CPU (C oracle, host_threads()):
    non-TOF MLEM (Mlem semantics) of 1e6 LORs on 64x64x32 x 3.4 mm, 10
    iterations, in full on both; the differentiable projector's forward and
    backward (adjoint backprojector) of 4e6 TOF LORs on 128x128x64 x 2 mm, all
    8 images on the GPU, one image in full on the CPU."""
    import numpy as np
    import torch
    from srf_amd import ops, phantom
    from srf_amd.autograd import Projector
    from srf_amd.functions import EmEngine, uniform_cylinder_lors
    dev = torch.device("cuda", 0)
    threads = host_threads()
    out = {}
    # --- MLEM, 64x64x32 x 3.4 mm, 1e6 LORs, 10 iterations
    act, center, size = phantom.cylinder_spheres_phantom((64, 64, 32), 3.4)
    lors = phantom.sample_events(act, center, size, 1_000_000, 300.0, 110.0, seed=21, device=dev)
    elors = phantom.sample_events(torch.ones_like(act), center, size, 1_000_000, 300.0, 110.0, seed=22, device=dev)
    geom = list(center) + list(size)
    emap = ops.from_kernel_layout(ops.backproject(elors, None, [64, 64, 32], geom, False, 0.0, 0.0, 0))
    y = torch.ones(lors.shape[0], device=dev)
    eps = float(np.float32(float(y.mean()) * 1e-8))
    eng = EmEngine(lors, y, emap, (64, 64, 32), geom, ratio_kind=0, eps=eps, zero_guard=1e8)
    eng.sub_iteration()                      # warm-up (code objects); then from x = 1 again
    eng.x.fill_(1.0)
    _, t_gpu = _timed(lambda: [eng.sub_iteration() for _ in range(10)])
    x_gpu = eng.image().cpu().numpy()
    res = {"workload": "non-TOF list-mode MLEM, R 300 mm, axial 110 mm, 64x64x32 x 3.4 mm, water cylinder + 2 "
                       "hot spheres, 1e6 LORs, 10 iterations (Mlem semantics)",
           "gpu_s": t_gpu, "gpu_events_per_s": 1e7 / t_gpu}
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        ln, en = lors.cpu().numpy(), emap.cpu().numpy()
        t = time.perf_counter()
        x_cpu = oracle.mlem(y.cpu().numpy(), ln, en, center, size, 10, driver="mlem", threads=threads)
        dt = time.perf_counter() - t
        m = en > 0
        res.update({"cpu_s": dt, "cpu_events_per_s": 1e7 / dt, "cores": threads, "kind": "port",
                    "image_rel_l2_gpu_vs_cpu": float(np.linalg.norm((x_gpu - x_cpu)[m]) / np.linalg.norm(x_cpu[m]))})
    out["mlem_64x64x32_1e6"] = res
    log(f"configs: MLEM GPU {t_gpu:.3f} s" + (f", CPU {res['cpu_s']:.1f} s" if "cpu_s" in res else ""))
    del eng, lors, elors
    # --- differentiable projector, 8 x 128x128x64 x 2 mm, 4e6 TOF LORs
    n = 4_000_000
    L = uniform_cylinder_lors(n, 300.0, 300.0, 21, dev)
    g = torch.Generator(device=dev).manual_seed(22)
    tofv = torch.randn(n, 1, device=dev, generator=g) * 2 * 25.48
    L = torch.cat([L, tofv], 1).contiguous()
    shape, c3, s3, tof = (128, 128, 64), [0.0, 0.0, 0.0], [256.0, 256.0, 128.0], (3.66, 649.2)
    imgs = phantom.smooth_lognormal_images(8, shape, 2.0, 5, device=dev)
    A, t_plan = _timed(lambda: Projector(L, shape, c3, s3, tof=tof))
    yv = torch.rand(n, device=dev, generator=g)

    def fwd_bwd():
        for b in range(8):
            x = imgs[b].clone().requires_grad_(True)
            (A(x) * yv).sum().backward()

    fwd_bwd()
    _, t_fb = _timed(fwd_bwd)
    res = {"workload": "8 images 128x128x64 x 2 mm (smoothed lognormal), 4e6 TOF LORs with uniform endpoint "
                       "pairs on a 300 mm cylinder, forward (projector) + backward (adjoint backprojector)",
           "gpu_s": t_fb, "gpu_events_per_s": 8 * n / t_fb, "plan_s": t_plan}
    if not args.no_cpu:
        ln = L.cpu().numpy()
        x0 = imgs[0].cpu().numpy()
        t = time.perf_counter()
        p = oracle.project(x0, ln, c3, s3, tof=tof, threads=threads)
        oracle.backproject(yv.cpu().numpy(), ln, shape, c3, s3, tof=tof, variant=oracle.FP, f64=False,
                           threads=threads)
        dt = time.perf_counter() - t
        del p
        res.update({"cpu_s_one_image": dt, "cpu_events_per_s": n / dt, "cores": threads, "kind": "port",
                    "cpu_sample": "image 0 of 8: forward + backward of all 4e6 LORs"})
    out["autograd_8x128x128x64_4e6"] = res
    log(f"configs: projector fwd+bwd GPU {t_fb:.3f} s for 8 images"
        + (f", CPU {res['cpu_s_one_image']:.1f} s for one" if "cpu_s_one_image" in res else ""))
    A.plan.close()
    return out


def cpu_baseline(eng, shape, center, size, tof_bin, tof_sigma2, sample):
    """The C oracle (oracle/srf_oracle.c, OpenMP) on the host cores: TOF
    projection + backprojection of the first `sample` events of subset 0."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = host_threads()
    l0, y0 = eng.subset(0)
    m = min(sample, l0.shape[0])
    lors = l0[:m].cpu().numpy()
    img = eng.image().cpu().numpy()
    log(f"cpu baseline: {m} events, {threads} threads")
    t = time.perf_counter()
    p = oracle.project(img, lors, center, size, tof=(tof_bin, tof_sigma2), threads=threads)
    q = np.where(p + np.float32(1e-5) == 0, 0, y0[:m].cpu().numpy() / (p + np.float32(1e-5))).astype(np.float32)
    oracle.backproject(q, lors, shape, center, size, tof=(tof_bin, tof_sigma2), f64=False, threads=threads)
    dt = time.perf_counter() - t
    return {"value": m / dt, "unit": "events/s", "cores": threads, "kind": "port",
            "sample": f"first {m} events of subset 0 of the same workload: TOF projection + ratio + "
                      f"backprojection, {dt:.1f} s"}


if __name__ == "__main__":
    main()
