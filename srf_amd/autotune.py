"""Autotuner for the batched reconstruction's launch schedule.

Candidates (BatchConfig): plan work-item size (segments per workgroup pass),
one workgroup per item or a persistent grid of 1-2 workgroups per CU, and the
work-item order.  Each candidate is timed on one OSEM subset of the actual
batch (plans built with its item size, a warm-up step, then the median of
`reps` timed steps); the fastest wins, ties going to the earlier candidate.
Decisions are cached in a JSON file under a signature of the device and the
workload, so the same workload gets the same schedule without re-timing.  A
cache file that cannot be read (truncated, not JSON, wrong structure) is
ignored with a warning (and replaced at the next save) and the workload is
timed again; an entry that names no
candidate is dropped and re-timed; if timing itself fails, the default
BatchConfig is used.  The file is written atomically (temporary file, then
os.replace), so an interrupted write never leaves a partial cache.

Every candidate computes the same reconstruction: the launch mode and order do
not change any sum, and the item size only changes how the backprojector's
per-tile partials are grouped (float rounding; the per-scan images agree to
about 1e-6).
"""
import json
import math
import os
import time
import warnings

import torch

from .batch import BatchConfig, BatchOsem

CANDIDATES = tuple(BatchConfig(item_records=ir, persistent=p, workgroups_per_cu=w, order=o)
                   for ir in (8192, 4096, 2048)
                   for (p, w) in ((0, 1), (1, 1), (1, 2))
                   for o in (1, 0))


def signature(scans, use_tof, device_name=None):
    """Workload key: device, TOF, scan count, log2 of the event total, and the
    sorted image grids rounded to 32 voxels."""
    if device_name is None:
        device_name = torch.cuda.get_device_name(scans[0].lors.device) if torch.cuda.is_available() else "cpu"
    ev = sum(int(s.lors.shape[0]) for s in scans)
    grids = sorted(tuple(int(math.ceil(v / 32.0)) for v in s.shape) for s in scans)
    return f"{device_name}|tof={int(bool(use_tof))}|scans={len(scans)}|ev2={int(round(math.log2(max(ev, 1))))}|" \
           f"grids={grids}"


def choose(timings, candidates=CANDIDATES):
    """Index of the fastest candidate (first one on ties) from {index: seconds}."""
    best = None
    for i in range(len(candidates)):
        if i in timings and (best is None or timings[i] < timings[best]):
            best = i
    if best is None:
        raise ValueError("no timings")
    return best


class Autotuner:
    def __init__(self, cache_path=None, candidates=CANDIDATES, reps=3):
        self.cache_path = cache_path
        self.candidates = tuple(candidates)
        self.reps = int(reps)
        self.cache = {}
        self.cache_error = None
        if cache_path and os.path.exists(cache_path):
            try:
                with open(cache_path) as f:
                    cache = json.load(f)
                if not isinstance(cache, dict):
                    raise ValueError("top level is not an object")
                self.cache = cache
            except (OSError, ValueError) as e:   # JSONDecodeError is a ValueError
                self.cache_error = f"{type(e).__name__}: {e}"
                warnings.warn(f"autotune cache {cache_path} unreadable ({self.cache_error}); timing again")

    def _save(self):
        if self.cache_path:
            path = os.path.abspath(self.cache_path)
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = f"{path}.tmp{os.getpid()}"
            with open(tmp, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)
            os.replace(tmp, path)

    def _cached(self, key):
        """The cached BatchConfig for key if the entry names a candidate."""
        e = self.cache.get(key)
        try:
            c = BatchConfig(**e["config"])
        except (TypeError, KeyError, ValueError):
            if e is not None:
                warnings.warn(f"autotune cache entry for {key!r} is malformed; timing again")
                self.cache.pop(key, None)
            return None
        if c not in self.candidates:
            warnings.warn(f"autotune cache entry for {key!r} names no candidate ({c}); timing again")
            self.cache.pop(key, None)
            return None
        return c

    def _time_step(self, b):
        b.step(0)
        torch.cuda.synchronize()
        ts = []
        for _ in range(self.reps):
            t = time.perf_counter()
            b.step(0)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return sorted(ts)[len(ts) // 2]

    def _gpu_timings(self, scans, n_sub, tof_bin, tof_sigma2, use_tof, log=None, **kw):
        """Median step time of every candidate on subset 0; the plans of one
        item size are built once and every launch schedule is bound to them."""
        timings = {}
        for ir in dict.fromkeys(c.item_records for c in self.candidates):
            idx = [i for i, c in enumerate(self.candidates) if c.item_records == ir]
            b = BatchOsem(scans, n_sub, tof_bin, tof_sigma2, use_tof, config=self.candidates[idx[0]], subsets=[0],
                          **kw)
            try:
                for i in idx:
                    b.set_schedule(self.candidates[i])
                    timings[i] = float(self._time_step(b))
                    if log:
                        log(f"autotune {self.candidates[i]}: {1e3 * timings[i]:.2f} ms")
            finally:
                b.close()
        return timings

    def tune(self, scans, n_sub, tof_bin, tof_sigma2, use_tof=True, timer=None, log=None, **kw):
        """BatchConfig for this workload (cached).  `timer(cand) -> seconds`
        replaces the GPU timing (tests)."""
        key = signature(scans, use_tof) + "".join(f"|{k}={kw[k]}" for k in sorted(kw))
        cached = self._cached(key)
        if cached is not None:
            return cached
        try:
            if timer is not None:
                timings = {i: float(timer(c)) for i, c in enumerate(self.candidates)}
            else:
                timings = self._gpu_timings(scans, n_sub, tof_bin, tof_sigma2, use_tof, log=log, **kw)
        except (RuntimeError, ValueError) as e:
            warnings.warn(f"autotune timing failed ({type(e).__name__}: {e}); using the default schedule")
            return BatchConfig()
        best = self.candidates[choose(timings, self.candidates)]
        self.cache[key] = {"config": best.__dict__.copy(),
                           "timings_ms": {str(self.candidates[i].__dict__): round(1e3 * v, 4)
                                          for i, v in timings.items()}}
        self._save()
        return best


__all__ = ["Autotuner", "CANDIDATES", "signature", "choose"]
