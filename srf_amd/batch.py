"""Batched TOF list-mode OSEM: many scans, one launch per direction per
OSEM subset, HIP-graph-captured iterations.

The reference reconstructs one scan per OsemFull call
(srfnef/functions/osem_full.py:49-87, one TF op launch per projection); here
every scan keeps OsemEngine's semantics (seeded permutation, contiguous
subsets of n // n_sub events, q = y / (A x + eps), x <- x * A^T q / emap) and
its result is bit-identical to running OsemEngine on that scan alone with the
same plan item size, because the batched kernels run the same per-item device
code and keep every reduction's order (srf_batch_* in include/srf_hip.h).
"""
import ctypes
import warnings
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from . import ops
from ._lib import SrfBatchCfg, SrfBatchScan, check, lib, stream_handle
from .functions import geometric_subset_order, image_from_kernel, image_to_kernel


def scan_permutation(n, seed, device):
    """Seeded event permutation generated on the GPU (torch.randperm): the
    subset order of a scan in both the batched and the per-scan runs."""
    g = torch.Generator(device=device).manual_seed(int(seed))
    return torch.randperm(int(n), generator=g, device=device)


@dataclass(frozen=True)
class BatchConfig:
    """Launch schedule of a batch: plan work-item size (segments), persistent
    grid or one workgroup per item, workgroups per CU, item order."""
    item_records: int = 8192
    persistent: int = 0
    workgroups_per_cu: int = 1
    order: int = 1

    def cfg(self):
        return SrfBatchCfg(int(self.persistent), int(self.workgroups_per_cu), int(self.order))


@dataclass
class Scan:
    """One reconstruction: LORs (N, 7) on the GPU, image grid, sensitivity
    image (x, y, z order), optional event values (default ones)."""
    lors: torch.Tensor
    shape: tuple
    center: tuple
    size: tuple
    emap: object
    values: Optional[torch.Tensor] = None
    seed: int = 0
    meta: dict = field(default_factory=dict)

    @property
    def geom(self):
        return [float(v) for v in self.center] + [float(v) for v in self.size]


class BatchOsem:
    """OSEM state for a list of scans; one srf_batch per subset."""

    def __init__(self, scans: List[Scan], n_sub, tof_bin, tof_sigma2, use_tof=True, config=BatchConfig(),
                 ratio_kind=0, eps=1e-8, zero_guard=0.0, subsets=None, tof_window=0.0):
        self.scans, self.n_sub, self.config = scans, int(n_sub), config
        self.ratio_kind, self.eps, self.zero_guard = int(ratio_kind), float(eps), float(zero_guard)
        self.dev = scans[0].lors.device
        self.x, self.emap, self.y, self.q, self.lors, self.perm, self.n_per_sub = [], [], [], [], [], [], []
        for sc in scans:
            n = sc.lors.shape[0]
            perm = scan_permutation(n, sc.seed, self.dev)
            nps = n // self.n_sub
            # OsemEngine's subsets: the same permutation and in-subset order
            pidx = geometric_subset_order(sc.lors, perm[: nps * self.n_sub], self.n_sub)
            self.lors.append(sc.lors.index_select(0, pidx).contiguous())
            vals = sc.values if sc.values is not None else torch.ones(n, dtype=torch.float32, device=self.dev)
            self.y.append(vals.reshape(-1).to(self.dev, torch.float32).index_select(0, pidx).contiguous())
            self.q.append(torch.empty(nps * self.n_sub, dtype=torch.float32, device=self.dev))
            nx, ny, nz = [int(v) for v in sc.shape]
            self.x.append(torch.ones((nz, ny, nx), dtype=torch.float32, device=self.dev))
            self.emap.append(image_to_kernel(sc.emap).contiguous())
            self.perm.append(pidx)
            self.n_per_sub.append(nps)
        subsets = range(self.n_sub) if subsets is None else subsets
        self.plans, self.batches = {}, {}
        for k in subsets:
            plans = []
            for i, sc in enumerate(scans):
                nps = self.n_per_sub[i]
                plans.append(ops.Plan(self.lors[i][k * nps:(k + 1) * nps], sc.shape, sc.geom, use_tof, tof_bin,
                                      tof_sigma2, item_records=config.item_records, tof_window=tof_window))
            self.plans[k] = plans
            self.batches[k] = self._make_batch(k, plans)
        self.graph = None
        self.graph_error = None    # why capture failed (then iterations run eager)
        self.done = 0              # iterations run so far
        self.values_set = False    # set_values was called: checkpoints carry the values
        self.tof = (bool(use_tof), float(tof_bin), float(tof_sigma2), float(tof_window))

    def _make_batch(self, k, plans):
        n = len(plans)
        P = (ctypes.c_void_p * n)(*[p.handle for p in plans])
        S = (SrfBatchScan * n)()
        for i in range(n):
            off = k * self.n_per_sub[i] * 4
            S[i] = SrfBatchScan(self.x[i].data_ptr(), self.emap[i].data_ptr(), self.y[i].data_ptr() + off,
                                self.q[i].data_ptr() + off)
        h = ctypes.c_void_p()
        cfg = self.config.cfg()
        check(lib().srf_batch_create(P, S, n, ctypes.byref(cfg), ctypes.byref(h), stream_handle(self.dev)),
              "srf_batch_create")
        return h

    def set_schedule(self, config):
        """Rebind the launch schedule (persistent / workgroups per CU / order)
        on the same plans; the item size is fixed by the plans."""
        if config.item_records != self.config.item_records:
            raise ValueError("item_records is fixed when the plans are built")
        torch.cuda.synchronize(self.dev)
        self.graph = None
        for h in self.batches.values():
            if h:
                lib().srf_batch_destroy(h)
        self.config = config
        self.batches = {k: self._make_batch(k, self.plans[k]) for k in self.plans}

    def info(self, k=0):
        ns, nw, gr, ne = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64()
        nb = ctypes.c_size_t()
        check(lib().srf_batch_info(self.batches[k], ctypes.byref(ns), ctypes.byref(nw), ctypes.byref(gr),
                                   ctypes.byref(ne), ctypes.byref(nb)), "srf_batch_info")
        return {"scans": ns.value, "work_items": nw.value, "grid": gr.value, "events": ne.value}

    def step(self, k):
        check(lib().srf_batch_em_step(self.batches[k], self.ratio_kind, self.eps, self.zero_guard,
                                      stream_handle(self.dev)), "srf_batch_em_step")

    def iteration(self):
        for k in sorted(self.batches):
            self.step(k)

    def capture(self):
        """Capture one full iteration (every subset) in a HIP graph.  Capture
        records the launches without running them, so a failed capture leaves
        every estimate as it was; the error is kept in graph_error and raised."""
        torch.cuda.synchronize(self.dev)
        g = torch.cuda.CUDAGraph()
        try:
            with torch.cuda.graph(g):
                self.iteration()
        except RuntimeError as e:
            self.graph = None
            self.graph_error = str(e)
            torch.cuda.synchronize(self.dev)
            raise
        self.graph = g
        return g

    def advance(self, graph=False):
        """One iteration: the captured graph's replay when `graph` (capturing
        it first), eager launches when not, or when capture failed (a warning
        names the reason once; the result is the same either way)."""
        if graph and self.graph is None and self.graph_error is None:
            try:
                self.capture()
            except RuntimeError as e:
                warnings.warn(f"HIP graph capture failed, running eager launches instead: {e}")
        if graph and self.graph is not None:
            self.graph.replay()
        else:
            self.iteration()
        self.done += 1

    def checkpoint_key(self):
        """What a checkpoint of this batch must match: every scan's events and
        grid, OSEM constants, TOF and plan options, and the launch schedule (the
        autotuned item size changes how partial sums are grouped)."""
        return {"kind": "batch_osem", "event_order": "geometric", "n_sub": self.n_sub,
                "ratio_kind": self.ratio_kind, "eps": self.eps,
                "zero_guard": self.zero_guard, "tof": list(self.tof), "config": dict(self.config.__dict__),
                "scans": [[int(sc.lors.shape[0]), int(sc.seed), [int(v) for v in sc.shape], sc.geom]
                          for sc in self.scans]}

    def run(self, n_iter, graph=False, checkpoint=None):
        """n_iter more iterations.  With `checkpoint` (a file path) a checkpoint
        of this batch found there is resumed and one is written after every
        iteration (estimates, and event values once set_values was used)."""
        self.run_to(self.done + int(n_iter), graph, checkpoint)

    def run_to(self, total, graph=False, checkpoint=None):
        from . import checkpoint as ckpt
        key = self.checkpoint_key() if checkpoint else None
        if checkpoint:
            state = ckpt.load(checkpoint, key)
            if state is not None and state["done"] > self.done:
                for x, v in zip(self.x, state["images"]):
                    x.copy_(v.to(self.dev))
                if "values" in state:
                    for y, v in zip(self.y, state["values"]):
                        y.copy_(v.to(self.dev))
                    self.values_set = True
                self.done = int(state["done"])
        while self.done < total:
            self.advance(graph)
            if checkpoint:
                ckpt.save(checkpoint, key, self.done, self.x, self.y if self.values_set else None)

    def reset(self):
        for x in self.x:
            x.fill_(1.0)

    def set_values(self, i, values):
        """Replace scan i's event values (original event order) in place; the
        batch reads them at every step, so a captured graph sees the change."""
        self.y[i].copy_(values.reshape(-1).to(self.dev, torch.float32).index_select(0, self.perm[i]))
        self.values_set = True

    def image(self, i):
        """Scan i's estimate in (x, y, z) order."""
        return image_from_kernel(self.x[i])

    def close(self, trim=True):
        """Free the batches and plans; with `trim`, the plan pool's freed
        memory goes back to the device (srf_plan_pool_trim) for other users."""
        torch.cuda.synchronize(self.dev)
        self.graph = None
        for k, h in list(self.batches.items()):
            if h:
                lib().srf_batch_destroy(h)
        self.batches = {}
        had_plans = bool(self.plans)
        for plans in self.plans.values():
            for p in plans:
                p.close()
        self.plans = {}
        if trim and had_plans:
            ops.trim_plan_pool(0, self.dev)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def osem_single(scan: Scan, n_sub, n_iter, tof_bin, tof_sigma2, use_tof=True, item_records=8192, ratio_kind=0,
                eps=1e-8, zero_guard=0.0, tof_window=0.0):
    """The same reconstruction for one scan with per-scan launches
    (OsemEngine): the reference point for the batch's per-scan equality."""
    from .functions import OsemEngine
    from .data import TofConfig
    n = scan.lors.shape[0]
    vals = scan.values if scan.values is not None else torch.ones(n, dtype=torch.float32, device=scan.lors.device)
    tof = TofConfig(tof_bin, tof_sigma2) if use_tof else None
    eng = OsemEngine(scan.lors, vals, scan.emap, scan.shape, scan.geom, n_sub,
                     scan_permutation(n, scan.seed, scan.lors.device), tof=tof, ratio_kind=ratio_kind, eps=eps,
                     zero_guard=zero_guard, item_records=item_records, tof_window=tof_window)
    eng.run(n_iter * n_sub)
    return eng


__all__ = ["BatchConfig", "Scan", "BatchOsem", "osem_single", "scan_permutation"]
