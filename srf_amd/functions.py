"""Drop-in operators and reconstruction drivers with srfnef's names and call
signatures, running on the MI355X kernels of libsrf_hip.so.

  Project(mode)(image, lors) -> Listmode              srfnef/functions/project.py:21-40
  ProjectTof(mode, tof)(image, lors) -> Listmode      srfnef/functions/project_tof.py:22-47
  BackProject(mode)(listmode, image) -> Image         srfnef/functions/back_project.py:22-49
  BackProjectTof(mode, tof)(listmode, image) -> Image srfnef/functions/back_project_tof.py:21-52
  Mlem(n_iter, emap)(listmode) -> Image               srfnef/functions/mlem.py:23-45
  MlemFull(...)() / MlemFullTof(...)() -> Image       srfnef/functions/mlem_full.py, mlem_full_tof.py
  OsemFull(...)() -> Image                            srfnef/functions/osem_full.py:30-87
  OsemTof(...)() -> Image                             TOF ops in OsemFull's subset scheme
  EmapGenerator(mode, scanner)(image_config) -> Emap  srfnef/functions/emap_generator.py:246-257

``mode``: 'tf-eager' returns host numpy data (the reference's .numpy(),
project.py:36); 'tf' and 'hip' return GPU tensors.  Every mode runs the HIP
kernels -- there is no CPU path.
"""
import math

import numpy as np
import torch

from . import ops
from .checkpoint import tensor_digest
from .data import Emap, Image, ImageConfig, Listmode, Lors, TofConfig, nef_class

_MODES = ("hip", "tf-eager", "tf")


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("srf_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def to_gpu(a, dtype=torch.float32):
    """numpy / tensor -> contiguous GPU tensor (no copy if already there)."""
    if isinstance(a, torch.Tensor):
        t = a
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a)))
    return t.to(device=device(), dtype=dtype).contiguous()


def _out(t, mode):
    if mode not in _MODES:
        raise NotImplementedError(f"mode {mode!r}")
    return t.detach().cpu().numpy() if mode == "tf-eager" else t


def _geom(obj):
    return [float(v) for v in obj.center] + [float(v) for v in obj.size]


def _shape(obj):
    return [int(v) for v in obj.shape]


def image_to_kernel(data):
    """(nx,ny,nz) -> GPU kernel layout (nz,ny,nx) (tf.transpose in project.py:27)."""
    return ops.to_kernel_layout(to_gpu(data))


def image_from_kernel(img_k):
    return ops.from_kernel_layout(img_k)


def _tof_args(tof):
    if tof is None:
        return False, 0.0, 0.0
    return True, float(tof.tof_bin), float(tof.tof_sigma2)


def project_data(image, lors_data, tof=None):
    """GPU projection of an Image along (N,6|7) LOR data; returns a GPU tensor."""
    use, tb, ts = _tof_args(tof)
    return ops.project(to_gpu(lors_data), image_to_kernel(image.data), _shape(image), _geom(image), use, tb, ts)


def backproject_data(values, lors_data, image, tof=None, adjoint=False):
    """GPU backprojection into an image shaped like `image`; returns (nx,ny,nz) GPU tensor."""
    use, tb, ts = _tof_args(tof)
    v = None if values is None else to_gpu(values).reshape(-1)
    img_k = ops.backproject(to_gpu(lors_data), v, _shape(image), _geom(image), use, tb, ts, int(adjoint))
    return image_from_kernel(img_k)


@nef_class
class Project:
    mode: str

    def __call__(self, image: Image, lors: Lors) -> Listmode:
        p = project_data(image, lors.data)
        return Listmode(_out(p, self.mode), lors)


@nef_class
class ProjectTof:
    mode: str
    tof: TofConfig

    def __call__(self, image: Image, lors: Lors) -> Listmode:
        p = project_data(image, lors.data, tof=self.tof or TofConfig())
        return Listmode(_out(p, self.mode), lors)


@nef_class
class BackProject:
    mode: str

    def __call__(self, listmode: Listmode, image: Image) -> Image:
        # the input image only supplies shape/centre/size (siddon.cc:123-128)
        bp = backproject_data(listmode.data, listmode.lors.data, image)
        return image.update(data=_out(bp, self.mode))


class BackProjectTof:
    """A plain class, as in srfnef/functions/back_project_tof.py:21-24."""

    def __init__(self, mode=None, tof=None):
        self.mode = mode
        self.tof = tof

    def __call__(self, listmode: Listmode, image: Image) -> Image:
        bp = backproject_data(listmode.data, listmode.lors.data, image, tof=self.tof or TofConfig())
        return image.update(data=_out(bp, self.mode or "hip"))


# --------------------------------------------------------------- EM engine --

class EmEngine:
    """Device-resident EM state for one scan: LORs, measured values, emap and
    the image estimate x, all in kernel layout on the GPU.

    One sub-iteration = project -> fused ratio+backproject -> multiplicative
    update (srf_em_update).  With ``tiled`` (default) the event list is binned
    once into an srf plan and both directions run the LDS-tiled kernels;
    otherwise the one-event-per-lane kernels run on the raw LOR array."""

    def __init__(self, lors_data, values, emap_data, shape, geom, tof=None, ratio_kind=0, eps=0.0,
                 zero_guard=0.0, tiled=True, plan=True, item_records=8192, tof_window=0.0):
        self.lors = to_gpu(lors_data)
        self.y = to_gpu(values).reshape(-1)
        self.shape = [int(v) for v in shape]
        self.geom = [float(v) for v in geom]
        self.use_tof, self.tof_bin, self.tof_sigma2 = _tof_args(tof)
        self.emap_k = image_to_kernel(emap_data)
        nx, ny, nz = self.shape
        self.x = torch.ones((nz, ny, nx), dtype=torch.float32, device=self.lors.device)
        self.ratio_kind = ratio_kind
        self.eps = float(eps)
        self.zero_guard = float(zero_guard)
        self.tiled = bool(tiled)
        self.item_records = int(item_records)
        self.tof_window = float(tof_window)
        self.plan = self.make_plan(self.lors) if (self.tiled and plan) else None

    def make_plan(self, lors):
        return ops.Plan(lors, self.shape, self.geom, self.use_tof, self.tof_bin, self.tof_sigma2,
                        item_records=self.item_records, tof_window=self.tof_window)

    def forward(self, lors=None, plan=None):
        if plan is None and lors is None:
            plan = self.plan
        if plan is not None:
            return plan.project(self.x)
        lors = self.lors if lors is None else lors
        return ops.project(lors, self.x, self.shape, self.geom, self.use_tof, self.tof_bin, self.tof_sigma2)

    def sub_iteration(self, lors=None, y=None, plan=None):
        if plan is None and lors is None:
            plan = self.plan
        y = self.y if y is None else y
        if plan is not None:
            q = plan.project_ratio(self.x, y, self.ratio_kind, self.eps)
            plan.backproject_update(q, self.x, self.emap_k, self.zero_guard)   # tile sums fused with the update
            return
        else:
            lors = self.lors if lors is None else lors
            p = ops.project(lors, self.x, self.shape, self.geom, self.use_tof, self.tof_bin, self.tof_sigma2)
            bp = ops.backproject_ratio(lors, y, p, self.ratio_kind, self.eps, self.shape, self.geom, self.use_tof,
                                       self.tof_bin, self.tof_sigma2, 0)
        ops.em_update(self.x, bp, self.emap_k, self.zero_guard)

    def image(self):
        return image_from_kernel(self.x)


@nef_class
class Mlem:
    n_iter: int
    emap: Emap

    def __call__(self, listmode: Listmode) -> Image:
        # mlem.py:30-43: x0 = 1, emap 0 -> 1e8, q = y / (p + mean(y)*1e-8)
        y = to_gpu(listmode.data).reshape(-1)
        eps = float(np.float32(float(y.mean().item()) * 1e-8))
        eng = EmEngine(listmode.lors.data, y, self.emap.data, self.emap.shape, _geom(self.emap),
                       ratio_kind=0, eps=eps, zero_guard=1e8)
        for _ in range(self.n_iter):
            eng.sub_iteration()
        return Image(eng.image().cpu().numpy(), list(self.emap.center), list(self.emap.size))


def _default_emap(scanner, image_config):
    """The all-crystal-pair sensitivity image for a crystal scanner (the
    reference's 'block-full', mlem_full.py:39-45, mlem_full_tof.py:39-41); an
    ideal cylinder (no crystals) gets the sampled one."""
    from .scanner import PetCylindricalScanner, PetEcatScanner
    crystals = isinstance(scanner, (PetEcatScanner, PetCylindricalScanner))
    return EmapGenerator("block-full" if crystals else "uniform", scanner)(image_config)


def _mlem_full(self, tof):
    """MlemFull (mlem_full.py:47-99) and MlemFullTof (mlem_full_tof.py:43-99).
    Only the TOF driver runs the scatter-corrected second reconstruction
    (mlem_full_tof.py:71-92); in mlem_full.py that block is commented out
    (:74-93), so MlemFull ignores scatter_corr, as the reference does."""
    if self.emap is None:
        object.__setattr__(self, "emap", _default_emap(self.scanner, self.image_config))
    lm = self.listmode
    if self.atten_corr is not None:
        lm = self.atten_corr(lm)
    # mlem_full.py:68 div_no_nan(y, p); mlem_full_tof.py:64 div_no_nan(y, p + 1e-5)
    eps = 1e-5 if tof is not None else 0.0
    eng = EmEngine(self.listmode.lors.data, lm.data, self.emap.data, self.image_config.shape,
                   _geom(self.image_config), tof=tof, ratio_kind=1, eps=eps, zero_guard=1e8)
    for _ in range(self.n_iter):
        eng.sub_iteration()
    x = Image(eng.image().cpu().numpy(), list(self.image_config.center), list(self.image_config.size))
    if tof is not None and self.scatter_corr is not None:
        assert self.atten_corr is not None
        lm2 = self.scatter_corr(x, self.atten_corr.u_map, self.scanner, self.listmode)
        eng = EmEngine(self.listmode.lors.data, lm2.data, self.emap.data, self.image_config.shape,
                       _geom(self.image_config), tof=tof, ratio_kind=1, eps=eps, zero_guard=1e8)
        for _ in range(self.n_iter):
            eng.sub_iteration()
        x = Image(eng.image().cpu().numpy(), list(self.image_config.center), list(self.image_config.size))
    if self.psf_corr is not None:
        x = self.psf_corr(x)
    return x


@nef_class
class MlemFull:
    n_iter: int
    image_config: ImageConfig
    scanner: object
    listmode: Listmode
    atten_corr: object
    scatter_corr: object
    psf_corr: object
    emap: Emap

    def __call__(self) -> Image:
        return _mlem_full(self, None)


@nef_class
class MlemFullTof:
    n_iter: int
    image_config: ImageConfig
    scanner: object
    listmode: Listmode
    atten_corr: object
    scatter_corr: object
    psf_corr: object
    emap: Emap

    def __call__(self) -> Image:
        return _mlem_full(self, getattr(self.scanner, "tof", None) or TofConfig())


def geometric_subset_order(lors, pidx, n_sub):
    """The subsets of ``pidx`` (n_sub contiguous chunks of event indices) with
    each subset's events ordered by LOR geometry: transaxial angle (128 bins
    over [0, pi)), axial midpoint (16 mm bins), then signed transaxial
    distance from the axis (1 mm bins), a stable sort (srf_event_order).
    Membership, and with it every EM update, is unchanged (the plans' sums do
    not depend on event order); events that cross the same image tile get
    nearby indices, so the tiled kernels' per-event reads and writes (ratios,
    partial sums) touch a few cache lines per tile instead of one per event."""
    from ._lib import check, lib, stream_handle
    idx = pidx.to(torch.long).contiguous()
    n = int(idx.numel())
    if n == 0:
        return idx
    L = lors.contiguous()
    if not L.is_cuda or L.dtype != torch.float32 or L.dim() != 2 or idx.device != L.device:
        raise ValueError("geometric_subset_order: float32 (N, 6|7) LORs and indices on one GPU")
    out = torch.empty_like(idx)
    with torch.cuda.device(L.device):
        check(lib().srf_event_order(L.data_ptr(), int(L.shape[0]), int(L.shape[1]), idx.data_ptr(), n, int(n_sub),
                                    out.data_ptr(), stream_handle(L.device)), "srf_event_order")
    return out


def subset_permutation(n, seed):
    """Seeded event permutation (the reference shuffles with the unseeded
    global RNG, osem_full.py:63-64)."""
    return np.random.default_rng(seed).permutation(n)


class OsemEngine(EmEngine):
    """OSEM over contiguous subsets of a once-permuted, device-resident LOR
    array (no per-sub-iteration host gather or upload, cf. osem_full.py:69-73);
    with ``tiled`` each subset is binned into its own plan at construction."""

    def __init__(self, lors_data, values, emap_data, shape, geom, n_sub, perm, tof=None, ratio_kind=0,
                 eps=1e-8, zero_guard=0.0, tiled=True, item_records=8192, tof_window=0.0, geometric_order=True):
        lors = to_gpu(lors_data)
        y = to_gpu(values).reshape(-1)
        n = lors.shape[0]
        self.n_sub = int(n_sub)
        self.n_per_sub = n // self.n_sub          # osem_full.py:65 (remainder dropped)
        if isinstance(perm, torch.Tensor):
            pidx = perm[: self.n_per_sub * self.n_sub].to(lors.device, torch.long)
        else:
            pidx = torch.as_tensor(np.asarray(perm[: self.n_per_sub * self.n_sub]), device=lors.device,
                                   dtype=torch.long)
        if geometric_order:
            pidx = geometric_subset_order(lors, pidx, self.n_sub)
        self.event_order = "geometric" if geometric_order else "permutation"
        # event_index[j]: the caller's index of this engine's event j (subset
        # k is j in [k n_per_sub, (k + 1) n_per_sub))
        self.event_index = pidx
        super().__init__(lors.index_select(0, pidx).contiguous(), y.index_select(0, pidx).contiguous(),
                         emap_data, shape, geom, tof=tof, ratio_kind=ratio_kind, eps=eps, zero_guard=zero_guard,
                         tiled=tiled, plan=False, item_records=item_records, tof_window=tof_window)
        self._perm_tag = None
        del lors, y
        self.plans = [self.make_plan(self.subset(k)[0]) for k in range(self.n_sub)] if self.tiled else None
        self.done = 0   # sub-iterations run so far (over every run call)

    def subset(self, k):
        a, b = k * self.n_per_sub, (k + 1) * self.n_per_sub
        return self.lors[a:b], self.y[a:b]

    @property
    def perm_tag(self):
        """The event order, named by a cheap fingerprint (checkpoint keys);
        computed on first use, so that construction does not wait for it."""
        if self._perm_tag is None:
            p = self.event_index
            step = max(1, p.numel() // 65536)
            self._perm_tag = f"{int(p.numel())}:{int(p[::step].sum())}:" \
                             f"{tensor_digest(torch.cat([p[:4096], p[-4096:]]))}"
        return self._perm_tag

    def checkpoint_key(self, reference_subset_order):
        """What a checkpoint of this engine must match to be resumed: the
        events (permutation, count, sum of values), grid, TOF, EM constants,
        plan options and subset order."""
        return {"kind": "osem", "n_sub": self.n_sub, "n_per_sub": self.n_per_sub, "perm": self.perm_tag,
                "event_order": self.event_order,
                "values_sum": float(self.y.double().sum()), "shape": list(self.shape), "geom": list(self.geom),
                "tof": [bool(self.use_tof), self.tof_bin, self.tof_sigma2], "ratio_kind": int(self.ratio_kind),
                "eps": self.eps, "zero_guard": self.zero_guard, "tiled": self.tiled,
                "item_records": self.item_records, "tof_window": self.tof_window,
                "reference_subset_order": bool(reference_subset_order)}

    def run(self, n_iter, reference_subset_order=False, checkpoint=None, every=None):
        """n_iter more sub-iterations (run_to(done + n_iter))."""
        self.run_to(self.done + int(n_iter), reference_subset_order, checkpoint, every)

    def run_to(self, total, reference_subset_order=False, checkpoint=None, every=None):
        """Sub-iterations until `total` have been run.  Sub-iteration t (counted
        over all calls) uses subset t % n_sub, or the reference's t // n_sub
        (osem_full.py:68, quirk Q14).  With `checkpoint` (a file path): a
        checkpoint of this workload found there is resumed (estimate and count
        restored, if it is ahead of this engine), and one is written after every
        `every` sub-iterations (default: each full iteration of n_sub, as the
        reference saves each iteration) and at the end."""
        from . import checkpoint as ckpt
        every = self.n_sub if every is None else max(1, int(every))
        key = self.checkpoint_key(reference_subset_order) if checkpoint else None
        if checkpoint:
            state = ckpt.load(checkpoint, key)
            if state is not None and state["done"] > self.done:
                self.x.copy_(state["images"][0].to(self.x.device))
                self.done = int(state["done"])
        while self.done < total:
            t = self.done
            k = (t // self.n_sub) if reference_subset_order else (t % self.n_sub)
            lors, y = self.subset(k)
            if k >= self.n_sub:
                # past the last subset the reference's slice of the permutation
                # is empty (osem_full.py:69): its backprojection is zero and the
                # update x * 0 / emap
                ops.em_update(self.x, torch.zeros_like(self.x), self.emap_k, self.zero_guard)
            elif self.plans is not None:
                self.sub_iteration(y=y, plan=self.plans[k])
            else:
                self.sub_iteration(lors, y)
            self.done += 1
            if checkpoint and (self.done % every == 0 or self.done == total):
                ckpt.save(checkpoint, key, self.done, [self.x])


@nef_class
class OsemFull:
    n_iter: int
    n_sub: int
    image_config: ImageConfig
    scanner: object
    listmode: Listmode
    atten_corr: object
    scatter_corr: object
    norm_corr: bool
    psf_corr: object
    emap: Emap
    image: Image
    seed: int = 0
    reference_subset_order: bool = False

    def __call__(self) -> Image:
        """osem_full.py:49-87.  The event values are those of the listmode
        snapped to crystal centres (PetListmodeTrans, :54) and attenuation
        corrected along the snapped LORs (:56-58); the projections run along
        the original LORs (:70-74).  Deviation: with emap None the reference
        asks EmapGenerator for a 'ring' mode no generator has (:45-47, raises
        NotImplementedError); here the default is MlemFull's all-pairs image."""
        if self.emap is None:
            object.__setattr__(self, "emap", _default_emap(self.scanner, self.image_config))
        from .data_transform import PetListmodeTrans
        lm = PetListmodeTrans(self.scanner)(self.listmode)
        if self.atten_corr is not None:
            lm = self.atten_corr(lm)
        perm = subset_permutation(len(self.listmode), self.seed)
        # osem_full.py:74-77: q = y / (p + 1e-8), emap unguarded
        eng = OsemEngine(self.listmode.lors.data, lm.data, self.emap.data, self.image_config.shape,
                         _geom(self.image_config), self.n_sub, perm, ratio_kind=0, eps=1e-8, zero_guard=0.0)
        eng.run(self.n_iter, self.reference_subset_order)
        x = Image(eng.image().cpu().numpy(), list(self.image_config.center), list(self.image_config.size))
        if self.psf_corr is not None:
            x = self.psf_corr(x)
        return x


@nef_class
class OsemTof:
    """TOF list-mode OSEM: ProjectTof/BackProjectTof in OsemFull's subset loop,
    seeded permutation, subset i % n_sub, MlemFullTof's div_no_nan(y, p+1e-5)
    ratio and 1e8 emap guard."""
    n_iter: int
    n_sub: int
    image_config: ImageConfig
    tof: TofConfig
    listmode: Listmode
    atten_corr: object
    emap: Emap
    seed: int = 0

    def __call__(self) -> Image:
        lm = self.listmode
        if self.atten_corr is not None:
            lm = self.atten_corr(lm)
        perm = subset_permutation(len(self.listmode), self.seed)
        eng = OsemEngine(self.listmode.lors.data, lm.data, self.emap.data, self.image_config.shape,
                         _geom(self.image_config), self.n_sub, perm, tof=self.tof or TofConfig(), ratio_kind=1,
                         eps=1e-5, zero_guard=1e8)
        eng.run(self.n_iter)
        return Image(eng.image().cpu().numpy(), list(self.image_config.center), list(self.image_config.size))


# ----------------------------------------------------------------- scanner --

@nef_class
class CylinderScanner:
    """Ideal cylindrical detector surface (radius, axial length in mm) with an
    optional TofConfig.  Stands in for srfnef's crystal-level scanner classes
    (srfnef/geometry/pet_cylindrical_scanner.py:32-49) on synthetic data."""
    inner_radius: float
    axial_length: float
    tof: TofConfig


def uniform_cylinder_lors(n, radius, axial_length, seed, device_=None):
    """n LORs with both endpoints uniform on the cylinder surface (torch, seeded)."""
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    u = torch.rand((n, 4), generator=g, dtype=torch.float64)
    phi1 = 2 * math.pi * u[:, 0]
    phi2 = 2 * math.pi * u[:, 1]
    z1 = (u[:, 2] - 0.5) * axial_length
    z2 = (u[:, 3] - 0.5) * axial_length
    lors = torch.stack([radius * torch.cos(phi1), radius * torch.sin(phi1), z1,
                        radius * torch.cos(phi2), radius * torch.sin(phi2), z2], 1).to(torch.float32)
    return lors if device_ is None else lors.to(device_)


@nef_class
class EmapGenerator:
    """Sensitivity image = backprojection of ones over a LOR set.

    For a crystal-level scanner (scanner.PetEcatScanner,
    scanner.PetCylindricalScanner) it is srfnef's sum over every ordered
    crystal pair (emap_generator.py:28-119 ECAT, :122-243 cylindrical,
    :246-257 dispatch): crystal_pair_emap.  Every mode gives that exact sum
    ('full', 'block-full', 'rsector-full' and 'auto' enumerate it; the
    reference's 'block' / 'rsector' approximate it by rotating block-pair
    images, whose interpolating rotation op is not part of this path).

    For an ideal cylinder (CylinderScanner) it is the backprojection of a
    seeded sample of uniform endpoint pairs (n_lors, default 4e6), scaled by
    1/n_lors -- the synthetic-data decision recorded for the configs without
    crystals."""
    mode: str
    scanner: object
    n_lors: int = 4_000_000
    seed: int = 12345

    def __call__(self, image_config) -> Emap:
        from .scanner import PetCylindricalScanner, PetEcatScanner
        if isinstance(self.scanner, (PetEcatScanner, PetCylindricalScanner)):
            if self.mode not in ("full", "block", "block-full", "rsector", "rsector-full", "auto"):
                raise NotImplementedError(self.mode)
            emap_k = crystal_pair_emap(self.scanner, image_config)
            return Emap(image_from_kernel(emap_k).cpu().numpy(), list(image_config.center),
                        list(image_config.size))
        sc = self.scanner
        lors = uniform_cylinder_lors(int(self.n_lors), float(sc.inner_radius), float(sc.axial_length), self.seed,
                                     device())
        img = Image(None, list(image_config.center), list(image_config.size))
        bp_k = ops.backproject(lors, None, [int(v) for v in image_config.shape], _geom(image_config), False, 0.0,
                               0.0, 0)
        emap = image_from_kernel(bp_k) / float(self.n_lors)
        del img
        return Emap(emap.cpu().numpy(), list(image_config.center), list(image_config.size))


def crystal_pair_emap(scanner, image_config, chunk_pairs=1 << 24, item_records=8192, stats=None):
    """Sum over every ordered crystal pair (a, b) of the reference
    backprojection of 1 along the LOR from crystal a to crystal b, in the GPU
    kernel layout (nz, ny, nx).

    The crystal-centre table is built once; the pair LORs are generated on the
    GPU chunk by chunk (srf_pair_lors, chunk_pairs pairs at a time), each chunk
    is binned into one tiled plan rebuilt in place (srf_plan_rebuild: no device
    allocation after the first chunk) and backprojected into the accumulator
    (srf_plan_backproject with accumulate).  Self pairs and pairs whose
    transaxial extent is below 10 voxels contribute nothing, as in the
    reference backprojector (siddon.cu.cc:364)."""
    from .scanner import crystal_positions
    from ._lib import check, lib, stream_handle
    dev = device()
    pos = torch.from_numpy(crystal_positions(scanner)).to(dev)
    ncrys = int(pos.shape[0])
    total = ncrys * ncrys
    shape = _shape(image_config)
    geom = _geom(image_config)
    nx, ny, nz = shape
    acc = torch.zeros((nz, ny, nx), dtype=torch.float32, device=dev)
    chunk = int(min(chunk_pairs, total))
    lors = torch.empty((chunk, 6), dtype=torch.float32, device=dev)
    nseg = 0
    plan = None
    with torch.cuda.device(dev):
        try:
            for p0 in range(0, total, chunk):
                n = min(chunk, total - p0)
                check(lib().srf_pair_lors(pos.data_ptr(), ncrys, p0, n, lors.data_ptr(), stream_handle(dev)),
                      "srf_pair_lors")
                # one plan for every chunk: its device memory is reused
                if plan is None:
                    plan = ops.Plan(lors[:n], shape, geom, False, item_records=item_records)
                else:
                    plan.rebuild(lors[:n])
                check(lib().srf_plan_backproject(plan.handle, None, 0, 1, acc.data_ptr(), stream_handle(dev)),
                      "srf_plan_backproject")
                if stats is not None:
                    nseg += plan.info()["segments"]
        finally:
            if plan is not None:
                plan.close()
    if stats is not None:
        stats.update({"crystals": ncrys, "pairs": total, "chunks": (total + chunk - 1) // chunk,
                      "segments": nseg})
    return acc


__all__ = ["Project", "ProjectTof", "BackProject", "BackProjectTof", "Mlem", "MlemFull", "MlemFullTof",
           "OsemFull", "OsemTof", "EmapGenerator", "CylinderScanner", "EmEngine", "OsemEngine",
           "uniform_cylinder_lors", "subset_permutation", "geometric_subset_order", "to_gpu", "project_data",
           "backproject_data", "crystal_pair_emap"]
