"""Learned scatter surrogate: a 3D encoder-decoder CNN that maps (emission,
mu-map) to the coarse SSS sinogram, trained and run by the HIP kernels of
srf_cnn.hip (BF16 storage, MFMA with FP32 accumulation, Adam on FP32 masters).

The reference has no learned part; this stands in for ScatterCorrect's SSS
estimate (srfnef/corrections/scattering/scatter.py:25-55) inside the EM loop,
with plain SSS as the fallback and the reference it is checked against:

  features   the emission and attenuation line integrals along every sinogram
             bin's LOR (the projector over the 131072 bin LORs), assembled per
             bin as [e, exp(-mu), e exp(-mu), mu/4] with e normalised by its
             mean m; the sinogram (rings, angles, radial) is the CNN's 3D grid
  network    U-Net: per level two 3x3x3 convolutions + leaky ReLU (16, 32, 64,
             128 channels), 2x2x2 average pooling down, nearest upsampling and
             skip concatenation up, a 1x1x1 linear head
  output     SSS ~= m / kappa * net(features): scatter is linear in the
             emission, so the network sees and predicts normalised quantities
  check      SSS evaluated on a few rings (their detectors only); when the
             surrogate's relative L2 error there exceeds the tolerance (2%), the
             full SSS is run and returned instead
"""
import math
import time

import numpy as np
import torch

from ._lib import check, lib, stream_handle
from . import ops
from .sss import SingleScatterSimulation, SssGeometry

SLOPE = 0.01


def _c8(n):
    return (n + 7) // 8 * 8


def _c32(n):
    return (n + 31) // 32 * 32


def unet_layers(in_ch, widths):
    """[(name, taps, cin, cout)] in parameter order."""
    L, c = [], in_ch
    for l, w in enumerate(widths):
        L += [(f"enc{l}a", 27, c, w), (f"enc{l}b", 27, w, w)]
        c = w
    for l in range(len(widths) - 2, -1, -1):
        L += [(f"dec{l}a", 27, widths[l + 1] + widths[l], widths[l]), (f"dec{l}b", 27, widths[l], widths[l])]
    L += [("head", 1, widths[0], 1)]
    return L


def bf16_bits_to_float(t):
    """int16 tensor of bf16 bit patterns -> float32 tensor (same device)."""
    return (t.to(torch.int32) << 16).view(torch.float32)


class _Layer:
    def __init__(self, name, taps, cin, cout, dev):
        self.name, self.taps, self.cin, self.cout = name, taps, cin, cout
        self.cin_p, self.cout_p = _c8(cin), _c32(cout)
        self.fcin_p, self.fcout_p = _c8(cout), _c32(cin)      # flipped (input-gradient) kernel
        wshape = (taps, self.cout_p, self.cin_p)
        z = lambda *s, dt=torch.float32: torch.zeros(s, dtype=dt, device=dev)
        self.w, self.gw, self.mw, self.vw = z(*wshape), z(*wshape), z(*wshape), z(*wshape)
        self.wbf = z(*wshape, dt=torch.int16)
        self.wf = z(taps, self.fcout_p, self.fcin_p, dt=torch.int16)
        self.b, self.gb, self.mb, self.vb = z(cout), z(cout), z(cout), z(cout)


class UNet:
    """The surrogate CNN on the GPU.  Every arithmetic step is a HIP kernel of
    libsrf_hip.so; torch provides device memory, zero fills and the stream."""

    def __init__(self, in_ch=4, widths=(16, 32, 64, 128), slope=SLOPE, seed=0, device=None):
        self.dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.in_ch, self.widths, self.slope = in_ch, tuple(widths), float(slope)
        if any(int(v) % 8 for v in self.widths):
            raise ValueError(f"widths {self.widths} must be multiples of 8 (the decoder reads its skip input in place)")
        self.layers = [_Layer(n, t, ci, co, self.dev) for n, t, ci, co in unet_layers(in_ch, widths)]
        self.L = {l.name: l for l in self.layers}
        self.step_count = 0
        self._bufs = {}
        self.init(seed)

    # ------------------------------------------------------------ parameters
    def init(self, seed):
        """He-normal weights (std sqrt(2 / (taps cin))), zero biases, head std 1e-2."""
        rng = np.random.default_rng(seed)
        p = {}
        for l in self.layers:
            std = math.sqrt(2.0 / (l.taps * l.cin)) if l.name != "head" else 1e-2
            p[l.name] = (rng.normal(0.0, std, (l.taps, l.cout, l.cin)), np.zeros(l.cout))
        self.set_params(p)

    def set_params(self, params):
        """params: {name: (w (taps, cout, cin), b (cout,))} (numpy or torch)."""
        for l in self.layers:
            w, b = params[l.name]
            l.w.zero_()
            l.w[:, :l.cout, :l.cin] = torch.as_tensor(np.asarray(w, np.float32)).to(self.dev)
            l.b.copy_(torch.as_tensor(np.asarray(b, np.float32)).to(self.dev))
            for t in (l.mw, l.vw, l.mb, l.vb):
                t.zero_()
        self.step_count = 0
        self._sync_weights()

    def params(self):
        return {l.name: (l.w[:, :l.cout, :l.cin].cpu().numpy().copy(), l.b.cpu().numpy().copy())
                for l in self.layers}

    def grads(self):
        return {l.name: (l.gw[:, :l.cout, :l.cin].cpu().numpy().copy(), l.gb.cpu().numpy().copy())
                for l in self.layers}

    def zero_grad(self):
        for l in self.layers:
            l.gw.zero_()
            l.gb.zero_()

    def _sync_weights(self):
        s = stream_handle(self.dev)
        for l in self.layers:
            check(lib().srf_cnn_cast(l.w.data_ptr(), l.w.numel(), 1, l.wbf.data_ptr(), s), "srf_cnn_cast")
            check(lib().srf_cnn_flip_weights(l.wbf.data_ptr(), l.taps, l.cout, l.cout_p, l.cin_p, l.cin,
                                             l.wf.data_ptr(), l.fcout_p, l.fcin_p, s), "srf_cnn_flip_weights")

    def save(self, path, **extra):
        """Compressed .npz of the inference weights: the bf16 values the
        convolutions use (as float32) and the float32 biases, plus metadata."""
        d = {}
        for l in self.layers:
            d[l.name + ".w"] = bf16_bits_to_float(l.wbf)[:, :l.cout, :l.cin].cpu().numpy().copy()
            d[l.name + ".b"] = l.b.cpu().numpy().copy()
        meta = dict(in_ch=self.in_ch, widths=list(self.widths), slope=self.slope, **extra)
        d["meta"] = np.frombuffer(repr(meta).encode(), dtype=np.uint8)
        np.savez_compressed(path, **d)

    @classmethod
    def load(cls, path, device=None):
        z = np.load(path, allow_pickle=False)
        meta = _parse_meta(bytes(z["meta"]).decode())
        net = cls(meta["in_ch"], tuple(meta["widths"]), meta["slope"], device=device)
        net.set_params({l.name: (z[l.name + ".w"], z[l.name + ".b"]) for l in net.layers})
        return net, meta

    # --------------------------------------------------------------- buffers
    def _buffers(self, D, H, W, N=1):
        """Activations of N samples of D x H x W, stacked along z (V[l] counts
        the voxels of all N at level l; dims are one sample's)."""
        key = (D, H, W, N)
        if key in self._bufs:
            return self._bufs[key]
        n = len(self.widths)
        if D % (1 << (n - 1)) or H % (1 << (n - 1)) or W % (1 << (n - 1)):
            raise ValueError(f"grid {key[:3]} must be divisible by {1 << (n - 1)}")
        dims = [(D >> l, H >> l, W >> l) for l in range(n)]
        V = [N * d[0] * d[1] * d[2] for d in dims]
        bf = lambda *s: torch.zeros(s, dtype=torch.int16, device=self.dev)
        f32 = lambda *s: torch.zeros(s, dtype=torch.float32, device=self.dev)
        w = self.widths
        B = {"dims": dims, "V": V, "N": N}
        for l in range(n):
            B[f"ea{l}"], B[f"e{l}"] = bf(V[l], w[l]), bf(V[l], w[l])
            if l + 1 < n:
                B[f"p{l}"] = bf(V[l + 1], w[l])
                B[f"da{l}"], B[f"d{l}"] = bf(V[l], w[l]), bf(V[l], w[l])
        B["out"] = f32(V[0])
        self._bufs[key] = B
        return B

    def _grad_buffers(self, D, H, W, N=1):
        """_buffers plus the gradient buffers (allocated on the first backward
        or loss of that size, so that inference holds activations only)."""
        B = self._buffers(D, H, W, N)
        if "gout" in B:
            return B
        n, w, V = len(self.widths), self.widths, B["V"]
        bf = lambda *s: torch.zeros(s, dtype=torch.int16, device=self.dev)
        f32 = lambda *s: torch.zeros(s, dtype=torch.float32, device=self.dev)
        for l in range(n):
            B[f"gea{l}"], B[f"ge{l}"] = f32(V[l], w[l]), f32(V[l], w[l])
            cmax = max(_c8(c) for c in [w[l], 1] + ([w[l + 1]] if l + 1 < n else []))
            B[f"dp{l}"] = bf(V[l], max(cmax, 8))
            if l + 1 < n:
                B[f"gp{l}"] = f32(V[l + 1], w[l])
                B[f"u{l}"] = bf(V[l], w[l + 1] + w[l])      # the decoder input, for its weight gradient
                B[f"gu{l}"] = f32(V[l], w[l + 1] + w[l])
                B[f"gda{l}"], B[f"gd{l}"] = f32(V[l], w[l]), f32(V[l], w[l])
        B["gout"] = f32(V[0])
        B["loss"] = f32(1)
        return B

    def release(self):
        """Drop every cached activation and gradient buffer."""
        self._bufs.clear()

    # --------------------------------------------------------------- kernels
    def _conv(self, L, x, dims, N, y_bf=None, y_f32=None, stride=None, off=0, act=True):
        D, H, W = dims
        check(lib().srf_cnn_conv_batch(x.data_ptr(), L.wbf.data_ptr(), L.b.data_ptr(), N, D, H, W, L.cin_p, L.cout,
                                       L.cout_p, L.taps, stride or L.cout, off, self.slope if act else -1.0,
                                       y_bf.data_ptr() if y_bf is not None else None,
                                       y_f32.data_ptr() if y_f32 is not None else None, stream_handle(self.dev)),
              "srf_cnn_conv_batch " + L.name)

    def _back(self, L, x, a, ga, ga_stride, dims, N, dp, gx=None, act=True):
        """Gradients of layer L (input x, output a, dL/da = ga); gx <- dL/dx.
        The weight-gradient kernel also accumulates the bias gradient."""
        D, H, W = dims
        V = N * D * H * W
        s = stream_handle(self.dev)
        check(lib().srf_cnn_act_bwd(ga.data_ptr(), ga_stride, a.data_ptr(), L.cout, V, L.cout, L.fcin_p,
                                    self.slope if act else -1.0, dp.data_ptr(), s), "srf_cnn_act_bwd")
        check(lib().srf_cnn_conv_wgrad_batch(x.data_ptr(), dp.data_ptr(), L.fcin_p, N, D, H, W, L.cin_p, L.cout,
                                             L.cout_p, L.taps, L.gw.data_ptr(), L.gb.data_ptr(), s),
              "srf_cnn_conv_wgrad_batch " + L.name)
        if gx is not None:
            check(lib().srf_cnn_conv_batch(dp.data_ptr(), L.wf.data_ptr(), None, N, D, H, W, L.fcin_p, L.cin,
                                           L.fcout_p, L.taps, L.cin, 0, -1.0, None, gx.data_ptr(), s),
                  "srf_cnn_conv_batch dgrad " + L.name)

    # ------------------------------------------------------------------ run
    def forward(self, x0, dims, n_samples=1):
        """x0: int16 (N V, 8) bf16 features (channels-last) of N = n_samples
        samples stacked along z; dims (D, H, W) of one sample.  Returns the
        float32 (N V,) output buffer (reused by the next call of that N)."""
        N = int(n_samples)
        B = self._buffers(*dims, N)
        n, w, s = len(self.widths), self.widths, stream_handle(self.dev)
        h = x0
        for l in range(n):
            d = B["dims"][l]
            if l > 0:
                pd = B["dims"][l - 1]
                check(lib().srf_cnn_pool(B[f"e{l - 1}"].data_ptr(), N * pd[0], pd[1], pd[2], w[l - 1], w[l - 1],
                                         B[f"p{l - 1}"].data_ptr(), w[l - 1], s), "srf_cnn_pool")
                h = B[f"p{l - 1}"]
            self._conv(self.L[f"enc{l}a"], h, d, N, y_bf=B[f"ea{l}"])
            self._conv(self.L[f"enc{l}b"], B[f"ea{l}"], d, N, y_bf=B[f"e{l}"])
        prev = B[f"e{n - 1}"]
        for l in range(n - 2, -1, -1):
            d = B["dims"][l]
            # dec_a reads [upsample(prev), e_l] in place (no concatenated copy)
            L = self.L[f"dec{l}a"]
            check(lib().srf_cnn_conv_up_batch(B[f"e{l}"].data_ptr(), prev.data_ptr(), w[l + 1], L.wbf.data_ptr(),
                                              L.b.data_ptr(), N, *d, L.cin_p, L.cout, L.cout_p, L.taps, L.cout, 0,
                                              self.slope, B[f"da{l}"].data_ptr(), None, s),
                  "srf_cnn_conv_up_batch " + L.name)
            Lb = self.L[f"dec{l}b"]
            if l == 0 and Lb.cout <= 64:
                # the 1x1 head in dec0b's epilogue
                H_ = self.L["head"]
                check(lib().srf_cnn_conv_head_batch(B["da0"].data_ptr(), Lb.wbf.data_ptr(), Lb.b.data_ptr(), N, *d,
                                                    Lb.cin_p, Lb.cout, Lb.cout_p, Lb.taps, self.slope,
                                                    B["d0"].data_ptr(), H_.wbf.data_ptr(), H_.b.data_ptr(),
                                                    B["out"].data_ptr(), s), "srf_cnn_conv_head_batch " + Lb.name)
                self._x0 = x0
                return B["out"]
            self._conv(Lb, B[f"da{l}"], d, N, y_bf=B[f"d{l}"])
            prev = B[f"d{l}"]
        self._conv(self.L["head"], prev, B["dims"][0], N, y_f32=B["out"], stride=1, act=False)
        self._x0 = x0
        return B["out"]

    def backward(self, gout, dims, n_samples=1):
        """Accumulate parameter gradients for dL/d(output) = gout (float32
        (N V,)) of the last forward of N = n_samples samples."""
        N = int(n_samples)
        B = self._grad_buffers(*dims, N)
        n, w, s = len(self.widths), self.widths, stream_handle(self.dev)
        dp = lambda l: B[f"dp{l}"]
        top = B["d0"] if n > 1 else B["e0"]
        gtop = B["gd0"] if n > 1 else B["ge0"]
        self._back(self.L["head"], top, gout, gout, 1, B["dims"][0], N, dp(0), gx=gtop, act=False)
        for l in range(0, n - 1):
            d, cd = B["dims"][l], B["dims"][l + 1]
            cu = w[l + 1] + w[l]
            # the decoder input the forward read in place, materialised for dW
            coarse = B[f"d{l + 1}"] if l + 1 < n - 1 else B[f"e{n - 1}"]
            check(lib().srf_cnn_upsample(coarse.data_ptr(), N * cd[0], cd[1], cd[2], w[l + 1], w[l + 1],
                                         B[f"u{l}"].data_ptr(), cu, 0, s), "srf_cnn_upsample")
            check(lib().srf_cnn_copy_channels(B[f"e{l}"].data_ptr(), B["V"][l], w[l], w[l], 0, B[f"u{l}"].data_ptr(),
                                              cu, w[l + 1], 0, 0, s), "srf_cnn_copy_channels")
            self._back(self.L[f"dec{l}b"], B[f"da{l}"], B[f"d{l}"], B[f"gd{l}"], w[l], d, N, dp(l),
                       gx=B[f"gda{l}"])
            self._back(self.L[f"dec{l}a"], B[f"u{l}"], B[f"da{l}"], B[f"gda{l}"], w[l], d, N, dp(l),
                       gx=B[f"gu{l}"])
            dst = B[f"gd{l + 1}"] if l + 1 < n - 1 else B[f"ge{n - 1}"]
            check(lib().srf_cnn_upsample_bwd(B[f"gu{l}"].data_ptr(), N * cd[0], cd[1], cd[2], w[l + 1], cu, 0,
                                             dst.data_ptr(), w[l + 1], s), "srf_cnn_upsample_bwd")
        for l in range(n - 1, -1, -1):
            d = B["dims"][l]
            if l < n - 1:
                cu = w[l + 1] + w[l]
                check(lib().srf_cnn_copy_channels(B[f"gu{l}"].data_ptr(), B["V"][l], w[l], cu, w[l + 1],
                                                  B[f"ge{l}"].data_ptr(), w[l], 0, 1, 1, s), "srf_cnn_copy_channels")
            xin = self._x0 if l == 0 else B[f"p{l - 1}"]
            self._back(self.L[f"enc{l}b"], B[f"ea{l}"], B[f"e{l}"], B[f"ge{l}"], w[l], d, N, dp(l),
                       gx=B[f"gea{l}"])
            self._back(self.L[f"enc{l}a"], xin, B[f"ea{l}"], B[f"gea{l}"], w[l], d, N, dp(l),
                       gx=B[f"gp{l - 1}"] if l > 0 else None)
            if l > 0:
                pd = B["dims"][l - 1]
                check(lib().srf_cnn_pool_bwd(B[f"gp{l - 1}"].data_ptr(), N * pd[0], pd[1], pd[2], w[l - 1],
                                             w[l - 1], B[f"ge{l - 1}"].data_ptr(), w[l - 1], s), "srf_cnn_pool_bwd")

    def loss_and_grad(self, out, target, dims, weight=1.0, n_samples=1):
        """Sum over the N = n_samples samples of weight[i] * MSE(out_i,
        target_i) into the loss buffer; returns dL/dout (N V,).  weight: one
        number for every sample or a sequence of N."""
        N = int(n_samples)
        B = self._grad_buffers(*dims, N)
        ws = [float(weight)] * N if np.isscalar(weight) else [float(v) for v in weight]
        if len(ws) != N:
            raise ValueError(f"{len(ws)} weights for {N} samples")
        V = B["V"][0] // N
        s = stream_handle(self.dev)
        for i, wi in enumerate(ws):
            check(lib().srf_cnn_mse(out[i * V:].data_ptr(), target.reshape(-1)[i * V:].data_ptr(), V, wi,
                                    B["gout"][i * V:].data_ptr(), B["loss"].data_ptr(), s), "srf_cnn_mse")
        return B["gout"]

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, gscale=1.0):
        self.step_count += 1
        s = stream_handle(self.dev)
        for l in self.layers:
            check(lib().srf_cnn_adam(l.w.data_ptr(), l.gw.data_ptr(), l.mw.data_ptr(), l.vw.data_ptr(), l.w.numel(),
                                     lr, b1, b2, eps, self.step_count, gscale, l.wbf.data_ptr(), s), "srf_cnn_adam")
            check(lib().srf_cnn_adam(l.b.data_ptr(), l.gb.data_ptr(), l.mb.data_ptr(), l.vb.data_ptr(), l.cout,
                                     lr, b1, b2, eps, self.step_count, gscale, None, s), "srf_cnn_adam")
            check(lib().srf_cnn_flip_weights(l.wbf.data_ptr(), l.taps, l.cout, l.cout_p, l.cin_p, l.cin,
                                             l.wf.data_ptr(), l.fcout_p, l.fcin_p, s), "srf_cnn_flip_weights")


def _parse_meta(text):
    import ast
    return ast.literal_eval(text)


def to_bf16(x):
    """float32 device tensor -> int16 bf16 bit patterns (HIP cast kernel)."""
    x = x.contiguous()
    y = torch.empty(x.shape, dtype=torch.int16, device=x.device)
    check(lib().srf_cnn_cast(x.data_ptr(), x.numel(), 1, y.data_ptr(), stream_handle(x.device)), "srf_cnn_cast")
    return y


# Scale of the coarse-SSS input channel (about the unit-mean normalisation of
# the targets, kappa, of the trained networks).
COARSE_KAPPA = 8.0e10


class SinogramFeatures:
    """Per-bin emission and attenuation line integrals along the coarse
    sinogram's LORs (the bins' detector pairs), assembled into the network's
    8-channel bf16 input by srf_cnn_features.  With coarse_spacing, channel 4
    is an SSS on a coarse scatter-point grid (cell-volume normalised to the
    fine grid's 8 mm cells), scaled like the targets."""

    def __init__(self, geometry, shape, center, size, device=None, coarse_spacing=None):
        self.g = geometry
        self.shape = [int(v) for v in shape]
        self.geom = [float(v) for v in center] + [float(v) for v in size]
        self.dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        det = geometry.detectors(self.dev)
        bins = torch.from_numpy(geometry.bin_detectors().astype(np.int64)).to(self.dev)
        self.lors = torch.cat([det[bins[:, 0], :3], det[bins[:, 1], :3]], 1).contiguous()
        L2 = ((self.lors[:, 3:6] - self.lors[:, 0:3]) ** 2).sum(1)
        dx_ = self.geom[3] / self.shape[0]
        self.to_integral = (L2 / (dx_ * dx_)).contiguous()
        self.dims = (geometry.n_rings, geometry.n_angles, geometry.n_radial)
        self.n = self.lors.shape[0]
        self.coarse = None
        if coarse_spacing:
            self.coarse = SingleScatterSimulation(geometry.update(spacing=float(coarse_spacing)), shape, center, size,
                                                  cell_normalized=True, cell_ref=geometry.spacing ** 3)

    def integrals(self, img_xyz):
        img_k = ops.to_kernel_layout(img_xyz.to(self.dev, torch.float32).contiguous())
        return ops.project(self.lors, img_k, self.shape, self.geom, False, 0.0, 0.0) * self.to_integral

    def _mu_integrals(self, mu_xyz):
        # the mu-map's integrals, kept while the same unmodified tensor comes back
        # (the cache holds a reference, so its address cannot be reused)
        c = getattr(self, "_mu_cache", None)
        if c is not None and c[0] is mu_xyz and c[1] == mu_xyz._version:
            return c[2]
        mu = self.integrals(mu_xyz)
        self._mu_cache = (mu_xyz, mu_xyz._version, mu)
        return mu

    def __call__(self, em_xyz, mu_xyz, out=None):
        """-> (x (n, 8) int16 bf16, m (1,) float32 mean emission integral);
        out: a contiguous (n, 8) int16 tensor to write x into (a slice of a
        batch's input)."""
        em = self.integrals(em_xyz)
        mu = self._mu_integrals(mu_xyz)
        m = em.mean().reshape(1).clamp_min(1e-30)
        scale = torch.cat([1.0 / m, COARSE_KAPPA / m]).contiguous()
        extra = None
        if self.coarse is not None:
            extra = self.coarse(em_xyz.to(self.dev, torch.float32), mu_xyz.to(self.dev, torch.float32))
            extra = extra.reshape(-1).contiguous()
        if out is not None:
            if out.shape != (self.n, 8) or out.dtype != torch.int16 or not out.is_contiguous():
                raise ValueError("out must be a contiguous (n, 8) int16 tensor")
            x = out
        else:
            x = torch.empty((self.n, 8), dtype=torch.int16, device=self.dev)
        check(lib().srf_cnn_features(em.data_ptr(), mu.data_ptr(), extra.data_ptr() if extra is not None else None,
                                     self.n, scale.data_ptr(), x.data_ptr(), stream_handle(self.dev)),
              "srf_cnn_features")
        return x, m


def relative_l2(a, b):
    return float(torch.linalg.vector_norm((a - b).double()) / torch.linalg.vector_norm(b.double()).clamp_min(1e-300))


class ScatterSurrogate:
    """SSS estimate by the trained CNN, checked against SSS on `check_rings`.

    estimate(em, mu) -> (sinogram (n_rings, n_angles, n_radial), info) where
    info = {"fallback": bool, "check_err": float}; on fallback the sinogram is
    the full SSS."""

    def __init__(self, net, kappa, geometry, shape, center, size, check_rings=(5, 16, 26), tol=0.02,
                 physical_kn=False, coarse_spacing=None):
        self.net, self.kappa, self.g, self.tol = net, float(kappa), geometry, float(tol)
        self.features = SinogramFeatures(geometry, shape, center, size, net.dev, coarse_spacing=coarse_spacing)
        self.sss_args = (geometry, shape, center, size)
        self.physical_kn = physical_kn
        self.check_rings = tuple(int(r) for r in check_rings) if check_rings else ()
        self._check = None
        self._full = None

    def predict(self, em_xyz, mu_xyz):
        x, m = self.features(em_xyz, mu_xyz)
        out = self.net.forward(x, self.features.dims)
        return (out * (m / self.kappa)).reshape(self.features.dims)

    def full_sss(self, em_xyz, mu_xyz):
        if self._full is None:
            self._full = SingleScatterSimulation(*self.sss_args, physical_kn=self.physical_kn)
        return self._full(em_xyz, mu_xyz)

    def estimate(self, em_xyz, mu_xyz):
        pred = self.predict(em_xyz, mu_xyz)
        info = {"fallback": False, "check_err": None}
        if self.check_rings:
            if self._check is None:
                self._check = SingleScatterSimulation(*self.sss_args, physical_kn=self.physical_kn,
                                                      rings=self.check_rings)
            ref = self._check(em_xyz, mu_xyz)
            err = relative_l2(pred[list(self.check_rings)], ref)
            info["check_err"] = err
            if not (err <= self.tol):
                info["fallback"] = True
                return self.full_sss(em_xyz, mu_xyz), info
        return pred, info


def train(net, X, T, dims, epochs, batch=4, lr=1e-3, seed=0, log=None, lr_min_frac=0.05, time_limit=None,
          relative=False):
    """Adam on the MSE between net(X[i]) and T[i] (both on the GPU:
    X (N, V, 8) int16 bf16, T (N, V) float32), each sample's MSE divided by
    mean(T[i]^2) with `relative` (the squared relative L2 error).  Cosine
    learning-rate decay to lr * lr_min_frac.  Returns the per-epoch mean losses."""
    N = X.shape[0]
    w = (1.0 / (T.double() ** 2).mean(1)).tolist() if relative else [1.0] * N
    g = torch.Generator().manual_seed(int(seed))
    nb = min(batch, N)
    B = net._grad_buffers(*dims, nb)
    steps_per_epoch = max(1, N // batch)
    total = epochs * steps_per_epoch
    hist, k, t0 = [], 0, time.time()
    for ep in range(epochs):
        perm = torch.randperm(N, generator=g)
        B["loss"].zero_()
        for sidx in range(steps_per_epoch):
            net.zero_grad()
            idx = perm[sidx * batch:(sidx + 1) * batch]
            ix = idx.to(X.device)
            out = net.forward(X.index_select(0, ix).reshape(-1, X.shape[-1]), dims, nb)
            gout = net.loss_and_grad(out, T.index_select(0, ix), dims, [w[i] for i in idx.tolist()], nb)
            net.backward(gout, dims, nb)
            frac = k / max(1, total - 1)
            cur = lr * (lr_min_frac + (1 - lr_min_frac) * 0.5 * (1 + math.cos(math.pi * frac)))
            net.adam(cur, gscale=1.0 / batch)
            k += 1
        hist.append(float(B["loss"].item()) / (steps_per_epoch * batch))
        if log:
            log(f"epoch {ep + 1}/{epochs} loss {hist[-1]:.4e} lr {cur:.2e} {time.time() - t0:.0f}s")
        if time_limit is not None and time.time() - t0 > time_limit:
            if log:
                log(f"time limit reached after {ep + 1} epochs")
            break
    return hist


def flops_per_pass(net, dims):
    """Dense MFMA FLOPs of one forward pass (real channels; 2 per multiply-add)."""
    D, H, W = dims
    total = 0
    for l in net.layers:
        lvl = int(l.name[3]) if l.name[:3] in ("enc", "dec") else 0
        V = (D >> lvl) * (H >> lvl) * (W >> lvl)
        total += 2 * V * l.taps * l.cin * l.cout
    return total


def _main():
    """Time forward and training passes of the surrogate network on random
    bf16 features at the sinogram grid (32 x 64 x 64), one sample and a batch
    of samples per pass."""
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--dims", type=int, nargs=3, default=(32, 64, 64))
    ap.add_argument("--batch", type=int, nargs="*", default=(1, 32))
    a = ap.parse_args()
    dev = torch.device("cuda", 0)
    dims = tuple(a.dims)
    V = dims[0] * dims[1] * dims[2]
    net = UNet(device=dev)
    f = flops_per_pass(net, dims)
    for N in a.batch:
        x = to_bf16(torch.randn((N * V, 8), device=dev))
        t_ = torch.rand(N * V, device=dev)
        for _ in range(2):
            net.forward(x, dims, N)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(a.steps):
            net.forward(x, dims, N)
        torch.cuda.synchronize()
        t_fwd = (time.perf_counter() - t0) / a.steps
        t0 = time.perf_counter()
        for _ in range(a.steps):
            net.zero_grad()
            out = net.forward(x, dims, N)
            net.backward(net.loss_and_grad(out, t_, dims, 1.0, N), dims, N)
            net.adam(1e-4)
        torch.cuda.synchronize()
        t_step = (time.perf_counter() - t0) / a.steps
        print(f"dims {dims} x {N}: forward {1e3 * t_fwd:.3f} ms ({N * f / t_fwd / 1e12:.1f} TFLOP/s dense), "
              f"training step {1e3 * t_step:.3f} ms ({3 * N * f / t_step / 1e12:.1f} TFLOP/s)", flush=True)


if __name__ == "__main__":
    _main()


__all__ = ["UNet", "SinogramFeatures", "ScatterSurrogate", "train", "relative_l2", "to_bf16",
           "bf16_bits_to_float", "unet_layers", "flops_per_pass"]
