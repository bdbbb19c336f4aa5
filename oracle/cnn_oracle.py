"""CPU float64 oracle of the scatter-surrogate CNN (test infrastructure only).

The CNN has no counterpart in the reference (it replaces the SSS estimate of
srfnef/corrections/scattering/scatter.py:25-55 inside the EM loop), so this
module restates the textbook operations the GPU kernels implement, pinned in
tests/test_cnn_oracle.py by brute-force loops, central finite differences and
closed forms -- not by retyping these formulas:

  conv3d            y[z,y,x,n] = b[n] + sum_t sum_c x[(z,y,x)+o_t, c] w[t,n,c],
                    3x3x3 (taps t = 9(dz+1) + 3(dy+1) + (dx+1)) or 1x1x1,
                    zero padding, stride 1; channels-last (D,H,W,C)
  leaky ReLU        y = x (x > 0), slope * x otherwise
  avg pool 2x2x2, nearest 2x upsampling, channel concatenation
  MSE               mean((p - t)^2)
  Adam              Kingma & Ba, "Adam: A Method for Stochastic Optimization",
                    ICLR 2015, Algorithm 1 (bias-corrected moments)

`bf16=True` rounds every tensor the GPU stores in bfloat16 (inputs, weights,
activations, activation gradients) to bfloat16 (round to nearest even) at the
same points, so that GPU and oracle differ only by accumulation order.

The U-Net here is the surrogate's architecture: encoder levels of two 3x3x3
convolutions with leaky ReLU (average pooling between levels), decoder levels
of nearest upsampling, concatenation with the encoder output of the same level
(upsampled channels first) and two convolutions, and a 1x1x1 linear head.
"""
import numpy as np


def bf16_round(a):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def tap_offsets(taps):
    if taps == 1:
        return [(0, 0, 0)]
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def _shift(x, o):
    """s[z,y,x] = x[z+oz, y+oy, x+ox] with zeros outside."""
    D, H, W = x.shape[:3]
    out = np.zeros_like(x)
    oz, oy, ox = o
    zs, ys, xs = slice(max(0, -oz), min(D, D - oz)), slice(max(0, -oy), min(H, H - oy)), slice(max(0, -ox),
                                                                                              min(W, W - ox))
    zt, yt, xt = slice(max(0, oz), min(D, D + oz)), slice(max(0, oy), min(H, H + oy)), slice(max(0, ox),
                                                                                            min(W, W + ox))
    out[zs, ys, xs] = x[zt, yt, xt]
    return out


def conv3d(x, w, b=None):
    """x (D,H,W,Cin), w (taps, Cout, Cin), b (Cout,) or None -> (D,H,W,Cout)."""
    taps = w.shape[0]
    y = np.zeros(x.shape[:3] + (w.shape[1],), dtype=np.float64)
    for t, o in enumerate(tap_offsets(taps)):
        y += _shift(x, o) @ w[t].T
    if b is not None:
        y += b
    return y


def conv3d_grad_input(dy, w):
    """d/dx of sum(dy * conv3d(x, w))."""
    taps = w.shape[0]
    dx = np.zeros(dy.shape[:3] + (w.shape[2],), dtype=np.float64)
    for t, o in enumerate(tap_offsets(taps)):
        dx += _shift(dy @ w[t], tuple(-v for v in o))
    return dx


def conv3d_grad_weight(x, dy, taps):
    """d/dw of sum(dy * conv3d(x, w)) -> (taps, Cout, Cin)."""
    Cin, Cout = x.shape[3], dy.shape[3]
    dw = np.zeros((taps, Cout, Cin), dtype=np.float64)
    dyf = dy.reshape(-1, Cout)
    for t, o in enumerate(tap_offsets(taps)):
        dw[t] = dyf.T @ _shift(x, o).reshape(-1, Cin)
    return dw


def leaky(x, slope):
    return np.where(x > 0, x, slope * x)


def leaky_grad(a, g, slope):
    """Gradient through leaky ReLU given its output a (sign(a) = sign(pre))."""
    return np.where(a > 0, g, slope * g)


def pool(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def pool_grad(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def upsample(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def upsample_grad(g):
    D, H, W, C = g.shape
    return g.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def mse(p, t):
    e = p - t
    return float(np.mean(e * e)), 2.0 * e / e.size


def adam(w, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step; returns (w, m, v) (Algorithm 1 of Kingma & Ba)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


def unet_layers(in_ch, widths):
    """[(name, taps, cin, cout)] of the U-Net, in parameter order."""
    L = []
    c = in_ch
    for l, w in enumerate(widths):
        L += [(f"enc{l}a", 27, c, w), (f"enc{l}b", 27, w, w)]
        c = w
    for l in range(len(widths) - 2, -1, -1):
        L += [(f"dec{l}a", 27, widths[l + 1] + widths[l], widths[l]), (f"dec{l}b", 27, widths[l], widths[l])]
    L += [("head", 1, widths[0], 1)]
    return L


class UNetOracle:
    """params: {name: (w (taps, cout, cin), b (cout,))} in float64."""

    def __init__(self, params, widths, slope, bf16=True):
        self.p = params
        self.widths = list(widths)
        self.slope = slope
        self.r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))

    def _w(self, name):
        w, b = self.p[name]
        return self.r(w), np.asarray(b, dtype=np.float64)

    def _conv(self, name, x, act=True):
        w, b = self._w(name)
        y = conv3d(x, w, b)
        return self.r(leaky(y, self.slope)) if act else y

    def forward(self, x):
        """x (D,H,W,in_ch) -> (D,H,W) float64 output; keeps what backward needs."""
        r, n = self.r, len(self.widths)
        c = {"x0": r(x)}
        h = c["x0"]
        for l in range(n):
            if l > 0:
                h = r(pool(c[f"e{l - 1}"]))
                c[f"p{l - 1}"] = h
            c[f"e{l}a"] = self._conv(f"enc{l}a", h)
            c[f"e{l}"] = self._conv(f"enc{l}b", c[f"e{l}a"])
        d = c[f"e{n - 1}"]
        for l in range(n - 2, -1, -1):
            c[f"u{l}"] = np.concatenate([upsample(d), c[f"e{l}"]], axis=3)
            c[f"d{l}a"] = self._conv(f"dec{l}a", c[f"u{l}"])
            d = self._conv(f"dec{l}b", c[f"d{l}a"])
            c[f"d{l}"] = d
        out = self._conv("head", d, act=False)[..., 0]
        self.cache = c
        return out

    def _conv_back(self, name, x, a, ga, grads, act=True, need_dx=True):
        """Gradient through conv(+leaky) with input x, output a, dL/da = ga."""
        w, _ = self._w(name)
        gp = self.r(leaky_grad(a, ga, self.slope) if act else ga)
        grads[name] = (conv3d_grad_weight(x, gp, w.shape[0]), gp.reshape(-1, gp.shape[3]).sum(0))
        return conv3d_grad_input(gp, w) if need_dx else None

    def backward(self, gout):
        """dL/d(output) (D,H,W) -> {name: (dw, db)}."""
        c, n, grads = self.cache, len(self.widths), {}
        g = self._conv_back("head", c["d0"], None, gout[..., None], grads, act=False)
        gskip = {}
        for l in range(0, n - 1):
            a_b, a_a = c[f"d{l}"], c[f"d{l}a"]
            g = self._conv_back(f"dec{l}b", a_a, a_b, g, grads)
            g = self._conv_back(f"dec{l}a", c[f"u{l}"], a_a, g, grads)
            wu = self.widths[l + 1]
            gskip[l] = g[..., wu:]
            g = upsample_grad(g[..., :wu])
        # g is now dL/d e_{n-1}
        for l in range(n - 1, -1, -1):
            if l < n - 1:
                g = g + gskip[l]
            xin = c["x0"] if l == 0 else c[f"p{l - 1}"]
            g = self._conv_back(f"enc{l}b", c[f"e{l}a"], c[f"e{l}"], g, grads)
            g = self._conv_back(f"enc{l}a", xin, c[f"e{l}a"], g, grads, need_dx=l > 0)
            if l > 0:
                g = pool_grad(g)
        return grads


def init_params(in_ch, widths, seed):
    """He-normal weights (fan_in = taps * cin), zero biases; head std 1e-2."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, taps, cin, cout in unet_layers(in_ch, widths):
        std = np.sqrt(2.0 / (taps * cin)) if name != "head" else 1e-2
        p[name] = (rng.normal(0.0, std, (taps, cout, cin)), np.zeros(cout))
    return p
