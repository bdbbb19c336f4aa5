"""Pins the CPU float64 CNN oracle (oracle/cnn_oracle.py) without a GPU:
brute-force loops for the convolution, exact adjointness of the linear maps,
central finite differences through the whole U-Net, closed forms of Adam and
known bfloat16 roundings.  Also checks the ABI validation of the CNN entry
points."""
import ctypes

import numpy as np
import pytest

import cnn_oracle as co


def _brute_conv(x, w, b):
    D, H, W, Cin = x.shape
    taps, Cout, _ = w.shape
    offs = co.tap_offsets(taps)
    y = np.zeros((D, H, W, Cout))
    for z in range(D):
        for yy in range(H):
            for xx in range(W):
                for n in range(Cout):
                    s = b[n]
                    for t, (dz, dy, dx) in enumerate(offs):
                        zz, y2, x2 = z + dz, yy + dy, xx + dx
                        if 0 <= zz < D and 0 <= y2 < H and 0 <= x2 < W:
                            s += np.dot(x[zz, y2, x2], w[t, n])
                    y[z, yy, xx, n] = s
    return y


@pytest.mark.parametrize("taps", [27, 1])
def test_conv_matches_brute_force(taps):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(3, 4, 5, 3))
    w = rng.normal(size=(taps, 2, 3))
    b = rng.normal(size=2)
    np.testing.assert_allclose(co.conv3d(x, w, b), _brute_conv(x, w, b), rtol=1e-12, atol=1e-12)


def test_tap_order_is_z_major():
    # a single tap at (dz,dy,dx) = (+1, 0, -1) reads x[z+1, y, x-1]
    t = 9 * 2 + 3 * 1 + 0
    assert co.tap_offsets(27)[t] == (1, 0, -1)
    x = np.zeros((3, 3, 3, 1))
    x[2, 1, 0, 0] = 1.0
    w = np.zeros((27, 1, 1))
    w[t] = 1.0
    assert co.conv3d(x, w)[1, 1, 1, 0] == 1.0


@pytest.mark.parametrize("taps", [27, 1])
def test_conv_gradients_are_adjoints(taps):
    # <conv(x, w), g> = <x, dgrad(g, w)> = <w, wgrad(x, g)>
    rng = np.random.default_rng(2)
    x = rng.normal(size=(4, 3, 5, 6))
    w = rng.normal(size=(taps, 7, 6))
    g = rng.normal(size=(4, 3, 5, 7))
    lhs = np.sum(co.conv3d(x, w) * g)
    assert np.sum(x * co.conv3d_grad_input(g, w)) == pytest.approx(lhs, rel=1e-12)
    assert np.sum(w * co.conv3d_grad_weight(x, g, taps)) == pytest.approx(lhs, rel=1e-12)


def test_pool_upsample_adjoints_and_values():
    rng = np.random.default_rng(3)
    x = rng.normal(size=(4, 6, 2, 3))
    y = rng.normal(size=(2, 3, 1, 3))
    assert np.sum(co.pool(x) * y) == pytest.approx(np.sum(x * co.pool_grad(y)), rel=1e-12)
    assert np.sum(co.upsample(y) * x) == pytest.approx(np.sum(y * co.upsample_grad(x)), rel=1e-12)
    assert co.pool(x)[1, 2, 0, 1] == pytest.approx(x[2:4, 4:6, 0:2, 1].mean())
    assert co.upsample(y)[3, 5, 1, 2] == y[1, 2, 0, 2]


def _unet_fd_check(widths, seed):
    rng = np.random.default_rng(seed)
    params = co.init_params(4, widths, seed)
    # non-zero biases and head so every path carries gradient
    params = {k: (w, rng.normal(0, 0.1, b.shape)) for k, (w, b) in params.items()}
    params["head"] = (rng.normal(0, 0.3, params["head"][0].shape), params["head"][1])
    x = rng.normal(size=(4, 4, 8, 4))
    t = rng.normal(size=(4, 4, 8))
    net = co.UNetOracle(params, widths, 0.01, bf16=False)

    def loss(p):
        n = co.UNetOracle(p, widths, 0.01, bf16=False)
        return co.mse(n.forward(x), t)[0]

    _, gout = co.mse(net.forward(x), t)
    grads = net.backward(gout)
    h = 1e-6
    for name in ("enc0a", "enc1b", "dec0a", "dec1b", "head"):
        for part in (0, 1):
            d = rng.normal(size=params[name][part].shape)
            pp = dict(params)
            pm = dict(params)
            lp, lm = list(params[name]), list(params[name])
            lp[part] = lp[part] + h * d
            lm[part] = lm[part] - h * d
            pp[name], pm[name] = tuple(lp), tuple(lm)
            fd = (loss(pp) - loss(pm)) / (2 * h)
            an = float(np.sum(grads[name][part] * d))
            assert an == pytest.approx(fd, rel=1e-5, abs=1e-10), (name, part)


def test_unet_backward_matches_finite_differences():
    _unet_fd_check((4, 8, 8), 5)


def test_adam_closed_forms():
    # constant gradient g: bias-corrected moments are g and g^2, so every step
    # moves by lr * g / (|g| + eps)
    w = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    m = np.zeros(3)
    v = np.zeros(3)
    w0 = w.copy()
    for k in range(1, 6):
        w, m, v = co.adam(w, g, m, v, k, lr=1e-2, eps=1e-8)
        np.testing.assert_allclose(w, w0 - k * 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    # zero gradient leaves zero state and the weight alone
    w2, m2, v2 = co.adam(np.ones(2), np.zeros(2), np.zeros(2), np.zeros(2), 1)
    assert np.all(w2 == 1.0) and np.all(m2 == 0) and np.all(v2 == 0)


def test_bf16_round_known_values():
    ulp = 2.0 ** -7
    vals = np.array([1.0, 1 + ulp / 2, 1 + 1.5 * ulp, 1 + 0.75 * ulp, -3.0, 2.0 ** -130, 65504.0])
    want = np.array([1.0, 1.0, 1 + 2 * ulp, 1 + ulp, -3.0, 2.0 ** -130, 65536.0])
    np.testing.assert_array_equal(co.bf16_round(vals), want)


def test_unet_layer_list():
    L = co.unet_layers(4, (16, 32, 64, 128))
    names = [n for n, *_ in L]
    assert names[:2] == ["enc0a", "enc0b"] and names[-1] == "head"
    assert dict((n, (ci, co_)) for n, _, ci, co_ in L)["dec2a"] == (128 + 64, 64)
    from srf_amd.surrogate import unet_layers
    assert unet_layers(4, (16, 32, 64, 128)) == L


def test_cnn_entry_points_validate_arguments():
    from srf_amd._lib import lib
    L = lib()
    fake = ctypes.c_void_p(64)
    # cin_p not a multiple of 8, cout_p not a multiple of 32, bad taps, out_off past stride
    assert L.srf_cnn_conv(fake, fake, None, 2, 2, 2, 12, 4, 32, 27, 4, 0, -1.0, fake, None, None) == -1
    assert L.srf_cnn_conv(fake, fake, None, 2, 2, 2, 8, 4, 16, 27, 4, 0, -1.0, fake, None, None) == -1
    assert L.srf_cnn_conv(fake, fake, None, 2, 2, 2, 8, 4, 32, 9, 4, 0, -1.0, fake, None, None) == -1
    assert L.srf_cnn_conv(fake, fake, None, 2, 2, 2, 8, 4, 32, 27, 4, 2, -1.0, fake, None, None) == -1
    assert L.srf_cnn_conv(fake, fake, None, 2, 2, 2, 8, 4, 32, 27, 4, 0, -1.0, None, None, None) == -1
    assert L.srf_cnn_conv_wgrad(fake, fake, 2, 2, 2, 2, 8, 4, 32, 27, fake, None) == -1
    assert L.srf_cnn_pool(fake, 3, 2, 2, 4, 4, fake, 4, None) == -1
    assert L.srf_cnn_adam(fake, fake, fake, fake, 4, 1e-3, 0.9, 0.999, 1e-8, 0, 1.0, None, None) == -1
    assert L.srf_cnn_flip_weights(fake, 27, 40, 64, 16, 16, fake, 32, 16, None) == -1
