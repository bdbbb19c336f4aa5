"""BF16 MFMA CNN kernels (srf_cnn.hip) against the CPU float64 oracle.

Single kernels use bf16-representable inputs, so the only difference from the
oracle is FP32 accumulation (relative L2 <= 1e-5).  The whole U-Net is
compared twice: with the oracle emulating bf16 storage at the same points
(differences come from accumulation order and the rare rounding flip it
causes: <= 1e-2) and with the exact float64 network, which measures what bf16
storage costs: output within 2e-2, every layer's gradient within 0.2 relative
L2 and at cosine >= 0.98 to the exact one (bias gradients that nearly cancel
over the volume amplify the 2^-9 rounding of each term)."""
import numpy as np
import pytest
import torch

import cnn_oracle as co

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _L():
    from srf_amd._lib import lib
    return lib()


def _s():
    from srf_amd._lib import stream_handle
    return stream_handle()


def _bf(a):
    """numpy float -> (int16 device tensor of bf16 bits, float64 rounded values)."""
    r = co.bf16_round(a)
    bits = (r.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    return torch.from_numpy(bits.copy()).to(DEV), r


def _f(t_bits):
    from srf_amd.surrogate import bf16_bits_to_float
    return bf16_bits_to_float(t_bits).double().cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def _pad_w(w, cout_p, cin_p):
    taps, cout, cin = w.shape
    out = np.zeros((taps, cout_p, cin_p))
    out[:, :cout, :cin] = w
    return out


@pytest.mark.parametrize("taps,dims,cin,cout", [(27, (4, 6, 10), 13, 40), (27, (2, 8, 8), 48, 16),
                                                (1, (3, 5, 7), 16, 1), (27, (8, 4, 4), 128, 128),
                                                (27, (4, 8, 16), 8, 16), (27, (2, 4, 40), 24, 96),
                                                (27, (1, 1, 3), 8, 16)])
def test_conv_forward(taps, dims, cin, cout):
    from srf_amd.surrogate import _c8, _c32
    rng = np.random.default_rng(taps + cin)
    cin_p, cout_p = _c8(cin), _c32(cout)
    V = int(np.prod(dims))
    x = np.zeros(dims + (cin_p,))
    x[..., :cin] = rng.normal(size=dims + (cin,))
    w = rng.normal(size=(taps, cout, cin)) / np.sqrt(taps * cin)
    b = rng.normal(size=cout).astype(np.float32)
    xd, xr = _bf(x)
    wd, wr = _bf(_pad_w(w, cout_p, cin_p))
    bd = torch.from_numpy(b).to(DEV)
    stride, off = cout + 8, 5
    yb = torch.full((V, stride), 0x1234, dtype=torch.int16, device=DEV)
    yf = torch.full((V, stride), -7.0, dtype=torch.float32, device=DEV)
    print(f"conv taps={taps} dims={dims} cin={cin} cout={cout}", flush=True)
    rc = _L().srf_cnn_conv(xd.data_ptr(), wd.data_ptr(), bd.data_ptr(), *dims, cin_p, cout, cout_p, taps, stride, off,
                           0.01, yb.data_ptr(), yf.data_ptr(), _s())
    assert rc == 0
    torch.cuda.synchronize()
    want = co.leaky(co.conv3d(xr, wr[:, :cout, :], b.astype(np.float64)), 0.01).reshape(V, cout)
    got = yf.double().cpu().numpy()
    assert _rel(got[:, off:off + cout], want) <= 1e-5
    # bf16 output is the rounding of the same fp32 value; other channels untouched
    gb = yb.cpu().numpy()
    np.testing.assert_array_equal(_f(yb)[:, off:off + cout], co.bf16_round(got[:, off:off + cout]))
    assert np.all(got[:, :off] == -7.0) and np.all(got[:, off + cout:] == -7.0)
    assert np.all(gb[:, :off] == 0x1234) and np.all(gb[:, off + cout:] == 0x1234)


@pytest.mark.parametrize("taps,dims,cin,cout", [(27, (4, 6, 10), 16, 40), (27, (8, 16, 16), 48, 16),
                                                (1, (3, 5, 7), 16, 1), (27, (4, 8, 32), 32, 64),
                                                (27, (4, 8, 8), 128, 128), (1, (4, 8, 16), 16, 1),
                                                (27, (2, 4, 40), 24, 40)])
def test_conv_weight_and_input_gradients(taps, dims, cin, cout):
    """Shapes with W a multiple of 8 take the LDS-staged weight-gradient kernel
    (W = 40: a partial second x tile), the others the element-load one."""
    from srf_amd.surrogate import _c8, _c32
    rng = np.random.default_rng(7 + taps)
    cin_p, cout_p = _c8(cin), _c32(cout)
    fcin_p, fcout_p = _c8(cout), _c32(cin)
    V = int(np.prod(dims))
    xd, xr = _bf(rng.normal(size=dims + (cin_p,)))
    gy = np.zeros(dims + (fcin_p,))
    gy[..., :cout] = rng.normal(size=dims + (cout,))
    gd, gr = _bf(gy)
    w = rng.normal(size=(taps, cout, cin))
    wd, wr = _bf(_pad_w(w, cout_p, cin_p))
    print(f"wgrad/dgrad taps={taps} dims={dims} cin={cin} cout={cout}", flush=True)
    # weight gradient accumulates into a pre-filled buffer
    dw0 = rng.normal(size=(taps, cout_p, cin_p)).astype(np.float32)
    dw = torch.from_numpy(dw0.copy()).to(DEV)
    assert _L().srf_cnn_conv_wgrad(xd.data_ptr(), gd.data_ptr(), fcin_p, *dims, cin_p, cout, cout_p, taps,
                                   dw.data_ptr(), _s()) == 0
    torch.cuda.synchronize()
    want = co.conv3d_grad_weight(xr, gr[..., :cout], taps)
    got = dw.double().cpu().numpy() - dw0
    assert _rel(got[:, :cout, :], want) <= 1e-5
    assert np.all(got[:, cout:, :] == 0)
    # input gradient = conv of gy with the flipped kernel
    wf = torch.zeros((taps, fcout_p, fcin_p), dtype=torch.int16, device=DEV)
    assert _L().srf_cnn_flip_weights(wd.data_ptr(), taps, cout, cout_p, cin_p, cin, wf.data_ptr(), fcout_p, fcin_p,
                                     _s()) == 0
    dx = torch.zeros((V, cin), dtype=torch.float32, device=DEV)
    assert _L().srf_cnn_conv(gd.data_ptr(), wf.data_ptr(), None, *dims, fcin_p, cin, fcout_p, taps, cin, 0, -1.0,
                             None, dx.data_ptr(), _s()) == 0
    torch.cuda.synchronize()
    want = co.conv3d_grad_input(gr[..., :cout], wr[:, :cout, :cin]).reshape(V, cin)
    assert _rel(dx.double().cpu().numpy(), want) <= 1e-5


@pytest.mark.parametrize("dims,cin,cout", [((4, 6, 16), 24, 40), ((4, 6, 10), 16, 16)])
def test_batched_conv_and_gradients_per_sample(dims, cin, cout):
    """Three samples stacked along z: the forward equals the oracle's
    per-sample convolution (no tap reaches into the neighbouring sample) and
    the weight and bias gradients equal the sums of the per-sample ones
    (W = 16: the LDS kernel with the folded bias sum; W = 10: the element
    path plus the bias kernel)."""
    from srf_amd.surrogate import _c8, _c32
    rng = np.random.default_rng(17)
    N, taps = 3, 27
    cin_p, cout_p, fcin_p = _c8(cin), _c32(cout), _c8(cout)
    V = int(np.prod(dims))
    x = np.zeros((N,) + dims + (cin_p,))
    x[..., :cin] = rng.normal(size=(N,) + dims + (cin,))
    xd, xr = _bf(x)
    w = rng.normal(size=(taps, cout, cin)) / np.sqrt(taps * cin)
    wd, wr = _bf(_pad_w(w, cout_p, cin_p))
    b = rng.normal(size=cout).astype(np.float32)
    y = torch.zeros((N * V, cout), dtype=torch.float32, device=DEV)
    print(f"batched conv N={N} dims={dims} cin={cin} cout={cout}", flush=True)
    assert _L().srf_cnn_conv_batch(xd.data_ptr(), wd.data_ptr(), torch.from_numpy(b).to(DEV).data_ptr(), N, *dims,
                                   cin_p, cout, cout_p, taps, cout, 0, -1.0, None, y.data_ptr(), _s()) == 0
    gy = np.zeros((N,) + dims + (fcin_p,))
    gy[..., :cout] = rng.normal(size=(N,) + dims + (cout,))
    gd, gr = _bf(gy)
    dw = torch.zeros((taps, cout_p, cin_p), device=DEV)
    db = torch.zeros(cout, device=DEV)
    assert _L().srf_cnn_conv_wgrad_batch(xd.data_ptr(), gd.data_ptr(), fcin_p, N, *dims, cin_p, cout, cout_p, taps,
                                         dw.data_ptr(), db.data_ptr(), _s()) == 0
    torch.cuda.synchronize()
    got = y.double().cpu().numpy().reshape(N, V, cout)
    for i in range(N):
        want = co.conv3d(xr[i], wr[:, :cout, :], b.astype(np.float64)).reshape(V, cout)
        assert _rel(got[i], want) <= 1e-5, i
    want_w = sum(co.conv3d_grad_weight(xr[i], gr[i][..., :cout], taps) for i in range(N))
    assert _rel(dw.double().cpu().numpy()[:, :cout, :], want_w) <= 1e-5
    assert np.all(dw.cpu().numpy()[:, cout:, :] == 0)
    want_b = gr[..., :cout].reshape(-1, cout).sum(0)
    assert _rel(db.double().cpu().numpy(), want_b) <= 1e-5


def test_conv_reads_upsampled_concatenation_in_place():
    """srf_cnn_conv_up_batch = the convolution of [upsample2x(coarse), fine]
    (two samples; 32 coarse + 16 fine channels, a chunk from each input)."""
    from srf_amd.surrogate import _c32
    rng = np.random.default_rng(21)
    N, dims, cc, cf, cout, taps = 2, (4, 6, 16), 32, 16, 24, 27
    cdims = tuple(v // 2 for v in dims)
    cin_p, cout_p = cc + cf, _c32(cout)
    V = int(np.prod(dims))
    fd, fr = _bf(rng.normal(size=(N,) + dims + (cf,)))
    cd_, cr = _bf(rng.normal(size=(N,) + cdims + (cc,)))
    w = rng.normal(size=(taps, cout, cin_p)) / np.sqrt(taps * cin_p)
    wd, wr = _bf(_pad_w(w, cout_p, cin_p))
    b = rng.normal(size=cout).astype(np.float32)
    y = torch.zeros((N * V, cout), dtype=torch.float32, device=DEV)
    print(f"conv of upsampled concatenation N={N} dims={dims}", flush=True)
    assert _L().srf_cnn_conv_up_batch(fd.data_ptr(), cd_.data_ptr(), cc, wd.data_ptr(),
                                      torch.from_numpy(b).to(DEV).data_ptr(), N, *dims, cin_p, cout, cout_p, taps,
                                      cout, 0, 0.01, None, y.data_ptr(), _s()) == 0
    # a coarse channel count that is not a multiple of 8 is refused
    assert _L().srf_cnn_conv_up_batch(fd.data_ptr(), cd_.data_ptr(), 12, wd.data_ptr(), None, N, *dims, cin_p, cout,
                                      cout_p, taps, cout, 0, 0.01, None, y.data_ptr(), _s()) != 0
    torch.cuda.synchronize()
    got = y.double().cpu().numpy().reshape(N, V, cout)
    for i in range(N):
        u = np.concatenate([co.upsample(cr[i]), fr[i]], -1)
        want = co.leaky(co.conv3d(u, wr[:, :cout, :], b.astype(np.float64)), 0.01).reshape(V, cout)
        assert _rel(got[i], want) <= 1e-5, i


def test_conv_with_head_in_the_epilogue():
    """srf_cnn_conv_head_batch: the layer's bf16 output as srf_cnn_conv_batch
    writes it, and the 1x1 one-channel head of that output within float
    rounding of the separate head convolution (and of the oracle)."""
    from srf_amd.surrogate import _c8, _c32
    rng = np.random.default_rng(23)
    N, dims, cin, cout, taps = 2, (4, 8, 32), 16, 16, 27
    cin_p, cout_p = _c8(cin), _c32(cout)
    V = int(np.prod(dims))
    xd, xr = _bf(rng.normal(size=(N,) + dims + (cin_p,)))
    w = rng.normal(size=(taps, cout, cin)) / np.sqrt(taps * cin)
    wd, wr = _bf(_pad_w(w, cout_p, cin_p))
    b = torch.from_numpy(rng.normal(size=cout).astype(np.float32)).to(DEV)
    hw = rng.normal(size=(1, 1, cout))
    hwd, hwr = _bf(_pad_w(hw, 32, cout))
    hb = torch.tensor([0.3], dtype=torch.float32, device=DEV)
    y1 = torch.zeros((N * V, cout), dtype=torch.int16, device=DEV)
    y2 = torch.zeros_like(y1)
    out = torch.zeros(N * V, dtype=torch.float32, device=DEV)
    ref = torch.zeros(N * V, dtype=torch.float32, device=DEV)
    print(f"conv + head epilogue N={N} dims={dims}", flush=True)
    assert _L().srf_cnn_conv_batch(xd.data_ptr(), wd.data_ptr(), b.data_ptr(), N, *dims, cin_p, cout, cout_p, taps,
                                   cout, 0, 0.01, y1.data_ptr(), None, _s()) == 0
    assert _L().srf_cnn_conv_head_batch(xd.data_ptr(), wd.data_ptr(), b.data_ptr(), N, *dims, cin_p, cout, cout_p,
                                        taps, 0.01, y2.data_ptr(), hwd.data_ptr(), hb.data_ptr(), out.data_ptr(),
                                        _s()) == 0
    assert _L().srf_cnn_conv_batch(y1.data_ptr(), hwd.data_ptr(), hb.data_ptr(), N, *dims, cout, 1, 32, 1, 1, 0,
                                   -1.0, None, ref.data_ptr(), _s()) == 0
    torch.cuda.synchronize()
    assert torch.equal(y1, y2)
    got, want = out.double().cpu().numpy(), ref.double().cpu().numpy()
    assert _rel(got, want) <= 1e-6
    orc = (_f(y1).reshape(N * V, cout) @ hwr[0, 0, :cout]) + 0.3
    assert _rel(got, orc) <= 1e-6


def test_unet_batch_equals_per_sample():
    """The U-Net over three stacked samples: outputs bit-identical to three
    single-sample passes, parameter gradients equal to their sum (1e-5)."""
    net, params, widths, xd, xr, t = _unet_case(12)
    dims = (8, 8, 16)
    V = int(np.prod(dims))
    rng = np.random.default_rng(4)
    N = 3
    xs = np.zeros((N, V, 8))
    xs[:, :, :4] = rng.normal(size=(N, V, 4))
    X, _ = _bf(xs)
    T = torch.from_numpy(rng.normal(size=(N, V)).astype(np.float32)).to(DEV)
    print("unet batch N=3 dims", dims, flush=True)
    singles, gsum = [], None
    for i in range(N):
        net.zero_grad()
        out = net.forward(X[i], dims)
        singles.append(out.clone())
        net.backward(net.loss_and_grad(out, T[i], dims, 0.5 + i), dims)
        g = net.grads()
        gsum = g if gsum is None else {k: (gsum[k][0] + g[k][0], gsum[k][1] + g[k][1]) for k in g}
    net.zero_grad()
    outb = net.forward(X.reshape(N * V, 8), dims, N)
    assert all(torch.equal(outb[i * V:(i + 1) * V], singles[i]) for i in range(N))
    net.backward(net.loss_and_grad(outb, T, dims, [0.5 + i for i in range(N)], N), dims, N)
    gb = net.grads()
    torch.cuda.synchronize()
    for k in gb:
        assert _rel(gb[k][0], gsum[k][0]) <= 1e-5 and _rel(gb[k][1], gsum[k][1]) <= 1e-5, k


def test_pool_upsample_copy_act_bias_mse_adam_cast():
    rng = np.random.default_rng(3)
    D, H, W, C = 4, 6, 8, 16
    V = D * H * W
    xd, xr = _bf(rng.normal(size=(D, H, W, C)))
    y = torch.zeros((V // 8, C), dtype=torch.int16, device=DEV)
    assert _L().srf_cnn_pool(xd.data_ptr(), D, H, W, C, C, y.data_ptr(), C, _s()) == 0
    torch.cuda.synchronize()
    np.testing.assert_allclose(_f(y), co.bf16_round(co.pool(xr)).reshape(V // 8, C), rtol=0, atol=0)
    # pool gradient and upsample (+ its gradient) against the oracle
    g = rng.normal(size=(D // 2, H // 2, W // 2, C)).astype(np.float32)
    gd = torch.from_numpy(g).to(DEV)
    dx = torch.zeros((V, C), device=DEV)
    assert _L().srf_cnn_pool_bwd(gd.data_ptr(), D, H, W, C, C, dx.data_ptr(), C, _s()) == 0
    up = torch.zeros((V, C + 8), dtype=torch.int16, device=DEV)
    assert _L().srf_cnn_upsample(y.data_ptr(), D // 2, H // 2, W // 2, C, C, up.data_ptr(), C + 8, 8, _s()) == 0
    gu = rng.normal(size=(D, H, W, C + 8)).astype(np.float32)
    gud = torch.from_numpy(gu).to(DEV)
    gp = torch.zeros((V // 8, C), device=DEV)
    assert _L().srf_cnn_upsample_bwd(gud.data_ptr(), D // 2, H // 2, W // 2, C, C + 8, 8, gp.data_ptr(), C, _s()) == 0
    torch.cuda.synchronize()
    np.testing.assert_allclose(dx.cpu().numpy(), co.pool_grad(g).reshape(V, C), rtol=1e-7)
    np.testing.assert_array_equal(_f(up)[:, 8:], co.upsample(_f(y).reshape(D // 2, H // 2, W // 2, C)).reshape(V, C))
    np.testing.assert_allclose(gp.cpu().numpy(), co.upsample_grad(gu[..., 8:]).reshape(V // 8, C), rtol=1e-6)
    # copy_channels (accumulate, f32)
    a = torch.from_numpy(rng.normal(size=(V, 24)).astype(np.float32)).to(DEV)
    b = torch.from_numpy(rng.normal(size=(V, 16)).astype(np.float32)).to(DEV)
    b0 = b.clone()
    assert _L().srf_cnn_copy_channels(a.data_ptr(), V, 8, 24, 16, b.data_ptr(), 16, 4, 1, 1, _s()) == 0
    torch.cuda.synchronize()
    want = b0.clone()
    want[:, 4:12] += a[:, 16:24]
    assert torch.equal(b, want)
    # activation backward + bias gradient
    ga = torch.from_numpy(rng.normal(size=(V, C)).astype(np.float32)).to(DEV)
    dp = torch.full((V, 24), 99, dtype=torch.int16, device=DEV)
    assert _L().srf_cnn_act_bwd(ga.data_ptr(), C, xd.data_ptr(), C, V, C, 24, 0.01, dp.data_ptr(), _s()) == 0
    db = torch.ones(C, device=DEV)
    assert _L().srf_cnn_bias_grad(dp.data_ptr(), V, C, 24, db.data_ptr(), _s()) == 0
    torch.cuda.synchronize()
    want_dp = co.bf16_round(co.leaky_grad(xr.reshape(V, C), ga.double().cpu().numpy(), 0.01))
    np.testing.assert_array_equal(_f(dp)[:, :C], want_dp)
    assert np.all(dp.cpu().numpy()[:, C:] == 0)
    np.testing.assert_allclose(db.double().cpu().numpy(), 1.0 + want_dp.sum(0), rtol=1e-5, atol=1e-5)
    # MSE
    p = torch.from_numpy(rng.normal(size=5000).astype(np.float32)).to(DEV)
    t = torch.from_numpy(rng.normal(size=5000).astype(np.float32)).to(DEV)
    grad = torch.zeros(5000, device=DEV)
    loss = torch.zeros(1, device=DEV)
    assert _L().srf_cnn_mse(p.data_ptr(), t.data_ptr(), 5000, 0.5, grad.data_ptr(), loss.data_ptr(), _s()) == 0
    torch.cuda.synchronize()
    l_want, g_want = co.mse(p.double().cpu().numpy(), t.double().cpu().numpy())
    assert float(loss) == pytest.approx(0.5 * l_want, rel=1e-5)
    np.testing.assert_allclose(grad.double().cpu().numpy(), 0.5 * g_want, rtol=1e-5, atol=1e-12)
    # Adam, three steps, with gscale
    w = rng.normal(size=1000)
    wd_ = torch.from_numpy(w.astype(np.float32)).to(DEV)
    m = torch.zeros(1000, device=DEV)
    v = torch.zeros(1000, device=DEV)
    wbf = torch.zeros(1000, dtype=torch.int16, device=DEV)
    wo, mo, vo = w.astype(np.float32).astype(np.float64), np.zeros(1000), np.zeros(1000)
    for k in range(1, 4):
        gk = rng.normal(size=1000).astype(np.float32)
        gdk = torch.from_numpy(gk).to(DEV)
        assert _L().srf_cnn_adam(wd_.data_ptr(), gdk.data_ptr(), m.data_ptr(), v.data_ptr(), 1000, 1e-2, 0.9, 0.999,
                                 1e-8, k, 0.5, wbf.data_ptr(), _s()) == 0
        wo, mo, vo = co.adam(wo, 0.5 * gk.astype(np.float64), mo, vo, k, lr=1e-2)
    torch.cuda.synchronize()
    np.testing.assert_allclose(wd_.double().cpu().numpy(), wo, rtol=1e-5, atol=1e-6)
    np.testing.assert_array_equal(_f(wbf), co.bf16_round(wd_.cpu().numpy()))
    # cast round trip
    z = torch.from_numpy(rng.normal(size=777).astype(np.float32) * 1e3).to(DEV)
    zb = torch.zeros(777, dtype=torch.int16, device=DEV)
    zf = torch.zeros(777, device=DEV)
    assert _L().srf_cnn_cast(z.data_ptr(), 777, 1, zb.data_ptr(), _s()) == 0
    assert _L().srf_cnn_cast(zb.data_ptr(), 777, 0, zf.data_ptr(), _s()) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(zf.double().cpu().numpy(), co.bf16_round(z.cpu().numpy()))


def _unet_case(seed, dims=(8, 8, 16)):
    from srf_amd.surrogate import UNet
    widths = (16, 32, 64, 128)
    rng = np.random.default_rng(seed)
    params = co.init_params(4, widths, seed)
    params = {k: (w, rng.normal(0, 0.05, b.shape)) for k, (w, b) in params.items()}
    params["head"] = (rng.normal(0, 0.2, params["head"][0].shape), params["head"][1])
    net = UNet(4, widths, 0.01, device=DEV)
    net.set_params(params)
    x = np.zeros(dims + (8,))
    x[..., :4] = rng.normal(size=dims + (4,))
    xd, xr = _bf(x)
    t = rng.normal(size=dims)
    return net, params, widths, xd, xr, t


def test_unet_forward_backward_against_oracle():
    net, params, widths, xd, xr, t = _unet_case(11)
    dims = (8, 8, 16)
    V = int(np.prod(dims))
    print("unet forward/backward dims", dims, flush=True)
    out = net.forward(xd.reshape(V, 8), dims)
    tt = torch.from_numpy(t.reshape(-1).astype(np.float32)).to(DEV)
    net.zero_grad()
    gout = net.loss_and_grad(out, tt, dims)
    net.backward(gout, dims)
    torch.cuda.synchronize()
    got_out = out.double().cpu().numpy().reshape(dims)
    grads = net.grads()
    for emulate, tol_out, tol_g in ((True, 1e-2, 1e-2), (False, 2e-2, 0.2)):
        orc = co.UNetOracle(params, widths, 0.01, bf16=emulate)
        o = orc.forward(xr[..., :4] if not emulate else xr[..., :4])
        assert _rel(got_out, o) <= tol_out, (emulate, _rel(got_out, o))
        _, g = co.mse(o, t)
        want = orc.backward(g)
        errs = {n: (_rel(grads[n][0], want[n][0]), _rel(grads[n][1], want[n][1])) for n in want}
        print("emulate", emulate, "out", _rel(got_out, o), "worst grad", max(max(e) for e in errs.values()),
              flush=True)
        for n, (ew, eb) in errs.items():
            assert ew <= tol_g and eb <= tol_g, (emulate, n, ew, eb)
            for k in (0, 1):
                a, b = grads[n][k].ravel(), want[n][k].ravel()
                cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-300))
                assert cos >= (0.9999 if emulate else 0.98), (emulate, n, k, cos)


def test_training_reduces_loss():
    from srf_amd.surrogate import UNet, train
    rng = np.random.default_rng(5)
    dims = (8, 16, 16)
    V = int(np.prod(dims))
    N = 4
    xs = np.zeros((N, V, 8))
    xs[:, :, :4] = rng.normal(size=(N, V, 4))
    X, _ = _bf(xs)
    # a smooth target that depends on the input: local mean of channel 0 plus a constant
    T = torch.from_numpy((0.5 * xs[:, :, 0] + 1.0).astype(np.float32)).to(DEV)
    net = UNet(4, (16, 32, 64, 128), seed=1, device=DEV)
    hist = train(net, X, T, dims, epochs=40, batch=2, lr=3e-3, seed=0)
    print("loss", hist[0], "->", hist[-1], flush=True)
    assert hist[-1] < 0.1 * hist[0]
