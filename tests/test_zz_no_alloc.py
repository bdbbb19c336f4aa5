"""The allocation-free entry point srf_smooth_noalloc (include/srf_hip.h)
against srf_smooth through the C-ABI: bit-identical outputs on two calls in a
row, the second on a smaller image (the static scratch keeps the first call's
contents, which must not leak into the second), with and without a mask;
sizes above the scratch's 2^23 floats are refused before any launch."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _smooth(x, sig, mask, noalloc):
    from srf_amd._lib import check, lib, stream_handle
    n0, n1, n2 = x.shape
    out = torch.full_like(x, float("nan"))
    mp = mask.data_ptr() if mask is not None else None
    s = stream_handle(x.device)
    if noalloc:
        rc = lib().srf_smooth_noalloc(x.data_ptr(), n0, n1, n2, sig[0], sig[1], sig[2], mp, out.data_ptr(), s)
        check(rc, "srf_smooth_noalloc")
    else:
        tmp = torch.empty_like(x)
        rc = lib().srf_smooth(x.data_ptr(), n0, n1, n2, sig[0], sig[1], sig[2], mp, tmp.data_ptr(),
                              out.data_ptr(), s)
        check(rc, "srf_smooth")
    return out


@pytest.mark.parametrize("masked", [False, True])
def test_noalloc_smooth_equals_smooth_twice(masked):
    g = torch.Generator(device="cuda").manual_seed(3)
    cases = [((256, 256, 128), (4.2, 4.2, 3.1)), ((96, 80, 40), (2.5, 1.5, 6.0))]   # second call smaller
    for shape, sig in cases:
        x = torch.rand(shape, generator=g, device="cuda") * 10.0
        mask = (torch.rand(shape, generator=g, device="cuda") > 0.3).float() if masked else None
        want = _smooth(x, sig, mask, noalloc=False)
        got = _smooth(x, sig, mask, noalloc=True)
        torch.cuda.synchronize()
        print(f"smooth {shape} masked={masked}: max |diff| {float((got - want).abs().max()):.3g}", flush=True)
        assert torch.isfinite(got).all()
        assert torch.equal(got, want), shape


def test_noalloc_smooth_refuses_oversize():
    from srf_amd._lib import lib, stream_handle
    x = torch.zeros(8, device="cuda")
    # 257 x 256 x 128 > 2^23 voxels: SRF_ERR_WORKSPACE, checked before any launch
    rc = lib().srf_smooth_noalloc(x.data_ptr(), 257, 256, 128, 1.0, 1.0, 1.0, None, x.data_ptr(),
                                  stream_handle(x.device))
    assert rc == -4
