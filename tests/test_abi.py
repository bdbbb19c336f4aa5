"""C-ABI checks that need no GPU: the library loads, exports every entry point
declared in include/srf_hip.h, and rejects bad arguments before any launch."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "srf_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(srf_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def L():
    from srf_amd import _lib
    return _lib.lib()


def test_every_declared_symbol_is_exported_and_bound(L):
    from srf_amd import _lib
    names = declared_symbols()
    assert len(names) >= 10
    bound = {n for n, _, _ in _lib.SIGNATURES}
    for n in names:
        assert hasattr(L, n), n
        assert n in bound, f"{n} has no ctypes signature"


def test_strerror_and_version(L):
    assert L.srf_version() >= 1
    for code in (0, -1, -2, -3, -4):
        assert len(L.srf_strerror(code)) > 0


def test_argument_validation_returns_errors(L):
    from srf_amd._lib import SrfGrid, SrfTof
    g = SrfGrid(8, 8, 8, 0, 0, 0, 16, 16, 16)
    bad_grid = SrfGrid(0, 8, 8, 0, 0, 0, 16, 16, 16)
    fake = ctypes.c_void_p(16)
    # negative n, null lors, short stride, bad grid, TOF without values
    assert L.srf_project(fake, -1, 6, None, 0, fake, g, None, fake, None) == -1
    assert L.srf_project(None, 5, 6, None, 0, fake, g, None, fake, None) == -1
    assert L.srf_project(fake, 5, 5, None, 0, fake, g, None, fake, None) == -1
    assert L.srf_project(fake, 5, 6, None, 0, fake, bad_grid, None, fake, None) == -1
    tof = ctypes.pointer(SrfTof(3.66, 1000.0))
    assert L.srf_project(fake, 5, 7, None, 0, fake, g, tof, fake, None) == -1
    bad_tof = ctypes.pointer(SrfTof(3.66, -1.0))
    assert L.srf_project(fake, 5, 7, fake, 7, fake, g, bad_tof, fake, None) == -1
    huge = SrfGrid(2048, 2048, 1024, 0, 0, 0, 1, 1, 1)
    assert L.srf_project(fake, 5, 6, None, 0, fake, huge, None, fake, None) == -2
    assert L.srf_backproject(fake, 5, 6, None, 0, None, g, None, 7, 0, fake, None) == -1
    assert L.srf_backproject(fake, 5, 6, None, 0, None, g, None, 0, 0, None, None) == -1
    assert L.srf_backproject_ratio(fake, 5, 6, None, 0, None, fake, 0, 0.0, g, None, 0, 0, fake, None) == -1
    assert L.srf_ratio(fake, fake, fake, 4, 9, 0.0, None) == -1
    assert L.srf_em_update(fake, fake, None, 4, 0.0, None) == -1
    assert L.srf_em_update(ctypes.c_void_p(18), fake, fake, 4, 0.0, None) == -1   # misaligned
    assert L.srf_lor_digest(fake, 5, 6, None, 0, g, None, 3, fake, fake, None) == -1
    assert L.srf_image_to_kernel_layout(None, fake, 4, 4, 4, None) == -1
    # plans: a slab index packs in 15 bits, so grids above 32767 cells along x or y are refused
    plan = ctypes.c_void_p()
    assert L.srf_plan_create(fake, 5, 6, None, 0, SrfGrid(32768, 8, 8, 0, 0, 0, 32768, 8, 8), None,
                             ctypes.byref(plan), None) == -2 and plan.value is None
    assert L.srf_plan_create(fake, 5, 6, None, 0, SrfGrid(8, 40000, 8, 0, 0, 0, 8, 40000, 8), None,
                             ctypes.byref(plan), None) == -2 and plan.value is None
    # event order: no subsets, short stride, output aliasing the input, too many entries
    assert L.srf_event_order(fake, 5, 6, fake, 5, 0, ctypes.c_void_p(32), None) == -1
    assert L.srf_event_order(fake, 5, 5, fake, 5, 1, ctypes.c_void_p(32), None) == -1
    assert L.srf_event_order(fake, 5, 6, fake, 5, 1, fake, None) == -1
    assert L.srf_event_order(fake, 5, 6, fake, 1 << 31, 1, ctypes.c_void_p(32), None) == -2


def test_empty_inputs_are_no_ops(L):
    from srf_amd._lib import SrfGrid
    g = SrfGrid(8, 8, 8, 0, 0, 0, 16, 16, 16)
    assert L.srf_project(None, 0, 6, None, 0, None, g, None, None, None) == 0
    assert L.srf_ratio(None, None, None, 0, 0, 0.0, None) == 0
    assert L.srf_em_update(None, None, None, 0, 0.0, None) == 0
    assert L.srf_lor_digest(None, 0, 6, None, 0, g, None, 0, None, None, None) == 0
    assert L.srf_event_order(None, 0, 6, None, 0, 4, None, None) == 0
