"""CPU oracle for the list-mode PET hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg may
import this module.  The product package ``srf_amd`` never does.

It wraps ``libsrf_oracle.so`` (srf_oracle.c, a C restatement of the reference
device functions in srfnef/cc/siddon/siddon.cu.cc and
srfnef/cc/siddon_tof/siddon_tof.cu.cc) and restates the reference's Python
drivers with numpy float32 arithmetic:

  Mlem          srfnef/functions/mlem.py:28-45
  MlemFull      srfnef/functions/mlem_full.py:47-70
  MlemFullTof   srfnef/functions/mlem_full_tof.py:43-67
  OsemFull      srfnef/functions/osem_full.py:49-77 (subset loop)
  attenuation   srfnef/corrections/attenuation/attenuation_correct.py:20-40

Images passed to these functions are in the Python layout (nx, ny, nz); like
the reference wrappers (srfnef/functions/project.py:26-27) they are transposed
to the kernel layout (x fastest) before the C marcher runs.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FP, BP = 0, 1


def build(force=False):
    so = os.path.join(_HERE, "libsrf_oracle.so")
    if force or not os.path.exists(so):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return so


def lib():
    global _LIB
    if _LIB is None:
        so = os.path.join(_HERE, "libsrf_oracle.so")
        if not os.path.exists(so):
            build()
        L = ctypes.CDLL(so)
        P = ctypes.c_void_p
        i64, i32, f32 = ctypes.c_int64, ctypes.c_int, ctypes.c_float
        L.og_project.argtypes = [P, i64, i64, P, i64, P, P, P, P, i32, f32, f32, i32, P, P, i32]
        L.og_backproject.argtypes = [P, i64, i64, P, i64, P, P, P, P, i32, f32, f32, i32, P, P, i32]
        L.og_digest.argtypes = [P, i64, i64, P, i64, P, P, P, i32, f32, f32, i32, P, P, i32]
        L.og_terms.argtypes = [P, f32, P, P, P, i32, f32, f32, i32, P, P, P, P, i64]
        L.og_terms.restype = i64
        L.og_num_threads.restype = i32
        for fn in (L.og_project, L.og_backproject, L.og_digest):
            fn.restype = i32
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _geom(shape, center, size):
    return (np.asarray(shape, np.int32), np.asarray(center, np.float32),
            np.asarray(size, np.float32))


def _lors(lors):
    lors = np.ascontiguousarray(lors, dtype=np.float32)
    assert lors.ndim == 2 and lors.shape[1] >= 6
    return lors


def to_kernel_layout(img):
    """(nx, ny, nz) C-order -> flat x-fastest, as tf.transpose in project.py:27."""
    return np.ascontiguousarray(np.transpose(np.asarray(img, np.float32), (2, 1, 0)))


def from_kernel_layout(flat, shape):
    nx, ny, nz = shape
    return np.ascontiguousarray(np.asarray(flat).reshape(nz, ny, nx).transpose(2, 1, 0))


def project(image, lors, center, size, tof=None, variant=FP, f64=False, threads=0):
    """Projection of `image` (nx,ny,nz) along `lors` (N,6|7).
    tof: None (siddon.cu.cc) or (tof_bin, tof_sigma2) (siddon_tof.cu.cc)."""
    image = np.asarray(image, np.float32)
    g, c, s = _geom(image.shape, center, size)
    lors = _lors(lors)
    n = lors.shape[0]
    kimg = to_kernel_layout(image)
    out32 = np.zeros(n, np.float32)
    out64 = np.zeros(n, np.float64)
    tb, ts = (tof if tof is not None else (0.0, 0.0))
    tofp = _p(lors[:, 6:7]) if tof is not None else None
    if tof is not None:
        assert lors.shape[1] >= 7
        tofp = ctypes.c_void_p(lors.ctypes.data + 6 * 4)
    lib().og_project(_p(lors), n, lors.shape[1], tofp, lors.shape[1], _p(kimg), _p(g), _p(c), _p(s),
                     int(tof is not None), float(tb), float(ts), variant, _p(out32), _p(out64), threads)
    return out64 if f64 else out32


def backproject(values, lors, shape, center, size, tof=None, variant=BP, f64=True, threads=0):
    """Backprojection of `values` (N,) along `lors` into a fresh (nx,ny,nz) image.
    f64=True accumulates in double (the parity target); False in float."""
    g, c, s = _geom(shape, center, size)
    lors = _lors(lors)
    n = lors.shape[0]
    values = np.ascontiguousarray(values, np.float32).reshape(-1)
    assert values.shape[0] == n
    nvox = int(np.prod(shape))
    tb, ts = (tof if tof is not None else (0.0, 0.0))
    tofp = ctypes.c_void_p(lors.ctypes.data + 6 * 4) if tof is not None else None
    if f64:
        img = np.zeros(nvox, np.float64)
        rc = lib().og_backproject(_p(lors), n, lors.shape[1], tofp, lors.shape[1], _p(values), _p(g), _p(c),
                                  _p(s), int(tof is not None), float(tb), float(ts), variant, None, _p(img),
                                  threads)
    else:
        img = np.zeros(nvox, np.float32)
        rc = lib().og_backproject(_p(lors), n, lors.shape[1], tofp, lors.shape[1], _p(values), _p(g), _p(c),
                                  _p(s), int(tof is not None), float(tb), float(ts), variant, _p(img), None,
                                  threads)
    assert rc == 0
    return from_kernel_layout(img, tuple(shape))


def digest(lors, shape, center, size, tof=None, variant=FP, threads=0):
    """Per-event (hash, count) of the (voxel, branch) terms the marcher visits."""
    g, c, s = _geom(shape, center, size)
    lors = _lors(lors)
    n = lors.shape[0]
    h = np.zeros(n, np.uint64)
    cnt = np.zeros(n, np.int32)
    tb, ts = (tof if tof is not None else (0.0, 0.0))
    tofp = ctypes.c_void_p(lors.ctypes.data + 6 * 4) if tof is not None else None
    lib().og_digest(_p(lors), n, lors.shape[1], tofp, lors.shape[1], _p(g), _p(c), _p(s),
                    int(tof is not None), float(tb), float(ts), variant, _p(h), _p(cnt), threads)
    return h, cnt


def terms(lor, shape, center, size, tof=None, variant=FP, cap=4096):
    """Terms of one LOR: (voxel flat index in kernel layout, branch slot, weight, w_tof)."""
    g, c, s = _geom(shape, center, size)
    lor = np.ascontiguousarray(lor, np.float32).reshape(-1)
    vox = np.zeros(cap, np.int64)
    slot = np.zeros(cap, np.int32)
    w = np.zeros(cap, np.float32)
    wt = np.zeros(cap, np.float32)
    tb, ts = (tof if tof is not None else (0.0, 0.0))
    tv = float(lor[6]) if (tof is not None and lor.shape[0] > 6) else 0.0
    cnt = lib().og_terms(_p(lor), tv, _p(g), _p(c), _p(s), int(tof is not None), float(tb), float(ts),
                         variant, _p(vox), _p(slot), _p(w), _p(wt), cap)
    assert cnt <= cap
    return vox[:cnt], slot[:cnt], w[:cnt], wt[:cnt]


def kernel_index_to_xyz(v, shape):
    nx, ny, nz = shape
    v = np.asarray(v)
    return v % nx, (v // nx) % ny, v // (nx * ny)


# ---------------------------------------------------------------- drivers --

def mlem(values, lors, emap, center, size, n_iter, tof=None, driver="mlem", threads=0):
    """EM loop restated from the reference drivers, in numpy float32.

    driver='mlem'      srfnef/functions/mlem.py:33-43   q = y / (p + mean(y)*1e-8)
    driver='full'      srfnef/functions/mlem_full.py    q = div_no_nan(y, p)
    driver='full_tof'  srfnef/functions/mlem_full_tof.py:62-67  q = div_no_nan(y, p + 1e-5)
    Each: x0 = 1; emap zeros -> 1e8; x <- x * BP(q) / emap.
    The backprojection is accumulated in double and rounded to float (the
    reference's atomicAdd order is unspecified)."""
    emap = np.array(emap, np.float32, copy=True)
    emap[emap == 0.0] = 1e8
    x = np.ones(emap.shape, np.float32)
    y = np.asarray(values, np.float32)
    for _ in range(n_iter):
        p = project(x, lors, center, size, tof=tof, threads=threads)
        q = ratio(y, p, driver)
        bp = backproject(q, lors, emap.shape, center, size, tof=tof, threads=threads).astype(np.float32)
        x = x * bp / emap
    return x


def ratio(y, p, driver):
    y = np.asarray(y, np.float32)
    p = np.asarray(p, np.float32)
    with np.errstate(divide="ignore", invalid="ignore"):
        if driver == "mlem":
            return y / (p + np.float32(np.mean(y) * 1e-8))
        if driver == "full":
            d = p
        elif driver == "full_tof":
            d = p + np.float32(1e-5)
        elif driver == "osem":
            return y / (p + np.float32(1e-8))
        else:
            raise ValueError(driver)
        return np.where(d == 0, np.float32(0), y / d).astype(np.float32)


def osem(values, lors, emap, center, size, n_iter, n_sub, perm, tof=None, driver="osem",
         reference_subset_order=False, guard_emap=False):
    """OSEM restated from srfnef/functions/osem_full.py:60-77.

    perm: the event permutation (the reference shuffles with the unseeded global
    RNG, osem_full.py:63-64; here it is passed in).  n_per_sub = N // n_sub
    (osem_full.py:65).  Sub-iteration i uses subset i % n_sub, or i // n_sub when
    reference_subset_order (osem_full.py:68).  The emap is neither guarded nor
    scaled per subset (osem_full.py:66,77) unless guard_emap."""
    emap = np.array(emap, np.float32, copy=True)
    if guard_emap:
        emap[emap == 0.0] = 1e8
    x = np.ones(emap.shape, np.float32)
    y = np.asarray(values, np.float32)
    lors = _lors(lors)
    n_per_sub = len(y) // n_sub
    for i in range(n_iter):
        ind = (i // n_sub) if reference_subset_order else (i % n_sub)
        idx = perm[ind * n_per_sub:(ind + 1) * n_per_sub]
        sl, sy = lors[idx], y[idx]
        p = project(x, sl, center, size, tof=tof)
        q = ratio(sy, p, driver)
        bp = backproject(q, sl, emap.shape, center, size, tof=tof).astype(np.float32)
        with np.errstate(divide="ignore", invalid="ignore"):
            x = x * bp / emap
    return x


def attenuation_factor(umap, lors, center, size, physical=False):
    """exp(-Project(u)*L^2/unit_size[0]) (attenuation_correct.py:21-28, 34-40)."""
    lors = _lors(lors)
    p = project(umap, lors[:, :6], center, size)
    d = lors[:, 0:3] - lors[:, 3:6]
    L = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2)
    unit = np.float32(size[0] / np.asarray(umap).shape[0])
    if physical:
        return np.exp(-(p * L * L / unit / unit)).astype(np.float32)
    return np.exp(-(p * L * L / unit)).astype(np.float32)


def num_threads():
    return int(lib().og_num_threads())
