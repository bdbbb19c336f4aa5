"""ctypes binding of libsrf_hip.so (include/srf_hip.h).

The library is loaded after ``import torch`` so that it binds to the HIP
runtime torch already loaded (both carry the SONAME libamdhip64.so.7): one
runtime, so torch's stream handles and device pointers are valid here.

There is no fallback: if the library is missing or fails to load, every op
raises.  Build it with ``python -m srf_amd.build``.
"""
import ctypes
import os

import torch  # noqa: F401  (must precede the CDLL load, see above)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SRF_HIP_LIB", os.path.join(_HERE, "libsrf_hip.so"))

SRF_BP_REFERENCE = 0
SRF_BP_ADJOINT = 1
SRF_RATIO_PLAIN = 0
SRF_RATIO_DIV_NO_NAN = 1


class SrfGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("cx", ctypes.c_float), ("cy", ctypes.c_float), ("cz", ctypes.c_float),
                ("sx", ctypes.c_float), ("sy", ctypes.c_float), ("sz", ctypes.c_float)]


class SrfTof(ctypes.Structure):
    _fields_ = [("tof_bin", ctypes.c_float), ("tof_sigma2", ctypes.c_float)]


class SrfPlanOpts(ctypes.Structure):
    _fields_ = [("item_records", ctypes.c_int32), ("tof_window", ctypes.c_float)]


class SrfBatchScan(ctypes.Structure):
    _fields_ = [("image", ctypes.c_void_p), ("emap", ctypes.c_void_p), ("y", ctypes.c_void_p),
                ("q", ctypes.c_void_p)]


class SrfBatchCfg(ctypes.Structure):
    _fields_ = [("persistent", ctypes.c_int32), ("workgroups_per_cu", ctypes.c_int32), ("order", ctypes.c_int32)]


class SrfEcat(ctypes.Structure):
    _fields_ = [("angle_per_block", ctypes.c_double), ("axial_length", ctypes.c_double), ("gap", ctypes.c_double),
                ("block_size_y", ctypes.c_double), ("block_size_z", ctypes.c_double), ("unit_y", ctypes.c_double),
                ("unit_z", ctypes.c_double), ("nb_blocks_per_ring", ctypes.c_int32), ("shape_y", ctypes.c_int32),
                ("nb_crystals_per_block", ctypes.c_int32), ("nb_crystals_per_ring", ctypes.c_int32)]


class SrfCyl(ctypes.Structure):
    _D2 = ctypes.c_double * 2
    _I2 = ctypes.c_int32 * 2
    _fields_ = [("angle_per_rsector", ctypes.c_double), ("sz_rsector", _D2), ("mv_module", _D2), ("sz_module", _D2),
                ("mv_submodule", _D2), ("sz_submodule", _D2), ("mv_crystal", _D2), ("nb_rsector", ctypes.c_int32),
                ("nb_crystal", _I2), ("nb_submodule", _I2), ("nb_module", _I2),
                ("nb_crystal_per_submodule", ctypes.c_int32), ("nb_crystal_per_module", ctypes.c_int32),
                ("nb_crystal_per_rsector", ctypes.c_int32)]


class SrfError(RuntimeError):
    pass


# (name, restype, argtypes) -- every symbol declared in include/srf_hip.h
_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_F = ctypes.c_float
_TOFP = ctypes.POINTER(SrfTof)
SIGNATURES = [
    ("srf_strerror", ctypes.c_char_p, [ctypes.c_int]),
    ("srf_version", ctypes.c_int, []),
    ("srf_project", ctypes.c_int, [_P, _I64, _I64, _P, _I64, _P, SrfGrid, _TOFP, _P, _P]),
    ("srf_backproject", ctypes.c_int, [_P, _I64, _I64, _P, _I64, _P, SrfGrid, _TOFP, ctypes.c_int, ctypes.c_int,
                                       _P, _P]),
    ("srf_backproject_ratio", ctypes.c_int, [_P, _I64, _I64, _P, _I64, _P, _P, ctypes.c_int, _F, SrfGrid, _TOFP,
                                             ctypes.c_int, ctypes.c_int, _P, _P]),
    ("srf_ratio", ctypes.c_int, [_P, _P, _P, _I64, ctypes.c_int, _F, _P]),
    ("srf_em_update", ctypes.c_int, [_P, _P, _P, _I64, _F, _P]),
    ("srf_lor_digest", ctypes.c_int, [_P, _I64, _I64, _P, _I64, SrfGrid, _TOFP, ctypes.c_int, _P, _P, _P]),
    ("srf_smooth", ctypes.c_int, [_P, _I32, _I32, _I32, _F, _F, _F, _P, _P, _P, _P]),
    ("srf_smooth_noalloc", ctypes.c_int, [_P, _I32, _I32, _I32, _F, _F, _F, _P, _P, _P]),
    ("srf_pair_lors", ctypes.c_int, [_P, _I64, _I64, _I64, _P, _P]),
    ("srf_psf_xy", ctypes.c_int, [_P, _I32, _I32, _I32, _P, _I32, _I32, _P, _F, _P, _P]),
    ("srf_crystal_index", ctypes.c_int, [_P, _I64, _I64, _I32, ctypes.POINTER(SrfEcat), ctypes.POINTER(SrfCyl), _P,
                                         _P]),
    ("srf_snap_lors", ctypes.c_int, [_P, _I64, _I64, _P, _P, _I64, _P, _P, _P]),
    ("srf_pair_sinogram", ctypes.c_int, [_P, _P, _I64, _I64, _P, _P, _P]),
    ("srf_sinogram_lors", ctypes.c_int, [_P, _I64, _P, _P, _P, _I64, ctypes.POINTER(_I64), _P]),
    ("srf_image_to_kernel_layout", ctypes.c_int, [_P, _P, _I32, _I32, _I32, _P]),
    ("srf_image_from_kernel_layout", ctypes.c_int, [_P, _P, _I32, _I32, _I32, _P]),
    ("srf_plan_create", ctypes.c_int, [_P, _I64, _I64, _P, _I64, SrfGrid, _TOFP, ctypes.POINTER(_P), _P]),
    ("srf_plan_create_ex", ctypes.c_int, [_P, _I64, _I64, _P, _I64, SrfGrid, _TOFP, ctypes.POINTER(SrfPlanOpts),
                                          ctypes.POINTER(_P), _P]),
    ("srf_plan_rebuild", ctypes.c_int, [_P, _P, _I64, _I64, _P, _I64, SrfGrid, _TOFP, ctypes.POINTER(SrfPlanOpts),
                                         _P]),
    ("srf_plan_destroy", ctypes.c_int, [_P]),
    ("srf_plan_pool_trim", ctypes.c_int, [ctypes.c_size_t]),
    ("srf_event_order", ctypes.c_int, [_P, _I64, _I64, _P, _I64, _I64, _P, _P]),
    ("srf_batch_create", ctypes.c_int, [ctypes.POINTER(_P), ctypes.POINTER(SrfBatchScan), _I32,
                                        ctypes.POINTER(SrfBatchCfg), ctypes.POINTER(_P), _P]),
    ("srf_batch_destroy", ctypes.c_int, [_P]),
    ("srf_batch_info", ctypes.c_int, [_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32), ctypes.POINTER(_I32),
                                      ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_size_t)]),
    ("srf_batch_project_ratio", ctypes.c_int, [_P, ctypes.c_int, _F, _P]),
    ("srf_batch_backproject_update", ctypes.c_int, [_P, _F, _P]),
    ("srf_batch_em_step", ctypes.c_int, [_P, ctypes.c_int, _F, _F, _P]),
    ("srf_plan_info", ctypes.c_int, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                                     ctypes.POINTER(ctypes.c_size_t)]),
    ("srf_plan_project", ctypes.c_int, [_P, _P, _P, _P]),
    ("srf_plan_backproject", ctypes.c_int, [_P, _P, ctypes.c_int, ctypes.c_int, _P, _P]),
    ("srf_plan_backproject_update", ctypes.c_int, [_P, _P, _P, _P, _F, _P]),
    ("srf_plan_backproject_ratio", ctypes.c_int, [_P, _P, _P, ctypes.c_int, _F, ctypes.c_int, ctypes.c_int, _P,
                                                  _P]),
    ("srf_plan_digest", ctypes.c_int, [_P, ctypes.c_int, _P, _P, _P]),
    ("srf_plan_time_next", ctypes.c_int, [_P, _P, _P]),
    ("srf_plan_project_ratio", ctypes.c_int, [_P, _P, _P, ctypes.c_int, _F, _P, _P]),
    ("srf_sss_rows_bytes", ctypes.c_size_t, [_I32, _I32]),
    ("srf_sss_prepare", ctypes.c_int, [_P, _I32, _P, _I32, _P, _P, _P, ctypes.c_size_t, _P]),
    ("srf_sss_sinogram", ctypes.c_int, [_P, _I32, _P, _I32, _P, _I32, _P, _F, _P, _P]),
    ("srf_scatter_correct", ctypes.c_int, [_P, _P, _P, _P, _I64, _I64, _F, _P, _P, _P]),
    ("srf_cnn_conv", ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _F, _P, _P,
                                    _P]),
    ("srf_cnn_conv_wgrad", ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _P, _P]),
    ("srf_cnn_conv_batch", ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32,
                                          _F, _P, _P, _P]),
    ("srf_cnn_conv_up_batch", ctypes.c_int, [_P, _P, _I32, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32,
                                             _I32, _I32, _F, _P, _P, _P]),
    ("srf_cnn_conv_head_batch", ctypes.c_int, [_P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _F, _P,
                                               _P, _P, _P, _P]),
    ("srf_cnn_conv_wgrad_batch", ctypes.c_int, [_P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _P,
                                                _P, _P]),
    ("srf_cnn_flip_weights", ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _P, _I32, _I32, _P]),
    ("srf_cnn_pool", ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _P, _I32, _P]),
    ("srf_cnn_pool_bwd", ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _P, _I32, _P]),
    ("srf_cnn_upsample", ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _P, _I32, _I32, _P]),
    ("srf_cnn_upsample_bwd", ctypes.c_int, [_P, _I32, _I32, _I32, _I32, _I32, _I32, _P, _I32, _P]),
    ("srf_cnn_copy_channels", ctypes.c_int, [_P, _I64, _I32, _I32, _I32, _P, _I32, _I32, _I32, _I32, _P]),
    ("srf_cnn_act_bwd", ctypes.c_int, [_P, _I32, _P, _I32, _I64, _I32, _I32, _F, _P, _P]),
    ("srf_cnn_bias_grad", ctypes.c_int, [_P, _I64, _I32, _I32, _P, _P]),
    ("srf_cnn_mse", ctypes.c_int, [_P, _P, _I64, _F, _P, _P, _P]),
    ("srf_cnn_adam", ctypes.c_int, [_P, _P, _P, _P, _I64, _F, _F, _F, _F, _I32, _F, _P, _P]),
    ("srf_cnn_features", ctypes.c_int, [_P, _P, _P, _I64, _P, _P, _P]),
    ("srf_cnn_cast", ctypes.c_int, [_P, _I64, _I32, _P, _P]),
]

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SrfError(
                f"libsrf_hip.so not found at {LIB_PATH}; build it with `python -m srf_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc, what):
    if rc != 0:
        msg = lib().srf_strerror(rc).decode()
        raise SrfError(f"{what} failed: {msg} ({rc})")


def stream_handle(device=None):
    """hipStream_t of torch's current stream on `device` (as an int)."""
    return torch.cuda.current_stream(device).cuda_stream
