"""Loader for libimexlb.so (the C ABI declared in include/imexlb.h).

The library is built in-tree (``imexlb-1.0_amd/libimexlb.so``).  There is no
fallback: if it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "libimexlb.so")
_LIB = None

OK = 0
E_ARG, E_HIP, E_OOM, E_STATE, E_UB, E_NODEV, E_NAN = -1, -2, -3, -4, -5, -6, -7
F64, F32, F16_F32, BF16_F32, BF16SR_F32 = 0, 1, 2, 3, 4
LAYOUT_REF, LAYOUT_INTERIOR = 0, 1
PDF_F, PDF_FB = 0, 1
LISTS = {"fluid_id": 0, "bl": 1, "br": 2, "bu": 3, "bd": 4, "bf": 5, "bb": 6, "b_user": 7}

# every symbol include/imexlb.h declares, with (restype, argtypes)
_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_I32P = C.POINTER(C.c_int32)
_DP = C.POINTER(C.c_double)
SIGNATURES = {
    "imexlb_strerror": (C.c_char_p, [C.c_int]),
    "imexlb_last_error": (C.c_char_p, []),
    "imexlb_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "imexlb_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "imexlb_destroy": (C.c_int, [_P]),
    "imexlb_set_params": (C.c_int, [_P, C.c_double, C.c_double, C.c_double]),
    "imexlb_set_geometry": (C.c_int, [_P, C.c_double, C.c_double]),
    "imexlb_set_geometry_shape": (C.c_int, [_P, C.c_double, C.c_double, C.c_double, C.c_double]),
    "imexlb_geometry_size": (C.c_int, [_P, C.c_int, _I64P]),
    "imexlb_geometry_list": (C.c_int, [_P, C.c_int, _I32P, C.c_int64]),
    "imexlb_geometry_geo": (C.c_int, [_P, _I32P, C.c_int64]),
    "imexlb_bc_table": (C.c_int, [_P, _I64P, _I64P, C.c_int64, _I64P]),
    "imexlb_init_up": (C.c_int, [_P, _DP]),
    "imexlb_upload_fields": (C.c_int, [_P, _DP, _DP, C.c_int]),
    "imexlb_init_fields_f32": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_double, C.c_double, _DP]),
    "imexlb_body_mask_f32": (C.c_int, [_P, _P, C.c_int, C.c_int]),
    "imexlb_init_pdf": (C.c_int, [_P]),
    "imexlb_collision": (C.c_int, [_P]),
    "imexlb_boundary_condition": (C.c_int, [_P]),
    "imexlb_propagation": (C.c_int, [_P]),
    "imexlb_post_processing": (C.c_int, [_P]),
    "imexlb_set_fused": (C.c_int, [_P, C.c_int]),
    "imexlb_advance": (C.c_int, [_P, C.c_int64]),
    "imexlb_monitor": (C.c_int, [_P, _DP]),
    "imexlb_forces": (C.c_int, [_P, _DP]),
    "imexlb_mass": (C.c_int, [_P, _DP]),
    "imexlb_download_fields": (C.c_int, [_P, _DP, _DP, C.c_int]),
    "imexlb_download_pdf": (C.c_int, [_P, C.c_int, _DP, C.c_int]),
    "imexlb_synchronize": (C.c_int, [_P]),
    "imexlb_profile": (C.c_int, [_P, C.c_int]),
    "imexlb_profile_read": (C.c_int, [_P, _I64P, _DP]),
    "imexlb_info": (C.c_int, [_P, _I64P]),
    "imexlb_batch_create": (C.c_int, [C.POINTER(_P), C.c_int, _I32P, _DP, C.c_int, C.c_int]),
    "imexlb_batch_create_shapes": (C.c_int, [C.POINTER(_P), C.c_int, _I32P, _DP, _DP, C.c_int, C.c_int]),
    "imexlb_batch_destroy": (C.c_int, [_P]),
    "imexlb_batch_init": (C.c_int, [_P, _DP]),
    "imexlb_batch_init_fields": (C.c_int, [_P, _DP, _DP]),
    "imexlb_batch_init_fields_f32": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_double, C.c_double, _DP]),
    "imexlb_batch_body_mask_f32": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int64]),
    "imexlb_batch_advance": (C.c_int, [_P, C.c_int64]),
    "imexlb_batch_set_config": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "imexlb_batch_tune": (C.c_int, [_P, C.c_int, _I32P, _DP]),
    "imexlb_batch_synchronize": (C.c_int, [_P]),
    "imexlb_batch_forces": (C.c_int, [_P, _DP]),
    "imexlb_batch_monitor": (C.c_int, [_P, _DP]),
    "imexlb_batch_download_fields": (C.c_int, [_P, C.c_int, _DP, _DP]),
    "imexlb_batch_info": (C.c_int, [_P, _I64P]),
    # warm-start network kernels (include/imexlb_nn.h); float pointers are device addresses
    "imexlb_nn_conv3x3": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_flip_weights": (C.c_int, [_P, _P, C.c_int, C.c_int, _P]),
    "imexlb_nn_conv3x3_dw": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_conv3x3_dw_work": (C.c_int64, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "imexlb_nn_conv3x3_dw_ordered": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                               _P]),
    "imexlb_nn_conv3x3_dw_ordered_noalloc": (C.c_int, [_P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                                       _P]),
    "imexlb_nn_bias_grad":(C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_relu_bwd": (C.c_int, [_P, _P, _P, C.c_int64, _P]),
    "imexlb_nn_pool2": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_pool2_bwd": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_up2": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_up2_bwd": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_copy_channels": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.c_int, _P]),
    "imexlb_nn_mse": (C.c_int, [_P, _P, _P, C.c_int64, _P, _P]),
    "imexlb_nn_adam": (C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                                 _P]),
    "imexlb_nn_pack_input_bf16": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_pack_weights_bf16": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_conv3x3_bf16": (C.c_int, [_P, _P, _P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                         C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_pool2_bf16": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "imexlb_nn_up2_bf16": (C.c_int, [_P, _P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
}


class ImexlbError(RuntimeError):
    def __init__(self, code, where):
        L = lib()
        msg = L.imexlb_strerror(code).decode()
        detail = L.imexlb_last_error().decode()
        super().__init__(f"{where}: {msg} ({code}): {detail}")
        self.code = code


def lib():
    """Load libimexlb.so, raising if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libimexlb.so not built at {LIB_PATH}: run __graft_entry__.build()")
        # PyTorch ships its own HIP runtime.  Loaded after this library has
        # started the system one, it finds no GPU; loaded first, both work
        # in one process.  So torch, when installed, is imported first.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def check(rc, where):
    if rc != OK:
        raise ImexlbError(rc, where)


def device_count():
    n = C.c_int(0)
    check(lib().imexlb_device_count(C.byref(n)), "imexlb_device_count")
    return n.value
