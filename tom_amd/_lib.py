"""ctypes binding of libtomhip.so (the C ABI declared in include/tomhip.h).

The library is loaded from the package directory only; if it is missing the
import fails loudly: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtomhip.so")

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_int32_p = ctypes.POINTER(ctypes.c_int32)
c_int64_p = ctypes.POINTER(ctypes.c_int64)
vp = ctypes.c_void_p

# name -> (restype, argtypes); every data pointer is passed as void*
SIGNATURES = {
    "tomhip_init": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(vp)]),
    "tomhip_destroy": (None, [vp]),
    "tomhip_last_error": (ctypes.c_char_p, [vp]),
    "tomhip_set_stream": (ctypes.c_int, [vp, vp]),
    "tomhip_synchronize": (ctypes.c_int, [vp]),
    "tomhip_rotate3d_f32": (ctypes.c_int, [vp, vp, vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, vp,
                                           ctypes.c_int, vp]),
    "tomhip_rotate2d_f32": (ctypes.c_int, [vp, vp, vp, ctypes.c_int64, ctypes.c_int64, vp, ctypes.c_int, vp]),
    "tomhip_rotate3d_batch_f32": (ctypes.c_int, [vp, vp, vp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, vp,
                                                 ctypes.c_int64, vp, vp]),
    "tomhip_peak_f32": (ctypes.c_int, [vp, vp, vp, vp, vp]),
    "tomhip_peak_f32_noalloc": (ctypes.c_int, [vp, vp, vp, vp, vp]),
    "tomhip_peak_all_f32": (ctypes.c_int, [vp, vp, vp, vp, vp, ctypes.c_int64, vp]),
    "tomhip_peak_radius_f32": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]),
    "tomhip_flcf_box_f32": (ctypes.c_int, [vp, vp, vp, vp, vp, vp]),
    "tomhip_os3_bestpeak_f32": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_int32, vp, vp, vp, vp, vp, vp, vp]),
    "tomhip_spheremask_os3_f32": (ctypes.c_int, [vp, vp, vp, vp]),
    "tomhip_spheremask_f32": (ctypes.c_int, [vp, vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, vp]),
    "tomhip_weight_f32": (ctypes.c_int, [vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, vp]),
    "tomhip_prepare_template_f32": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp, vp]),
    "tomhip_search_create": (ctypes.c_int, [vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]),
    "tomhip_search_set_template": (ctypes.c_int, [vp, vp, vp, ctypes.c_double, ctypes.c_double, ctypes.c_int]),
    "tomhip_search_run": (ctypes.c_int, [vp, vp, ctypes.c_int64, ctypes.c_int64]),
    "tomhip_search_finalize": (ctypes.c_int, [vp, vp, vp]),
    "tomhip_search_profile": (ctypes.c_int, [vp, ctypes.c_int, vp, vp]),
    "tomhip_search_set_graph": (ctypes.c_int, [vp, ctypes.c_int]),
    "tomhip_search_destroy": (None, [vp]),
    "tomhip_flcf_search_f32": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int, vp,
                                              vp]),
    "tomhip_os3_corr_f32": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, vp]),
    "tomhip_oscar_f32": (ctypes.c_int, [vp, vp, vp, vp, vp, vp, ctypes.c_int64, vp, vp]),
    "tomhip_picklist_f32": (ctypes.c_int, [vp, vp, vp, vp, vp, ctypes.c_int64, ctypes.c_int, vp, vp]),
    "tomhip_search_refine": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, ctypes.c_int64, vp, vp, vp,
                                             vp]),
    "tomhip_local_maxima_f32": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_int64, vp, vp]),
    "tomhip_psr_f32": (ctypes.c_int, [vp, vp, vp, vp, vp]),
    "tomhip_bin_f32": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, vp]),
    "tomhip_search_set_tiles": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int]),
    "tomhip_search_set_zchunk": (ctypes.c_int, [vp, ctypes.c_int]),
    "tomhip_snap_maxima_f32": (ctypes.c_int, [vp, vp, vp, ctypes.c_int, vp, ctypes.c_int64, ctypes.c_int]),
    "tomhip_unet_create": (ctypes.c_int, [vp, vp, ctypes.c_int, ctypes.c_uint64, ctypes.c_int, vp]),
    "tomhip_unet_destroy": (None, [vp]),
    "tomhip_unet_num_params": (ctypes.c_int, [vp, vp]),
    "tomhip_unet_get_params": (ctypes.c_int, [vp, vp]),
    "tomhip_unet_set_params": (ctypes.c_int, [vp, vp]),
    "tomhip_unet_get_optimizer": (ctypes.c_int, [vp, vp, vp, vp]),
    "tomhip_unet_set_optimizer": (ctypes.c_int, [vp, vp, vp, ctypes.c_int64]),
    "tomhip_unet_forward": (ctypes.c_int, [vp, vp, vp]),
    "tomhip_unet_get_logits": (ctypes.c_int, [vp, vp]),
    "tomhip_unet_backward": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_float, vp]),
    "tomhip_unet_get_grads": (ctypes.c_int, [vp, vp]),
    "tomhip_unet_adam": (ctypes.c_int, [vp, ctypes.c_float]),
    "tomhip_unet_predict": (ctypes.c_int, [vp, vp, vp, vp]),
    "tomhip_unet_infer": (ctypes.c_int, [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, vp, vp, vp]),
    "tomhip_em_read_header": (ctypes.c_int, [ctypes.c_char_p, vp]),
    "tomhip_em_read_f32": (ctypes.c_int, [ctypes.c_char_p, vp, ctypes.c_int64, vp]),
    "tomhip_em_write_f32": (ctypes.c_int, [ctypes.c_char_p, vp, vp, vp]),
    "tomhip_em_last_error": (ctypes.c_char_p, []),
    "tomhip_subtomo_score_batch": (ctypes.c_int, [vp, ctypes.c_int, vp, vp, vp, vp, vp, ctypes.c_int64, vp, vp, vp,
                                                  vp, vp, ctypes.c_int64, vp, vp, vp]),
}

_lib = None


def load() -> ctypes.CDLL:
    """Load libtomhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m tom_amd.build` "
                              "(there is no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


class TomHipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"tomhip error {code}: {msg}")
        self.code = code
