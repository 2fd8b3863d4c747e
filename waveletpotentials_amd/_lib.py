"""ctypes binding of libwhhd.so (the HIP library; C ABI declared in include/whhd.h).

The library is built in-tree by ``waveletpotentials_amd.build``.  Loading it
does not need a GPU; calling a compute entry point does.  There is no CPU
fallback: if the library is missing, :func:`lib` raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwhhd.so")

STATUS = {0: "WHHD_OK", 1: "WHHD_INVALID_ARG", 2: "WHHD_NOT_POW2", 3: "WHHD_UNSUPPORTED_WAVELET",
          4: "WHHD_HIP_ERROR", 5: "WHHD_OOM", 6: "WHHD_NONCONVERGED"}

# every function include/whhd.h declares, with (restype, argtypes)
_P, _I, _C, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_char, ctypes.c_double
_S = ctypes.c_char_p
SIGNATURES = {
    "whhd_write_txt": (_I, [_S, _P, _I, _I, _I, _I, _I]),
    "whhd_read_txt": (_I, [_S, _P, _I, _I, _I]),
    "whhd_write_txt_sparse": (_I, [_S, _P, _I, _I, _I, _D]),
    "whhd_write_vector_txt_sparse": (_I, [_S, _P, _P, _P, _I, _I, _I, _D]),
    "whhd_write_binary": (_I, [_S, _P, ctypes.c_longlong]),
    "whhd_read_binary": (_I, [_S, _P, ctypes.c_longlong]),
    "whhd_write_ply": (_I, [_S, _P, _I, _I, _I, _D, _D, _D, _D, _I, ctypes.POINTER(ctypes.c_longlong)]),
    "whhd_set_slice_zero": (_I, [_P, _I, _I, _I, ctypes.c_char]),
    "whhd_sweep2d_create": (_I, [_I, _I, _D, _D, ctypes.c_char, ctypes.POINTER(_P)]),
    "whhd_sweep2d_solve": (_I, [_P, _P, _P, _P, _P]),
    "whhd_gradient_stream3d": (_I, [_P] * 6 + [_I, _I, _I, _D, _D, _D]),
    "whhd_gradient_stream2d": (_I, [_P] * 3 + [_I, _I, _D, _D]),
    "whhd_error_l2": (_I, [ctypes.POINTER(_P), ctypes.POINTER(_P),
                           ctypes.POINTER(ctypes.c_longlong), _I, ctypes.POINTER(_D)]),
    "whhd_status_string": (_S, [_I]),
    "whhd_last_error": (_S, []),
    "whhd_version": (_S, []),
    "whhd_proj_bc": (_C, [_C]),
    "whhd_proj_wavelet": (_I, [_I]),
    "whhd_wavelet_supported": (_I, [_I, _C]),
    "whhd_set_line_kernel": (_I, [_I]),
    "whhd_transform1d": (_I, [_P, _P, _I, _I, _I, _I, _C, _C, _I, _I, _P]),
    "whhd_transform1d_f64": (_I, [_P, _P, _I, _I, _I, _I, _C, _C, _I, _I, _P]),
    "whhd_gram_table": (_I, [_P, _I, _C, _I, _P]),
    "whhd_project3d_create": (_I, [_I, _I, _I, _D, _D, _D, _I, _C, ctypes.POINTER(_P)]),
    "whhd_project3d_run": (_I, [_P] * 10 + [_P]),
    "whhd_project3d_destroy": (None, [_P]),
    "whhd_project2d_create": (_I, [_I, _I, _D, _D, _I, _C, ctypes.POINTER(_P)]),
    "whhd_project2d_run": (_I, [_P, _P, _P, _P, _P]),
    "whhd_project2d_destroy": (None, [_P]),
    "whhd_recover3d_create": (_I, [_I, _I, _I, _D, _D, _D, _C, _I, _I, ctypes.POINTER(_P)]),
    "whhd_recover3d_scratch_bytes": (_I, [_I, _I, _I, ctypes.POINTER(ctypes.c_ulonglong)]),
    "whhd_recover3d_solve": (_I, [_P] * 7 + [_P]),
    "whhd_recover3d_solve_proj": (_I, [_P] * 10 + [_P]),
    "whhd_recover3d_cut": (_I, [_P] * 7 + [_I, _P]),
    "whhd_recover3d_trace_begin": (_I, [_P, _I]),
    "whhd_recover3d_trace_end": (_I, [_P, _P, _I, _P, _P]),
    "whhd_recover3d_trace_label": (_S, [_P, _I]),
    "whhd_recover3d_destroy": (None, [_P]),
    "whhd_batch3d_create": (_I, [_I, _P, _P, _C, _I, _I, ctypes.POINTER(_P)]),
    "whhd_batch3d_layout": (_I, [_P, _P, _P]),
    "whhd_batch3d_solve": (_I, [_P, _P, _P]),
    "whhd_batch3d_capture": (_I, [_P, _P]),
    "whhd_batch3d_replay": (_I, [_P, _P]),
    "whhd_batch3d_autotune": (_I, [_P, _P, _I, _P]),
    "whhd_batch3d_num_classes": (_I, [_P]),
    "whhd_batch3d_class_info": (_I, [_P, _I, _P]),
    "whhd_batch3d_set_tiles": (_I, [_P, _I, _I, _I]),
    "whhd_batch3d_destroy": (None, [_P]),
    "whhd_recover2d_create": (_I, [_I, _I, _D, _D, _C, _I, _I, ctypes.POINTER(_P)]),
    "whhd_recover2d_solve": (_I, [_P, _P, _P, _P, _P]),
    "whhd_recover2d_destroy": (None, [_P]),
    "whhd_levelset_spheres": (_I, [_P, _I, _I, _I, _D, _D, _D, _P, _P, _I, _P]),
    "whhd_frac_from_ls": (_I, [_P, _P, _P, _P, _I, _I, _I, _D, _D, _D, _P]),
    "whhd_frac_divide": (_I, [_P, _P, _P, _P, _I, _I, _I, _D, _D, _D, _P]),
    "whhd_eikonal_sweep": (_I, [_P, _P, _I, _I, _I, _D, _I, _P]),
    "whhd_mgfrac3d_create": (_I, [_P, _P, _P, _I, _I, _I, _D, _D, _D, _C, ctypes.POINTER(_P)]),
    "whhd_mgfrac3d_levels": (_I, [_P, _P, _I]),
    "whhd_mgfrac3d_solve": (_I, [_P, _P, _P, _P, _I, _I, _I, ctypes.c_float, _P]),
    "whhd_mgfrac3d_solve_tol": (_I, [_P, _P, _P, _P, _P, _I, _D, _I, _I, ctypes.c_float, _P, _P,
                                     _P]),
    "whhd_mgfrac3d_solve_pcg": (_I, [_P, _P, _P, _P, _P, _I, _D, _I, _I, ctypes.c_float, _P, _P,
                                     _P]),
    "whhd_mgfrac3d_residual": (_I, [_P, _P, _P, ctypes.POINTER(_D), _P]),
    "whhd_mgfrac3d_destroy": (None, [_P]),
    "whhd_unet_create": (_I, [_I, _I, _I, _I, ctypes.c_ulonglong, ctypes.POINTER(_P)]),
    "whhd_unet_num_params": (ctypes.c_longlong, [_P]),
    "whhd_unet_get_params": (_I, [_P, _P, _I]),
    "whhd_unet_set_params": (_I, [_P, _P]),
    "whhd_unet_predict": (_I, [_P, _P, _P, _P, _D, _P]),
    "whhd_unet_loss_grad": (_I, [_P, _P, _P, _P, _D, ctypes.POINTER(_D), _P]),
    "whhd_unet_adam": (_I, [_P, ctypes.c_float, _P]),
    "whhd_unet_set_output_scale": (_I, [_P, _D]),
    "whhd_unet_output_scale": (_D, [_P]),
    "whhd_unet_destroy": (None, [_P]),
    "whhd_gradient3d_frac": (_I, [_P] * 7 + [_I, _I, _I, _D, _D, _D, _C, _P]),
    "whhd_subtract_gradient_frac": (_I, [_P] * 7 + [_I, _I, _I, _D, _D, _D, _C, _P]),
    "whhd_subtract_gradient_frac2": (_I, [_P] * 8 + [_I, _I, _I, _D, _D, _D, _C, _P]),
    "whhd_divergence3d_frac": (_I, [_P] * 7 + [_I, _I, _I, _D, _D, _D, _P]),
    "whhd_axpy": (_I, [_P, _P, ctypes.c_longlong, ctypes.c_float, _P]),
    "whhd_mul": (_I, [_P, _P, _P, ctypes.c_longlong, _P]),
    "whhd_reduce": (_I, [_P, ctypes.c_longlong, _I, _P, ctypes.POINTER(_D), _P]),
    "whhd_reduce_scratch_doubles": (_I, []),
    "whhd_reduce_noalloc": (_I, [_P, ctypes.c_longlong, _I, ctypes.POINTER(_D), _P]),
    "whhd_dst3d_create": (_I, [_I, _I, _I, _D, _D, _D, ctypes.POINTER(_P)]),
    "whhd_dst3d_solve": (_I, [_P, _P, _P, _P]),
    "whhd_dst3d_destroy": (None, [_P]),
    "whhd_dst3d_create_f64": (_I, [_I, _I, _I, _D, _D, _D, ctypes.POINTER(_P)]),
    "whhd_dst2d_create_f64": (_I, [_I, _I, _D, _D, ctypes.POINTER(_P)]),
    "whhd_dst3d_solve_f64": (_I, [_P, _P, _P, _P]),
    "whhd_sweep3d_create_f64": (_I, [_I, _I, _I, _D, _D, _D, ctypes.c_char, ctypes.POINTER(_P)]),
    "whhd_sweep3d_solve_f64": (_I, [_P] * 8),
    "whhd_sweep3d_project_f64": (_I, [_P] * 5),
    "whhd_sweep2d_create_f64": (_I, [_I, _I, _D, _D, ctypes.c_char, ctypes.POINTER(_P)]),
    "whhd_sweep2d_solve_f64": (_I, [_P] * 5),
    "whhd_dst2d_create": (_I, [_I, _I, _D, _D, ctypes.POINTER(_P)]),
    "whhd_sweep3d_create": (_I, [_I, _I, _I, _D, _D, _D, _C, ctypes.POINTER(_P)]),
    "whhd_sweep3d_solve": (_I, [_P] * 7 + [_P]),
    "whhd_sweep3d_project": (_I, [_P] * 4 + [_P]),
    "whhd_sweep3d_destroy": (None, [_P]),
}

_lib = None


class WhhdError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = ""
        if _lib is not None:
            msg = (_lib.whhd_last_error() or b"").decode(errors="replace")
        super().__init__(f"{where}: {STATUS.get(status, status)}: {msg}")


def lib():
    """Load libwhhd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m waveletpotentials_amd.build` "
                               "(or __graft_entry__.build()) first; there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != 0:
        raise WhhdError(status, where)
    return status
