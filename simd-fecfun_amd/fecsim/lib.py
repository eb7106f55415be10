"""ctypes binding of libfecsim.so (the C ABI of include/fecsim.h).

The HIP library is the only implementation: if it is missing this module raises,
there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfecsim.so")

I8, I16, I32, F32, F64 = 0, 1, 2, 3, 4
MIN, PROD = 0, 1
KERNEL_AUTO, KERNEL_GENERIC, KERNEL_PAIR, KERNEL_CODE, KERNEL_FLOAT, KERNEL_PAIR_RT = 0, 1, 2, 3, 4, 5

# (name, restype, argtypes) of every exported entry point
_vp, _u32, _i32, _int, _u64 = C.c_void_p, C.c_uint32, C.c_int32, C.c_int, C.c_uint64
SIGNATURES = [
    ("fecsim_abi_version", _int, []),
    ("fecsim_strerror", C.c_char_p, [_int]),
    ("fecsim_last_error", C.c_char_p, []),
    ("fecsim_hdec_load", _int, [C.c_char_p, C.POINTER(_vp)]),
    ("fecsim_hdec_from_words", _int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    ("fecsim_hdec_info", _int, [_vp, _vp]),
    ("fecsim_hdec_free", None, [_vp]),
    ("fecsim_decoder_create", _int, [_vp, C.c_char_p, _vp, C.POINTER(_vp)]),
    ("fecsim_decoder_select_kernel", _int, [_vp, _int]),
    ("fecsim_decoder_kernel", _int, [_vp]),
    ("fecsim_decoder_info", _int, [_vp, C.POINTER(_int), C.POINTER(_int), C.POINTER(_int)]),
    ("fecsim_decoder_free", None, [_vp]),
    ("fecsim_decode_set", _int, [_vp, _u32, _int, _vp, _vp, _vp, _vp, _vp]),
    ("fecsim_decode", _int, [_vp, _u32, _int, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("fecsim_decode_set_host", _int, [_vp, _u32, _int, _vp, _vp, C.POINTER(C.c_ulonglong)]),
    ("fecsim_henc_load", _int, [C.c_char_p, C.POINTER(_vp)]),
    ("fecsim_henc_from_words", _int, [_vp, C.c_size_t, C.POINTER(_vp)]),
    ("fecsim_henc_info", _int, [_vp, _vp]),
    ("fecsim_henc_masks", _int, [_vp, _vp, _vp]),
    ("fecsim_henc_free", None, [_vp]),
    ("fecsim_gen_codewords", _int, [_vp, _u64, _u64, _u32, _int, _vp, _vp]),
    ("fecsim_encode", _int, [_vp, _u32, _int, _vp, _vp]),
    ("fecsim_gen_llr", _int, [_vp, _vp, _u64, _u64, _u32, _int, _vp, _int, _int, _vp, _vp, _vp]),
    ("fecsim_chan_sigma", _int, [_vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("fecsim_count_errors", _int, [_vp, _u32, _int, _vp, _int, _int, _vp, _vp, _vp, _vp]),
    ("fecsim_hard_decisions", _int, [_vp, _u32, _int, _int, _int, _vp, _vp, _vp]),
    ("fecsim_iter_histogram", _int, [_vp, _u32, _u32, _vp, _vp]),
    ("fecsim_sim_iter_hist", _int, [_vp, _vp, _u32]),
    ("fecsim_refrng_create", _int, [_vp, C.POINTER(_vp)]),
    ("fecsim_refrng_free", None, [_vp]),
    ("fecsim_refrng_normdist2", _int, [_vp, _vp, _u32, _vp]),
    ("fecsim_refrng_databits", _int, [_vp, _vp, _u32, _int, _vp, _vp]),
    ("fecsim_gen_llr_from_noise", _int, [_vp, _vp, _vp, _u32, _int, _vp, _int, _int, _vp, _vp, _vp]),
    ("fecsim_sim_create", _int, [_vp, _vp, _vp, C.POINTER(_vp)]),
    ("fecsim_sim_run", _int, [_vp, _u32, _vp]),
    ("fecsim_sim_stats_get", _int, [_vp, _vp]),
    ("fecsim_sim_reset", _int, [_vp, _vp]),
    ("fecsim_sim_free", None, [_vp]),
    ("fecsim_decoder_set_offsets", _int, [_vp, _vp]),
    ("fecsim_multi_create", _int, [_vp, _int, _u32, C.POINTER(_vp)]),
    ("fecsim_multi_decode", _int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("fecsim_multi_decode_static", _int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("fecsim_sp_math", _int, [_int, _vp, _vp, _u32, _vp]),
    ("fecsim_multi_autotune", _int, [_vp, _u32, _vp, _vp, _vp, _vp, _int, _vp]),
    ("fecsim_multi_shape", _int, [_vp, _int, C.POINTER(_int), C.POINTER(_int), C.POINTER(_int)]),
    ("fecsim_multi_select_kernel", _int, [_vp, _int]),
    ("fecsim_multi_kernel", _int, [_vp]),
    ("fecsim_multi_free", None, [_vp]),
    ("fecsim_msim_create", _int, [_vp, _vp, _vp, C.POINTER(_vp)]),
    ("fecsim_msim_run", _int, [_vp, _u32, _vp]),
    ("fecsim_msim_stats", _int, [_vp, _vp, C.POINTER(C.c_ulonglong)]),
    ("fecsim_msim_queue", _int, [_vp, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    ("fecsim_msim_llr_bytes", _int, [_vp, C.POINTER(C.c_ulonglong)]),
    ("fecsim_msim_fill", _int, [_vp, _vp]),
    ("fecsim_msim_reset", _int, [_vp, _vp]),
    ("fecsim_msim_free", None, [_vp]),
    ("fecsim_gen_llr_var", _int, [_vp, _vp, C.c_float, _u64, _u64, _u32, _int, _vp, _int, _int, _vp, _vp, _vp]),
    ("fecsim_trainer_create", _int, [_vp, _vp, C.POINTER(_vp)]),
    ("fecsim_train_fwd", _int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    ("fecsim_train_bwd", _int, [_vp, _u32, _vp, _vp, _vp, _vp, _vp]),
    ("fecsim_trainer_free", None, [_vp]),
]


class FecsimError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = "%s failed: %s (%d)" % (where, _lib.fecsim_strerror(status).decode(), status)
        detail = _lib.fecsim_last_error().decode()
        if detail:
            msg += ": " + detail
        super().__init__(msg)


_lib = None


def lib():
    """Load libfecsim.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libfecsim.so not built at %s (run simd-fecfun_amd/build.py)" % LIB_PATH)
        # torch ships its own HIP runtime with the same soname (libamdhip64.so.7) as the
        # system one libfecsim.so links.  Loading torch first makes libfecsim.so bind to
        # that runtime, so one runtime owns the device, the streams and the memory that
        # both use; loaded the other way round torch would open a second runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != 0:
        raise FecsimError(status, where)
