"""Mixed-code queue decoder over the C ABI (fecsim_multi_*): one persistent launch
decodes codewords of several codes, each with its own decoder parameters.  No
counterpart in the reference."""
from __future__ import annotations

import ctypes as C

from . import lib as _L
from .decoder import Decoder, _ptr


def _stream(stream):
    import torch
    return C.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)


class MultiDecoder:
    def __init__(self, decoders, max_cw: int):
        self.decoders = list(decoders)          # the handles must outlive this object
        arr = (C.c_void_p * len(self.decoders))(*[d._d.value for d in self.decoders])
        m = C.c_void_p()
        _L.check(_L.lib().fecsim_multi_create(arr, len(self.decoders), max_cw, C.byref(m)), "fecsim_multi_create")
        self._m = m
        self.max_cw = max_cw

    def decode(self, code, off, llr_in, post_out, iters=None, tot_iter=None, stream=None):
        """code: int32 [n], off: int64 [n] element offsets, llr_in/post_out: flat packed buffers."""
        n = code.numel()
        _L.check(_L.lib().fecsim_multi_decode(self._m, n, _ptr(code), _ptr(off), _ptr(llr_in), _ptr(post_out),
                                              _ptr(iters), _ptr(tot_iter), _stream(stream)), "fecsim_multi_decode")

    def decode_static(self, code, off, llr_in, post_out, iters=None, tot_iter=None, stream=None):
        """decode() through fecsim_multi_decode_static: the queue workspace is the library's
        static device array (no allocation; one call at a time per GPU)."""
        n = code.numel()
        _L.check(_L.lib().fecsim_multi_decode_static(self._m, n, _ptr(code), _ptr(off), _ptr(llr_in),
                                                     _ptr(post_out), _ptr(iters), _ptr(tot_iter), _stream(stream)),
                 "fecsim_multi_decode_static")

    def autotune(self, code, off, llr_in, post_out, rounds: int = 3, stream=None):
        _L.check(_L.lib().fecsim_multi_autotune(self._m, code.numel(), _ptr(code), _ptr(off), _ptr(llr_in),
                                                _ptr(post_out), rounds, _stream(stream)), "fecsim_multi_autotune")

    def shape(self, c: int):
        """(rows per lane, codewords per wave, waves per block) used for code c."""
        k, cpw, w = C.c_int(), C.c_int(), C.c_int()
        _L.check(_L.lib().fecsim_multi_shape(self._m, c, C.byref(k), C.byref(cpw), C.byref(w)), "fecsim_multi_shape")
        return k.value, cpw.value, w.value

    def select_kernel(self, kernel: int) -> None:
        """lib.KERNEL_AUTO, KERNEL_GENERIC or KERNEL_CODE (the codes compiled in)."""
        _L.check(_L.lib().fecsim_multi_select_kernel(self._m, kernel), "fecsim_multi_select_kernel")

    @property
    def kernel(self) -> int:
        """lib.KERNEL_GENERIC or KERNEL_CODE: the kernel decode() runs now."""
        k = _L.lib().fecsim_multi_kernel(self._m)
        if k < 0:
            _L.check(k, "fecsim_multi_kernel")
        return k

    def __del__(self):
        if getattr(self, "_m", None) and _L._lib is not None:
            _L.lib().fecsim_multi_free(self._m)
            self._m = None


class MsimParams(C.Structure):
    _fields_ = [("num_cw", C.c_uint32), ("ebn0_lo", C.c_float), ("ebn0_hi", C.c_float), ("llr_scale", C.c_float),
                ("max_abs_llr", C.c_float), ("seed", C.c_uint64), ("use_graph", C.c_int)]


class MixedSim:
    """Mixed-code simulation loop (fecsim_msim_*): encoders hencs[c] for the queue
    decoder's codes, a fixed queue of num_cw entries with uniform codes and Eb/N0."""

    def __init__(self, hencs, multi: MultiDecoder, num_cw: int, ebn0=(1.0, 4.0), llr_scale: float = 2.0,
                 max_abs_llr: float = 24.0, seed: int = 1, use_graph: bool = True):
        self.hencs, self.multi = list(hencs), multi
        arr = (C.c_void_p * len(self.hencs))(*[h._h.value for h in self.hencs])
        p = MsimParams(num_cw, ebn0[0], ebn0[1], llr_scale, max_abs_llr, seed, int(use_graph))
        s = C.c_void_p()
        _L.check(_L.lib().fecsim_msim_create(arr, multi._m, C.byref(p), C.byref(s)), "fecsim_msim_create")
        self._s = s
        self.num_cw = num_cw

    def run(self, loops: int = 1, stream=None):
        _L.check(_L.lib().fecsim_msim_run(self._s, loops, _stream(stream)), "fecsim_msim_run")

    def reset(self, stream=None):
        _L.check(_L.lib().fecsim_msim_reset(self._s, _stream(stream)), "fecsim_msim_reset")

    def fill(self, stream=None):
        """The next loop's channel LLRs into the queue's buffer, without decoding
        (fecsim_msim_fill); llr() then views them."""
        _L.check(_L.lib().fecsim_msim_fill(self._s, _stream(stream)), "fecsim_msim_fill")

    def llr(self):
        """torch int8 view of the packed LLR / posterior buffer (owned by the library)"""
        import torch
        llr = C.c_void_p()
        nb = C.c_ulonglong()
        _L.check(_L.lib().fecsim_msim_queue(self._s, None, None, C.byref(llr), None), "fecsim_msim_queue")
        _L.check(_L.lib().fecsim_msim_llr_bytes(self._s, C.byref(nb)), "fecsim_msim_llr_bytes")
        return torch.as_tensor(_DeviceArray(llr.value, nb.value, "|i1"), device="cuda")

    def stats(self):
        """(per-code list of dicts cw/uncoded_bit_err/bit_err/cw_err, summed iterations)"""
        from .sim import SimStats
        arr = (SimStats * len(self.hencs))()
        it = C.c_ulonglong()
        _L.check(_L.lib().fecsim_msim_stats(self._s, arr, C.byref(it)), "fecsim_msim_stats")
        return [{n: getattr(x, n) for n, _ in SimStats._fields_ if n != "iters"} for x in arr], it.value

    def queue(self):
        """torch views (code ids, LLR offsets, packed LLR bytes, (sigma, factor) pairs in code order)"""
        import torch
        code, off, llr, sf = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
        _L.check(_L.lib().fecsim_msim_queue(self._s, C.byref(code), C.byref(off), C.byref(llr), C.byref(sf)),
                 "fecsim_msim_queue")
        n = self.num_cw
        torch.cuda.synchronize()
        views = [torch.as_tensor(_DeviceArray(p.value, k, t), device="cuda").clone()
                 for p, k, t in ((code, n, "<i4"), (off, n, "<i8"), (sf, 2 * n, "<f4"))]
        return views[0], views[1], None, views[2]

    def __del__(self):
        if getattr(self, "_s", None) and _L._lib is not None:
            _L.lib().fecsim_msim_free(self._s)
            self._s = None


class _DeviceArray:
    """A device buffer owned by the library, exposed to torch (read-only use)."""

    def __init__(self, ptr, n, typestr):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (ptr, False), "version": 3,
                                         "strides": None}
