"""Host-side mirror of the reference's decoder interface over the C ABI.

  Hdec.load(path)            <- decode_matrix_allocate()  (src/LDPCDec/decode_matrix_allocate.cpp:12)
  Decoder(hdec, "Vec64c_min", max_iter, max_row_metric, beta_offset)
                             <- the executable's compile-time decoder selection
                                (build/Makefile:11-14, include/simd_defs.h:34-54)
  Decoder.decode_set(...)    <- LDPCdecode_set()          (src/LDPCDec/LDPCdecode.cpp:203-245)

Device buffers are torch tensors on the GPU (torch is used only for device memory
and streams); decode_set_host works on numpy arrays in host memory.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _L
from .codes import DecTables

NP_OF_ELEM = {_L.I8: np.int8, _L.I16: np.int16, _L.I32: np.int32, _L.F32: np.float32, _L.F64: np.float64}


class HdecInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("mrow", "mcol", "z", "total_hc_ones", "max_row_weight", "len_f2b",
                                         "len_b2f")]


class Hdec:
    """Decoder tables (LDPCdecode.raw.dat) held by the library."""

    def __init__(self, handle):
        self._h = handle
        info = HdecInfo()
        _L.check(_L.lib().fecsim_hdec_info(self._h, C.byref(info)), "fecsim_hdec_info")
        self.info = {n: getattr(info, n) for n, _ in HdecInfo._fields_}

    @classmethod
    def load(cls, path: str) -> "Hdec":
        h = C.c_void_p()
        _L.check(_L.lib().fecsim_hdec_load(path.encode(), C.byref(h)), "fecsim_hdec_load")
        return cls(h)

    @classmethod
    def from_bytes(cls, b: bytes) -> "Hdec":
        words = np.frombuffer(b, dtype="<i4").copy()
        h = C.c_void_p()
        _L.check(_L.lib().fecsim_hdec_from_words(words.ctypes.data, words.size, C.byref(h)), "fecsim_hdec_from_words")
        return cls(h)

    @classmethod
    def from_tables(cls, dec: DecTables) -> "Hdec":
        return cls.from_bytes(dec.to_bytes())

    @property
    def n(self):
        return self.info["mcol"] * self.info["z"]

    @property
    def e(self):
        return self.info["total_hc_ones"] * self.info["z"]

    def __del__(self):
        if getattr(self, "_h", None) and _L._lib is not None:
            _L.lib().fecsim_hdec_free(self._h)
            self._h = None


class DecParams(C.Structure):
    _fields_ = [("max_iter", C.c_uint32), ("max_row_metric", C.c_float), ("beta_offset", C.c_float)]


def _row_elems(hdec, layout_p):
    """Elements per codeword in the dense (0) or reference (P > 0) layout."""
    z, mcol = hdec.info["z"], hdec.info["mcol"]
    if layout_p == 0:
        return z * mcol
    return mcol * ((z + layout_p - 1) // layout_p + 2) * layout_p


def _ptr(t):
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


class Decoder:
    """GPU layered decoder of one kind ("Vec64c_min", "Vec16f_prod", ...)."""

    def __init__(self, hdec: Hdec, kind: str = "Vec64c_min", max_iter: int = 10, max_row_metric: float = 31.0,
                 beta_offset: float = 1.0):
        self.hdec = hdec
        self.kind = kind
        self.max_iter = max_iter
        p = DecParams(max_iter, max_row_metric, beta_offset)
        d = C.c_void_p()
        _L.check(_L.lib().fecsim_decoder_create(hdec._h, kind.encode(), C.byref(p), C.byref(d)),
                 "fecsim_decoder_create")
        self._d = d
        e, m, b = C.c_int(), C.c_int(), C.c_int()
        _L.check(_L.lib().fecsim_decoder_info(self._d, C.byref(e), C.byref(m), C.byref(b)), "fecsim_decoder_info")
        self.elem, self.mode, self.elem_bytes = e.value, m.value, b.value
        self.np_dtype = NP_OF_ELEM[self.elem]

    def select_kernel(self, kernel: int) -> None:
        """lib.KERNEL_AUTO / KERNEL_GENERIC / KERNEL_PAIR / KERNEL_CODE / KERNEL_FLOAT /
        KERNEL_PAIR_RT (results are identical)."""
        _L.check(_L.lib().fecsim_decoder_select_kernel(self._d, kernel), "fecsim_decoder_select_kernel")

    def set_offsets(self, beta) -> None:
        """Per-(iteration, layer) offsets [max_iter, mrow] (converted like beta_offset),
        or None for the constant beta_offset."""
        if beta is None:
            _L.check(_L.lib().fecsim_decoder_set_offsets(self._d, None), "fecsim_decoder_set_offsets")
            return
        b = np.ascontiguousarray(beta, dtype=np.float32)
        if b.shape != (self.max_iter, self.hdec.info["mrow"]):
            raise ValueError("offsets must have shape (max_iter, mrow) = (%d, %d)" % (self.max_iter, self.hdec.info["mrow"]))
        _L.check(_L.lib().fecsim_decoder_set_offsets(self._d, b.ctypes.data), "fecsim_decoder_set_offsets")

    @property
    def kernel(self) -> int:
        return _L.lib().fecsim_decoder_kernel(self._d)

    @property
    def torch_dtype(self):
        import torch
        return {_L.I8: torch.int8, _L.I16: torch.int16, _L.I32: torch.int32, _L.F32: torch.float32,
                _L.F64: torch.float64}[self.elem]

    def decode_set(self, metrics, iters=None, c2v=None, tot_iter=None, layout_p: int = 0, stream=None,
                   num_cw: int | None = None):
        """In-place decode of device tensors (asynchronous on `stream`, default the
        current torch stream)."""
        import torch
        if metrics.dtype != self.torch_dtype or not metrics.is_cuda or not metrics.is_contiguous():
            raise TypeError("metrics must be a contiguous %s CUDA tensor" % self.torch_dtype)
        if num_cw is None:
            num_cw = metrics.shape[0]
        if metrics.numel() < num_cw * _row_elems(self.hdec, layout_p):
            raise ValueError("metrics holds fewer than num_cw codewords of this layout")
        if iters is not None and (iters.dtype != torch.int32 or iters.numel() < num_cw):
            raise TypeError("iters must be an int32 tensor of num_cw entries")
        if c2v is not None and (c2v.dtype != self.torch_dtype or c2v.numel() < num_cw * self.hdec.e):
            raise TypeError("c2v must hold num_cw*E elements of the decoder's type")
        if tot_iter is not None and tot_iter.dtype != torch.int64:
            raise TypeError("tot_iter must be an int64 tensor")
        if stream is None:
            stream = torch.cuda.current_stream()
        st = C.c_void_p(stream.cuda_stream)
        _L.check(_L.lib().fecsim_decode_set(self._d, num_cw, layout_p, _ptr(metrics), _ptr(iters), _ptr(c2v),
                                            _ptr(tot_iter), st), "fecsim_decode_set")

    def decode(self, llr_in, post_out, iters=None, c2v=None, tot_iter=None, layout_p: int = 0, stream=None,
               num_cw: int | None = None):
        """Out-of-place decode: llr_in is left untouched."""
        import torch
        for t in (llr_in, post_out):
            if t.dtype != self.torch_dtype or not t.is_cuda or not t.is_contiguous():
                raise TypeError("llr_in/post_out must be contiguous %s CUDA tensors" % self.torch_dtype)
        if num_cw is None:
            num_cw = llr_in.shape[0]
        need = num_cw * _row_elems(self.hdec, layout_p)
        if llr_in.numel() < need or post_out.numel() < need:
            raise ValueError("llr_in/post_out hold fewer than num_cw codewords of this layout")
        if iters is not None and (iters.dtype != torch.int32 or iters.numel() < num_cw):
            raise TypeError("iters must be an int32 tensor of num_cw entries")
        if c2v is not None and (c2v.dtype != self.torch_dtype or c2v.numel() < num_cw * self.hdec.e):
            raise TypeError("c2v must hold num_cw*E elements of the decoder's type")
        if stream is None:
            stream = torch.cuda.current_stream()
        _L.check(_L.lib().fecsim_decode(self._d, num_cw, layout_p, _ptr(llr_in), _ptr(post_out), _ptr(iters),
                                        _ptr(c2v), _ptr(tot_iter), C.c_void_p(stream.cuda_stream)), "fecsim_decode")

    def decode_set_host(self, metrics: np.ndarray, layout_p: int = 0, want_iters: bool = True):
        """Synchronous decode of a host numpy array in place; returns (tot_iter, iters)."""
        if metrics.dtype != self.np_dtype or not metrics.flags.c_contiguous:
            raise TypeError("metrics must be a contiguous %s array" % self.np_dtype)
        ncw = metrics.shape[0]
        if metrics.size < ncw * _row_elems(self.hdec, layout_p):
            raise ValueError("metrics holds fewer than its codewords of this layout")
        iters = np.zeros(ncw, np.uint32) if want_iters else None
        tot = C.c_ulonglong(0)
        _L.check(_L.lib().fecsim_decode_set_host(self._d, ncw, layout_p, metrics.ctypes.data,
                                                 None if iters is None else iters.ctypes.data, C.byref(tot)),
                 "fecsim_decode_set_host")
        return tot.value, iters

    def __del__(self):
        if getattr(self, "_d", None) and _L._lib is not None:
            _L.lib().fecsim_decoder_free(self._d)
            self._d = None
