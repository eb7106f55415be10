"""Host-side mirror of the reference's encoder / channel / error-count interface and
of its simulation driver loop, over the C ABI.

  Henc.load(path)         <- encode_matrix_allocate()  (src/LDPCEnc/encode_matrix_allocate.cpp:38-92)
  Henc.gen_codewords(...) <- gen_rand_databits() + LDPCencode()
                             (src/Top/gen_rand_databits.cpp:11-40, src/LDPCEnc/LDPCencode.cpp:40-100)
  Henc.encode(...)        <- LDPCencode()
  Henc.gen_llr(...)       <- normdist2() + gen_metrics()
                             (src/RNGfunc/prng_functions.cpp:148-186, src/Top/gen_metrics.cpp:18-121)
  Henc.count_errors(...)  <- LDPC_count_errors()       (src/Top/LDPC_count_errors.cpp:40-136)
  Sim(...).run(loops)     <- the inner/SNR loops of main() (src/Top/LDPCtop.cpp:86-121)

Codeword bits are int32 tensors [num_cw, mcol, col_words] (row r of a column at word
r // 32, bit r % 32); col_words = 16 is the reference's cw_bits layout.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as _L
from .codes import EncTables
from .decoder import Decoder, _ptr, _row_elems


def _torch_dtype(elem):
    import torch
    return {_L.I8: torch.int8, _L.I16: torch.int16, _L.I32: torch.int32, _L.F32: torch.float32,
            _L.F64: torch.float64}[elem]


class HencInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("z", "mrow", "mcol", "num_par", "punc_cnt", "fill_cnt", "punc_bits",
                                         "fill_bits", "col_words", "num_src")]


class ChanParams(C.Structure):
    _fields_ = [("snr_db", C.c_float), ("llr_scale", C.c_float), ("max_abs_llr", C.c_float)]


class SimParams(C.Structure):
    _fields_ = [("num_cw", C.c_uint32), ("num_snr", C.c_uint32), ("snr_db", C.POINTER(C.c_float)),
                ("llr_scale", C.c_float), ("max_abs_llr", C.c_float), ("seed", C.c_uint64), ("use_graph", C.c_int)]


class SimStats(C.Structure):
    _fields_ = [(n, C.c_ulonglong) for n in ("cw", "uncoded_bit_err", "iters", "bit_err", "cw_err")]


def chan_sigma(snr_db: float, llr_scale: float = 2.0):
    """(sigma, LLR factor) of gen_metrics.cpp:28-30 as the library computes them."""
    p = ChanParams(snr_db, llr_scale, 0.0)
    s, f = C.c_float(), C.c_float()
    _L.check(_L.lib().fecsim_chan_sigma(C.byref(p), C.byref(s), C.byref(f)), "fecsim_chan_sigma")
    return s.value, f.value


def _stream(stream):
    import torch
    return C.c_void_p((stream or torch.cuda.current_stream()).cuda_stream)


class Henc:
    """Encoder tables (LDPCencode.raw.dat) held by the library."""

    def __init__(self, handle):
        self._h = handle
        info = HencInfo()
        _L.check(_L.lib().fecsim_henc_info(self._h, C.byref(info)), "fecsim_henc_info")
        self.info = {n: getattr(info, n) for n, _ in HencInfo._fields_}

    @classmethod
    def load(cls, path: str) -> "Henc":
        h = C.c_void_p()
        _L.check(_L.lib().fecsim_henc_load(path.encode(), C.byref(h)), "fecsim_henc_load")
        return cls(h)

    @classmethod
    def from_bytes(cls, b: bytes) -> "Henc":
        words = np.frombuffer(b, dtype="<i4").copy()
        h = C.c_void_p()
        _L.check(_L.lib().fecsim_henc_from_words(words.ctypes.data, words.size, C.byref(h)), "fecsim_henc_from_words")
        return cls(h)

    @classmethod
    def from_tables(cls, enc: EncTables) -> "Henc":
        return cls.from_bytes(enc.to_bytes())

    @property
    def z(self):
        return self.info["z"]

    @property
    def mcol(self):
        return self.info["mcol"]

    @property
    def n(self):
        return self.info["z"] * self.info["mcol"]

    @property
    def col_words(self):
        return self.info["col_words"]

    def masks(self):
        """(puncture, filler) masks as uint32 arrays [mcol, col_words]."""
        shape = (self.info["mcol"], self.info["col_words"])
        p = np.zeros(shape, np.uint32)
        f = np.zeros(shape, np.uint32)
        _L.check(_L.lib().fecsim_henc_masks(self._h, p.ctypes.data, f.ctypes.data), "fecsim_henc_masks")
        return p, f

    def new_bits(self, num_cw, col_words=None):
        import torch
        cw = col_words or self.col_words
        return torch.zeros((num_cw, self.mcol, cw), dtype=torch.int32, device="cuda")

    def _cw(self, bits, col_words):
        if bits.dtype.itemsize != 4 or not bits.is_cuda or not bits.is_contiguous():
            raise TypeError("bits must be a contiguous 32-bit CUDA tensor")
        cw = col_words or bits.shape[-1]
        num_cw = bits.numel() // (self.mcol * cw)
        return num_cw, cw

    def gen_codewords(self, bits, seed: int, cw0: int = 0, col_words=None, stream=None):
        num_cw, cw = self._cw(bits, col_words)
        _L.check(_L.lib().fecsim_gen_codewords(self._h, seed, cw0, num_cw, cw, _ptr(bits), _stream(stream)),
                 "fecsim_gen_codewords")

    def encode(self, bits, col_words=None, stream=None):
        num_cw, cw = self._cw(bits, col_words)
        _L.check(_L.lib().fecsim_encode(self._h, num_cw, cw, _ptr(bits), _stream(stream)), "fecsim_encode")

    def gen_llr(self, bits, llr, elem: int, snr_db: float, seed: int, cw0: int = 0, llr_scale: float = 2.0,
                max_abs_llr: float = 24.0, layout_p: int = 0, uncoded=None, col_words=None, stream=None):
        num_cw, cw = self._cw(bits, col_words)
        if llr.dtype != _torch_dtype(elem) or not llr.is_cuda or not llr.is_contiguous():
            raise TypeError("llr must be a contiguous CUDA tensor of the element kind")
        need = num_cw * _row_elems(_Info(self), layout_p)
        if llr.numel() < need:
            raise ValueError("llr holds fewer than num_cw codewords of this layout")
        p = ChanParams(snr_db, llr_scale, max_abs_llr)
        _L.check(_L.lib().fecsim_gen_llr(self._h, C.byref(p), seed, cw0, num_cw, cw, _ptr(bits), elem, layout_p,
                                         _ptr(llr), _ptr(uncoded), _stream(stream)), "fecsim_gen_llr")

    def gen_llr_from_noise(self, bits, wgn, llr, elem: int, snr_db: float, llr_scale: float = 2.0,
                           max_abs_llr: float = 24.0, layout_p: int = 0, uncoded=None, col_words=None, stream=None):
        """gen_metrics on a noise buffer in normdist2's order (RefRng.normdist2)"""
        num_cw, cw = self._cw(bits, col_words)
        if llr.dtype != _torch_dtype(elem) or not llr.is_cuda or not llr.is_contiguous():
            raise TypeError("llr must be a contiguous CUDA tensor of the element kind")
        if llr.numel() < num_cw * _row_elems(_Info(self), layout_p):
            raise ValueError("llr holds fewer than num_cw codewords of this layout")
        if wgn.numel() < num_cw * self.mcol * ((self.z + 7) // 8) * 8 or not wgn.is_cuda:
            raise ValueError("wgn holds fewer values than num_cw codewords need")
        p = ChanParams(snr_db, llr_scale, max_abs_llr)
        _L.check(_L.lib().fecsim_gen_llr_from_noise(self._h, C.byref(p), _ptr(wgn), num_cw, cw, _ptr(bits), elem,
                                                    layout_p, _ptr(llr), _ptr(uncoded), _stream(stream)),
                 "fecsim_gen_llr_from_noise")

    def count_errors(self, bits, post, elem: int, counts, layout_p: int = 0, cw_bit_err=None, col_words=None,
                     stream=None):
        num_cw, cw = self._cw(bits, col_words)
        if post.dtype != _torch_dtype(elem) or not post.is_cuda or not post.is_contiguous():
            raise TypeError("post must be a contiguous CUDA tensor of the element kind")
        if post.numel() < num_cw * _row_elems(_Info(self), layout_p):
            raise ValueError("post holds fewer than num_cw codewords of this layout")
        _L.check(_L.lib().fecsim_count_errors(self._h, num_cw, cw, _ptr(bits), elem, layout_p, _ptr(post),
                                              _ptr(counts), _ptr(cw_bit_err), _stream(stream)),
                 "fecsim_count_errors")

    def hard_decisions(self, post, elem: int, num_cw: int, bits_out=None, layout_p: int = 0, col_words=None,
                       stream=None):
        """hard decisions of every column packed like new_bits(): a uint32 CUDA tensor
        [num_cw][mcol][col_words] (bits_out, or a new one)"""
        import torch
        cw = int(col_words or self.col_words)
        if post.dtype != _torch_dtype(elem) or not post.is_cuda or not post.is_contiguous():
            raise TypeError("post must be a contiguous CUDA tensor of the element kind")
        if post.numel() < num_cw * _row_elems(_Info(self), layout_p):
            raise ValueError("post holds fewer than num_cw codewords of this layout")
        if bits_out is None:
            bits_out = torch.empty((num_cw, self.mcol, cw), dtype=torch.int32, device=post.device)
        if bits_out.numel() < num_cw * self.mcol * cw or not bits_out.is_cuda or not bits_out.is_contiguous():
            raise ValueError("bits_out must be a contiguous CUDA tensor of num_cw * mcol * col_words words")
        _L.check(_L.lib().fecsim_hard_decisions(self._h, num_cw, cw, elem, layout_p, _ptr(post), _ptr(bits_out),
                                                _stream(stream)), "fecsim_hard_decisions")
        return bits_out

    def __del__(self):
        if getattr(self, "_h", None) and _L._lib is not None:
            _L.lib().fecsim_henc_free(self._h)
            self._h = None


class _Info:
    """Adapter so decoder._row_elems can size buffers of an encoder's code."""

    def __init__(self, henc):
        self.info = {"z": henc.z, "mcol": henc.mcol}


class Sim:
    """The simulation driver loop on the device: per inner loop, new codewords, then
    for every SNR point LLRs (same noise), decode, error count."""

    def __init__(self, henc: Henc, dec: Decoder, num_cw: int, snr_db, llr_scale: float = 2.0,
                 max_abs_llr: float = 24.0, seed: int = 100, use_graph: bool = True):
        self.henc, self.dec = henc, dec          # the handles must outlive the simulation
        self.snr = np.asarray(snr_db, np.float32)
        p = SimParams(num_cw, len(self.snr), self.snr.ctypes.data_as(C.POINTER(C.c_float)), llr_scale, max_abs_llr,
                      seed, int(use_graph))
        s = C.c_void_p()
        _L.check(_L.lib().fecsim_sim_create(henc._h, dec._d, C.byref(p), C.byref(s)), "fecsim_sim_create")
        self._s = s

    def run(self, loops: int = 1, stream=None):
        _L.check(_L.lib().fecsim_sim_run(self._s, loops, _stream(stream)), "fecsim_sim_run")

    def reset(self, stream=None):
        _L.check(_L.lib().fecsim_sim_reset(self._s, _stream(stream)), "fecsim_sim_reset")

    def stats(self):
        """list of dicts per SNR point: cw, uncoded_bit_err, iters, bit_err, cw_err"""
        arr = (SimStats * len(self.snr))()
        _L.check(_L.lib().fecsim_sim_stats_get(self._s, arr), "fecsim_sim_stats_get")
        return [{n: getattr(x, n) for n, _ in SimStats._fields_} for x in arr]

    def iter_hist(self):
        """per SNR point the histogram of iteration counts since the last reset:
        uint64 array [num_snr][max_iter + 1], bin i = codewords that stopped after i iterations"""
        nb = min(int(self.dec.max_iter) + 1, 1024)
        arr = np.zeros((len(self.snr), nb), np.uint64)
        _L.check(_L.lib().fecsim_sim_iter_hist(self._s, arr.ctypes.data, nb), "fecsim_sim_iter_hist")
        return arr

    def __del__(self):
        if getattr(self, "_s", None) and _L._lib is not None:
            _L.lib().fecsim_sim_free(self._s)
            self._s = None


def iter_histogram(iters, nbins: int, hist=None, stream=None):
    """hist[min(iters[i], nbins - 1)] += 1 over a CUDA tensor of per-codeword iteration
    counts (int32/uint32 as the decode calls write them); returns the int64 CUDA
    histogram (hist, or a new zeroed one)"""
    import torch
    if not iters.is_cuda or iters.dtype != torch.int32 or not iters.is_contiguous():
        raise TypeError("iters must be a contiguous int32 CUDA tensor")
    if hist is None:
        hist = torch.zeros(nbins, dtype=torch.int64, device=iters.device)
    if hist.numel() < nbins or hist.dtype != torch.int64 or not hist.is_cuda:
        raise ValueError("hist must be an int64 CUDA tensor of nbins entries")
    _L.check(_L.lib().fecsim_iter_histogram(_ptr(iters), iters.numel(), nbins, _ptr(hist), _stream(stream)),
             "fecsim_iter_histogram")
    return hist


class RefRng:
    """The reference's own sequential noise and data-bit streams (sfc64 / xoshiro256++
    seeded from its 32 rand_state values), replayed on the device for bit-exact replays
    of one of its runs; one wave, not a throughput path."""

    def __init__(self, rand_state):
        st = np.ascontiguousarray(np.asarray(rand_state, np.uint64))
        if st.size != 32:
            raise ValueError("rand_state holds 32 values")
        g = C.c_void_p()
        _L.check(_L.lib().fecsim_refrng_create(st.ctypes.data, C.byref(g)), "fecsim_refrng_create")
        self._g = g

    def normdist2(self, wgn, simd_count: int, stream=None):
        """simd_count (even) Vec8f of normdist2 into the float32 CUDA tensor wgn"""
        if wgn.numel() < 8 * simd_count or not wgn.is_cuda or not wgn.is_contiguous():
            raise ValueError("wgn must be a contiguous float32 CUDA tensor of 8 * simd_count values")
        _L.check(_L.lib().fecsim_refrng_normdist2(self._g, _ptr(wgn), simd_count, _stream(stream)),
                 "fecsim_refrng_normdist2")

    def databits(self, henc, bits, num_cw: int, col_words=None, stream=None):
        """gen_rand_databits into the data columns of bits ([num_cw][mcol][col_words] int32)"""
        cw = int(col_words or henc.col_words)
        if bits.numel() < num_cw * henc.mcol * cw or not bits.is_cuda or not bits.is_contiguous():
            raise ValueError("bits must be a contiguous CUDA tensor of num_cw * mcol * col_words words")
        _L.check(_L.lib().fecsim_refrng_databits(self._g, henc._h, num_cw, cw, _ptr(bits), _stream(stream)),
                 "fecsim_refrng_databits")

    def __del__(self):
        if getattr(self, "_g", None) and _L._lib is not None:
            _L.lib().fecsim_refrng_free(self._g)
            self._g = None
