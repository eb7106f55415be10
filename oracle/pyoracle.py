"""ctypes view of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker; the product never imports it.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
TREE_FIXTURE = os.path.join(os.path.dirname(HERE), "tests", "golden", "minprod_tree.json")

I8, I16, I32, F32, F64 = 0, 1, 2, 3, 4
MIN, PROD = 0, 1
NP = {I8: np.int8, I16: np.int16, I32: np.int32, F32: np.float32, F64: np.float64}
ELEM_OF_CHAR = {"c": I8, "s": I16, "w": I32, "f": F32, "d": F64}


class Hdec(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("mrow", "mcol", "z", "total_hc_ones", "max_row_weight")] + \
               [(n, C.POINTER(C.c_int32)) for n in ("row_weights", "row_active_cols", "row_offset",
                                                    "f2b_copy_cnt", "f2b_cols", "f2b_offsets",
                                                    "b2f_copy_cnt", "b2f_cols", "b2f_offsets")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        vp, u32, i64 = C.c_void_p, C.c_uint32, C.c_int64
        L.orc_decode_ideal.restype = i64
        L.orc_decode_ideal.argtypes = [C.POINTER(Hdec), C.c_int, C.c_int, u32, u32, C.c_float, C.c_float, vp, vp, vp, C.c_int]
        L.orc_decode_literal.restype = i64
        L.orc_decode_literal.argtypes = [C.POINTER(Hdec), C.c_int, C.c_int, C.c_int, u32, u32, C.c_float, C.c_float, vp, vp, vp]
        L.orc_decode_vec64c.restype = i64
        L.orc_decode_vec64c.argtypes = [C.POINTER(Hdec), u32, u32, C.c_float, C.c_float, vp, vp, C.c_int, C.POINTER(C.c_int)]
        L.orc_set_tree.restype = C.c_int
        L.orc_set_tree.argtypes = [C.c_int, C.c_int, vp]
        L.orc_fastm2log.restype = C.c_float
        L.orc_fastm2log.argtypes = [C.c_float]
        L.orc_sincos_norm.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_normdist2.argtypes = [vp, C.c_int]
        L.orc_init_randstate.argtypes = [vp]
        L.orc_sfc64_N.argtypes = [C.c_int, vp]
        L.orc_xoshiro_N.argtypes = [C.c_int, vp]
        L.orc_gen_rand_databits.argtypes = [u32, C.c_int, C.c_int, C.c_int, vp, vp]
        L.orc_encode.argtypes = [u32, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.orc_gen_metrics.restype = u32
        L.orc_gen_metrics.argtypes = [u32, C.c_float, vp, C.c_int, C.c_int, vp, vp, C.c_int, C.c_int, C.c_float, C.c_float, vp, vp]
        L.orc_count_errors.argtypes = [u32, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp, C.c_int, C.POINTER(u32), C.POINTER(u32)]
        L.orc_snr_params.argtypes = [C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_splitmix64.restype = C.c_uint64
        L.orc_splitmix64.argtypes = [C.c_uint64]
        L.orc_noise_sample.restype = C.c_float
        L.orc_noise_sample.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_noise_llr.restype = u32
        L.orc_noise_llr.argtypes = [C.c_uint64, C.c_uint64, u32, vp, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, vp, vp, C.c_int, vp]
        L.orc_gpu_databits.argtypes = [C.c_uint64, C.c_uint64, u32, C.c_int, C.c_int, C.c_int, vp, vp]
        L.orc_set_beta_table.argtypes = [vp]
        L.orc_train_fwd_f32.argtypes = [C.POINTER(Hdec), u32, u32, vp, vp, C.c_float, vp, vp]
        L.orc_train_fwd_f64.argtypes = [C.POINTER(Hdec), u32, u32, vp, vp, C.c_double, vp, vp]
        L.orc_train_grad_f64.argtypes = [C.POINTER(Hdec), u32, u32, vp, vp, C.c_double, vp, vp, vp, vp, vp]
        for f in ("orc_tanhf", "orc_atanhf"):
            getattr(L, f).restype = C.c_float
            getattr(L, f).argtypes = [C.c_float]
        for f in ("orc_tanh", "orc_atanh"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double]
        with open(TREE_FIXTURE) as fh:
            trees = json.load(fh)
        for w, s in trees.items():
            ops = np.asarray(s["ops"], dtype=np.int64).reshape(-1)
            rc = L.orc_set_tree(int(w), len(s["ops"]), ops.ctypes.data)
            if rc != 0:
                raise RuntimeError("orc_set_tree(%s) = %d" % (w, rc))
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


class OracleHdec:
    """Holds the int32 arrays alive for the Hdec struct view."""

    def __init__(self, dec):
        self.dec = dec
        self.arrs = {}
        self.s = Hdec()
        for n in ("mrow", "mcol", "z", "total_hc_ones", "max_row_weight"):
            setattr(self.s, n, int(getattr(dec, n)))
        for n in ("row_weights", "row_active_cols", "row_offset", "f2b_copy_cnt", "f2b_cols", "f2b_offsets",
                  "b2f_copy_cnt", "b2f_cols", "b2f_offsets"):
            a = np.ascontiguousarray(np.asarray(getattr(dec, n)), dtype=np.int32)
            if a.size == 0:
                a = np.zeros(1, np.int32)
            self.arrs[n] = a
            setattr(self.s, n, a.ctypes.data_as(C.POINTER(C.c_int32)))

    @property
    def ref(self):
        return C.byref(self.s)


def decode_ideal(dec, llr, elem=I8, mode=MIN, max_iter=10, max_row_metric=31.0, beta=1.0, use_tree=False,
                 want_c2v=True, beta_table=None):
    """llr: [num_cw, N] array of the element type.  Returns (post, c2v, iters, tot).
    beta_table: optional [max_iter, mrow] per-(iteration, layer) offsets."""
    if beta_table is not None:
        tab = np.ascontiguousarray(beta_table, dtype=np.float32)
        assert tab.shape == (max_iter, dec.mrow)
        lib().orc_set_beta_table(tab.ctypes.data)
        try:
            return decode_ideal(dec, llr, elem, mode, max_iter, max_row_metric, beta, use_tree, want_c2v)
        finally:
            lib().orc_set_beta_table(None)
    h = OracleHdec(dec)
    post = np.ascontiguousarray(llr, dtype=NP[elem]).copy()
    ncw = post.shape[0]
    c2v = np.zeros((ncw, dec.total_hc_ones * dec.z), dtype=NP[elem]) if want_c2v else None
    iters = np.zeros(ncw, np.uint32)
    tot = lib().orc_decode_ideal(h.ref, elem, mode, ncw, max_iter, max_row_metric, beta, _p(post), _p(c2v),
                                 _p(iters), int(use_tree))
    if tot < 0:
        raise RuntimeError("orc_decode_ideal error %d" % tot)
    return post, c2v, iters, tot


def dense_to_ref(dense, z, mcol, P, fill=None, rng=None):
    """[ncw, mcol*z] -> reference layout [ncw, mcol*(SPZ+2)*P] (guards filled with
    garbage from rng to show they do not matter)."""
    spz = (z + P - 1) // P
    ncw = dense.shape[0]
    out = np.zeros((ncw, mcol, (spz + 2) * P), dtype=dense.dtype)
    if rng is not None:
        if np.issubdtype(dense.dtype, np.integer):
            info = np.iinfo(dense.dtype)
            out[:] = rng.integers(info.min, info.max + 1, size=out.shape, dtype=np.int64).astype(dense.dtype)
        else:
            out[:] = rng.normal(0, 30, size=out.shape).astype(dense.dtype)
    out[:, :, P:P + z] = dense.reshape(ncw, mcol, z)
    return out.reshape(ncw, -1)


def ref_to_dense(ref, z, mcol, P):
    spz = (z + P - 1) // P
    ncw = ref.shape[0]
    return ref.reshape(ncw, mcol, (spz + 2) * P)[:, :, P:P + z].reshape(ncw, mcol * z).copy()


def decode_literal(dec, llr, elem=I8, mode=MIN, P=64, max_iter=10, max_row_metric=31.0, beta=1.0, rng=None):
    """Literal reference decoder; returns (post dense, c2v, iters, tot)."""
    h = OracleHdec(dec)
    dense = np.ascontiguousarray(llr, dtype=NP[elem])
    ref = dense_to_ref(dense, dec.z, dec.mcol, P, rng=rng)
    ncw = dense.shape[0]
    c2v = np.zeros((ncw, dec.total_hc_ones * dec.z), dtype=NP[elem])
    iters = np.zeros(ncw, np.uint32)
    tot = lib().orc_decode_literal(h.ref, elem, mode, P, ncw, max_iter, max_row_metric, beta, _p(ref), _p(c2v),
                                   _p(iters))
    if tot < 0:
        raise RuntimeError("orc_decode_literal error %d" % tot)
    return ref_to_dense(ref, dec.z, dec.mcol, P), c2v, iters, tot


def decode_vec64c(dec, ref_metrics, max_iter=10, max_row_metric=31.0, beta=1.0, nthreads=1, iters=None):
    """In place on the reference layout (P=64).  Returns (tot_iter, isa)."""
    h = OracleHdec(dec)
    isa = C.c_int(0)
    tot = lib().orc_decode_vec64c(h.ref, ref_metrics.shape[0], max_iter, max_row_metric, beta, _p(ref_metrics),
                                  _p(iters), nthreads, C.byref(isa))
    if tot < 0:
        raise RuntimeError("orc_decode_vec64c error %d" % tot)
    return tot, isa.value


def snr_params(snr, llr_scale):
    s, f = C.c_float(), C.c_float()
    lib().orc_snr_params(snr, llr_scale, C.byref(s), C.byref(f))
    return s.value, f.value


def noise_llr(seed, cw0, bits, z, mcol, sigma, factor, max_abs, punct=None, fill=None, elem=I8):
    """Dense LLRs of the product's counter-based noise scheme.  bits: [ncw, N] 0/1."""
    bits = np.ascontiguousarray(bits, dtype=np.uint8)
    ncw = bits.shape[0]
    out = np.zeros((ncw, mcol * z), dtype=NP[elem])
    pu = None if punct is None else np.ascontiguousarray(punct, np.uint8).reshape(-1)
    fi = None if fill is None else np.ascontiguousarray(fill, np.uint8).reshape(-1)
    err = lib().orc_noise_llr(seed, cw0, ncw, _p(bits), z, mcol, sigma, factor, max_abs, _p(pu), _p(fi), elem,
                              _p(out))
    return out, err


def train_fwd(dec, llr, alpha, beta, cap=1e30, dtype=np.float32, want_c2v=False):
    """Differentiable min-sum forward (T = alpha.shape[0] iterations).  Returns (post, c2v)."""
    h = OracleHdec(dec)
    post = np.ascontiguousarray(llr, dtype=dtype).copy()
    a = np.ascontiguousarray(alpha, dtype=dtype)
    b = np.ascontiguousarray(beta, dtype=dtype)
    T = a.shape[0]
    ncw = post.shape[0]
    c2v = np.zeros((ncw, dec.total_hc_ones * dec.z), dtype) if want_c2v else None
    f = lib().orc_train_fwd_f32 if dtype == np.float32 else lib().orc_train_fwd_f64
    rc = f(h.ref, ncw, T, _p(a), _p(b), cap, _p(post), _p(c2v))
    if rc:
        raise RuntimeError("orc_train_fwd error %d" % rc)
    return post, c2v


def train_grad(dec, llr, alpha, beta, gpost, cap=1e30):
    """Reverse mode of train_fwd in double: (galpha, gbeta, gllr)."""
    h = OracleHdec(dec)
    x = np.ascontiguousarray(llr, dtype=np.float64)
    a = np.ascontiguousarray(alpha, dtype=np.float64)
    b = np.ascontiguousarray(beta, dtype=np.float64)
    g = np.ascontiguousarray(gpost, dtype=np.float64)
    T = a.shape[0]
    ga, gb = np.zeros_like(a), np.zeros_like(b)
    gl = np.zeros_like(x)
    rc = lib().orc_train_grad_f64(h.ref, x.shape[0], T, _p(a), _p(b), cap, _p(x), _p(g), _p(ga), _p(gb), _p(gl))
    if rc:
        raise RuntimeError("orc_train_grad_f64 error %d" % rc)
    return ga, gb, gl
