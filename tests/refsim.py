"""Reference simulation chain on the CPU oracle (test infrastructure).

Restates the worker loop of src/Top/LDPCtop.cpp:86-123 with the oracle's functions:
normdist2 (WGN) -> gen_rand_databits -> LDPCencode -> per SNR gen_metrics -> decode
-> LDPC_count_errors, and the seeds run_LDPCsim.pl:74-77 packs (srand(100); 64 x
2**56*rand, Perl's drand48).
"""
from __future__ import annotations

import os
import sys

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (os.path.join(REPO, "oracle"), os.path.join(REPO, "simd-fecfun_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

import pyoracle as O  # noqa: E402

WB = 64  # bytes per encoder column word


def perl_seeds(seed: int = 100, count: int = 64):
    """run_LDPCsim.pl:74-77: srand($seed); 2**56 * rand, packed as Q (uint64).
    Perl's rand is drand48: x <- (0x5DEECE66D x + 0xB) mod 2^48 from
    x0 = (seed << 16) | 0x330E, rand = x / 2^48, so 2^56*rand = x << 8 exactly."""
    x = ((seed & 0xFFFFFFFF) << 16) | 0x330E
    out = []
    for _ in range(count):
        x = (0x5DEECE66D * x + 0xB) & ((1 << 48) - 1)
        out.append(x << 8)
    return np.array(out, dtype=np.uint64)


def enc_arrays(enc):
    sched = enc.schedule()
    cnt = np.array([len(s) for _, s in sched], np.int32)
    colnum = np.array([c for c, _ in sched], np.int32)
    src_col = np.array([c for _, s in sched for c, _ in s] or [0], np.int32)
    src_shift = np.array([sh for _, s in sched for _, sh in s] or [0], np.int32)
    return cnt, colnum, src_col, src_shift


def word_masks(enc):
    """puncture/filler masks as [mcol][64] bytes (bit i of column at byte i>>3)."""
    pm, fm = enc.masks()
    out = []
    for m in (pm, fm):
        w = np.zeros((enc.mcol, WB), np.uint8)
        for c in range(enc.mcol):
            bits = np.zeros(WB * 8, np.uint8)
            bits[:enc.z] = m[c]
            w[c] = np.packbits(bits, bitorder="little")
        out.append(w)
    return out


def unpack_bits(cw_bits, z, mcol):
    """[ncw][mcol][64] byte words -> dense [ncw, mcol*z] 0/1."""
    ncw = cw_bits.shape[0]
    b = np.unpackbits(cw_bits.reshape(ncw, mcol, WB), axis=2, bitorder="little")[:, :, :z]
    return b.reshape(ncw, mcol * z)


class RefStream:
    """The reference's sequential random streams (sfc64 noise, xoshiro data)."""

    def __init__(self, cs, seeds=None):
        self.cs = cs
        seeds = perl_seeds() if seeds is None else np.asarray(seeds, np.uint64)
        O.lib().orc_init_randstate(seeds.ctypes.data)
        self.enc = enc_arrays(cs.enc)
        self.punct, self.fill = word_masks(cs.enc)

    def inner_loop(self, num_cw):
        """One inner loop: (wgn floats, cw_bits words) as LDPCtop.cpp:90-92 makes them."""
        z, mcol, mrow = self.cs.z, self.cs.enc.mcol, self.cs.enc.mrow
        spf = (z + 7) // 8
        wgn_count = num_cw * mcol * spf
        wgn_count += wgn_count & 1
        wgn = np.zeros(wgn_count * 8, np.float32)
        O.lib().orc_normdist2(wgn.ctypes.data, wgn_count)
        bits = np.zeros((num_cw, mcol, WB), np.uint8)
        O.lib().orc_gen_rand_databits(num_cw, z, mrow, mcol, self.fill.ctypes.data, bits.ctypes.data)
        cnt, colnum, sc, ss = self.enc
        O.lib().orc_encode(num_cw, z, mrow, mcol, cnt.ctypes.data, colnum.ctypes.data, sc.ctypes.data,
                           ss.ctypes.data, bits.ctypes.data)
        return wgn, bits

    def metrics(self, wgn, bits, snr, elem=O.I8, P=0, llr_scale=2.0, max_abs=24.0):
        """gen_metrics into the dense (P=0) or reference (P>0) layout; (metrics, uncoded errors)."""
        z, mcol = self.cs.z, self.cs.enc.mcol
        ncw = bits.shape[0]
        if P:
            spz = (z + P - 1) // P
            m = np.zeros((ncw, mcol * (spz + 2) * P), O.NP[elem])
        else:
            m = np.zeros((ncw, mcol * z), O.NP[elem])
        err = O.lib().orc_gen_metrics(ncw, snr, bits.ctypes.data, z, mcol, wgn.ctypes.data, m.ctypes.data, elem, P,
                                      llr_scale, max_abs, self.punct.ctypes.data, self.fill.ctypes.data)
        return m, err


def count_errors(cs, bits, post, elem=O.I8, num_par=None):
    import ctypes as C
    ce, be = C.c_uint32(), C.c_uint32()
    npar = cs.num_par if num_par is None else num_par
    post = np.ascontiguousarray(post)
    O.lib().orc_count_errors(bits.shape[0], cs.z, cs.enc.mrow, cs.enc.mcol, npar, bits.ctypes.data, post.ctypes.data,
                             elem, C.byref(ce), C.byref(be))
    return ce.value, be.value
