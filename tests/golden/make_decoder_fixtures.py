"""Writes tests/golden/dec_<case>.npz: decoder golden vectors produced by the CPU
oracle (oracle/liboracle.so, the C restatement of src/LDPCDec/LDPCdecode.cpp).

Each file holds the LDPCdecode.raw.dat bytes the case uses, the decoder kind and
parameters, the channel LLRs, and the oracle's posteriors, final C2V messages and
per-codeword iteration counts.  LLR sources:
  ref   - the reference's own chain (sfc64 WGN, xoshiro data bits, QC encoder,
          gen_metrics) with the run_LDPCsim.pl seeds, at the given SNR
  rand  - uniform random elements (saturation / wrap corners)
  zero  - all-zero codeword with noise from numpy at the given SNR
Run from the repo root: python tests/golden/make_decoder_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import refsim  # noqa: E402
from refsim import O  # noqa: E402
from fecsim import codes  # noqa: E402

# name: (MatrixSet, num_par, exec kind, source, snr, ncw, max_iter, max_row_metric, beta, llr_scale, max_abs)
CASES = {
    "11ad12_c64_ref_m100": ("11ad 1/2", 1, "Vec64c_min", "ref", -1.0, 48, 10, 31.0, 1.0, 2.0, 24.0),
    "11ad12_c64_ref_p000": ("11ad 1/2", 1, "Vec64c_min", "ref", 0.0, 48, 10, 31.0, 1.0, 2.0, 24.0),
    "11ad12_c64_rand": ("11ad 1/2", 1, "Vec64c_min", "rand", 0.0, 32, 10, 31.0, 1.0, 2.0, 24.0),
    "11ad12_c64_rand_beta0_cap127": ("11ad 1/2", 1, "Vec64c_min", "rand", 0.0, 32, 6, 127.0, 0.0, 2.0, 24.0),
    "11ad12_c64_npar6": ("11ad 1/2", 6, "Vec64c_min", "ref", -0.5, 8, 10, 31.0, 1.0, 2.0, 24.0),
    "11ad1316_c64_ref": ("11ad 13/16", 1, "Vec64c_min", "ref", 3.0, 32, 10, 31.0, 1.0, 2.0, 24.0),
    "11n12z27_c64_ref": ("11n 1/2 z27", 1, "Vec64c_min", "ref", -0.5, 32, 10, 31.0, 1.0, 2.0, 24.0),
    "11n34z54_c64_ref": ("11n 3/4 z54", 1, "Vec64c_min", "ref", 2.0, 16, 10, 31.0, 1.0, 2.0, 24.0),
    "11n56z81_c64_ref": ("11n 5/6 z81", 1, "Vec64c_min", "ref", 3.5, 16, 10, 31.0, 1.0, 2.0, 24.0),
    "16e23A_c64_z96": ("16e 2/3A", 1, "Vec64c_min", "zero", 1.5, 8, 12, 31.0, 1.0, 2.0, 24.0),
    "11ad12_s32_ref": ("11ad 1/2", 1, "Vec32s_min", "ref", -0.5, 32, 10, 200.0, 4.0, 8.0, 96.0),
    "11ad12_w16_rand": ("11ad 1/2", 1, "Vec16w_min", "rand", 0.0, 16, 5, 1000.0, 2.0, 2.0, 24.0),
    "16e12_z48_f16_prod": ("16e 1/2", 1, "Vec16f_prod", "ref", -1.0, 16, 10, 6e-8, 0.0, 1.0, 1e9),
    "11n12z81_f16_min": ("11n 1/2 z81", 1, "Vec16f_min", "ref", -1.5, 8, 10, 31.0, 0.5, 1.0, 1e9),
    "11ad12_d8_prod": ("11ad 1/2", 1, "Vec8d_prod", "ref", -1.0, 8, 10, 1e-16, 0.0, 1.0, 1e9),
}
Z_OVER = {"16e12_z48_f16_prod": 48, "16e23A_c64_z96": 96}


def make(name, spec):
    ms, npar, kind, src, snr, ncw, max_iter, mrm, beta, scale, max_abs = spec
    p, ch, mode, simd_size, elem_size = codes.parse_exec_name(kind)
    cs = codes.gen_structs(ms, num_par=npar, simd_size=simd_size, elem_size=elem_size, z_over=Z_OVER.get(name))
    elem = O.ELEM_OF_CHAR[ch]
    omode = O.PROD if mode == "prod" else O.MIN
    rng = np.random.default_rng(abs(hash(name)) % (2 ** 32))
    if src == "ref":
        rs = refsim.RefStream(cs)
        wgn, bits = rs.inner_loop(ncw)
        llr, _ = rs.metrics(wgn, bits, snr, elem=elem, P=0, llr_scale=scale, max_abs=max_abs)
    elif src == "rand":
        info = np.iinfo(O.NP[elem])
        llr = rng.integers(info.min, info.max + 1, size=(ncw, cs.n), dtype=np.int64).astype(O.NP[elem])
    else:
        s, f = O.snr_params(snr, scale)
        y = (1.0 + s * rng.standard_normal((ncw, cs.n))).astype(np.float32)
        l = np.clip(y * np.float32(f), -max_abs, max_abs)
        llr = np.rint(l).astype(O.NP[elem]) if elem <= O.I32 else l.astype(O.NP[elem])
    post, c2v, iters, tot = O.decode_ideal(cs.dec, llr, elem=elem, mode=omode, max_iter=max_iter,
                                           max_row_metric=mrm, beta=beta)
    np.savez_compressed(os.path.join(HERE, "dec_%s.npz" % name), dat=np.frombuffer(cs.dec.to_bytes(), np.uint8),
                        kind=np.array(kind), params=np.array([max_iter, mrm, beta], np.float64),
                        llr=llr, post=post, c2v=c2v, iters=iters)
    print("%-28s iters %s" % (name, np.bincount(iters, minlength=max_iter + 1)[1:]), file=sys.stderr)


if __name__ == "__main__":
    only = sys.argv[1:]
    for name, spec in CASES.items():
        if not only or name in only:
            make(name, spec)
