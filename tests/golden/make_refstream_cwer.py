"""Writes tests/golden/refstream_cwer_11ad12.json: bit and codeword error counts of the
reference's own chain -- its sequential streams (sfc64 normdist2, xoshiro256++
gen_rand_databits; src/RNGfunc/prng_functions.cpp:148-186, src/Top/gen_rand_databits.cpp)
seeded as run_LDPCsim.pl:74-77 seeds them, the encoder, gen_metrics and LDPCtop_Vec64c_min
(int8, beta 1, cap 31, max 10 iterations, num_par 1) -- restated by the oracle, at Es/N0
-1.00, -0.75 and -0.50 dB over 2e6 codewords per point (20 inner loops of 1e5, as
LDPCtop.cpp:86-121 runs them: the same noise and codewords at every SNR of a loop).

The GPU simulation loop's error rates are compared with these counts within binomial
bounds (tests/test_gpu_noise_parity.py).

    python tests/golden/make_refstream_cwer.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(HERE), os.path.join(os.path.dirname(HERE), "..", "oracle"),
                os.path.join(os.path.dirname(HERE), "..", "simd-fecfun_amd")]

import numpy as np  # noqa: E402

from fecsim import codes  # noqa: E402
import refsim  # noqa: E402
from refsim import O  # noqa: E402

SNRS = (-1.0, -0.75, -0.5)
LOOPS, PER = 20, 100000


def main():
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    out = {str(s): dict(cw=0, cw_err=0, bit_err=0, iters=0) for s in SNRS}
    for loop in range(LOOPS):
        wgn, bits = rs.inner_loop(PER)
        for s in SNRS:
            m, _ = rs.metrics(wgn, bits, s, P=64)
            tot, _ = O.decode_vec64c(cs.dec, m, nthreads=os.cpu_count() or 1)
            ce, be = refsim.count_errors(cs, bits, O.ref_to_dense(m, cs.z, cs.dec.mcol, 64))
            d = out[str(s)]
            d["cw"] += PER
            d["cw_err"] += ce
            d["bit_err"] += be
            d["iters"] += int(tot)
        print("loop", loop, {k: v["cw_err"] for k, v in out.items()}, flush=True)
    doc = dict(code="11ad 1/2", decoder="Vec64c_min beta 1 cap 31 max_iter 10 num_par 1",
               llr="scale 2, max |LLR| 24", seeds="run_LDPCsim.pl seed 100", counts=out,
               made_by="python tests/golden/make_refstream_cwer.py")
    with open(os.path.join(HERE, "refstream_cwer_11ad12.json"), "w") as fh:
        json.dump(doc, fh, indent=1)


if __name__ == "__main__":
    main()
