"""Statistical parity of the product's noise and data streams with the reference's.

The product draws data bits and noise from counter-based streams (orc_gpu_databits,
orc_noise_llr restate them; the GPU equals them bit for bit), where the reference runs
sequential sfc64 / xoshiro256++ streams (prng_functions.cpp:148-186,
gen_rand_databits.cpp:11-40).  The two cannot agree sample for sample, so the claim
is statistical: decoded with the same int8 decoder, both chains give codeword error
rates that agree within binomial bounds (two-sample |z| < 3) at Es/N0 -1.00, -0.75 and
-0.50 dB.  The reference chain's counts over 2e6 codewords per point are committed
(tests/golden/refstream_cwer_11ad12.json, tests/golden/make_refstream_cwer.py); its
first inner loop is recomputed here to show the fixture is what the oracle gives.
"""
import json
import os

import numpy as np
import pytest

from fecsim import codes
import refsim
from refsim import O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "refstream_cwer_11ad12.json")
SNRS = (-1.0, -0.75, -0.5)


def two_sample_z(e1, n1, e2, n2):
    p = (e1 + e2) / (n1 + n2)
    se = np.sqrt(p * (1 - p) * (1.0 / n1 + 1.0 / n2))
    return (e1 / n1 - e2 / n2) / se


def refstream_counts():
    with open(GOLDEN) as fh:
        return json.load(fh)["counts"]


def _chan(snr, scale=2.0):
    return O.snr_params(snr, scale)


def product_chain(cs, seed, cw0, ncw):
    """codeword errors per SNR of the product's streams decoded by the SIMD CPU decoder"""
    z, mrow, mcol = cs.z, cs.enc.mrow, cs.enc.mcol
    _, fillw = refsim.word_masks(cs.enc)
    words = np.zeros((ncw, mcol, 64), np.uint8)
    O.lib().orc_gpu_databits(seed, cw0, ncw, z, mrow, mcol, fillw.ctypes.data, words.ctypes.data)
    cnt, colnum, sc, ss = refsim.enc_arrays(cs.enc)
    O.lib().orc_encode(ncw, z, mrow, mcol, cnt.ctypes.data, colnum.ctypes.data, sc.ctypes.data, ss.ctypes.data,
                       words.ctypes.data)
    dense = refsim.unpack_bits(words, z, mcol)
    pm, fm = cs.enc.masks()
    out = {}
    for s in SNRS:
        sigma, f = _chan(s)
        llr, _ = O.noise_llr(seed, cw0, dense, z, mcol, sigma, f, 24.0, pm.reshape(-1).astype(np.uint8),
                             fm.reshape(-1).astype(np.uint8), O.I8)
        m = O.dense_to_ref(llr, z, mcol, 64)
        O.decode_vec64c(cs.dec, m, nthreads=os.cpu_count() or 1)
        ce, _ = refsim.count_errors(cs, words, O.ref_to_dense(m, z, mcol, 64))
        out[s] = ce
    return out


def test_refstream_fixture_first_loop_reproduces(oracle):
    """the committed reference-chain counts start with what the oracle computes now
    (inner loop 0 of 1e5 codewords, seeds of run_LDPCsim.pl)"""
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(100000)
    fx = refstream_counts()
    for s in SNRS:
        m, _ = rs.metrics(wgn, bits, s, P=64)
        O.decode_vec64c(cs.dec, m, nthreads=os.cpu_count() or 1)
        ce, _ = refsim.count_errors(cs, bits, O.ref_to_dense(m, cs.z, cs.dec.mcol, 64))
        d = fx[str(s)]
        # loop 0 is 1/20 of the fixture: its count sits inside the binomial band of the total
        p = d["cw_err"] / d["cw"]
        assert abs(ce - p * 100000) < 5 * np.sqrt(100000 * p * (1 - p)), (s, ce, p)
    assert all(fx[str(s)]["cw"] == 2000000 for s in SNRS)


@pytest.mark.slow
def test_product_streams_match_reference_streams_cwer(oracle):
    """4e5 codewords of the product chain per SNR against the reference chain's 2e6"""
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    fx = refstream_counts()
    err = {s: 0 for s in SNRS}
    n, per = 0, 100000
    for k in range(4):
        for s, ce in product_chain(cs, 100, k * per, per).items():
            err[s] += ce
        n += per
    for s in SNRS:
        d = fx[str(s)]
        z = two_sample_z(err[s], n, d["cw_err"], d["cw"])
        assert abs(z) < 3.0, (s, err[s] / n, d["cw_err"] / d["cw"], z)
