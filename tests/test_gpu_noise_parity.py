"""The GPU simulation loop's codeword error rates against the reference's own chain
(sequential sfc64 / xoshiro256++ streams restated by the oracle, 2e6 codewords per
point, tests/golden/refstream_cwer_11ad12.json): two-sample |z| < 3 at Es/N0 -1.00,
-0.75 and -0.50 dB, 802.11ad rate 1/2, int8 offset min-sum beta 1, cap 31, max 10
iterations."""
import sys

import numpy as np
import pytest

from fecsim import codes
from fecsim.decoder import Decoder, Hdec
from fecsim.sim import Henc, Sim
from test_noise_parity import SNRS, refstream_counts, two_sample_z

pytestmark = pytest.mark.gpu


def test_gpu_sim_loop_cwer_matches_reference_streams(cuda):
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    dec = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 10, 31.0, 1.0)
    sim = Sim(Henc.from_tables(cs.enc), dec, 400000, list(SNRS), seed=100, use_graph=True)
    sim.run(5)
    got = sim.stats()
    fx = refstream_counts()
    for s, g in zip(SNRS, got):
        d = fx[str(s)]
        z = two_sample_z(g["cw_err"], g["cw"], d["cw_err"], d["cw"])
        print("snr %.2f gpu cwer %.4e ref-stream cwer %.4e z %.2f" % (s, g["cw_err"] / g["cw"], d["cw_err"] / d["cw"], z),
              file=sys.stderr, flush=True)
        assert g["cw"] == 2000000
        assert abs(z) < 3.0, (s, g, d, z)
