"""The drop-in worker executable (bin/LDPCtop_<kind>) driven through the reference's
process protocol: the parameter blob of run_LDPCsim.pl:79-82 on stdin, the .dat files
written by the table generator, and the output lines run_LDPCsim.pl:159-170 parses.
Its counts equal the device simulation loop's on the same seed and codeword indices."""
import os
import re
import struct
import subprocess
import sys

import numpy as np
import pytest

from fecsim import codes
from fecsim.decoder import Decoder, Hdec
from fecsim.sim import Henc, Sim

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "simd-fecfun_amd", "bin")
M64 = (1 << 64) - 1


def _mix64(z):
    z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M64
    z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M64
    return z ^ (z >> 31)


def _blob(outer, loops, ncw, max_iter, mrm, scale, max_abs, beta, s0, s1, step, seeds, enc, dec):
    """pack('L6 f7 Q64 Z128 Z128') as run_LDPCsim.pl:79-82 writes it."""
    b = struct.pack("<6I", outer, loops, ncw, max_iter, 0, 1)
    b += struct.pack("<7f", mrm, scale, max_abs, beta, s0, s1, step)
    b += struct.pack("<64Q", *seeds)
    b += enc.encode().ljust(128, b"\0") + dec.encode().ljust(128, b"\0")
    assert len(b) == 820
    return b


@pytest.mark.parametrize("name,kind,scale,max_abs,mrm,beta,snrs", [
    ("11ad 1/2", "Vec64c_min", 2.0, 24.0, 31.0, 1.0, (-1.0, 0.0, 0.5)),
    ("16e 1/2", "Vec16f_prod", 1.0, 1e9, 6e-8, 0.0, (-1.5, -1.0, 0.5)),
])
def test_gpu_dropin_worker_protocol(name, kind, scale, max_abs, mrm, beta, snrs, tmp_path, cuda):
    import torch
    exe = os.path.join(BIN, "LDPCtop_" + kind)
    assert os.path.exists(exe), "build the drop-in executables (simd-fecfun_amd/build.py)"
    enc_p, dec_p = str(tmp_path / "LDPCencode.raw.dat"), str(tmp_path / "LDPCdecode.raw.dat")
    cs = codes.gen_structs(name, simd_size=512, elem_size=8 if kind[5] == "c" else 32, enc_path=enc_p,
                           dec_path=dec_p)
    rng = np.random.default_rng(3)
    seeds = [int(x) for x in rng.integers(0, 2 ** 63, 64, dtype=np.int64)]
    outer, loops, ncw, mi = 2, 2, 3000, 10
    s0, s1 = snrs[0], snrs[-1]
    step = snrs[1] - snrs[0]
    blob = _blob(outer, loops, ncw, mi, mrm, scale, max_abs, beta, s0, s1, step, seeds, enc_p, dec_p)
    r = subprocess.run([exe], input=blob, capture_output=True, timeout=120)
    out = r.stdout.decode()
    print(out, r.stderr.decode(), file=sys.stderr)
    assert r.returncode == 0
    # the lines run_LDPCsim.pl:159-170 matches
    assert re.search(r"CW Count: %d codewords in inner loop \* num_codeword" % (loops * ncw), out)
    m = re.search(r"Matrix Params: Z=(\d+)\s+Hcol=(\d+)\s+Hrow=(\d+) num_par_cw=(\d+) fill_cnt=(\d+)", out)
    assert m and tuple(int(x) for x in m.groups()) == (cs.z, cs.enc.mcol, cs.enc.mrow, 1, 0)
    rows = re.findall(r"loop (\d+): SNR=([-\d.]+)\s+inp_bit_err=(\d+) tot_iter=(\d+) tot_bit_err=(\d+) "
                      r"tot_cw_err=(\d+) tot_dec_ticks=(\d+)", out)
    npts = int((s1 - s0 + 1e-6) / step) + 1
    assert len(rows) == outer * npts
    # the same stream through the device simulation loop: seed folded from the 64 seeds
    seed = 0
    for x in seeds:
        seed = _mix64(seed ^ x)
    dec = Decoder(Hdec.from_tables(cs.dec), kind, mi, mrm, beta)
    pts = [np.float32(k * np.float32(step) + np.float32(s0)) for k in range(npts)]
    sim = Sim(Henc.from_tables(cs.enc), dec, ncw, [float(p) for p in pts], llr_scale=scale, max_abs_llr=max_abs,
              seed=seed, use_graph=False)
    for o in range(outer):
        sim.reset()
        sim.run(loops)
        st = sim.stats()
        for k, x in enumerate(st):
            row = rows[o * npts + k]
            assert int(row[0]) == o and abs(float(row[1]) - float(pts[k])) < 1e-5
            assert (int(row[2]), int(row[3]), int(row[4]), int(row[5])) == (
                x["uncoded_bit_err"], x["iters"], x["bit_err"], x["cw_err"]), (o, k)
            assert int(row[6]) > 0
