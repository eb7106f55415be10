"""Pins of the oracle's restatement of the product's counter-based data and noise
scheme (orc_splitmix64, orc_gpu_databits, orc_noise_sample, orc_noise_llr): a
published known-answer stream, closed forms and the statistics of the noise."""
import numpy as np
import pytest

import refsim
from fecsim import codes

O = refsim.O
M64 = (1 << 64) - 1
GAMMA = 0x9e3779b97f4a7c15
DATA_KEY = 0xd1b54a32d192ed03


class SplitMix64:
    """Stateful SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c): the state
    advances by GAMMA before every output."""

    def __init__(self, seed):
        self.x = seed & M64

    def next(self):
        self.x = (self.x + GAMMA) & M64
        z = self.x
        z = ((z ^ (z >> 30)) * 0xbf58476d1ce4e5b9) & M64
        z = ((z ^ (z >> 27)) * 0x94d049bb133111eb) & M64
        return z ^ (z >> 31)


def test_splitmix64_known_answers(oracle):
    # published first outputs of SplitMix64 seeded with 0
    g = SplitMix64(0)
    assert [g.next() for _ in range(3)] == [0xe220a8397b1dcdaf, 0x6e789e6aa1b965f4, 0x06c45d188009454f]
    # the counter form: output q+1 of the generator seeded with s is splitmix64(s + (q+1)*GAMMA)
    for seed in (0, 12345, M64):
        g = SplitMix64(seed)
        for q in range(20):
            assert oracle.lib().orc_splitmix64((seed + (q + 1) * GAMMA) & M64) == g.next()


def test_gpu_databits_closed_form(oracle):
    cs = codes.gen_structs("11n 1/2 z81", simd_size=512, elem_size=8)
    z, mrow, mcol = cs.z, cs.enc.mrow, cs.enc.mcol
    _, fill = refsim.word_masks(cs.enc)
    seed, cw0, ncw = 77, 1000, 3
    bits = np.zeros((ncw, mcol, 64), np.uint8)
    oracle.lib().orc_gpu_databits(seed, cw0, ncw, z, mrow, mcol, fill.ctypes.data, bits.ctypes.data)
    for cw in range(ncw):
        g = SplitMix64(seed ^ DATA_KEY)
        # skip to the stream position of codeword cw0+cw: q = ((cw0+cw)*mcol + c)*8 + k
        g.x = (g.x + ((cw0 + cw) * mcol * 8) * GAMMA) & M64
        for c in range(mcol):
            words = [g.next() for _ in range(8)]
            if c >= mcol - mrow:
                assert not bits[cw, c].any()       # parity columns are left zero
                continue
            want = np.frombuffer(np.array(words, "<u8").tobytes(), np.uint8).copy()
            allbits = np.unpackbits(want, bitorder="little")
            allbits[z:] = 0
            assert np.array_equal(bits[cw, c], np.packbits(allbits, bitorder="little"))


def test_noise_samples_are_standard_normal(oracle):
    L = oracle.lib()
    z, mcol = 42, 16
    xs = np.array([L.orc_noise_sample(5, cw, c, r, z, mcol) for cw in range(200) for c in range(mcol)
                   for r in range(z)], np.float64)
    assert abs(xs.mean()) < 1.2e-2 and abs(xs.var() - 1.0) < 1.5e-2
    from scipy import stats
    assert stats.kstest(xs, "norm").statistic < 6e-3
    # the two samples of a pair (even/odd column) are uncorrelated
    a = xs.reshape(200, mcol // 2, 2, z)
    assert abs(np.corrcoef(a[:, :, 0].ravel(), a[:, :, 1].ravel())[0, 1]) < 1.5e-2
    # distinct seeds give distinct streams
    assert L.orc_noise_sample(5, 0, 0, 0, z, mcol) != L.orc_noise_sample(6, 0, 0, 0, z, mcol)


@pytest.mark.parametrize("elem,scale,mx", [(O.I8, 2.0, 24.0), (O.F32, 1.0, 1e9), (O.I16, 2.0, 300.0)])
def test_noise_llr_closed_form(elem, scale, mx, oracle):
    L = oracle.lib()
    z, mcol, ncw, seed, cw0 = 27, 24, 4, 9, 50
    rng = np.random.default_rng(1)
    bits = rng.integers(0, 2, (ncw, mcol * z)).astype(np.uint8)
    punct = np.zeros(mcol * z, np.uint8)
    punct[:2 * z] = 1
    fill = np.zeros(mcol * z, np.uint8)
    fill[5 * z + 3:6 * z] = 1
    sigma, f = O.snr_params(-0.5, scale)
    out, err = O.noise_llr(seed, cw0, bits, z, mcol, sigma, f, mx, punct, fill, elem)
    n = np.array([[L.orc_noise_sample(seed, cw0 + cw, v // z, v % z, z, mcol) for v in range(mcol * z)]
                  for cw in range(ncw)], np.float32)
    x = np.where(bits == 1, np.float32(-1), np.float32(1))
    y = (n.astype(np.float64) * np.float64(np.float32(sigma)) + x).astype(np.float32)   # fma: one rounding
    want = np.clip(y * np.float32(f), -mx, mx).astype(np.float32)
    want[:, punct == 1] = 0
    want[:, fill == 1] = mx
    if elem != O.F32:
        want = np.rint(want).astype(O.NP[elem])
    assert np.array_equal(out, want)
    assert err == int(((y < 0) != (x < 0)).sum())
