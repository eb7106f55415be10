"""Pins of the CPU oracle (oracle/liboracle.so) before it is trusted as the checker.

* two restatements of different structure agree bit for bit: the ideal decoder
  (modular addressing, min1/min2 or the min_find tree) and the literal one (P-lane
  words, guard words with garbage in them, f2b/b2f copies of the .dat);
* a one-layer code decoded for one iteration equals a numpy brute force of the
  check-row update with the x86 int8 semantics (sat, wrapping abs(-128), beta wrap);
* the SIMD CPU decoder (CPU baseline) equals the ideal decoder;
* PRNGs equal independent Python implementations of sfc64 / xoshiro256++;
  the Box-Muller pieces are within their stated accuracy of -2 ln x and sin/cos;
* gen_metrics and count_errors equal numpy closed forms;
* the full reference chain reproduces the README's CWER within loose bounds.
"""
import glob
import os

import numpy as np
import pytest

from fecsim import codes
import refsim
from refsim import O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _bytes_eq(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def _noisy(cs, ncw, snr, elem, rng, scale=2.0, max_abs=24.0):
    s, f = O.snr_params(snr, scale)
    y = (1.0 + s * rng.standard_normal((ncw, cs.n))).astype(np.float32)
    l = np.clip(y * np.float32(f), -max_abs, max_abs)
    return np.rint(l).astype(O.NP[elem]) if elem <= O.I32 else l.astype(O.NP[elem])


CROSS = [
    # MatrixSet, num_par, P, elem, mode, max_row_metric, beta, z_over, snr
    ("11ad 1/2", 1, 64, O.I8, O.MIN, 31.0, 1.0, None, -0.5),
    ("11ad 1/2", 6, 64, O.I8, O.MIN, 31.0, 1.0, None, 0.0),
    ("11ad 1/2", 1, 16, O.I8, O.MIN, 31.0, 0.0, None, -0.5),
    ("11n 1/2 z27", 2, 64, O.I8, O.MIN, 31.0, 1.0, None, -0.5),
    ("11n 5/6 z81", 1, 64, O.I8, O.MIN, 31.0, 1.0, None, 3.5),
    ("11ad 13/16", 1, 16, O.I8, O.MIN, 20.0, 2.0, None, 3.0),
    ("16e 2/3A", 1, 32, O.I16, O.MIN, 200.0, 3.0, 96, 1.0),
    ("11ad 1/2", 1, 16, O.I32, O.MIN, 100.0, 2.0, None, -0.5),
    ("16e 1/2", 1, 16, O.F32, O.PROD, 6e-8, 0.0, 48, -1.0),
    ("16e 1/2", 1, 16, O.F32, O.PROD, 6e-8, 0.0, 24, -0.5),
    ("11n 1/2 z81", 1, 16, O.F32, O.MIN, 31.0, 0.5, None, -1.0),
    ("11ad 1/2", 1, 8, O.F64, O.PROD, 1e-16, 0.0, None, -1.0),
]


@pytest.mark.parametrize("case", CROSS, ids=lambda c: "%s-np%d-P%d-e%d-m%d" % c[:5])
@pytest.mark.parametrize("src", ["noisy", "rand"])
def test_ideal_equals_literal(case, src, oracle):
    ms, npar, P, elem, mode, mrm, beta, zo, snr = case
    es = {O.I8: 8, O.I16: 16, O.I32: 32, O.F32: 32, O.F64: 64}[elem]
    cs = codes.gen_structs(ms, num_par=npar, simd_size=P * es, elem_size=es, z_over=zo)
    rng = np.random.default_rng(11)
    scale = 1.0 if elem >= O.F32 else (8.0 if elem != O.I8 else 2.0)
    if src == "noisy":
        llr = _noisy(cs, 16, snr, elem, rng, scale=scale, max_abs=1e9 if elem >= O.F32 else 24.0 * scale / 2)
    elif elem <= O.I32:
        info = np.iinfo(O.NP[elem])
        llr = rng.integers(info.min, info.max + 1, size=(16, cs.n), dtype=np.int64).astype(O.NP[elem])
    else:
        llr = (rng.standard_normal((16, cs.n)) * 6).astype(O.NP[elem])
    kw = dict(elem=elem, mode=mode, max_row_metric=mrm, beta=beta)
    a = O.decode_ideal(cs.dec, llr, **kw)
    b = O.decode_ideal(cs.dec, llr, use_tree=True, **kw)
    c = O.decode_literal(cs.dec, llr, P=P, rng=rng, **kw)
    for i in range(3):
        assert _bytes_eq(a[i], b[i]), ("tree", i)
        assert _bytes_eq(a[i], c[i]), ("literal", i)
    assert a[3] == int(a[2].sum())


def _brute_row_update(llr, cols, shifts, z, beta8, cap8):
    """One layer, one iteration, C2V starting at zero: the check-row update of
    decode_functions.cpp:39-102 written out per row with explicit exclusion."""
    def wrap8(x):
        return ((x + 128) % 256) - 128

    post = llr.astype(np.int64).copy()
    w = len(cols)
    c2v = np.zeros((w, z), np.int64)
    for r in range(z):
        v = [cols[j] * z + (shifts[j] + r) % z for j in range(w)]
        q = [int(np.clip(post[x], -128, 127)) for x in v]            # sat(L - 0)
        a = []
        for x in q:
            ab = -128 if x == -128 else abs(x)                          # abs_epi8
            a.append(max(wrap8(ab - beta8), 0))
        for j in range(w):
            m = min([a[k] for k in range(w) if k != j] + [127])
            m = min(cap8, m)
            neg = sum(1 for k in range(w) if k != j and q[k] < 0) % 2
            rn = wrap8(-m) if neg else m
            post[v[j]] = np.clip(q[j] + rn, -128, 127)
            c2v[j, r] = rn
    return post.astype(np.int8), c2v.reshape(-1).astype(np.int8)


@pytest.mark.parametrize("beta,cap", [(1.0, 31.0), (0.0, 127.0), (2.9, 5.0)])
def test_single_layer_brute_force(beta, cap, oracle):
    """A one-layer code (w = 5 blocks) at max_iter = 1 reduces to one textbook offset
    min-sum check update; the numpy brute force fixes signs, exclusion, the cap, the
    beta truncation and abs(-128) = -128."""
    z, cols, shifts = 7, [0, 1, 2, 3, 4], [3, 0, 6, 2, 5]
    dec = codes.DecTables(1, 5, z, 5, 5, np.array([5]), np.array(cols), np.array(shifts), np.array([0]),
                          np.zeros(0, np.int64), np.zeros(0, np.int64), np.array([0, 0]), np.zeros(0, np.int64),
                          np.zeros(0, np.int64))
    rng = np.random.default_rng(3)
    llr = rng.integers(-128, 128, size=(20, 5 * z)).astype(np.int8)
    llr[:, :3] = -128
    post, c2v, iters, _ = O.decode_ideal(dec, llr, max_iter=1, max_row_metric=cap, beta=beta)
    for i in range(llr.shape[0]):
        ep, ec = _brute_row_update(llr[i], cols, shifts, z, int(beta), int(cap))
        assert np.array_equal(post[i], ep)
        assert np.array_equal(c2v[i], ec)
    assert (iters == 1).all()


def test_all_zero_codeword_converges_in_one_iteration(oracle):
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    llr = np.full((3, cs.n), 20, np.int8)
    post, c2v, iters, tot = O.decode_ideal(cs.dec, llr)
    assert (iters == 1).all() and tot == 3 and (post > 0).all()


@pytest.mark.parametrize("ms,npar", [("11ad 1/2", 1), ("11ad 1/2", 6), ("11n 1/2 z81", 1), ("11n 5/6 z81", 1)])
def test_simd_cpu_decoder_equals_ideal(ms, npar, oracle):
    cs = codes.gen_structs(ms, num_par=npar, simd_size=512, elem_size=8)
    rng = np.random.default_rng(5)
    llr = np.concatenate([_noisy(cs, 24, -0.5, O.I8, rng),
                          rng.integers(-128, 128, size=(8, cs.n)).astype(np.int8)])
    post, _, iters, tot = O.decode_ideal(cs.dec, llr)
    ref = O.dense_to_ref(llr, cs.z, cs.dec.mcol, 64, rng=rng)
    it = np.zeros(llr.shape[0], np.uint32)
    t2, isa = O.decode_vec64c(cs.dec, ref, iters=it, nthreads=3)
    assert isa in (256, 512)
    assert np.array_equal(O.ref_to_dense(ref, cs.z, cs.dec.mcol, 64), post)
    assert np.array_equal(it, iters) and t2 == tot


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "dec_*.npz"))), ids=os.path.basename)
def test_golden_fixtures_reproduce(path, oracle):
    f = np.load(path)
    dec = codes.DecTables.from_bytes(f["dat"].tobytes())
    p, ch, mode, _, _ = codes.parse_exec_name(str(f["kind"]))
    mi, mrm, beta = f["params"]
    post, c2v, iters, _ = O.decode_ideal(dec, f["llr"], elem=O.ELEM_OF_CHAR[ch],
                                         mode=O.PROD if mode == "prod" else O.MIN, max_iter=int(mi),
                                         max_row_metric=float(mrm), beta=float(beta))
    assert _bytes_eq(post, f["post"]) and _bytes_eq(c2v, f["c2v"]) and np.array_equal(iters, f["iters"])


# ---------------------------------------------------------------------------
# RNG, WGN, metrics, error counts
# ---------------------------------------------------------------------------

M64 = (1 << 64) - 1


def _rotl(x, k):
    return ((x << k) | (x >> (64 - k))) & M64


def _sfc64(s):  # published sfc64 step (PractRand), state [a, b, c, counter]
    a, b, c, d = s
    t = (a + b + d) & M64
    s[3] = (d + 1) & M64
    s[0] = b ^ (b >> 11)
    s[1] = (c + (c << 3)) & M64
    s[2] = (_rotl(c, 24) + t) & M64
    return t


def _xoshiro(s):  # published xoshiro256++ (prng.di.unimi.it)
    res = (_rotl((s[0] + s[3]) & M64, 23) + s[0]) & M64
    t = (s[1] << 17) & M64
    s[2] ^= s[0]
    s[3] ^= s[1]
    s[1] ^= s[2]
    s[0] ^= s[3]
    s[2] ^= t
    s[3] = _rotl(s[3], 45)
    return res


def test_prngs_match_published_algorithms(oracle):
    seeds = refsim.perl_seeds(100)
    oracle.lib().orc_init_randstate(seeds.ctypes.data)
    st_sfc = [[int(seeds[4 * i + lane]) for i in range(4)] for lane in range(4)]
    st_xo = [[int(seeds[16 + 4 * i + lane]) for i in range(4)] for lane in range(4)]
    out = np.zeros(4, np.uint64)
    for _ in range(50):
        oracle.lib().orc_sfc64_N(0, out.ctypes.data)
        assert [int(x) for x in out] == [_sfc64(st_sfc[l]) for l in range(4)]
        oracle.lib().orc_xoshiro_N(1, out.ctypes.data)
        assert [int(x) for x in out] == [_xoshiro(st_xo[l]) for l in range(4)]
    # xoshiro256++ from state {1,2,3,4}: first output rotl(1+4, 23) + 1
    assert _xoshiro([1, 2, 3, 4]) == (5 << 23) + 1


# xoshiro256++ from the state {1, 2, 3, 4}: the published test vector of the reference
# implementation (prng.di.unimi.it/xoshiro256plusplus.c), as the rand_xoshiro crate's
# test_vector checks it
XOSHIRO256PP_1234 = [41943041, 58720359, 3588806011781223, 3591011842654386, 9228616714210784205,
                     9973669472204895162, 14011001112246962877, 12406186145184390807, 15849039046786891736,
                     10450023813501588000]


def test_oracle_xoshiro_known_answer(oracle):
    """orc_xoshiro_N (prng_functions.cpp:44-61) seeded with {1, 2, 3, 4} in every lane
    (rand_state[16 + 4 i + lane] = s_i, init_randstate :22-26) gives the published vector."""
    rs = np.zeros(64, np.uint64)
    for i in range(4):
        rs[16 + 4 * i: 16 + 4 * i + 4] = i + 1
    oracle.lib().orc_init_randstate(rs.ctypes.data)
    out = np.zeros(4, np.uint64)
    got = []
    for _ in range(len(XOSHIRO256PP_1234)):
        oracle.lib().orc_xoshiro_N(1, out.ctypes.data)
        assert len(set(int(x) for x in out)) == 1
        got.append(int(out[0]))
    assert got == XOSHIRO256PP_1234


def test_oracle_sfc64_equals_numpy_sfc64(oracle):
    """orc_sfc64_N (prng_functions.cpp:33-41) against numpy's own SFC64 bit generator
    (PractRand's sfc64, an independent implementation) from the same per-lane states
    (a, b, c, counter) = rand_state[4 i + lane], i = 0..3."""
    seeds = refsim.perl_seeds(100)
    oracle.lib().orc_init_randstate(seeds.ctypes.data)
    gens = []
    for lane in range(4):
        bg = np.random.SFC64()
        st = bg.state
        st["state"]["state"] = np.array([seeds[4 * i + lane] for i in range(4)], dtype=np.uint64)
        st["has_uint32"], st["uinteger"] = 0, 0
        bg.state = st
        gens.append(bg.random_raw(300))
    out = np.zeros(4, np.uint64)
    for k in range(300):
        oracle.lib().orc_sfc64_N(0, out.ctypes.data)
        assert [int(x) for x in out] == [int(g[k]) for g in gens], k


def test_box_muller_pieces_accuracy(oracle):
    L = oracle.lib()
    xs = np.concatenate([np.geomspace(1e-19, 1.0, 2000, endpoint=False), np.linspace(0.01, 0.999, 500)])
    for x in xs:
        got = L.orc_fastm2log(float(np.float32(x)))
        want = -2.0 * np.log(np.float32(x))
        assert abs(got - want) <= 1.5e-2, x   # quadratic log2 fit: |error| < 9e-3 * 2 ln 2 (measured 1.24e-2)
    import ctypes as C
    s, c = C.c_float(), C.c_float()
    for x in np.linspace(-1.0, 0.999, 4001):
        L.orc_sincos_norm(float(np.float32(x)), C.byref(s), C.byref(c))
        assert abs(s.value - np.sin(np.pi * x)) < 1.5e-3, x
        assert abs(c.value - np.cos(np.pi * x)) < 1.5e-3, x


def test_normdist2_is_standard_normal(oracle):
    oracle.lib().orc_init_randstate(refsim.perl_seeds(7).ctypes.data)
    n = 1 << 20
    x = np.zeros(n, np.float32)
    oracle.lib().orc_normdist2(x.ctypes.data, n // 8)
    assert abs(x.mean()) < 5e-3 and abs(x.var() - 1.0) < 5e-3
    from scipy import stats
    assert stats.kstest(x.astype(np.float64), "norm").statistic < 3e-3
    assert np.abs(x).max() < 9.5


def test_gen_metrics_closed_form(oracle):
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(20)
    for elem, scale, mx in ((O.I8, 2.0, 24.0), (O.F32, 1.0, 1e9)):
        m, err = rs.metrics(wgn, bits, -0.5, elem=elem, llr_scale=scale, max_abs=mx)
        sigma, f = O.snr_params(-0.5, scale)
        assert np.isclose(sigma, 10 ** (-(-0.5 + 10 * np.log10(2)) / 20), rtol=1e-6)
        b = refsim.unpack_bits(bits, cs.z, cs.enc.mcol).reshape(20, cs.enc.mcol, cs.z)
        n = wgn[:20 * cs.enc.mcol * 48].reshape(20, cs.enc.mcol, 48)[:, :, :cs.z]
        x = np.where(b == 1, np.float32(-1), np.float32(1))
        y = (n.astype(np.float64) * np.float64(np.float32(sigma)) + x).astype(np.float32)  # single rounding of fma
        want = np.clip(y * np.float32(f), -mx, mx).astype(np.float32)
        if elem == O.I8:
            want = np.rint(want).astype(np.int8)
        assert np.array_equal(m.reshape(want.shape), want)
        assert err == int(((y < 0) != (x < 0)).sum())


def test_count_errors_closed_form(oracle):
    cs = codes.gen_structs("11ad 1/2", num_par=3, simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    _, bits = rs.inner_loop(10)
    b = refsim.unpack_bits(bits, cs.z, cs.enc.mcol)
    post = np.where(b == 1, -5, 5).astype(np.int8)
    post[2, 7] = -post[2, 7]        # bit 7 -> slot 7 % 3 = 1
    post[2, 8] = -post[2, 8]        # slot 2
    post[5, 3 * cs.z + 4] = -post[5, 3 * cs.z + 4]
    post[6, cs.k + 1] = -post[6, cs.k + 1]   # parity bits are not counted
    ce, be = refsim.count_errors(cs, bits, post)
    assert (ce, be) == (3, 3)


@pytest.mark.slow
def test_reference_chain_reproduces_readme_cwer(oracle):
    """README.md:43-47 (num_par=1, 1e6 cw): CWER 1.05e-2 at -0.75 dB, 2.01e-3 at
    -0.50 dB.  The table comes from an older script version and other seeds, so the
    bound is loose: within 4 binomial sigma plus 15 %."""
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    n, per = 0, 20000
    cwe = {-0.75: 0, -0.5: 0}
    for _ in range(5):
        wgn, bits = rs.inner_loop(per)
        for snr in cwe:
            m, _ = rs.metrics(wgn, bits, snr, P=64)
            O.decode_vec64c(cs.dec, m, nthreads=os.cpu_count() or 1)
            ce, _ = refsim.count_errors(cs, bits, O.ref_to_dense(m, cs.z, cs.dec.mcol, 64))
            cwe[snr] += ce
        n += per
    for snr, ref in ((-0.75, 1.05e-2), (-0.5, 2.01e-3)):
        p = cwe[snr] / n
        sig = np.sqrt(ref * (1 - ref) / n)
        assert abs(p - ref) < 4 * sig + 0.15 * ref, (snr, p, ref)
