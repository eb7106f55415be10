"""Float sum-product parity on the CPU: the oracle's own tanh/atanh against libm, the
product schedule against plain exclusive products, and the oracle's sum-product
decoder against an independent numpy restatement that uses libm (tests/sp_ref.py)
over the whole 802.16e rate-1/2 lifting sweep (Z = 24..96 step 4).

These pin the float sum-product path of the oracle; the GPU decoder is held to the
same tolerance in test_gpu_sp_sweep.py."""
import numpy as np
import pytest

from fecsim import codes
import refsim
import sp_ref
from refsim import O

CAP_F32 = np.float32(1.0 - float(np.float32(6e-8)))   # decode_functions.cpp:28-32, Vec16f


def _ulps32(a, b):
    """distance in float32 units in the last place (same-sign finite values)"""
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def _f32_grid():
    """float32 arguments of tanh(A/2) in the decoder: A >= 0 up to beyond saturation"""
    rng = np.random.default_rng(5)
    x = np.concatenate([np.geomspace(1e-30, 1e-3, 400), np.linspace(0, 20, 40001),
                        rng.uniform(0, 12, 20000), np.array([0.25, 0.34657359, 0.5, 0.625, 8.66, 9.0, 30.0, 1e6])])
    return x.astype(np.float32)


def test_own_tanhf_within_4_ulps_of_libm(oracle):
    L = oracle.lib()
    x = _f32_grid()
    got = np.array([L.orc_tanhf(float(v)) for v in x], np.float32)
    want = np.tanh(x.astype(np.float64)).astype(np.float32)
    d = _ulps32(got, want)
    assert d.max() <= 4, (x[d.argmax()], got[d.argmax()], want[d.argmax()])
    neg = np.array([L.orc_tanhf(float(-v)) for v in x[::50]], np.float32)
    assert np.array_equal(neg, -got[::50])       # odd


def test_own_atanhf_within_4_ulps_of_libm(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(6)
    p = np.concatenate([np.geomspace(1e-30, 1e-3, 400), np.linspace(0, 0.9999, 40001), rng.uniform(0, 1, 20000),
                        1.0 - np.arange(1, 2000) * 2.0 ** -24]).astype(np.float32)
    p = np.unique(np.concatenate([p, [CAP_F32, np.float32(0.25), np.float32(0.5)]]))
    p = p[(p >= 0) & (p <= CAP_F32)]
    got = np.array([L.orc_atanhf(float(v)) for v in p], np.float32)
    want = np.arctanh(p.astype(np.float64)).astype(np.float32)
    d = _ulps32(got, want)
    assert d.max() <= 4, (p[d.argmax()], got[d.argmax()], want[d.argmax()])
    # the largest message the float decoder can produce: 2 atanh(cap) = 17.33
    assert abs(2.0 * L.orc_atanhf(float(CAP_F32)) - 2.0 * np.arctanh(np.float64(CAP_F32))) < 1e-5


def test_own_tanh_atanh_double_within_1e15(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(7)
    x = np.concatenate([np.geomspace(1e-300, 1e-3, 300), np.linspace(0, 40, 20001), rng.uniform(0, 20, 5000)])
    got = np.array([L.orc_tanh(float(v)) for v in x])
    want = np.tanh(x)
    nz = want != 0
    assert np.max(np.abs(got[nz] - want[nz]) / want[nz]) <= 1e-15
    p = np.concatenate([np.geomspace(1e-300, 1e-3, 300), np.linspace(0, 1 - 1e-6, 20001), 1 - np.geomspace(1e-6, 1e-16, 200)])
    got = np.array([L.orc_atanh(float(v)) for v in p])
    want = np.arctanh(p)
    nz = want != 0
    assert np.max(np.abs(got[nz] - want[nz]) / want[nz]) <= 1e-15


@pytest.mark.parametrize("w", list(range(2, 21)))
def test_product_schedule_is_the_exclusive_product(w):
    """min_find.pl's product schedule (its printed output, tests/golden/minprod_tree.json)
    evaluated in float32 gives min(cap, prod_{k != j} t_k) to within a few units: the
    schedule names the right factors for every output and caps only products."""
    rng = np.random.default_rng(w)
    t = [rng.uniform(0.0, 1.0, 4000).astype(np.float32) for _ in range(w)]
    for k in range(w):              # some factors saturated at 1 - ulp and 1
        t[k][: 400] = np.float32(1.0) - np.float32(2.0 ** -24) * rng.integers(0, 3, 400).astype(np.float32)
    out = sp_ref.exclusive_products(t, CAP_F32, np.float32)
    T = np.stack([x.astype(np.float64) for x in t])
    for j in range(w):
        exact = np.prod(np.delete(T, j, axis=0), axis=0)
        want = np.minimum(float(CAP_F32), exact)
        assert np.all(out[j] <= CAP_F32)
        assert np.max(np.abs(out[j] - want) / np.maximum(want, 1e-30)) <= (w + 1) * 2.0 ** -24, (w, j)


def _sweep_case(z, ebn0, max_iter, ncw=48):
    cs = codes.gen_structs("16e 1/2", simd_size=512, elem_size=32, z_over=z)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(ncw)
    # run_LDPCsim.pl:56-63 sum-product settings: scale 1, max |LLR| 1e9, beta 0, max_row_metric 6e-8
    llr, _ = rs.metrics(wgn, bits, ebn0 - 10 * np.log10(2.0), elem=O.F32, llr_scale=1.0, max_abs=1e9)
    return cs, llr


SWEEP = [(z, e, mi) for z in range(24, 97, 4) for (e, mi) in ((0.5, 10), (2.5, 10), (2.5, 1))]


@pytest.mark.parametrize("case", SWEEP, ids=lambda c: "z%d-%.1fdB-it%d" % c)
def test_oracle_sum_product_matches_libm_restatement(case, oracle):
    z, ebn0, mi = case
    cs, llr = _sweep_case(z, ebn0, mi)
    post, _, iters, _ = O.decode_ideal(cs.dec, llr, elem=O.F32, mode=O.PROD, max_iter=mi, max_row_metric=6e-8,
                                       beta=0.0, want_c2v=False)
    want, want_it, peak = sp_ref.decode_sp(cs.dec, llr, max_iter=mi)
    rep = sp_ref.sp_parity(post, iters, want, want_it, peak)
    assert rep["ok"], rep
    if ebn0 == 0.5 or mi == 1:
        # the 1e-4 comparison is not vacuous: most codewords stay well conditioned here
        assert rep["well_conditioned"] >= rep["codewords"] // 2, rep


@pytest.mark.parametrize("z", [24, 48, 96])
def test_oracle_double_sum_product_matches_libm_restatement(z, oracle):
    """Vec8d_prod: the double decoder against the double restatement (libm tanh/atanh)."""
    cs = codes.gen_structs("16e 1/2", simd_size=512, elem_size=64, z_over=z)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(32)
    llr, _ = rs.metrics(wgn, bits, 0.5 - 10 * np.log10(2.0), elem=O.F64, llr_scale=1.0, max_abs=1e9)
    post, _, iters, _ = O.decode_ideal(cs.dec, llr, elem=O.F64, mode=O.PROD, max_row_metric=1e-16, beta=0.0,
                                       want_c2v=False)
    want, want_it, peak = sp_ref.decode_sp(cs.dec, llr, max_row_metric=1e-16, dtype=np.float64)
    rep = sp_ref.sp_parity(post, iters, want, want_it, peak, rtol=1e-12)
    assert rep["ok"], rep
    assert rep["well_conditioned"] >= rep["codewords"] // 2, rep
