"""GPU float sum-product (Vec16f_prod, Vec8d_prod) against the independent libm
restatement (tests/sp_ref.py) over the 802.16e rate-1/2 lifting sweep Z = 24..96 step
4 at Eb/N0 0.5 and 2.5 dB: equal iteration counts and hard decisions, and posteriors
within 1e-4 relative on every well-conditioned codeword (sp_ref.sp_parity)."""
import sys

import numpy as np
import pytest

from fecsim import codes
from fecsim.decoder import Decoder, Hdec
import refsim
import sp_ref
from refsim import O

pytestmark = pytest.mark.gpu

SWEEP = [(z, e, mi) for z in range(24, 97, 4) for (e, mi) in ((0.5, 10), (2.5, 10), (2.5, 1))]


def _gpu(dec_tables, kind, llr, max_iter, mrm):
    import torch
    d = Decoder(Hdec.from_tables(dec_tables), kind, max_iter, mrm, 0.0)
    x = torch.from_numpy(np.ascontiguousarray(llr)).cuda()
    out = torch.empty_like(x)
    iters = torch.zeros(llr.shape[0], dtype=torch.int32, device="cuda")
    d.decode(x, out, iters=iters)
    torch.cuda.synchronize()
    return out.cpu().numpy(), iters.cpu().numpy().astype(np.uint32)


@pytest.mark.parametrize("case", SWEEP, ids=lambda c: "z%d-%.1fdB-it%d" % c)
def test_gpu_sum_product_sweep_matches_libm_restatement(case, cuda):
    z, ebn0, mi = case
    print("sp sweep", case, file=sys.stderr, flush=True)
    cs = codes.gen_structs("16e 1/2", simd_size=512, elem_size=32, z_over=z)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(64)
    llr, _ = rs.metrics(wgn, bits, ebn0 - 10 * np.log10(2.0), elem=O.F32, llr_scale=1.0, max_abs=1e9)
    post, iters = _gpu(cs.dec, "Vec16f_prod", llr, mi, 6e-8)
    want, want_it, peak = sp_ref.decode_sp(cs.dec, llr, max_iter=mi)
    rep = sp_ref.sp_parity(post, iters, want, want_it, peak)
    print("  ", rep, file=sys.stderr, flush=True)
    assert rep["ok"], rep
    if ebn0 == 0.5 or mi == 1:
        assert rep["well_conditioned"] >= rep["codewords"] // 2, rep


@pytest.mark.parametrize("z", [24, 60, 96])
def test_gpu_double_sum_product_matches_libm_restatement(z, cuda):
    cs = codes.gen_structs("16e 1/2", simd_size=512, elem_size=64, z_over=z)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(32)
    llr, _ = rs.metrics(wgn, bits, 0.5 - 10 * np.log10(2.0), elem=O.F64, llr_scale=1.0, max_abs=1e9)
    post, iters = _gpu(cs.dec, "Vec8d_prod", llr, 10, 1e-16)
    want, want_it, peak = sp_ref.decode_sp(cs.dec, llr, max_row_metric=1e-16, dtype=np.float64)
    rep = sp_ref.sp_parity(post, iters, want, want_it, peak, rtol=1e-12)
    assert rep["ok"], rep
    assert rep["well_conditioned"] >= rep["codewords"] // 2, rep


def _ulps32(a, b):
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ia - ib)


def test_gpu_sp_device_functions_within_4_ulps(cuda):
    """The float32 decoder's tanh(a/2) and 2 atanh(p) (fecsim_sp_math) against libm
    rounded to float32, over the arguments the decoder feeds them: a = |Q| up to far
    beyond saturation, p from 0 to the cap 1 - 6e-8 (and every float32 within 3000
    units of 1 below the cap)."""
    import torch
    from fecsim import lib as L
    rng = np.random.default_rng(3)
    a = np.concatenate([np.geomspace(1e-30, 1e-2, 500), np.linspace(0, 40, 400001), rng.uniform(0, 20, 100000),
                        [1.5, np.nextafter(np.float32(1.5), np.float32(0)), 1e6, 1e9]]).astype(np.float32)
    cap = np.float32(1.0 - float(np.float32(6e-8)))
    p = np.concatenate([np.geomspace(1e-30, 1e-2, 500), np.linspace(0, 0.9999, 400001), rng.uniform(0, 1, 100000),
                        1.0 - np.arange(1, 3000) * 2.0 ** -24, [0.5, np.nextafter(np.float32(0.5), np.float32(0))]])
    p = p.astype(np.float32)
    p = p[(p >= 0) & (p <= cap)]
    for fn, x, want in ((0, a, np.tanh(a.astype(np.float64) * 0.5).astype(np.float32)),
                        (1, p, (2.0 * np.arctanh(p.astype(np.float64))).astype(np.float32))):
        xt = torch.from_numpy(x).cuda()
        yt = torch.empty_like(xt)
        L.check(L.lib().fecsim_sp_math(fn, xt.data_ptr(), yt.data_ptr(), x.size, None), "fecsim_sp_math")
        torch.cuda.synchronize()
        got = yt.cpu().numpy()
        d = _ulps32(got, want)
        print("fn", fn, "max ulp", int(d.max()), "at", float(x[d.argmax()]), file=sys.stderr, flush=True)
        assert d.max() <= 4, (fn, x[d.argmax()], got[d.argmax()], want[d.argmax()])
        assert np.all(got >= 0)
