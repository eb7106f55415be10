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
