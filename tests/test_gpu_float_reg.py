"""The register-C2V float32 kernel (FECSIM_KERNEL_FLOAT, csrc/float_reg.hip) against the
generic float kernels on the same inputs: posteriors, C2V messages, per-codeword
iteration counts and the iteration total identical bit for bit -- the kernel runs the
generic kernels' row arithmetic with the C2V values held in registers, so any
difference is a bug.  Covers min-sum and sum-product, every workgroup plan of the
802.16e rate-1/2 lifting sweep (one, two or four waves; one to nine codewords per block), the
other compiled shapes (802.11ad 1/2, 802.11n 1/2 at Z = 27 and 54), per-(iteration,
layer) offset tables, in-place decoding, batch sizes that leave slots empty, and the
automatic choice.  The generic kernels are themselves pinned to the oracle and to the
libm sum-product restatement by test_gpu_decode.py and test_gpu_sp_sweep.py."""
import sys

import numpy as np
import pytest

from fecsim import codes
from fecsim import lib as L
from fecsim.decoder import Decoder, Hdec
import refsim
from refsim import O

pytestmark = pytest.mark.gpu


def _llrs(cs, ncw, ebn0):
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(ncw)
    llr, _ = rs.metrics(wgn, bits, ebn0 - 10 * np.log10(2.0), elem=O.F32, llr_scale=1.0, max_abs=1e9)
    return np.ascontiguousarray(llr)


def _run(dec, llr, kernel, table=None, inplace=False):
    import torch
    dec.select_kernel(kernel)
    dec.set_offsets(table)
    x = torch.from_numpy(llr).cuda()
    out = x if inplace else torch.empty_like(x)
    n = llr.shape[0]
    iters = torch.zeros(n, dtype=torch.int32, device="cuda")
    c2v = torch.zeros(n * dec.hdec.e, dtype=torch.float32, device="cuda")
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    if inplace:
        dec.decode_set(x, iters=iters, c2v=c2v, tot_iter=tot)
    else:
        dec.decode(x, out, iters=iters, c2v=c2v, tot_iter=tot)
    torch.cuda.synchronize()
    return out.cpu().numpy(), iters.cpu().numpy(), c2v.cpu().numpy(), int(tot.item())


def _same(a, b):
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            assert x.shape == y.shape
            assert np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                                  y.view(np.uint32) if y.dtype == np.float32 else y)
        else:
            assert x == y


SWEEP_Z = list(range(24, 97, 4))


@pytest.mark.parametrize("z", SWEEP_Z)
@pytest.mark.parametrize("kind", ["Vec16f_prod", "Vec16f_min"])
def test_gpu_float_reg_equals_generic_16e_sweep(z, kind, cuda):
    print("float_reg 16e 1/2", z, kind, file=sys.stderr, flush=True)
    cs = codes.gen_structs("16e 1/2", simd_size=512, elem_size=32, z_over=z)
    dec = Decoder(Hdec.from_tables(cs.dec), kind, 10, 6e-8 if kind.endswith("prod") else 1e9, 0.0)
    assert dec.kernel == L.KERNEL_FLOAT
    llr = _llrs(cs, 37, 1.0)   # 37 codewords: the last block's slots are partly empty
    _same(_run(dec, llr, L.KERNEL_FLOAT), _run(dec, llr, L.KERNEL_GENERIC))


@pytest.mark.parametrize("name,z,kind,mrm,beta", [
    ("11ad 1/2", 0, "Vec16f_prod", 6e-8, 0.0),
    ("11ad 1/2", 0, "Vec16f_min", 31.0, 0.5),
    ("11n 1/2 z27", 0, "Vec16f_min", 1e9, 0.25),
    ("11n 1/2 z54", 0, "Vec16f_prod", 6e-8, 0.0),
    ("16e 1/2", 96, "Vec16f_prod", 0.9, 0.1),       # cap of a product row below 1
])
def test_gpu_float_reg_equals_generic_other_shapes(name, z, kind, mrm, beta, cuda):
    print("float_reg", name, z, kind, file=sys.stderr, flush=True)
    kw = {"z_over": z} if z else {}
    cs = codes.gen_structs(name, simd_size=512, elem_size=32, **kw)
    dec = Decoder(Hdec.from_tables(cs.dec), kind, 8, mrm, beta)
    assert dec.kernel == L.KERNEL_FLOAT
    llr = _llrs(cs, 101, 0.5)
    _same(_run(dec, llr, L.KERNEL_FLOAT), _run(dec, llr, L.KERNEL_GENERIC))


@pytest.mark.parametrize("kind", ["Vec16f_prod", "Vec16f_min"])
def test_gpu_float_reg_offset_table_and_in_place(kind, cuda):
    cs = codes.gen_structs("16e 1/2", simd_size=512, elem_size=32, z_over=52)
    dec = Decoder(Hdec.from_tables(cs.dec), kind, 9, 6e-8 if kind.endswith("prod") else 1e9, 0.0)
    rng = np.random.default_rng(8)
    tab = rng.uniform(0.0, 0.6, (9, cs.dec.mrow)).astype(np.float32)
    llr = _llrs(cs, 64, 0.75)
    _same(_run(dec, llr, L.KERNEL_FLOAT, table=tab), _run(dec, llr, L.KERNEL_GENERIC, table=tab))
    assert dec.kernel == L.KERNEL_GENERIC   # select_kernel(GENERIC) sticks
    dec.select_kernel(L.KERNEL_AUTO)
    assert dec.kernel == L.KERNEL_FLOAT      # the table keeps the float kernel
    _same(_run(dec, llr, L.KERNEL_FLOAT, inplace=True), _run(dec, llr, L.KERNEL_GENERIC, inplace=True))


def test_gpu_float_reg_not_offered_elsewhere(cuda):
    from fecsim.lib import FecsimError
    cs = codes.gen_structs("16e 1/2", simd_size=512, elem_size=64, z_over=48)
    dec = Decoder(Hdec.from_tables(cs.dec), "Vec8d_prod", 10, 1e-16, 0.0)   # double: not compiled
    assert dec.kernel == L.KERNEL_GENERIC
    with pytest.raises(FecsimError):
        dec.select_kernel(L.KERNEL_FLOAT)
    cs = codes.gen_structs("16e 2/3A", simd_size=512, elem_size=32)
    dec = Decoder(Hdec.from_tables(cs.dec), "Vec16f_prod", 10, 6e-8, 0.0)   # shape not compiled
    assert dec.kernel == L.KERNEL_GENERIC
