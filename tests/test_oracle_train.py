"""Pins of the oracle's differentiable min-sum decoder (oracle/train.c): its forward
reduces to the reference-restating min-sum decoder, and its reverse mode agrees with
FP64 central finite differences and with torch autograd of an independent forward."""
import numpy as np
import pytest

import refsim
import train_ref
from fecsim import codes

O = refsim.O


def _llrs(cs, ncw, snr, seed=3):
    rs = refsim.RefStream(cs, seeds=refsim.perl_seeds(seed))
    wgn, bits = rs.inner_loop(ncw)
    llr, _ = rs.metrics(wgn, bits, snr, elem=O.F32, llr_scale=1.0, max_abs=1e9)
    return llr, bits


@pytest.mark.parametrize("name,T,beta,cap", [("11ad 1/2", 4, 0.5, 31.0), ("11n 1/2 z27", 3, 0.0, 1e30),
                                             ("16e 1/2", 5, 0.25, 6.0)])
def test_train_forward_reduces_to_min_sum(name, T, beta, cap, oracle):
    cs = codes.gen_structs(name, simd_size=512, elem_size=8)
    llr, _ = _llrs(cs, 60, -2.0)
    post, c2v, iters, _ = O.decode_ideal(cs.dec, llr, elem=O.F32, max_iter=T, max_row_metric=cap, beta=beta)
    al = np.ones((T, cs.dec.mrow), np.float32)
    be = np.full((T, cs.dec.mrow), beta, np.float32)
    p2, c2 = O.train_fwd(cs.dec, llr, al, be, cap=cap, want_c2v=True)
    full = iters == T                       # codewords that ran all T iterations
    assert full.sum() > 30
    assert np.array_equal(post[full].view(np.uint32), p2[full].view(np.uint32))
    assert np.array_equal(c2v[full].view(np.uint32), c2[full].view(np.uint32))


def _params(cs, T, rng):
    al = rng.uniform(0.6, 1.0, (T, cs.dec.mrow))
    be = rng.uniform(0.1, 0.9, (T, cs.dec.mrow))
    return al, be


@pytest.mark.parametrize("name,T", [("11n 1/2 z27", 4), ("11ad 1/2", 3)])
def test_train_gradient_matches_finite_differences(name, T, oracle):
    cs = codes.gen_structs(name, simd_size=512, elem_size=8)
    llr, _ = _llrs(cs, 3, -1.0)
    x = llr.astype(np.float64)
    rng = np.random.default_rng(11)
    al, be = _params(cs, T, rng)
    gpost = rng.normal(size=x.shape)
    cap = 8.0

    def loss(a, b, xx=x):
        p, _ = O.train_fwd(cs.dec, xx, a, b, cap=cap, dtype=np.float64)
        return float((p * gpost).sum())

    ga, gb, gl = O.train_grad(cs.dec, x, al, be, gpost, cap=cap)
    eps = 1e-6
    fa, fb = np.zeros_like(al), np.zeros_like(be)
    for t in range(T):
        for l in range(cs.dec.mrow):
            for arr, out in ((al, fa), (be, fb)):
                d = np.zeros_like(arr)
                d[t, l] = eps
                out[t, l] = (loss(al + d if arr is al else al, be + d if arr is be else be) -
                             loss(al - d if arr is al else al, be - d if arr is be else be)) / (2 * eps)
    scale = max(1.0, np.abs(fa).max(), np.abs(fb).max())
    np.testing.assert_allclose(ga, fa, rtol=1e-5, atol=1e-6 * scale)
    np.testing.assert_allclose(gb, fb, rtol=1e-5, atol=1e-6 * scale)
    assert np.abs(ga).max() > 0 and np.abs(gb).max() > 0
    # input-LLR gradient on a few coordinates
    for i in rng.choice(x.size, 12, replace=False):
        d = np.zeros(x.size)
        d[i] = eps
        d = d.reshape(x.shape)
        fd = (loss(al, be, x + d) - loss(al, be, x - d)) / (2 * eps)
        assert abs(fd - gl.reshape(-1)[i]) <= 1e-5 * max(1.0, abs(fd)), i


def test_train_gradient_matches_torch_autograd(oracle):
    import torch
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    T = 5
    llr, _ = _llrs(cs, 4, -0.5)
    x = llr.astype(np.float64)
    rng = np.random.default_rng(5)
    al, be = _params(cs, T, rng)
    gpost = rng.normal(size=x.shape)
    ta = torch.tensor(al, requires_grad=True)
    tb = torch.tensor(be, requires_grad=True)
    tx = torch.tensor(x, requires_grad=True)
    p = train_ref.forward(cs.dec, tx, ta, tb, cap=6.0)
    (p * torch.from_numpy(gpost)).sum().backward()
    pf, _ = O.train_fwd(cs.dec, x, al, be, cap=6.0, dtype=np.float64)
    np.testing.assert_allclose(p.detach().numpy(), pf, rtol=0, atol=1e-12)
    ga, gb, gl = O.train_grad(cs.dec, x, al, be, gpost, cap=6.0)
    np.testing.assert_allclose(ga, ta.grad.numpy(), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(gb, tb.grad.numpy(), rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(gl, tx.grad.numpy(), rtol=1e-10, atol=1e-10)


def test_register_trainer_shapes_cover_their_sets():
    """Every matrix set named for the register-C2V trainer has its decoder's layer weights
    (in gen_structs' layer order, which the trainer's tables follow) among the compiled
    shapes of the generated header, each shape once."""
    import os
    from fecsim import gen_pair_shapes
    from fecsim.codes import gen_structs
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(here, "simd-fecfun_amd", "csrc", "train_shapes.h")).read()
    shapes = gen_pair_shapes.train_shapes()
    assert len({w for w, _ in shapes}) == len(shapes)
    for ms in gen_pair_shapes.TRAIN_SETS:
        d = gen_structs(ms, simd_size=512, elem_size=8).dec
        w = tuple(int(x) for x in d.row_weights)
        i = [x for x, _ in shapes].index(w)
        assert "static constexpr int mrow = %d, ne = %d, maxw = %d;" % (len(w), sum(w), max(w)) in hdr
        cond = " && ".join("w[%d] == %d" % (k, x) for k, x in enumerate(w))
        assert "if (mrow == %d && %s) return %d;" % (len(w), cond, i) in hdr
        assert d.z <= 128   # one or two rows per lane
