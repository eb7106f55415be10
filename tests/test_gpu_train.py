"""GPU parity of the differentiable decoder and of per-(iteration, layer) offset
tables: forward bit-exact with the oracle, adjoint against the oracle's FP64 reverse
mode, FP64 central finite differences and torch autograd."""
import sys

import numpy as np
import pytest

from fecsim import codes
from fecsim.decoder import Decoder, Hdec
import refsim
import train_ref
from refsim import O

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _cs(name):
    return codes.gen_structs(name, simd_size=512, elem_size=8)


def _llrs(cs, ncw, snr, elem=O.F32, scale=1.0, mx=1e9, seed=3):
    rs = refsim.RefStream(cs, seeds=refsim.perl_seeds(seed))
    wgn, bits = rs.inner_loop(ncw)
    llr, _ = rs.metrics(wgn, bits, snr, elem=elem, llr_scale=scale, max_abs=mx)
    return llr


@pytest.mark.parametrize("kind,snr", [("Vec64c_min", -0.75), ("Vec32s_min", -0.5), ("Vec16f_min", -0.5)])
def test_gpu_offset_table_matches_oracle(kind, snr, cuda):
    import torch
    from fecsim import lib as L
    cs = _cs("11ad 1/2")
    elem = O.ELEM_OF_CHAR[kind[kind.index("Vec") + 3:].lstrip("0123456789")[0]]
    scale, mx = (1.0, 1e9) if elem == O.F32 else (2.0, 24.0)
    llr = _llrs(cs, 400, snr, elem, scale, mx)
    rng = np.random.default_rng(2)
    if elem == O.F32:
        tab = rng.uniform(0.0, 1.5, (10, cs.dec.mrow)).astype(np.float32)
    else:
        tab = rng.integers(0, 4, (10, cs.dec.mrow)).astype(np.float32) + 0.7   # truncates to 0..3
    want = O.decode_ideal(cs.dec, llr, elem=elem, beta=1.0, beta_table=tab)
    d = Decoder(Hdec.from_tables(cs.dec), kind, 10, 31.0, 1.0)
    d.set_offsets(tab)
    # int8 tables with entries 0..127 run on the row-pair kernel, float32 ones on the
    # register-C2V float kernel, the others on the generic one
    assert d.kernel == {O.I8: L.KERNEL_PAIR, O.F32: L.KERNEL_FLOAT}.get(elem, L.KERNEL_GENERIC)
    x = torch.from_numpy(llr).cuda()
    out = torch.empty_like(x)
    it = torch.zeros(400, dtype=torch.int32, device="cuda")
    c2v = torch.zeros(400 * d.hdec.e, dtype=x.dtype, device="cuda")
    d.decode(x, out, iters=it, c2v=c2v)
    torch.cuda.synchronize()
    assert np.array_equal(it.cpu().numpy().astype(np.uint32), want[2])
    assert np.array_equal(out.cpu().numpy().view(np.uint8), want[0].view(np.uint8))
    assert np.array_equal(c2v.cpu().numpy().reshape(400, -1).view(np.uint8), want[1].view(np.uint8))
    # a table equal to beta_offset everywhere is the plain decoder; clearing restores the pair kernel
    d.set_offsets(np.ones((10, cs.dec.mrow), np.float32))
    out2 = torch.empty_like(x)
    d.decode(x, out2)
    d.set_offsets(None)
    out3 = torch.empty_like(x)
    d.decode(x, out3)
    torch.cuda.synchronize()
    assert torch.equal(out2, out3)
    if elem == O.I8:
        assert d.kernel == L.KERNEL_PAIR


def _trainer(cs, T, batch, cap, double):
    from fecsim.train import Trainer
    return Trainer(Hdec.from_tables(cs.dec), T, batch, cap=cap, double=double)


@pytest.mark.parametrize("name,T,cap,double", [("11ad 1/2", 10, 1e30, False), ("11n 1/2 z81", 10, 6.0, False),
                                               ("11n 1/2 z27", 4, 1e30, True)])
def test_gpu_train_forward_matches_oracle(name, T, cap, double, cuda):
    import torch
    cs = _cs(name)
    dt = np.float64 if double else np.float32
    llr = _llrs(cs, 37, -1.0).astype(dt)
    rng = np.random.default_rng(8)
    al = rng.uniform(0.6, 1.0, (T, cs.dec.mrow)).astype(dt)
    be = rng.uniform(0.0, 0.9, (T, cs.dec.mrow)).astype(dt)
    want, _ = O.train_fwd(cs.dec, llr, al, be, cap=cap, dtype=dt)
    tr = _trainer(cs, T, 64, cap, double)
    got = tr.forward(torch.from_numpy(llr).cuda(), torch.from_numpy(al).cuda(), torch.from_numpy(be).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint8), want.view(np.uint8))


@pytest.mark.parametrize("name,T,double", [("11n 1/2 z27", 5, True), ("11ad 1/2", 4, True), ("11ad 1/2", 4, False),
                                          ("5GNR_n25344_k8448", 2, True)])
def test_gpu_train_backward_matches_oracle(name, T, double, cuda):
    """Adjoint against the oracle's FP64 reverse mode (the BG1 case keeps its
    posteriors in global memory)."""
    import torch
    cs = _cs(name)
    ncw = 2 if cs.z > 300 else 9
    x = _llrs(cs, ncw, -0.5 if cs.z < 300 else -4.2).astype(np.float64)
    rng = np.random.default_rng(9)
    al = rng.uniform(0.6, 1.0, (T, cs.dec.mrow))
    be = rng.uniform(0.1, 0.9, (T, cs.dec.mrow))
    gpost = rng.normal(size=x.shape)
    cap = 8.0
    ga, gb, gl = O.train_grad(cs.dec, x, al, be, gpost, cap=cap)
    dt = torch.float64 if double else torch.float32
    tr = _trainer(cs, T, ncw, cap, double)
    t = lambda a: torch.tensor(a, dtype=dt, device="cuda")   # noqa: E731
    tr.forward(t(x), t(al), t(be))
    gl2, ga2, gb2 = tr.backward(t(gpost))
    torch.cuda.synchronize()
    rtol, atol = (1e-9, 1e-9) if double else (5e-3, 5e-3)
    sa = max(1.0, np.abs(ga).max())
    sb = max(1.0, np.abs(gb).max())
    np.testing.assert_allclose(ga2.cpu().numpy(), ga, rtol=rtol, atol=atol * sa)
    np.testing.assert_allclose(gb2.cpu().numpy(), gb, rtol=rtol, atol=atol * sb)
    np.testing.assert_allclose(gl2.cpu().numpy(), gl, rtol=rtol, atol=atol * max(1.0, np.abs(gl).max()))
    _say("grad", name, T, double, "max|ga|", sa, "max|gb|", sb)


def test_gpu_train_gradient_finite_differences_fp64(cuda):
    """GPU FP64 adjoint against central finite differences (eps 1e-6) of the GPU FP64 forward."""
    import torch
    cs = _cs("11n 1/2 z81")
    T = 10
    x = torch.tensor(_llrs(cs, 6, -0.5).astype(np.float64), device="cuda")
    rng = np.random.default_rng(4)
    al = torch.tensor(rng.uniform(0.6, 1.0, (T, cs.dec.mrow)), device="cuda")
    be = torch.tensor(rng.uniform(0.1, 0.9, (T, cs.dec.mrow)), device="cuda")
    w = torch.tensor(rng.normal(size=tuple(x.shape)), device="cuda")
    tr = _trainer(cs, T, 8, 1e30, True)

    def loss(a, b):
        return float((tr.forward(x, a, b) * w).sum().item())

    tr.forward(x, al, be)
    _, ga, gb = tr.backward(w)
    eps = 1e-6
    picks = [(int(t), int(l)) for t, l in zip(rng.integers(0, T, 16), rng.integers(0, cs.dec.mrow, 16))]
    for t_, l_ in picks:
        for which in ("a", "b"):
            d = torch.zeros_like(al)
            d[t_, l_] = eps
            if which == "a":
                fd = (loss(al + d, be) - loss(al - d, be)) / (2 * eps)
                g = float(ga[t_, l_])
            else:
                fd = (loss(al, be + d) - loss(al, be - d)) / (2 * eps)
                g = float(gb[t_, l_])
            assert abs(fd - g) <= 1e-5 * max(1.0, abs(fd)), (which, t_, l_, fd, g)


def test_gpu_autograd_function_matches_torch_reference(cuda):
    """The torch.autograd.Function over the C ABI against autograd of the independent
    torch forward (FP64)."""
    import torch
    from fecsim.train import min_sum_function
    cs = _cs("11ad 1/2")
    T = 6
    x = _llrs(cs, 5, -0.5).astype(np.float64)
    rng = np.random.default_rng(6)
    al = rng.uniform(0.6, 1.0, (T, cs.dec.mrow))
    be = rng.uniform(0.1, 0.9, (T, cs.dec.mrow))
    bits = (rng.random(x.shape) < 0.5).astype(np.float64)
    tr = _trainer(cs, T, 8, 1e30, True)
    fn = min_sum_function()
    ta = torch.tensor(al, device="cuda", requires_grad=True)
    tb = torch.tensor(be, device="cuda", requires_grad=True)
    tx = torch.tensor(x, device="cuda", requires_grad=True)
    loss = torch.nn.functional.binary_cross_entropy_with_logits(-fn.apply(tx, ta, tb, tr), torch.tensor(bits, device="cuda"))
    loss.backward()
    ra = torch.tensor(al, requires_grad=True)
    rb = torch.tensor(be, requires_grad=True)
    rx = torch.tensor(x, requires_grad=True)
    rl = torch.nn.functional.binary_cross_entropy_with_logits(-train_ref.forward(cs.dec, rx, ra, rb), torch.tensor(bits))
    rl.backward()
    assert abs(loss.item() - rl.item()) < 1e-12
    np.testing.assert_allclose(ta.grad.cpu().numpy(), ra.grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(tb.grad.cpu().numpy(), rb.grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(tx.grad.cpu().numpy(), rx.grad.numpy(), rtol=1e-9, atol=1e-12)


def test_gpu_training_reduces_loss(cuda):
    from fecsim.sim import Henc
    from fecsim.train import train
    cs = _cs("11ad 1/2")
    henc = Henc.from_tables(cs.enc)
    al, be, losses = train(cs, henc, T=5, batch=1024, steps=60, ebn0=(2.0, 3.0), lr=2e-2, log=_say, log_every=20)
    assert np.mean(losses[-10:]) < np.mean(losses[:10])
    assert al.shape == (5, cs.dec.mrow) and be.shape == (5, cs.dec.mrow)


def test_gpu_offset_search(cuda):
    """search_offsets: group rows move together, values stay in 1..6, the kept table has
    no more codeword errors than the start on the search batch, and the table it returns
    decodes like the oracle with that table."""
    import torch
    from fecsim.sim import Henc
    from fecsim.train import search_offsets
    cs = _cs("11ad 1/2")
    henc = Henc.from_tables(cs.enc)
    start = np.ones((6, cs.dec.mrow), np.float32)
    tab, e0, e1, v0, v1 = search_offsets(cs, henc, start, -1.0, [0, 2, 6], ncw=8192, sweeps=1, max_iter=6, log=_say)
    assert tab.shape == start.shape and np.array_equal(tab, np.rint(tab))
    assert tab.min() >= 1 and tab.max() <= 6
    assert np.array_equal(tab[0], tab[1]) and all(np.array_equal(tab[2], tab[t]) for t in range(3, 6))
    assert e1 <= e0 and e0 > 0 and v0 > 0
    llr = _llrs(cs, 200, -0.5, O.I8, 2.0, 24.0)
    want = O.decode_ideal(cs.dec, llr, max_iter=6, beta=1.0, beta_table=tab)
    d = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 6, 31.0, 1.0)
    d.set_offsets(tab)
    x = torch.from_numpy(llr).cuda()
    out = torch.empty_like(x)
    it = torch.zeros(200, dtype=torch.int32, device="cuda")
    d.decode(x, out, iters=it)
    torch.cuda.synchronize()
    assert np.array_equal(it.cpu().numpy().astype(np.uint32), want[2])
    assert np.array_equal(out.cpu().numpy(), want[0])


@pytest.mark.parametrize("name,T,snr", [("11ad 1/2", 6, -1.5), ("11n 1/2 z81", 8, -0.5)])
def test_gpu_train_int8_forward_matches_int8_decoder(name, T, snr, cuda):
    """The int8-emulating forward (FECSIM_I8 trainer) on int8 LLRs with an integer offset
    table equals the bit-exact int8 decoder (the oracle, same table) on every codeword
    that decoder ran for all T iterations."""
    import torch
    from fecsim.train import Trainer
    cs = _cs(name)
    llr = _llrs(cs, 300, snr, O.I8, 2.0, 24.0)
    rng = np.random.default_rng(12)
    tab = rng.integers(0, 4, (T, cs.dec.mrow)).astype(np.float32)
    want = O.decode_ideal(cs.dec, llr, max_iter=T, max_row_metric=31.0, beta=1.0, beta_table=tab)
    full = want[2] == T
    assert full.sum() >= 20, int(full.sum())
    tr = Trainer(Hdec.from_tables(cs.dec), T, 300, cap=31.0, int8=True)
    x = torch.from_numpy(llr.astype(np.float32)).cuda()
    al = torch.ones((T, cs.dec.mrow), device="cuda")
    got = tr.forward(x, al, torch.from_numpy(tab + 0.3).cuda())   # rounds to the table
    torch.cuda.synchronize()
    g = got.cpu().numpy()
    assert np.array_equal(g, np.rint(g))
    assert np.array_equal(g[full].astype(np.int8), want[0][full])


def test_gpu_train_int8_gradient_matches_torch_reference(cuda):
    """Gradients of the int8-emulating decoder (saturation, rounded offsets with a
    straight-through gradient, the offset-0 wrap) against autograd of the independent
    torch restatement in the same mode."""
    import torch
    from fecsim.train import Trainer, min_sum_function
    cs = _cs("11ad 1/2")
    T = 5
    x = _llrs(cs, 6, 1.0, O.I8, 2.0, 24.0).astype(np.float32) * 4   # large LLRs: saturation binds
    x = np.clip(x, -128, 127)
    rng = np.random.default_rng(21)
    be = rng.integers(0, 4, (T, cs.dec.mrow)).astype(np.float64) + rng.uniform(-0.3, 0.3, (T, cs.dec.mrow))
    be = np.clip(be, 0.0, None)
    bits = (rng.random(x.shape) < 0.5).astype(np.float64)
    tr = Trainer(Hdec.from_tables(cs.dec), T, 8, cap=31.0, int8=True)
    fn = min_sum_function()
    tb = torch.tensor(be, dtype=torch.float32, device="cuda", requires_grad=True)
    tx = torch.tensor(x, device="cuda", requires_grad=True)
    ta = torch.ones((T, cs.dec.mrow), device="cuda")
    post = fn.apply(tx, ta, tb, tr)
    loss = torch.nn.functional.binary_cross_entropy_with_logits(-post / 2, torch.tensor(bits, dtype=torch.float32, device="cuda"))
    loss.backward()
    rb = torch.tensor(be, requires_grad=True)
    rx = torch.tensor(x.astype(np.float64), requires_grad=True)
    rp = train_ref.forward(cs.dec, rx, torch.ones((T, cs.dec.mrow), dtype=torch.float64), rb, cap=31.0, int8=True)
    assert np.array_equal(post.detach().cpu().numpy(), rp.detach().numpy().astype(np.float32))
    assert (np.abs(rp.detach().numpy()) == 127).any() or (rp.detach().numpy() == -128).any()   # saturation occurred
    rl = torch.nn.functional.binary_cross_entropy_with_logits(-rp / 2, torch.tensor(bits))
    rl.backward()
    np.testing.assert_allclose(tb.grad.cpu().numpy(), rb.grad.numpy(), rtol=1e-4, atol=1e-6)
    np.testing.assert_allclose(tx.grad.cpu().numpy(), rx.grad.numpy(), rtol=1e-4, atol=1e-7)


@pytest.mark.parametrize("greedy,loss_iters", [(True, None), (False, [1, 2])])
def test_gpu_train_int8_schedules(greedy, loss_iters, cuda):
    """train_int8 (greedy, and joint with a multi-iteration loss) on 11ad 1/2 at T = 3:
    one finite loss per step, an integer table of shape [T, mrow] within [0, 20], and
    the bit-exact decoder with that table decodes like the oracle with it."""
    import torch
    from fecsim.sim import Henc
    from fecsim.train import train_int8
    cs = _cs("11ad 1/2")
    henc = Henc.from_tables(cs.enc)
    T = 3
    tab, losses = train_int8(cs, henc, T=T, batch=64, steps=6, ebn0=(0.0, 1.0), lr=0.5, greedy=greedy,
                             loss_iters=loss_iters)
    assert len(losses) == 6 and np.isfinite(losses).all()
    assert tab.shape == (T, cs.dec.mrow) and np.array_equal(tab, np.rint(tab))
    assert tab.min() >= 0 and tab.max() <= 20
    with pytest.raises(ValueError):
        train_int8(cs, henc, T=T, batch=64, steps=1, loss_iters=[T + 1])
    llr = _llrs(cs, 200, -0.5, O.I8, 2.0, 24.0)
    want = O.decode_ideal(cs.dec, llr, max_iter=T, max_row_metric=31.0, beta=1.0, beta_table=tab)
    d = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", T, 31.0, 1.0)
    d.set_offsets(tab)
    x = torch.from_numpy(llr).cuda()
    out = torch.empty_like(x)
    d.decode(x, out)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), want[0])


def test_gpu_train_backward_keeps_its_parameters(cuda):
    """The backward uses the forward's alpha and beta even when the caller's tensors were
    temporaries freed (and their memory reused) in between; a second backward without a
    new forward is refused."""
    import torch
    from fecsim.lib import FecsimError
    from fecsim.train import Trainer
    cs = _cs("11ad 1/2")
    T = 5
    x = torch.tensor(_llrs(cs, 16, 0.0).astype(np.float32), device="cuda")
    rng = np.random.default_rng(5)
    al = torch.tensor(rng.uniform(0.6, 1.0, (T, cs.dec.mrow)), dtype=torch.float32, device="cuda")
    be = torch.tensor(rng.uniform(0.1, 0.9, (T, cs.dec.mrow)), dtype=torch.float32, device="cuda")
    w = torch.tensor(rng.normal(size=tuple(x.shape)), dtype=torch.float32, device="cuda")
    tr = Trainer(Hdec.from_tables(cs.dec), T, 16, cap=8.0)
    tr.forward(x, al, be)
    _, ga, gb = tr.backward(w)
    tr.forward(x, al.clone(), be.clone())   # temporaries, freed when the call returns
    junk = [torch.full_like(al, 1e9) for _ in range(64)]   # takes their memory back
    _, ga2, gb2 = tr.backward(w)
    torch.cuda.synchronize()
    del junk
    assert np.array_equal(ga.cpu().numpy(), ga2.cpu().numpy())
    assert np.array_equal(gb.cpu().numpy(), gb2.cpu().numpy())
    with pytest.raises(FecsimError):
        tr.backward(w)


@pytest.mark.parametrize("name,T,cap,int8", [("11n 1/2 z81", 10, 6.0, False), ("11n 1/2 z54", 7, 1e30, False),
                                            ("11n 1/2 z27", 4, 1e30, False), ("11ad 1/2", 9, 1e30, False),
                                            ("16e 1/2", 5, 8.0, False), ("11ad 1/2", 6, 31.0, True),
                                            ("11n 1/2 z81", 5, 31.0, True), ("11ad 1/2", 1, 1e30, False),
                                            ("11ad 1/2", 2, 4.0, False), ("11n 1/2 z81", 1, 1e30, False),
                                            ("11n 1/2 z81", 20, 6.0, False), ("11ad 1/2", 13, 31.0, True)])
def test_gpu_train_register_kernels_match_global_c2v(name, T, cap, int8, cuda):
    """The register-C2V kernels (compiled shapes, one or two waves per codeword) against
    the global-C2V kernels forced by FECSIM_TRAIN_GLOBAL_C2V: posteriors and all three
    gradients equal bit for bit, from one segment (T = 1, 2: nothing recomputed) to several
    (T = 20: segments of 10), and a batch that is not a multiple of the global kernels'
    waves per block."""
    import torch
    from fecsim.train import Trainer
    cs = _cs(name)
    B = 45
    if int8:
        x = np.clip(_llrs(cs, B, 0.0, O.I8, 2.0, 24.0).astype(np.float32) * 3, -128, 127)
    else:
        x = _llrs(cs, B, -0.5).astype(np.float32)
    rng = np.random.default_rng(31)
    al = np.ones((T, cs.dec.mrow), np.float32) if int8 else rng.uniform(0.6, 1.0, (T, cs.dec.mrow)).astype(np.float32)
    be = rng.uniform(0.0, 3.0 if int8 else 0.9, (T, cs.dec.mrow)).astype(np.float32)
    w = rng.normal(size=x.shape).astype(np.float32)
    t = lambda a: torch.from_numpy(a).cuda()   # noqa: E731
    res = []
    for glob in (False, True):
        tr = Trainer(Hdec.from_tables(cs.dec), T, 64, cap=cap, int8=int8, global_c2v=glob)
        post = tr.forward(t(x), t(al), t(be))
        gl, ga, gb = tr.backward(t(w))
        torch.cuda.synchronize()
        res.append([v.cpu().numpy() for v in (post, gl, ga, gb)])
    for a, b, what in zip(res[0], res[1], ("post", "grad_llr", "grad_alpha", "grad_beta")):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), what
    assert np.abs(res[0][2]).max() > 0 and np.abs(res[0][3]).max() > 0
