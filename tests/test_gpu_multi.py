"""GPU parity of the mixed-code queue decoder: every queue entry decoded exactly as
its own code's single decoder (and the oracle) would, before and after autotuning,
eagerly and in HIP-graph replays."""
import sys

import numpy as np
import pytest

from fecsim import codes
from fecsim.decoder import Decoder, Hdec
from fecsim.multi import MultiDecoder
from fecsim.sim import Henc
import refsim
from refsim import O

pytestmark = pytest.mark.gpu

# the mixed queue: (matrix set, z_over, max iterations)
CODES = [("11n 1/2 z27", None, 8), ("11n 3/4 z54", None, 12), ("11n 5/6 z81", None, 16),
         ("11ad 13/16", None, 20), ("16e 2/3A", 96, 10)]


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _queue(n, seed, kind="Vec64c_min"):
    import torch
    from fecsim import lib as L
    rng = np.random.default_rng(seed)
    code = rng.integers(0, len(CODES), n).astype(np.int32)
    css, decs, parts, offs = [], [], [], np.zeros(n, np.int64)
    base = 0
    for c, (name, zo, mi) in enumerate(CODES):
        cs = codes.gen_structs(name, simd_size=512, elem_size=8, z_over=zo)
        css.append(cs)
        decs.append(Decoder(Hdec.from_tables(cs.dec), kind, mi, 31.0, 1.0))
        idx = np.nonzero(code == c)[0]
        henc = Henc.from_tables(cs.enc)
        bits = henc.new_bits(max(1, len(idx)))
        henc.gen_codewords(bits, seed=seed, cw0=0)
        rate = (cs.enc.mcol - cs.enc.mrow) / cs.enc.mcol
        snr = float(rng.uniform(1.0, 4.0) + 10 * np.log10(rate))
        llr = torch.empty((max(1, len(idx)), henc.n), dtype=torch.int8, device="cuda")
        henc.gen_llr(bits, llr, L.I8, snr, seed=seed)
        parts.append(llr[:len(idx)])
        offs[idx] = base + np.arange(len(idx)) * henc.n
        base += len(idx) * henc.n
    packed = torch.cat([p.reshape(-1) for p in parts]).contiguous()
    return code, offs, packed, parts, css, decs


def _expected(parts, decs):
    import torch
    outs, its = [], []
    for p, d in zip(parts, decs):
        o = torch.empty_like(p)
        it = torch.zeros(max(1, p.shape[0]), dtype=torch.int32, device="cuda")
        if p.shape[0]:
            d.decode(p, o, iters=it)
        outs.append(o)
        its.append(it[:p.shape[0]])
    return outs, its


def _check(code, offs, post, iters, outs, its):
    post = post.cpu().numpy()
    iters = iters.cpu().numpy()
    for c in range(len(CODES)):
        idx = np.nonzero(code == c)[0]
        o = outs[c].cpu().numpy()
        n = o.shape[1] if o.ndim == 2 else 0
        for k, i in enumerate(idx):
            assert np.array_equal(post[offs[i]:offs[i] + n], o[k]), (c, k)
        assert np.array_equal(iters[idx], its[c].cpu().numpy()), c


def test_gpu_multi_queue_matches_single_decoders_and_oracle(cuda):
    import torch
    n = 3000
    code, offs, packed, parts, css, decs = _queue(n, 5)
    outs, its = _expected(parts, decs)
    # the single decoders against the oracle on a slice of every code
    for c, (cs, p, o) in enumerate(zip(css, parts, outs)):
        k = min(25, p.shape[0])
        want = O.decode_ideal(cs.dec, p[:k].cpu().numpy(), max_iter=CODES[c][2], want_c2v=False)
        assert np.array_equal(o[:k].cpu().numpy(), want[0]), c
        assert np.array_equal(its[c][:k].cpu().numpy().astype(np.uint32), want[2]), c
    m = MultiDecoder(decs, n)
    tcode = torch.from_numpy(code).cuda()
    toff = torch.from_numpy(offs).cuda()
    post = torch.zeros_like(packed)
    iters = torch.zeros(n, dtype=torch.int32, device="cuda")
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    m.decode(tcode, toff, packed, post, iters=iters, tot_iter=tot)
    torch.cuda.synchronize()
    _check(code, offs, post, iters, outs, its)
    assert int(tot.item()) == int(iters.sum().item())
    shapes0 = [m.shape(c) for c in range(len(CODES))]
    # autotune on the same queue, then decode again: identical results
    scratch = torch.zeros_like(packed)
    m.autotune(tcode, toff, packed, scratch, rounds=2)
    shapes1 = [m.shape(c) for c in range(len(CODES))]
    _say("tile shapes before", shapes0, "after autotune", shapes1)
    post2 = torch.zeros_like(packed)
    iters2 = torch.zeros(n, dtype=torch.int32, device="cuda")
    m.decode(tcode, toff, packed, post2, iters=iters2)
    torch.cuda.synchronize()
    _check(code, offs, post2, iters2, outs, its)


def test_gpu_multi_queue_graph_replays(cuda):
    import torch
    n = 1500
    code, offs, packed, parts, css, decs = _queue(n, 9)
    outs, its = _expected(parts, decs)
    m = MultiDecoder(decs, n)
    tcode = torch.from_numpy(code).cuda()
    toff = torch.from_numpy(offs).cuda()
    post = torch.zeros_like(packed)
    iters = torch.zeros(n, dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    g = torch.cuda.CUDAGraph()
    torch.cuda.synchronize()
    with torch.cuda.graph(g, stream=s):
        m.decode(tcode, toff, packed, post, iters=iters, stream=s)
    for r in range(3):
        post.zero_()
        iters.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        torch.zeros(1000, device="cuda").sum()   # eager work between replays
        _check(code, offs, post, iters, outs, its)


@pytest.mark.parametrize("graph", [False, True])
def test_gpu_mixed_sim_loop_counts(graph, cuda):
    """The mixed-code simulation loop: per-code counters equal a recount of one loop
    done with the separate steps (encoder, per-entry channel, queue decoder, counter)
    on the same codeword indices; graph replays advance the codewords."""
    import ctypes as C
    import torch
    from fecsim import lib as L
    from fecsim.decoder import _ptr
    from fecsim.multi import MixedSim
    n = 2000
    decs, hencs, css = [], [], []
    for name, zo, mi in CODES:
        cs = codes.gen_structs(name, simd_size=512, elem_size=8, z_over=zo)
        css.append(cs)
        decs.append(Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", mi, 31.0, 1.0))
        hencs.append(Henc.from_tables(cs.enc))
    m = MultiDecoder(decs, n)
    sim = MixedSim(hencs, m, n, ebn0=(1.0, 4.0), seed=77, use_graph=graph)
    sim.run(2)
    st, iters = sim.stats()
    assert sum(s["cw"] for s in st) == 2 * n
    # recount loop 1 (codeword indices n .. 2n-1) with the separate entry points
    code, off, _, sf = sim.queue()
    code = code.cpu().numpy()
    sf = sf.cpu().numpy().reshape(-1, 2)
    counts = np.bincount(code, minlength=len(CODES))
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    tot_cw_err = tot_bit_err = 0
    for c, (cs, h, d) in enumerate(zip(css, hencs, decs)):
        k = int(counts[c])
        if not k:
            continue
        bits = h.new_bits(k)
        h.gen_codewords(bits, seed=77, cw0=n + int(first[c]))
        llr = torch.empty((k, h.n), dtype=torch.int8, device="cuda")
        sft = torch.from_numpy(np.ascontiguousarray(sf[first[c]:first[c] + k])).cuda()
        L.check(L.lib().fecsim_gen_llr_var(h._h, _ptr(sft), C.c_float(24.0), 77, n + int(first[c]), k, h.col_words,
                                           _ptr(bits), L.I8, 0, _ptr(llr), None,
                                           C.c_void_p(torch.cuda.current_stream().cuda_stream)), "gen_llr_var")
        post = torch.empty_like(llr)
        d.decode(llr, post)
        cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
        h.count_errors(bits, post, L.I8, cnt)
        be, ce = (int(v) for v in cnt.tolist())
        tot_bit_err += be
        tot_cw_err += ce
    sim.reset()
    sim.run(1)   # loop 2: indices 2n .. 3n-1, so compare against a fresh count of loop 1 below
    st2, _ = sim.stats()
    assert sum(s["cw"] for s in st2) == n
    # loop 1 alone = (two loops) - (loop 0); recompute loop 0 the same way
    tot0_cw = tot0_bit = 0
    for c, (cs, h, d) in enumerate(zip(css, hencs, decs)):
        k = int(counts[c])
        if not k:
            continue
        bits = h.new_bits(k)
        h.gen_codewords(bits, seed=77, cw0=int(first[c]))
        llr = torch.empty((k, h.n), dtype=torch.int8, device="cuda")
        sft = torch.from_numpy(np.ascontiguousarray(sf[first[c]:first[c] + k])).cuda()
        L.check(L.lib().fecsim_gen_llr_var(h._h, _ptr(sft), C.c_float(24.0), 77, int(first[c]), k, h.col_words,
                                           _ptr(bits), L.I8, 0, _ptr(llr), None,
                                           C.c_void_p(torch.cuda.current_stream().cuda_stream)), "gen_llr_var")
        post = torch.empty_like(llr)
        d.decode(llr, post)
        cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
        h.count_errors(bits, post, L.I8, cnt)
        be, ce = (int(v) for v in cnt.tolist())
        tot0_bit += be
        tot0_cw += ce
    assert sum(s["cw_err"] for s in st) == tot0_cw + tot_cw_err
    assert sum(s["bit_err"] for s in st) == tot0_bit + tot_bit_err
    _say("mixed sim", graph, "cw errors", [s["cw_err"] for s in st], "iterations", iters)


def test_gpu_held_decoder_refuses_offset_and_kernel_changes(cuda):
    """A queue decoder and a simulation loop copy their decoders' tables and kernel
    choice when made; while they exist the decoder refuses set_offsets / select_kernel
    (FECSIM_E_INVAL) and their results stay those of the decoder as it was."""
    import torch
    from fecsim import lib as L
    from fecsim.sim import Sim
    n = 600
    code, offs, packed, parts, css, decs = _queue(n, seed=77)
    outs, its = _expected(parts, decs)
    m = MultiDecoder(decs, max_cw=n)
    tab = np.full((decs[0].max_iter, css[0].dec.mrow), 2.0, np.float32)
    for bad in (lambda: decs[0].set_offsets(tab), lambda: decs[0].set_offsets(None),
                lambda: decs[1].select_kernel(L.KERNEL_GENERIC)):
        with pytest.raises(L.FecsimError) as ei:
            bad()
        assert ei.value.status == -1
    post = torch.zeros_like(packed)
    iters = torch.zeros(n, dtype=torch.int32, device="cuda")
    m.decode(torch.from_numpy(code).cuda(), torch.from_numpy(offs).cuda(), packed, post, iters=iters)
    torch.cuda.synchronize()
    _check(code, offs, post, iters, outs, its)
    del m
    decs[0].set_offsets(tab)          # free again
    decs[0].set_offsets(None)
    cs = css[0]
    sim = Sim(Henc.from_tables(cs.enc), decs[0], 64, [1.0], seed=3)
    with pytest.raises(L.FecsimError):
        decs[0].set_offsets(tab)
    del sim
    decs[0].set_offsets(tab)
