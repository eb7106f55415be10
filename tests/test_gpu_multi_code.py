"""GPU parity of the mixed-code queue kernel with the codes compiled in (k_multi_code):
the same queue decoded by it and by the generic queue kernel, and by every code's own
single decoder, gives identical posteriors, iteration counts and totals, also for
queues that leave codes out, tiny queues, long ones (slots refilled many times, work
groups moving between codes), offset tables and HIP-graph replays."""
import sys

import numpy as np
import pytest

from fecsim import codes
from fecsim import lib as L
from fecsim.decoder import Decoder, Hdec
from fecsim.multi import MultiDecoder
from fecsim.sim import Henc

pytestmark = pytest.mark.gpu

# the codes compiled into k_multi_code: (matrix set, z_over, max iterations)
CODES = [("11n 1/2 z27", None, 8), ("11n 3/4 z54", None, 12), ("11n 5/6 z81", None, 16),
         ("11ad 13/16", None, 20), ("16e 2/3A", 96, 10)]


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


# other codes compiled into k_multi_code (the rest of the 802.11n and 802.11ad families)
CODES_B = [("11n 2/3 z27", None, 10), ("11n 5/6 z54", None, 12), ("11ad 1/2", None, 10), ("11ad 3/4", None, 8),
           ("11n 1/2 z81", None, 6), ("11n 3/4 z27", None, 9), ("11ad 5/8", None, 10)]


def _decoders(tables=False, codes_=CODES):
    css, decs = [], []
    for name, zo, mi in codes_:
        cs = codes.gen_structs(name, simd_size=512, elem_size=8, z_over=zo)
        d = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", mi, 31.0, 1.0)
        if tables:   # per-(iteration, layer) offsets 1..3
            rng = np.random.default_rng(len(css))
            d.set_offsets(rng.integers(1, 4, (mi, cs.dec.mrow)).astype(np.float32))
        css.append(cs)
        decs.append(d)
    return css, decs


def _queue(css, n, seed, weights=None, ebn0=(1.0, 4.0)):
    """n entries, code ids drawn with `weights`; each code's codewords at one Eb/N0."""
    import torch
    rng = np.random.default_rng(seed)
    p = None if weights is None else np.asarray(weights, float) / np.sum(weights)
    code = rng.choice(len(css), n, p=p).astype(np.int32)
    parts, offs, base = [], np.zeros(n, np.int64), 0
    for c, cs in enumerate(css):
        idx = np.nonzero(code == c)[0]
        henc = Henc.from_tables(cs.enc)
        k = max(1, len(idx))
        bits = henc.new_bits(k)
        henc.gen_codewords(bits, seed=seed, cw0=0)
        rate = (cs.enc.mcol - cs.enc.mrow) / cs.enc.mcol
        snr = float(rng.uniform(*ebn0) + 10 * np.log10(rate))
        llr = torch.empty((k, henc.n), dtype=torch.int8, device="cuda")
        henc.gen_llr(bits, llr, L.I8, snr, seed=seed)
        parts.append(llr[:len(idx)])
        offs[idx] = base + np.arange(len(idx)) * henc.n
        base += len(idx) * henc.n
    packed = torch.cat([q.reshape(-1) for q in parts]).contiguous()
    return code, offs, packed, parts


def _run(m, code, offs, packed):
    import torch
    n = len(code)
    post = torch.zeros_like(packed)
    iters = torch.zeros(n, dtype=torch.int32, device="cuda")
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    m.decode(torch.from_numpy(code).cuda(), torch.from_numpy(offs).cuda(), packed, post, iters=iters, tot_iter=tot)
    torch.cuda.synchronize()
    return post.cpu().numpy(), iters.cpu().numpy(), int(tot.item())


def _single(decs, parts, code, offs, post, iters):
    """every entry equals its code's single decoder"""
    import torch
    for c, (d, p) in enumerate(zip(decs, parts)):
        if not p.shape[0]:
            continue
        o = torch.empty_like(p)
        it = torch.zeros(p.shape[0], dtype=torch.int32, device="cuda")
        d.decode(p, o, iters=it)
        o = o.cpu().numpy()
        idx = np.nonzero(code == c)[0]
        for k, i in enumerate(idx):
            assert np.array_equal(post[offs[i]:offs[i] + o.shape[1]], o[k]), (c, k)
        assert np.array_equal(iters[idx], it.cpu().numpy()), c


def test_gpu_multi_code_kernel_is_chosen_for_compiled_codes(cuda):
    css, decs = _decoders()
    m = MultiDecoder(decs, 16)
    assert m.kernel == L.KERNEL_CODE
    m.select_kernel(L.KERNEL_GENERIC)
    assert m.kernel == L.KERNEL_GENERIC
    m.select_kernel(L.KERNEL_AUTO)
    assert m.kernel == L.KERNEL_CODE
    with pytest.raises(L.FecsimError):
        m.select_kernel(L.KERNEL_PAIR)
    # a code that is not compiled in: the generic kernel only
    cs = codes.gen_structs("ITU G.hn 1/2 z14", simd_size=512, elem_size=8)
    other = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 10, 31.0, 1.0)
    m2 = MultiDecoder(decs[:2] + [other], 16)
    assert m2.kernel == L.KERNEL_GENERIC
    with pytest.raises(L.FecsimError) as ei:
        m2.select_kernel(L.KERNEL_CODE)
    assert ei.value.status == -6   # FECSIM_E_UNSUPPORTED
    # beta 0 (the reference's abs(-128) wrap applies): generic kernel
    d0 = Decoder(Hdec.from_tables(css[0].dec), "Vec64c_min", 8, 31.0, 0.0)
    assert MultiDecoder([d0] + decs[1:], 16).kernel == L.KERNEL_GENERIC


@pytest.mark.parametrize("n,seed,weights", [
    (1, 1, None), (7, 2, None), (300, 3, [1, 0, 0, 0, 0]), (300, 4, [0, 0, 1, 0, 1]),
    (2000, 5, None), (20000, 6, [1, 1, 4, 1, 4]),
])
def test_gpu_multi_code_matches_generic_and_single(n, seed, weights, cuda):
    css, decs = _decoders()
    code, offs, packed, parts = _queue(css, n, seed, weights)
    _say("multi code kernel: n", n, "seed", seed, "per code", np.bincount(code, minlength=len(CODES)).tolist())
    m = MultiDecoder(decs, n)
    assert m.kernel == L.KERNEL_CODE
    post, iters, tot = _run(m, code, offs, packed)
    m.select_kernel(L.KERNEL_GENERIC)
    post_g, iters_g, tot_g = _run(m, code, offs, packed)
    assert np.array_equal(post, post_g)
    assert np.array_equal(iters, iters_g)
    assert tot == tot_g == int(iters.sum())
    if n <= 2000:
        _single(decs, parts, code, offs, post, iters)


def test_gpu_multi_code_other_compiled_codes(cuda):
    """a queue of seven more compiled codes (11n 2/3, 5/6 z54, 1/2 z81, 3/4 z27; 11ad 1/2,
    3/4, 5/8): the compiled-code kernel equals the generic one and the single decoders"""
    css, decs = _decoders(codes_=CODES_B)
    n = 4000
    code, offs, packed, parts = _queue(css, n, 41, ebn0=(0.5, 4.5))
    _say("multi code kernel, other codes: per code", np.bincount(code, minlength=len(CODES_B)).tolist())
    m = MultiDecoder(decs, n)
    assert m.kernel == L.KERNEL_CODE
    post, iters, tot = _run(m, code, offs, packed)
    m.select_kernel(L.KERNEL_GENERIC)
    post_g, iters_g, tot_g = _run(m, code, offs, packed)
    assert np.array_equal(post, post_g) and np.array_equal(iters, iters_g) and tot == tot_g
    _single(decs, parts, code, offs, post, iters)


def test_gpu_multi_code_offset_tables(cuda):
    css, decs = _decoders(tables=True)
    n = 1500
    code, offs, packed, parts = _queue(css, n, 11)
    m = MultiDecoder(decs, n)
    assert m.kernel == L.KERNEL_CODE
    post, iters, tot = _run(m, code, offs, packed)
    m.select_kernel(L.KERNEL_GENERIC)
    post_g, iters_g, tot_g = _run(m, code, offs, packed)
    assert np.array_equal(post, post_g) and np.array_equal(iters, iters_g) and tot == tot_g
    _single(decs, parts, code, offs, post, iters)


def test_gpu_multi_code_graph_replays_and_static(cuda):
    import torch
    css, decs = _decoders()
    n = 3000
    code, offs, packed, parts = _queue(css, n, 21)
    m = MultiDecoder(decs, n)
    want_post, want_iters, want_tot = _run(m, code, offs, packed)
    tcode, toff = torch.from_numpy(code).cuda(), torch.from_numpy(offs).cuda()
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
        assert np.array_equal(post.cpu().numpy(), want_post), r
        assert np.array_equal(iters.cpu().numpy(), want_iters), r
    for k in (n, 500):   # the static workspace, then a shorter queue
        post.zero_()
        tot = torch.zeros(1, dtype=torch.int64, device="cuda")
        it = torch.zeros(k, dtype=torch.int32, device="cuda")
        m.decode_static(tcode[:k], toff[:k], packed, post, iters=it, tot_iter=tot)
        torch.cuda.synchronize()
        p = post.cpu().numpy()
        for i in range(0, k, 37):
            ln = css[code[i]].dec.mcol * css[code[i]].dec.z
            assert np.array_equal(p[offs[i]:offs[i] + ln], want_post[offs[i]:offs[i] + ln]), (k, i)
        assert np.array_equal(it.cpu().numpy(), want_iters[:k]), k
        assert int(tot.item()) == int(want_iters[:k].sum())


def test_gpu_mixed_sim_fill_is_the_next_loops_input(cuda):
    """fecsim_msim_fill writes exactly the LLRs the next loop decodes: the queue decoder on
    a copy of them gives the loop's in-place posteriors and its iteration total."""
    import torch
    from fecsim.multi import MixedSim
    css, decs = _decoders()
    hencs = [Henc.from_tables(cs.enc) for cs in css]
    n = 3000
    m = MultiDecoder(decs, n)
    sim = MixedSim(hencs, m, n, ebn0=(1.0, 4.0), seed=31, use_graph=False)
    sim.run(1)   # the fill is for the loop at the current codeword counter, not the first one
    sim.reset()
    sim.fill()
    fresh = sim.llr().clone()
    code, off, _, _ = sim.queue()
    post = torch.zeros_like(fresh)
    it = torch.zeros(n, dtype=torch.int32, device="cuda")
    m.decode(code, off, fresh, post, iters=it)
    sim.run(1)
    torch.cuda.synchronize()
    assert torch.equal(sim.llr(), post)
    st, iters = sim.stats()
    assert sum(s["cw"] for s in st) == n
    assert iters == int(it.sum().item())
