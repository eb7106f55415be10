"""GPU parity of the simulation kernels (through the C ABI) against the CPU oracle:
data bits + encoder, noise + quantized LLRs, error counting, and the device
simulation loop (eager and HIP-graph replays) -- all bit-exact."""
import sys

import numpy as np
import pytest

from fecsim import codes
from fecsim.decoder import Decoder, Hdec
from fecsim.sim import Henc, Sim, chan_sigma
import refsim
from refsim import O

pytestmark = pytest.mark.gpu

# (matrix set, num_par, z_over): Z < 32, Z multiple of 32, odd word counts, lifted
# num_par, 11ay lifting, 5G NR with puncturing and filler bits, BG1 Z=384
ENC_CASES = [
    ("11ad 1/2", 1, None),
    ("11ad 13/16", 1, None),
    ("11n 1/2 z27", 1, None),
    ("11n 5/6 z81", 1, None),
    ("16e 2/3A", 1, 96),
    ("16e 1/2", 1, 64),
    ("ITU G.hn 1/2 z14", 1, None),
    ("11ay 7/8", 1, None),
    ("11ad 1/2", 3, None),
    ("5GNR_n600_k300", 1, None),
    ("5GNR_n500_k220", 1, None),
    ("5GNR_n25344_k8448", 1, None),
    ("802.3ca", 1, None),
    ("802.3ca", 2, None),
]


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _cs(name, npar=1, zo=None):
    return codes.gen_structs(name, num_par=npar, simd_size=512, elem_size=8, z_over=zo)


def _oracle_codewords(cs, seed, cw0, ncw):
    """[ncw][mcol][64] byte words: the oracle's counter-based data bits + encoder."""
    z, mrow, mcol = cs.z, cs.enc.mrow, cs.enc.mcol
    _, fill = refsim.word_masks(cs.enc)
    bits = np.zeros((ncw, mcol, 64), np.uint8)
    O.lib().orc_gpu_databits(seed, cw0, ncw, z, mrow, mcol, fill.ctypes.data, bits.ctypes.data)
    cnt, colnum, sc, ss = refsim.enc_arrays(cs.enc)
    O.lib().orc_encode(ncw, z, mrow, mcol, cnt.ctypes.data, colnum.ctypes.data, sc.ctypes.data, ss.ctypes.data,
                       bits.ctypes.data)
    return bits


def _words(t, mcol):
    """GPU bits [ncw, mcol, col_words] int32 -> [ncw][mcol][64] byte words."""
    a = t.cpu().numpy().view(np.uint32)
    out = np.zeros((a.shape[0], mcol, 16), np.uint32)
    k = min(16, a.shape[2])
    out[:, :, :k] = a[:, :, :k]
    return out.view(np.uint8)


def _dense_masks(cs):
    pm, fm = cs.enc.masks()
    return pm.reshape(-1).astype(np.uint8), fm.reshape(-1).astype(np.uint8)


def _syndrome_zero(cs, dense):
    """H c = 0 over GF(2) from the base matrix (block (i, j) with shift s:
    check row i*Z + r sees bit j*Z + (r + s) mod Z)."""
    z = cs.z
    hc = cs.hc
    b = dense.reshape(dense.shape[0], hc.shape[1], z)
    r = np.arange(z)
    for i in range(hc.shape[0]):
        acc = np.zeros((dense.shape[0], z), np.uint8)
        for j in range(hc.shape[1]):
            if hc[i, j] >= 0:
                acc ^= b[:, j, (r + hc[i, j]) % z]
        if acc.any():
            return False
    return True


@pytest.mark.parametrize("name,npar,zo", ENC_CASES)
def test_gpu_gen_codewords_matches_oracle(name, npar, zo, cuda):
    import torch
    cs = _cs(name, npar, zo)
    h = Henc.from_tables(cs.enc)
    ncw = 37 if cs.z < 300 else 5
    _say("gen_codewords", name, npar, zo, "z", cs.z, "ncw", ncw)
    want = _oracle_codewords(cs, 1234, 555, ncw)
    for cwords in (h.col_words, 16):
        bits = torch.full((ncw, cs.enc.mcol, cwords), -1, dtype=torch.int32, device="cuda")
        h.gen_codewords(bits, seed=1234, cw0=555)
        torch.cuda.synchronize()
        got = _words(bits, cs.enc.mcol)
        assert np.array_equal(got, want), (name, cwords)
        if cwords > h.col_words:   # words past ceil(z/32) are cleared
            assert not bits[:, :, h.col_words:].any().item()
    dense = refsim.unpack_bits(want, cs.z, cs.enc.mcol)
    assert _syndrome_zero(cs, dense)


@pytest.mark.parametrize("name,npar,zo", [c for c in ENC_CASES if c[0] != "5GNR_n25344_k8448"])
def test_gpu_encode_reference_data_stream(name, npar, zo, cuda):
    """fecsim_encode on the reference's own data bits (xoshiro256++ stream of
    gen_rand_databits) in the reference's 64-byte word layout."""
    import torch
    cs = _cs(name, npar, zo)
    h = Henc.from_tables(cs.enc)
    rs = refsim.RefStream(cs)
    _, want = rs.inner_loop(23)
    data = want.copy()
    data[:, cs.enc.mcol - cs.enc.mrow:, :] = 0
    bits = torch.from_numpy(data.view(np.int32).reshape(23, cs.enc.mcol, 16).copy()).cuda()
    h.encode(bits)
    torch.cuda.synchronize()
    assert np.array_equal(_words(bits, cs.enc.mcol), want)


@pytest.mark.parametrize("name,npar,zo", [("11ad 1/2", 1, None), ("5GNR_n600_k300", 1, None),
                                         ("11n 1/2 z27", 2, None), ("16e 1/2", 1, 92)])
@pytest.mark.parametrize("elem", [O.I8, O.F32, O.I16])
def test_gpu_llr_matches_oracle(name, npar, zo, elem, cuda):
    import torch
    cs = _cs(name, npar, zo)
    h = Henc.from_tables(cs.enc)
    ncw, seed, cw0, snr = 29, 99, 7000, -0.75
    scale, mx = (1.0, 1e9) if elem == O.F32 else (2.0, 24.0)
    _say("llr", name, npar, zo, "elem", elem)
    bits = h.new_bits(ncw)
    h.gen_codewords(bits, seed=seed, cw0=cw0)
    dense = refsim.unpack_bits(_words(bits, cs.enc.mcol), cs.z, cs.enc.mcol)
    punct, fill = _dense_masks(cs)
    sigma, f = chan_sigma(snr, scale)
    want, werr = O.noise_llr(seed, cw0, dense, cs.z, cs.enc.mcol, sigma, f, mx, punct, fill, elem)
    tdt = {O.I8: torch.int8, O.I16: torch.int16, O.F32: torch.float32}[elem]
    llr = torch.zeros((ncw, cs.enc.mcol * cs.z), dtype=tdt, device="cuda")
    unc = torch.zeros(1, dtype=torch.int64, device="cuda")
    h.gen_llr(bits, llr, elem, snr, seed=seed, cw0=cw0, llr_scale=scale, max_abs_llr=mx, uncoded=unc)
    torch.cuda.synchronize()
    assert np.array_equal(llr.cpu().numpy().view(np.uint8), want.view(np.uint8))
    assert int(unc.item()) == werr
    # the reference layout (P = 64 lanes): same values at +P + c*(ceil(z/P)+2)*P
    P = 64
    spz = (cs.z + P - 1) // P
    ref = torch.zeros((ncw, cs.enc.mcol * (spz + 2) * P), dtype=tdt, device="cuda")
    h.gen_llr(bits, ref, elem, snr, seed=seed, cw0=cw0, llr_scale=scale, max_abs_llr=mx, layout_p=P)
    torch.cuda.synchronize()
    assert np.array_equal(O.ref_to_dense(ref.cpu().numpy(), cs.z, cs.enc.mcol, P).view(np.uint8),
                          want.view(np.uint8))


@pytest.mark.parametrize("name,npar", [("11ad 1/2", 1), ("11ad 1/2", 3), ("5GNR_n600_k300", 1)])
def test_gpu_count_errors_matches_oracle(name, npar, cuda):
    import torch
    cs = _cs(name, npar)
    h = Henc.from_tables(cs.enc)
    ncw, seed = 301, 5
    bits = h.new_bits(ncw)
    h.gen_codewords(bits, seed=seed, cw0=0)
    words = _words(bits, cs.enc.mcol)
    dense = refsim.unpack_bits(words, cs.z, cs.enc.mcol)
    for snr in (-3.0, 1.0):
        llr = torch.zeros((ncw, cs.enc.mcol * cs.z), dtype=torch.int8, device="cuda")
        h.gen_llr(bits, llr, O.I8, snr, seed=seed)
        counts = torch.zeros(2, dtype=torch.int64, device="cuda")
        per = torch.zeros(ncw, dtype=torch.int32, device="cuda")
        h.count_errors(bits, llr, O.I8, counts, cw_bit_err=per)
        torch.cuda.synchronize()
        post = llr.cpu().numpy()
        ce, be = refsim.count_errors(cs, words, post)
        assert (int(counts[1].item()), int(counts[0].item())) == (ce, be)
        k = (cs.enc.mcol - cs.enc.mrow) * cs.z
        assert np.array_equal(per.cpu().numpy(), ((post[:, :k] < 0) != (dense[:, :k] == 1)).sum(1))
        _say("count", name, npar, snr, "bit errors", be, "cw errors", ce)
    # float posteriors in the reference layout; -0.0 is a 0 decision
    P = 16
    spz = (cs.z + P - 1) // P
    postf = np.where(dense == 1, np.float32(-1.5), np.float32(0.0)).astype(np.float32)
    postf[3, 5] = np.float32(2.0) if dense[3, 5] == 1 else np.float32(-2.0)   # one error
    if dense[4, 6] == 0:
        postf[4, 6] = np.float32(-0.0)                                       # no error
    refl = np.zeros((ncw, cs.enc.mcol, (spz + 2) * P), np.float32)
    refl[:, :, P:P + cs.z] = postf.reshape(ncw, cs.enc.mcol, cs.z)
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    h.count_errors(bits, torch.from_numpy(refl.reshape(ncw, -1)).cuda(), O.F32, counts, layout_p=P)
    torch.cuda.synchronize()
    ce, be = refsim.count_errors(cs, words, postf, elem=O.F32)
    assert (int(counts[1].item()), int(counts[0].item())) == (ce, be) == (1, 1)


def _oracle_sim(cs, kind, snrs, seed, num_cw, loops, max_iter=10, mrm=31.0, beta=1.0, scale=2.0, mx=24.0, hist=None):
    """the oracle chain's per-SNR counters; `hist` ([num_snr][max_iter + 1] uint64), when
    given, receives the iteration histograms"""
    elem = O.ELEM_OF_CHAR[kind[kind.index("Vec") + 3:].lstrip("0123456789")[0]]
    punct, fill = _dense_masks(cs)
    st = [dict(cw=0, uncoded_bit_err=0, iters=0, bit_err=0, cw_err=0) for _ in snrs]
    for l in range(loops):
        words = _oracle_codewords(cs, seed, l * num_cw, num_cw)
        dense = refsim.unpack_bits(words, cs.z, cs.enc.mcol)
        for i, snr in enumerate(snrs):
            sigma, f = chan_sigma(snr, scale)
            llr, err = O.noise_llr(seed, l * num_cw, dense, cs.z, cs.enc.mcol, sigma, f, mx, punct, fill, elem)
            post, _, its, tot = O.decode_ideal(cs.dec, llr, elem=elem, max_iter=max_iter, max_row_metric=mrm,
                                               beta=beta, want_c2v=False)
            if hist is not None:
                hist[i] += np.bincount(np.asarray(its, np.int64), minlength=max_iter + 1).astype(np.uint64)
            ce, be = refsim.count_errors(cs, words, post, elem=elem)
            s = st[i]
            s["cw"] += num_cw
            s["uncoded_bit_err"] += err
            s["iters"] += tot
            s["bit_err"] += be
            s["cw_err"] += ce
    return st


@pytest.mark.parametrize("name,kind,graph", [("11ad 1/2", "Vec64c_min", False), ("11ad 1/2", "Vec64c_min", True),
                                             ("11n 1/2 z27", "Vec64c_min", True),
                                             ("5GNR_n600_k300", "Vec32s_min", True),
                                             ("16e 1/2", "Vec16f_min", False)])
def test_gpu_sim_loop_matches_oracle_chain(name, kind, graph, cuda):
    import torch
    cs = _cs(name)
    henc = Henc.from_tables(cs.enc)
    dec = Decoder(Hdec.from_tables(cs.dec), kind, 10, 31.0, 1.0)
    snrs = [-1.0, 0.0, 1.5]
    num_cw, loops, seed = 400, 3, 31
    _say("sim", name, kind, "graph", graph)
    sim = Sim(henc, dec, num_cw, snrs, seed=seed, use_graph=graph)
    sim.run(1)
    torch.cuda.synchronize()
    torch.zeros(1000, device="cuda").sum()   # eager work between graph replays
    sim.run(loops - 1)
    got = sim.stats()
    hist = np.zeros((len(snrs), 11), np.uint64)
    want = _oracle_sim(cs, kind, snrs, seed, num_cw, loops, hist=hist)
    assert got == want, (got, want)
    gh = sim.iter_hist()
    assert np.array_equal(gh, hist), (gh, hist)
    assert all(int(gh[i].sum()) == w["cw"] and int((gh[i] * np.arange(11)).sum()) == w["iters"]
               for i, w in enumerate(want))
    # reset clears the counters; the codeword index keeps advancing (loop 3 next)
    sim.reset()
    sim.run(1)
    got = sim.stats()
    full = _oracle_sim(cs, kind, snrs, seed, num_cw, loops + 1)
    for g, w4, w3 in zip(got, full, want):
        assert g == {k: w4[k] - w3[k] for k in g}


@pytest.mark.parametrize("kind,punct", [("Vec64c_min", False), ("Vec64c_min", True), ("Vec16f_min", True)])
def test_gpu_8023ca_sim_loop_matches_oracle_chain(kind, punct, cuda):
    """802.3ca (Z = 256, 12 x 69 blocks, rows of weight 22-23), with and without the
    puncture of its last two parity block columns: encode, LLR, decode and count on
    the device equal the oracle chain's counters and iteration histograms."""
    import torch
    cs = codes.gen_structs("802.3ca", simd_size=512, elem_size=8, puncture_8023ca=punct)
    henc = Henc.from_tables(cs.enc)
    dec = Decoder(Hdec.from_tables(cs.dec), kind, 10, 31.0, 1.0)
    snrs = [2.0, 2.25, 3.0, 5.0]   # Es/N0 dB: the waterfall lies near 2.2 dB, 2.8 dB when punctured
    num_cw, loops, seed = 48, 2, 77
    _say("sim 802.3ca", kind, "punctured", punct)
    sim = Sim(henc, dec, num_cw, snrs, seed=seed, use_graph=True)
    sim.run(loops)
    torch.cuda.synchronize()
    got = sim.stats()
    hist = np.zeros((len(snrs), 11), np.uint64)
    want = _oracle_sim(cs, kind, snrs, seed, num_cw, loops, hist=hist)
    assert got == want, (got, want)
    assert np.array_equal(sim.iter_hist(), hist)
    # the points span the waterfall: errors at the lowest SNR, none at the highest
    assert want[0]["cw_err"] > 0 and want[-1]["cw_err"] == 0, want


def test_gpu_llr_per_codeword_channel_matches_oracle(cuda):
    """fecsim_gen_llr_var: a (sigma, factor) pair per codeword."""
    import torch
    from fecsim import lib as L
    cs = _cs("11n 1/2 z27")
    h = Henc.from_tables(cs.enc)
    ncw, seed, cw0 = 40, 21, 300
    bits = h.new_bits(ncw)
    h.gen_codewords(bits, seed=seed, cw0=cw0)
    dense = refsim.unpack_bits(_words(bits, cs.enc.mcol), cs.z, cs.enc.mcol)
    rng = np.random.default_rng(3)
    sf = np.array([chan_sigma(float(s), 2.0) for s in rng.uniform(-2.0, 2.0, ncw)], np.float32)
    llr = torch.zeros((ncw, cs.enc.mcol * cs.z), dtype=torch.int8, device="cuda")
    unc = torch.zeros(1, dtype=torch.int64, device="cuda")
    import ctypes as C
    from fecsim.decoder import _ptr
    sft = torch.from_numpy(sf).cuda()
    L.check(L.lib().fecsim_gen_llr_var(h._h, _ptr(sft), C.c_float(24.0), seed, cw0, ncw, h.col_words, _ptr(bits),
                                       L.I8, 0, _ptr(llr), _ptr(unc),
                                       C.c_void_p(torch.cuda.current_stream().cuda_stream)), "gen_llr_var")
    torch.cuda.synchronize()
    got = llr.cpu().numpy()
    tot = 0
    for i in range(ncw):
        want, err = O.noise_llr(seed, cw0 + i, dense[i:i + 1], cs.z, cs.enc.mcol, float(sf[i, 0]), float(sf[i, 1]),
                                24.0, None, None, O.I8)
        assert np.array_equal(got[i], want[0]), i
        tot += err
    assert int(unc.item()) == tot


@pytest.mark.parametrize("name,kind,layout_p", [("11ad 1/2", "Vec64c_min", 0), ("11ad 1/2", "Vec64c_min", 64),
                                                ("11n 1/2 z27", "Vec16f_min", 0), ("5GNR_n600_k300", "Vec32s_min", 0)])
def test_gpu_hard_decisions_and_iteration_histogram(name, kind, layout_p, cuda):
    """fecsim_hard_decisions packs value < 0 of every column like the codeword bits
    (float -0.0 -> 0); fecsim_iter_histogram equals a bincount of the iterations."""
    import torch
    from fecsim import lib as L
    from fecsim.sim import iter_histogram
    cs = _cs(name)
    henc = Henc.from_tables(cs.enc)
    p, ch, mode, ss, es = codes.parse_exec_name(kind)
    elem = O.ELEM_OF_CHAR[ch]
    ncw = 300
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(ncw)
    scale, mx = (1.0, 1e9) if elem == O.F32 else (2.0, 24.0)
    llr, _ = rs.metrics(wgn, bits, 0.0, elem=elem, llr_scale=scale, max_abs=mx, P=layout_p)
    if elem == O.F32:
        llr.reshape(ncw, -1)[0, :8] = np.float32(-0.0)    # -0.0 is no error (LDPC_count_errors.cpp:84)
    dec = Decoder(Hdec.from_tables(cs.dec), kind, 10, 31.0, 0.5 if elem == O.F32 else 1.0)
    x = torch.from_numpy(np.ascontiguousarray(llr)).cuda()
    it = torch.zeros(ncw, dtype=torch.int32, device="cuda")
    dec.decode_set(x, iters=it, layout_p=layout_p)
    hb = henc.hard_decisions(x, elem, ncw, layout_p=layout_p)
    h = iter_histogram(it, 11)
    torch.cuda.synchronize()
    post = x.cpu().numpy()
    dense = O.ref_to_dense(post, cs.z, cs.dec.mcol, layout_p) if layout_p else post.reshape(ncw, -1)
    hard = (dense.reshape(ncw, cs.enc.mcol, cs.z) < 0).astype(np.uint64)
    cw = hb.shape[2]
    want = np.zeros((ncw, cs.enc.mcol, cw), np.uint64)
    for r in range(cs.z):
        want[:, :, r >> 5] |= hard[:, :, r] << np.uint64(r & 31)
    assert np.array_equal(hb.cpu().numpy().view(np.uint32).astype(np.uint64), want)
    assert np.array_equal(h.cpu().numpy(), np.bincount(it.cpu().numpy(), minlength=11))
    with pytest.raises(Exception):
        iter_histogram(it, 2000)


@pytest.mark.parametrize("name,elem,P", [("11ad 1/2", O.I8, 0), ("11n 1/2 z27", O.I8, 64), ("16e 1/2", O.F32, 0)])
def test_gpu_reference_stream_replay(name, elem, P, cuda):
    """The reference's sequential streams replayed on the device equal the oracle's
    restatement of prng_functions.cpp bit for bit: normdist2 noise, gen_rand_databits +
    encoder, gen_metrics on that noise, over two inner loops (the state carries on);
    and a state whose first draw is tiny takes normdist2's extra draw."""
    import torch
    from fecsim.sim import RefRng
    cs = _cs(name, zo=48 if name.startswith("16e") else None)
    henc = Henc.from_tables(cs.enc)
    seeds = refsim.perl_seeds(77)
    rs = refsim.RefStream(cs, seeds=seeds)
    g = RefRng(seeds[:32])
    ncw, mcol, z = 50, cs.enc.mcol, cs.z
    spf = (z + 7) // 8
    scale, mx = (1.0, 1e9) if elem == O.F32 else (2.0, 24.0)
    for loop in range(2):
        wgn, words = rs.inner_loop(ncw)
        count = ncw * mcol * spf
        count += count & 1
        gw = torch.empty(count * 8, dtype=torch.float32, device="cuda")
        g.normdist2(gw, count)
        bits = henc.new_bits(ncw)
        g.databits(henc, bits, ncw)
        henc.encode(bits)
        torch.cuda.synchronize()
        assert np.array_equal(gw.cpu().numpy().view(np.uint32), wgn.view(np.uint32)), loop
        assert np.array_equal(_words(bits, mcol), words), loop
        want, err = rs.metrics(wgn, words, -0.5, elem=elem, P=P, llr_scale=scale, max_abs=mx)
        llr = torch.zeros(want.shape, dtype=torch.from_numpy(want).dtype, device="cuda")
        unc = torch.zeros(1, dtype=torch.int64, device="cuda")
        henc.gen_llr_from_noise(bits, gw, llr, elem, -0.5, llr_scale=scale, max_abs_llr=mx, layout_p=P, uncoded=unc)
        torch.cuda.synchronize()
        got = llr.cpu().numpy()
        if P:
            got, want = O.ref_to_dense(got, z, mcol, P), O.ref_to_dense(want, z, mcol, P)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), loop
        assert int(unc.item()) == err
    # lane 0's first sfc64 output 0x7fffffff: its 31-bit value is 0, so the extra draw
    tiny = seeds.copy()
    tiny[0], tiny[4], tiny[12] = 0x7FFFFFFF, 0, 0
    rs2 = refsim.RefStream(cs, seeds=tiny)
    w2, _ = rs2.inner_loop(4)
    g2 = RefRng(tiny[:32])
    gw2 = torch.empty(w2.size, dtype=torch.float32, device="cuda")
    g2.normdist2(gw2, w2.size // 8)
    torch.cuda.synchronize()
    assert np.array_equal(gw2.cpu().numpy().view(np.uint32), w2.view(np.uint32))
