"""GPU decoder parity: the HIP layered decoders (through the C ABI) against the CPU
oracle -- bit-exact posteriors, C2V messages and iteration counts."""
import glob
import os
import sys

import numpy as np
import pytest

from fecsim import codes
from fecsim.decoder import Decoder, Hdec
import refsim
import sp_ref
from refsim import O

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _say(*a):
    print(*a, file=sys.stderr, flush=True)


def _bytes_eq(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint8), np.ascontiguousarray(b).view(np.uint8))


def _f32_prod(kind):
    """float32 sum-product kinds: their tanh/atanh are the hardware-based device functions,
    so they are held to the sum-product tolerance against the libm restatement rather
    than to the oracle's bits"""
    _, ch, mode, _, _ = codes.parse_exec_name(kind)
    return ch == "f" and mode == "prod"


def _check_f32_prod(dec_tables, llr, got, max_iter, mrm, beta):
    want, want_it, peak = sp_ref.decode_sp(dec_tables, llr, max_iter=max_iter, max_row_metric=mrm, beta=beta)
    rep = sp_ref.sp_parity(got[0], got[2], want, want_it, peak)
    _say("  sum-product parity", rep)
    assert rep["ok"], rep
    assert got[3] == int(want_it.sum())


def _gpu_decode(dec_tables, kind, llr, max_iter, mrm, beta, layout_p=0, want_c2v=True, inplace=False):
    import torch
    h = Hdec.from_tables(dec_tables)
    d = Decoder(h, kind, max_iter, mrm, beta)
    x = torch.from_numpy(np.ascontiguousarray(llr)).cuda()
    ncw = llr.shape[0]
    iters = torch.zeros(ncw, dtype=torch.int32, device="cuda")
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    c2v = torch.zeros(ncw * h.e, dtype=d.torch_dtype, device="cuda") if want_c2v else None
    if inplace:
        d.decode_set(x, iters=iters, c2v=c2v, tot_iter=tot, layout_p=layout_p)
        out = x
    else:
        out = torch.empty_like(x)
        d.decode(x, out, iters=iters, c2v=c2v, tot_iter=tot, layout_p=layout_p)
        assert torch.equal(x.cpu(), torch.from_numpy(np.ascontiguousarray(llr)))  # input untouched
    torch.cuda.synchronize()
    c = None if c2v is None else c2v.cpu().numpy().reshape(ncw, -1)
    return out.cpu().numpy(), c, iters.cpu().numpy().astype(np.uint32), int(tot.item())


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "dec_*.npz"))), ids=os.path.basename)
def test_gpu_matches_golden(path, cuda):
    f = np.load(path)
    _say("golden", os.path.basename(path), str(f["kind"]), f["llr"].shape)
    dec = codes.DecTables.from_bytes(f["dat"].tobytes())
    mi, mrm, beta = f["params"]
    post, c2v, iters, tot = _gpu_decode(dec, str(f["kind"]), f["llr"], int(mi), float(mrm), float(beta))
    if _f32_prod(str(f["kind"])):
        _check_f32_prod(dec, f["llr"], (post, c2v, iters, tot), int(mi), float(mrm), float(beta))
        return
    assert np.array_equal(iters, f["iters"])
    assert _bytes_eq(post, f["post"])
    assert _bytes_eq(c2v, f["c2v"])
    assert tot == int(f["iters"].sum())


LIVE = [
    # MatrixSet, num_par, kind, snr, ncw, z_over
    ("11ad 1/2", 1, "Vec64c_min", -1.0, 600, None),
    ("11ad 1/2", 1, "Vec64c_min", 0.0, 600, None),
    ("11ad 1/2", 6, "Vec64c_min", -0.5, 100, None),
    ("11n 1/2 z27", 1, "Vec64c_min", -0.5, 300, None),
    ("11n 3/4 z54", 1, "Vec64c_min", 2.0, 200, None),
    ("11n 5/6 z81", 1, "Vec64c_min", 3.5, 200, None),
    ("11ad 13/16", 1, "Vec64c_min", 3.0, 300, None),
    ("16e 2/3A", 1, "Vec64c_min", 1.0, 100, 96),
    ("11n 1/2 z81", 1, "Vec64c_min", -0.5, 200, None),
    ("16e 1/2", 1, "Vec16f_prod", -1.0, 60, 48),
    ("16e 1/2", 1, "Vec16f_prod", -0.5, 60, 96),   # float codewords this large: workgroup per codeword
    ("16e 1/2", 1, "Vec16f_min", -0.5, 60, 68),
    ("11n 1/2 z81", 1, "Vec16f_min", -1.0, 60, None),
    ("11ad 1/2", 1, "Vec32s_min", -0.5, 100, None),
]


@pytest.mark.parametrize("case", LIVE, ids=lambda c: "%s-np%d-%s-%s" % c[:4])
def test_gpu_matches_oracle_reference_chain(case, cuda):
    """LLRs from the reference chain (sfc64 WGN, xoshiro bits, encoder, gen_metrics)."""
    ms, npar, kind, snr, ncw, zo = case
    _say("live", case)
    p, ch, mode, ss, es = codes.parse_exec_name(kind)
    cs = codes.gen_structs(ms, num_par=npar, simd_size=ss, elem_size=es, z_over=zo)
    elem = O.ELEM_OF_CHAR[ch]
    fl = elem >= O.F32
    mrm, beta, scale, mx = (6e-8, 0.0, 1.0, 1e9) if mode == "prod" else ((31.0, 0.5, 1.0, 1e9) if fl else (31.0, 1.0, 2.0, 24.0))
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(ncw)
    llr, _ = rs.metrics(wgn, bits, snr, elem=elem, llr_scale=scale, max_abs=mx)
    ref = O.decode_ideal(cs.dec, llr, elem=elem, mode=O.PROD if mode == "prod" else O.MIN, max_row_metric=mrm,
                         beta=beta)
    got = _gpu_decode(cs.dec, kind, llr, 10, mrm, beta)
    if _f32_prod(kind):
        _check_f32_prod(cs.dec, llr, got, 10, mrm, beta)
        return
    assert np.array_equal(got[2], ref[2])
    assert _bytes_eq(got[0], ref[0])
    assert _bytes_eq(got[1], ref[1])
    assert got[3] == ref[3]


def test_gpu_random_int8_saturation_corners(cuda):
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    rng = np.random.default_rng(2)
    llr = rng.integers(-128, 128, size=(257, cs.n)).astype(np.int8)
    llr[::3, ::5] = -128
    for beta, cap in ((1.0, 31.0), (0.0, 127.0), (3.0, 9.0)):
        ref = O.decode_ideal(cs.dec, llr, max_iter=7, max_row_metric=cap, beta=beta)
        got = _gpu_decode(cs.dec, "Vec64c_min", llr, 7, cap, beta)
        assert np.array_equal(got[2], ref[2]) and np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1])


def test_gpu_reference_layout_and_inplace(cuda):
    """layout_p = 64 decodes the reference's padded buffer in place; guard words are
    not touched and the data equals the dense decode."""
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(150)
    dense, _ = rs.metrics(wgn, bits, -0.75, P=0)
    padded, _ = rs.metrics(wgn, bits, -0.75, P=64)
    garbage = np.random.default_rng(9).integers(-128, 128, size=padded.shape).astype(np.int8)
    spz = 1
    mask = np.zeros((150, cs.dec.mcol, (spz + 2) * 64), bool)
    mask[:, :, 64:64 + cs.z] = True
    padded = np.where(mask.reshape(150, -1), padded, garbage)
    ref = O.decode_ideal(cs.dec, dense)
    got_p = _gpu_decode(cs.dec, "Vec64c_min", padded, 10, 31.0, 1.0, layout_p=64, inplace=True, want_c2v=False)
    got_d = _gpu_decode(cs.dec, "Vec64c_min", dense, 10, 31.0, 1.0, inplace=True)
    assert np.array_equal(O.ref_to_dense(got_p[0], cs.z, cs.dec.mcol, 64), ref[0])
    assert np.array_equal(got_p[0].reshape(150, -1)[~mask.reshape(150, -1)], garbage[~mask.reshape(150, -1)])
    assert np.array_equal(got_d[0], ref[0]) and np.array_equal(got_d[2], ref[2])


def test_gpu_decode_set_host(cuda):
    cs = codes.gen_structs("11n 1/2 z27", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(64)
    llr, _ = rs.metrics(wgn, bits, -0.5, P=64)
    want = O.decode_ideal(cs.dec, O.ref_to_dense(llr, cs.z, cs.dec.mcol, 64))
    d = Decoder(Hdec.from_tables(cs.dec), "LDPCtop_Vec64c_min", 10, 31.0, 1.0)
    m = llr.copy()
    tot, iters = d.decode_set_host(m, layout_p=64)
    assert tot == want[3] and np.array_equal(iters, want[2])
    assert np.array_equal(O.ref_to_dense(m, cs.z, cs.dec.mcol, 64), want[0])
    assert d.decode_set_host(m[:0], layout_p=64)[0] == 0


def test_gpu_large_batch_properties(cuda):
    """Full-size batch (11ad 1/2, 200k codewords at -1 dB): every codeword that
    stopped early has a zero syndrome; a sample equals the oracle."""
    import torch
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    ncw = 200_000
    s, f = O.snr_params(-1.0, 2.0)
    g = torch.Generator(device="cuda").manual_seed(5)
    y = 1.0 + s * torch.randn(ncw, cs.n, device="cuda", generator=g)
    llr = torch.clamp(y * f, -24, 24).round().to(torch.int8)
    h = Hdec.from_tables(cs.dec)
    d = Decoder(h, "Vec64c_min", 10, 31.0, 1.0)
    out = torch.empty_like(llr)
    iters = torch.zeros(ncw, dtype=torch.int32, device="cuda")
    d.decode(llr, out, iters=iters)
    torch.cuda.synchronize()
    hd = (out < 0).to(torch.int32).cpu().numpy()
    H = codes.expand_h(cs.hc, cs.z).astype(np.int32)
    it = iters.cpu().numpy()
    early = it < 10
    syn = (hd[early][:5000] @ H.T) & 1
    assert not syn.any()
    idx = np.random.default_rng(0).choice(ncw, 3000, replace=False)
    ref = O.decode_ideal(cs.dec, llr.cpu().numpy()[idx])
    assert np.array_equal(out.cpu().numpy()[idx], ref[0]) and np.array_equal(it[idx].astype(np.uint32), ref[2])
    _say("mean iterations at -1 dB:", it.mean())


PAIR_CASES = [
    # MatrixSet, num_par, z_over, snr, ncw, beta, cap
    ("11ad 1/2", 1, None, -1.0, 1000, 1.0, 31.0),
    ("11ad 1/2", 1, None, 0.0, 7, 1.0, 31.0),
    ("11ad 1/2", 1, None, -0.5, 1, 2.0, 20.0),
    ("11ad 13/16", 1, None, 3.0, 300, 1.0, 31.0),
    ("11ad 3/4", 1, None, 2.0, 300, 1.0, 31.0),
    ("11n 3/4 z54", 1, None, 2.0, 200, 1.0, 31.0),
    ("11n 1/2 z27", 2, None, -0.5, 200, 1.0, 31.0),   # Z = 54
    ("16e 1/2", 1, 48, -0.5, 200, 1.0, 31.0),
    ("ITU G.hn 1/2 z14", 1, None, 0.0, 500, 1.0, 31.0),
]


@pytest.mark.parametrize("case", PAIR_CASES, ids=lambda c: "%s-np%d-z%s-%g-n%d" % c[:5])
def test_gpu_pair_kernel_matches_oracle(case, cuda):
    """The row-pair int8 kernel (even Z) against the oracle and the generic kernel:
    posteriors, C2V, iteration counts and the summed iterations, bit for bit."""
    import torch
    from fecsim import lib as L
    ms, npar, zo, snr, ncw, beta, cap = case
    _say("pair", case)
    cs = codes.gen_structs(ms, num_par=npar, simd_size=512, elem_size=8, z_over=zo)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(ncw)
    llr, _ = rs.metrics(wgn, bits, snr)
    rng = np.random.default_rng(4)
    llr[: max(1, ncw // 10)] = rng.integers(-128, 128, size=(max(1, ncw // 10), cs.n))   # saturation corners
    ref = O.decode_ideal(cs.dec, llr, max_row_metric=cap, beta=beta)
    h = Hdec.from_tables(cs.dec)
    for kern in (L.KERNEL_PAIR, L.KERNEL_PAIR_RT, L.KERNEL_GENERIC):
        d = Decoder(h, "Vec64c_min", 10, cap, beta)
        d.select_kernel(kern)
        assert d.kernel == kern
        x = torch.from_numpy(llr).cuda()
        out = torch.empty_like(x)
        iters = torch.zeros(ncw, dtype=torch.int32, device="cuda")
        c2v = torch.zeros(ncw * h.e, dtype=torch.int8, device="cuda")
        tot = torch.zeros(1, dtype=torch.int64, device="cuda")
        d.decode(x, out, iters=iters, c2v=c2v, tot_iter=tot)
        torch.cuda.synchronize()
        assert np.array_equal(iters.cpu().numpy().astype(np.uint32), ref[2]), kern
        assert np.array_equal(out.cpu().numpy(), ref[0]), kern
        assert np.array_equal(c2v.cpu().numpy().reshape(ncw, -1), ref[1]), kern
        assert int(tot.item()) == ref[3]


def test_gpu_pair_kernel_concurrent_streams(cuda):
    """Twelve decodes of one 802.11ad 1/2 decoder issued on two streams without waiting
    in between (the work-queue kernel takes one of the decoder's 8 counter sets per call,
    in turn): every batch equals the oracle."""
    import torch
    from fecsim import lib as L
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    batches = []
    for snr, ncw in ((-1.0, 3000), (-0.5, 2000), (0.0, 1000)):
        wgn, bits = rs.inner_loop(ncw)
        llr, _ = rs.metrics(wgn, bits, snr)
        batches.append((llr, O.decode_ideal(cs.dec, llr)))
    d = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 10, 31.0, 1.0)
    assert d.kernel == L.KERNEL_PAIR
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    ins = [torch.from_numpy(b[0]).cuda() for b in batches]
    torch.cuda.synchronize()
    outs, its = [], []
    for k in range(12):
        b = k % len(batches)
        st = streams[k % 2]
        with torch.cuda.stream(st):   # outputs made (and zeroed) on the stream that decodes
            out = torch.empty_like(ins[b])
            it = torch.zeros(ins[b].shape[0], dtype=torch.int32, device="cuda")
            d.decode(ins[b], out, iters=it, stream=st)
        outs.append((b, out, it))
    torch.cuda.synchronize()
    for b, out, it in outs:
        ref = batches[b][1]
        assert np.array_equal(out.cpu().numpy(), ref[0]), b
        assert np.array_equal(it.cpu().numpy().astype(np.uint32), ref[2]), b


@pytest.mark.parametrize("ms", ["11ad 1/2", "11n 1/2 z54", "11ad 13/16"])
def test_gpu_pair_kernel_extreme_rows(ms, cuda):
    """Rows whose Q values are mostly -128 with one 0 (|Q| = 128 beside |Q| = 0, the
    widest spread the exclusive-min select sees) and cap + beta both below and above 128:
    the row-pair kernel against the oracle, bit for bit."""
    import torch
    from fecsim import lib as L
    cs = codes.gen_structs(ms, simd_size=512, elem_size=8)
    rng = np.random.default_rng(11)
    ncw = 400
    llr = rng.choice(np.array([-128, 0, 127, -1, 1], dtype=np.int8), p=[0.8, 0.1, 0.04, 0.03, 0.03],
                     size=(ncw, cs.n))
    h = Hdec.from_tables(cs.dec)
    for beta, cap in [(1.0, 31.0), (1.0, 127.0), (2.0, 127.0), (3.0, 5.0), (127.0, 0.5), (64.0, 100.0)]:
        _say("extreme rows", ms, beta, cap)
        ref = O.decode_ideal(cs.dec, llr, max_iter=6, max_row_metric=cap, beta=beta)
        for kern in (L.KERNEL_PAIR, L.KERNEL_PAIR_RT):
            d = Decoder(h, "Vec64c_min", 6, cap, beta)
            assert d.kernel == L.KERNEL_PAIR
            d.select_kernel(kern)
            x = torch.from_numpy(llr).cuda()
            out = torch.empty_like(x)
            iters = torch.zeros(ncw, dtype=torch.int32, device="cuda")
            c2v = torch.zeros(ncw * h.e, dtype=torch.int8, device="cuda")
            d.decode(x, out, iters=iters, c2v=c2v)
            torch.cuda.synchronize()
            assert np.array_equal(c2v.cpu().numpy().reshape(ncw, -1), ref[1]), (kern, beta, cap)
            assert np.array_equal(out.cpu().numpy(), ref[0]), (kern, beta, cap)
            assert np.array_equal(iters.cpu().numpy().astype(np.uint32), ref[2]), (kern, beta, cap)


def test_gpu_pair_kernel_selected_for_11ad(cuda):
    from fecsim import lib as L
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    h = Hdec.from_tables(cs.dec)
    assert Decoder(h, "Vec64c_min", 10, 31.0, 1.0).kernel == L.KERNEL_PAIR
    assert Decoder(h, "Vec64c_min", 10, 31.0, 0.0).kernel == L.KERNEL_GENERIC   # beta 0: abs(-128) wrap path
    assert Decoder(h, "Vec32s_min", 10, 31.0, 1.0).kernel == L.KERNEL_GENERIC
    odd = codes.gen_structs("11n 1/2 z27", simd_size=512, elem_size=8)
    assert Decoder(Hdec.from_tables(odd.dec), "Vec64c_min", 10, 31.0, 1.0).kernel == L.KERNEL_GENERIC


CODE_CASES = [
    # (Eb/N0 dB, codewords, max_iter, beta, cap) on 5G NR base graph 1, Z = 384, rate 1/3;
    # codewords 0 and 1 are replaced by full-range and +-saturated bytes
    (0.2, 6, 20, 1.0, 31.0),    # most codewords need all 20 iterations
    (0.6, 8, 20, 1.0, 31.0),    # most converge
    (1.0, 8, 20, 2.0, 127.0),
    (0.6, 6, 4, 3.0, 5.0),
    (0.8, 4, 20, 127.0, 0.5),   # beta 127, cap 0: every C2V message is 0
]


def _bg1_llrs(ebn0, ncw, seed=7):
    cs = codes.gen_structs("5GNR_n25344_k8448", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(ncw)
    llr, _ = rs.metrics(wgn, bits, ebn0 - 10 * np.log10(3.0), elem=O.I8, llr_scale=2.0, max_abs=24.0)
    rng = np.random.default_rng(seed)
    llr[0] = rng.integers(-128, 128, cs.n)
    llr[1] = np.where(rng.random(cs.n) < 0.5, -128, 127)
    return cs, llr


@pytest.mark.parametrize("case", CODE_CASES, ids=lambda c: "ebn0%g-n%d-it%d-b%g-c%g" % c)
def test_gpu_code_kernel_matches_oracle(case, cuda):
    """The code-specialized int8 kernel (BG1 Z=384 compiled in, one codeword per
    workgroup) against the oracle and the generic kernel: posteriors, C2V, iteration
    counts and the summed iterations, bit for bit; out of place and in place."""
    import torch
    from fecsim import lib as L
    ebn0, ncw, mi, beta, cap = case
    _say("code kernel", case)
    cs, llr = _bg1_llrs(ebn0, ncw)
    ref = O.decode_ideal(cs.dec, llr, max_iter=mi, max_row_metric=cap, beta=beta)
    h = Hdec.from_tables(cs.dec)
    for kern in (L.KERNEL_CODE, L.KERNEL_GENERIC):
        d = Decoder(h, "Vec64c_min", mi, cap, beta)
        d.select_kernel(kern)
        assert d.kernel == kern
        for inplace in (False, True):
            x = torch.from_numpy(llr).cuda()
            out = x if inplace else torch.empty_like(x)
            iters = torch.zeros(ncw, dtype=torch.int32, device="cuda")
            c2v = torch.zeros(ncw * h.e, dtype=torch.int8, device="cuda")
            tot = torch.zeros(1, dtype=torch.int64, device="cuda")
            if inplace:
                d.decode_set(x, iters=iters, c2v=c2v, tot_iter=tot)
            else:
                d.decode(x, out, iters=iters, c2v=c2v, tot_iter=tot)
            torch.cuda.synchronize()
            assert np.array_equal(iters.cpu().numpy().astype(np.uint32), ref[2]), (kern, inplace)
            assert np.array_equal(out.cpu().numpy(), ref[0]), (kern, inplace)
            assert np.array_equal(c2v.cpu().numpy().reshape(ncw, -1), ref[1]), (kern, inplace)
            assert int(tot.item()) == ref[3]
    _say("  iterations", ref[2])


@pytest.mark.parametrize("case", [(2.3, 6, 10, 1.0, 31.0, False), (2.1, 4, 10, 2.0, 127.0, False),
                                  (2.4, 4, 10, 127.0, 0.5, False), (2.3, 6, 6, 1.0, 31.0, True)],
                         ids=lambda c: "esn0%g-n%d-it%d-b%g-c%g-tab%d" % c)
def test_gpu_code_kernel_8023ca_matches_oracle(case, cuda):
    """802.3ca (Z = 256, rows of weight 22-23) on the code-specialized int8 kernel against
    the oracle: posteriors, C2V and iteration counts bit for bit, with plain offsets, wide
    and zero caps, a per-(iteration, layer) offset table, and full-range / saturated
    codewords."""
    import torch
    from fecsim import lib as L
    esn0, ncw, mi, beta, cap, with_tab = case
    _say("code kernel 802.3ca", case)
    cs = codes.gen_structs("802.3ca", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(ncw)
    llr, _ = rs.metrics(wgn, bits, esn0, elem=O.I8, llr_scale=2.0, max_abs=24.0)
    rng = np.random.default_rng(13)
    llr[0] = rng.integers(-128, 128, cs.n)
    llr[1] = np.where(rng.random(cs.n) < 0.5, -128, 127)
    tab = (rng.integers(1, 4, (mi, cs.dec.mrow)).astype(np.float32) if with_tab else None)
    ref = O.decode_ideal(cs.dec, llr, max_iter=mi, max_row_metric=cap, beta=beta, beta_table=tab)
    d = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", mi, cap, beta)
    if with_tab:
        d.set_offsets(tab)
    assert d.kernel == L.KERNEL_CODE
    x = torch.from_numpy(llr).cuda()
    out = torch.empty_like(x)
    iters = torch.zeros(ncw, dtype=torch.int32, device="cuda")
    c2v = torch.zeros(ncw * cs.dec.total_hc_ones * cs.z, dtype=torch.int8, device="cuda")
    d.decode(x, out, iters=iters, c2v=c2v)
    torch.cuda.synchronize()
    assert np.array_equal(iters.cpu().numpy().astype(np.uint32), ref[2])
    assert np.array_equal(out.cpu().numpy(), ref[0])
    assert np.array_equal(c2v.cpu().numpy().reshape(ncw, -1), ref[1])
    _say("  iterations", ref[2])


def test_gpu_code_kernel_offsets_and_fallbacks(cuda):
    """Per-(iteration, layer) offset tables run on the code kernel when every entry is
    1..127 and match the oracle; a table with a 0 entry, beta 0, 16-bit elements and
    buffers that are not 16-byte aligned fall back to the generic kernel, same results."""
    import torch
    from fecsim import lib as L
    mi, ncw = 8, 6
    cs, llr = _bg1_llrs(0.5, ncw, seed=9)
    h = Hdec.from_tables(cs.dec)
    assert Decoder(h, "Vec64c_min", mi, 31.0, 1.0).kernel == L.KERNEL_CODE
    assert Decoder(h, "Vec64c_min", mi, 31.0, 0.0).kernel == L.KERNEL_GENERIC   # beta 0: abs(-128) wrap path
    assert Decoder(h, "Vec32s_min", mi, 31.0, 1.0).kernel == L.KERNEL_GENERIC
    rng = np.random.default_rng(5)
    tab = rng.integers(1, 4, (mi, cs.dec.mrow)).astype(np.float32) + 0.6      # truncates to 1..3
    want = O.decode_ideal(cs.dec, llr, max_iter=mi, beta=1.0, beta_table=tab)
    d = Decoder(h, "Vec64c_min", mi, 31.0, 1.0)
    d.set_offsets(tab)
    assert d.kernel == L.KERNEL_CODE
    x = torch.from_numpy(llr).cuda()
    out = torch.empty_like(x)
    it = torch.zeros(ncw, dtype=torch.int32, device="cuda")
    c2v = torch.zeros(ncw * h.e, dtype=torch.int8, device="cuda")
    d.decode(x, out, iters=it, c2v=c2v)
    torch.cuda.synchronize()
    assert np.array_equal(it.cpu().numpy().astype(np.uint32), want[2])
    assert np.array_equal(out.cpu().numpy(), want[0])
    assert np.array_equal(c2v.cpu().numpy().reshape(ncw, -1), want[1])
    tab0 = tab.copy()
    tab0[2, 5] = 0.0
    want0 = O.decode_ideal(cs.dec, llr, max_iter=mi, beta=1.0, beta_table=tab0)
    d.set_offsets(tab0)
    assert d.kernel == L.KERNEL_GENERIC
    out0 = torch.empty_like(x)
    d.decode(x, out0)
    torch.cuda.synchronize()
    assert np.array_equal(out0.cpu().numpy(), want0[0])
    # 1-byte offset buffers: generic kernel, plain offsets
    d.set_offsets(None)
    ref = O.decode_ideal(cs.dec, llr, max_iter=mi, beta=1.0)
    n = llr.size
    buf_in = torch.zeros(n + 1, dtype=torch.int8, device="cuda")
    buf_out = torch.zeros(n + 1, dtype=torch.int8, device="cuda")
    xin = buf_in[1:].view(ncw, cs.n)
    xin.copy_(x)
    xout = buf_out[1:].view(ncw, cs.n)
    it2 = torch.zeros(ncw, dtype=torch.int32, device="cuda")
    d.decode(xin, xout, iters=it2)
    torch.cuda.synchronize()
    assert np.array_equal(xout.cpu().numpy(), ref[0])
    assert np.array_equal(it2.cpu().numpy().astype(np.uint32), ref[2])


@pytest.mark.parametrize("case", [
    # (matrix set, kind, max_iter, SNR dB, codewords): codewords whose C2V does not fit the
    # wave's LDS slice keep it in a global scratch; 3000 x 11n z81 double exceeds the
    # resident grid, so the kernel's stride loop runs several rounds
    ("5GNR_n25344_k8448", "Vec64c_min", 20, 0.2 - 10 * np.log10(3.0), 6),   # Eb/N0 0.2 dB: 20 iterations
    ("5GNR_n25344_k8448", "Vec64c_min", 20, 0.6 - 10 * np.log10(3.0), 6),   # Eb/N0 0.6 dB: converges
    ("5GNR_n25344_k8448", "Vec32s_min", 20, 0.6 - 10 * np.log10(3.0), 3),
    ("11n 1/2 z81", "Vec8d_min", 10, -0.5, 3000),
], ids=lambda c: "%s-%s-%.2f-%d" % (c[0], c[1], c[3], c[4]))
def test_gpu_global_c2v_codes_match_oracle(case, cuda):
    """Large codes (5G NR base graph 1, Z=384, rate 1/3 with 2Z punctured columns):
    posteriors, C2V and iterations against the oracle."""
    ms, kind, mi, snr, ncw = case
    _say("global c2v", case)
    cs = codes.gen_structs(ms, simd_size=512, elem_size=8)
    elem = O.ELEM_OF_CHAR[kind[kind.index("Vec") + 3:].lstrip("0123456789")[0]]
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(ncw)
    scale, mx = (2.0, 24.0) if elem in (O.I8, O.I16) else (1.0, 1e9)
    llr, _ = rs.metrics(wgn, bits, snr, elem=elem, llr_scale=scale, max_abs=mx)
    ref = O.decode_ideal(cs.dec, llr, elem=elem, max_iter=mi)
    post, c2v, iters, tot = _gpu_decode(cs.dec, kind, llr, mi, 31.0, 1.0)
    assert np.array_equal(iters, ref[2])
    if elem in (O.F32, O.F64):
        np.testing.assert_allclose(post, ref[0], rtol=1e-6, atol=1e-6)   # same op order: equal in practice
        assert _bytes_eq(post, ref[0])
    else:
        assert _bytes_eq(post, ref[0])
    assert _bytes_eq(c2v, ref[1])
    assert tot == ref[3]
    ce, _ = refsim.count_errors(cs, bits, post, elem=elem)
    _say("  iterations", iters.mean(), "cw errors", ce, "of", ncw)


@pytest.mark.parametrize("case", [(-0.5, 10, 31.0, 600), (0.0, 6, 9.0, 600), (-1.0, 10, 127.0, 300),
                                  (-0.5, 20, 0.5, 300)],
                         ids=lambda c: "snr%g-it%d-cap%g-n%d" % c)
def test_gpu_pair_kernel_offset_tables(case, cuda):
    """Per-(iteration, layer) offset tables on the row-pair kernel (802.11ad rate 1/2,
    entries 1..4), both instances (edge table compiled in, and read at run time):
    posteriors, C2V and iteration counts equal the oracle's with the same table; the same
    decoder with a constant table equals the plain decoder; entries of 0 stay on the pair
    kernel, a negative entry falls back to the generic kernel."""
    import torch
    from fecsim import lib as L
    snr, mi, cap, ncw = case
    _say("pair kernel table", case)
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(ncw)
    llr, _ = rs.metrics(wgn, bits, snr)
    rng = np.random.default_rng(int(mi * 7 + cap))
    llr[0] = rng.integers(-128, 128, cs.n)
    llr[1] = np.where(rng.random(cs.n) < 0.5, -128, 127)
    tab = rng.integers(1, 5, (mi, cs.dec.mrow)).astype(np.float32)
    want = O.decode_ideal(cs.dec, llr, max_iter=mi, max_row_metric=cap, beta=1.0, beta_table=tab)
    plain = O.decode_ideal(cs.dec, llr, max_iter=mi, max_row_metric=cap, beta=1.0)
    for inst in (L.KERNEL_PAIR, L.KERNEL_PAIR_RT):
        d = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", mi, cap, 1.0)
        d.select_kernel(inst)
        d.set_offsets(tab)
        assert d.kernel == inst
        x = torch.from_numpy(llr).cuda()
        out = torch.empty_like(x)
        it = torch.zeros(ncw, dtype=torch.int32, device="cuda")
        c2v = torch.zeros(ncw * cs.dec.total_hc_ones * cs.z, dtype=torch.int8, device="cuda")
        tot = torch.zeros(1, dtype=torch.int64, device="cuda")
        d.decode(x, out, iters=it, c2v=c2v, tot_iter=tot)
        torch.cuda.synchronize()
        assert np.array_equal(it.cpu().numpy().astype(np.uint32), want[2]), inst
        assert np.array_equal(out.cpu().numpy(), want[0]), inst
        assert np.array_equal(c2v.cpu().numpy().reshape(ncw, -1), want[1]), inst
        assert int(tot.item()) == want[3]
        # a constant table of 1 is the plain beta-1 decoder
        d.set_offsets(np.ones((mi, cs.dec.mrow), np.float32))
        assert d.kernel == inst
        d.decode(x, out, iters=it)
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), plain[0]) and np.array_equal(it.cpu().numpy().astype(np.uint32), plain[2])
        # entries of 0 (the abs(-128) wrap path: a saturated -128 gets magnitude 0) on the
        # pair kernel; a negative entry falls back to the generic kernel; both equal the oracle
        for bad, kern in ((0.0, inst), (-1.0, L.KERNEL_GENERIC)):
            tab0 = tab.copy()
            tab0[0, 3] = bad
            tab0[mi - 1, :] = bad
            d.set_offsets(tab0)
            assert d.kernel == kern
            want0 = O.decode_ideal(cs.dec, llr, max_iter=mi, max_row_metric=cap, beta=1.0, beta_table=tab0)
            d.decode(x, out, iters=it, c2v=c2v)
            torch.cuda.synchronize()
            assert np.array_equal(out.cpu().numpy(), want0[0]), (inst, bad)
            assert np.array_equal(it.cpu().numpy().astype(np.uint32), want0[2]), (inst, bad)
            assert np.array_equal(c2v.cpu().numpy().reshape(ncw, -1), want0[1]), (inst, bad)


def test_gpu_pair_kernel_table_lifted_z_falls_back(cuda):
    """num_par = 2 lifts 11ad 1/2 to Z = 84 (no row-pair kernel); with a table the generic
    kernel decodes it, equal to the oracle."""
    import torch
    from fecsim import lib as L
    cs = codes.gen_structs("11ad 1/2", num_par=2, simd_size=512, elem_size=8)
    rs = refsim.RefStream(cs)
    wgn, bits = rs.inner_loop(64)
    llr, _ = rs.metrics(wgn, bits, -0.5)
    tab = np.random.default_rng(4).integers(1, 4, (10, cs.dec.mrow)).astype(np.float32)
    d = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 10, 31.0, 1.0)
    d.set_offsets(tab)
    assert d.kernel == L.KERNEL_GENERIC
    want = O.decode_ideal(cs.dec, llr, max_iter=10, beta=1.0, beta_table=tab)
    x = torch.from_numpy(llr).cuda()
    out = torch.empty_like(x)
    it = torch.zeros(64, dtype=torch.int32, device="cuda")
    d.decode(x, out, iters=it)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), want[0]) and np.array_equal(it.cpu().numpy().astype(np.uint32), want[2])
