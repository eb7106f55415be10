"""Measurements of the workloads beside the headline one (bench.py --config N).

  2  802.16e rate 1/2, Z = 24..96 step 4, float32 layered sum-product (Vec16f_prod:
     scale 1.0, |LLR| <= 1e9, beta 0, max_row_metric 6e-8, run_LDPCsim.pl:56-63), max 10
     iterations, Eb/N0 0.5..2.5 dB, 2^18 codewords per (Z, Eb/N0) point
  3  mixed-code queue: 65536 codewords per loop, code uniform over 11n 1/2 z27,
     11n 3/4 z54, 11n 5/6 z81, 11ad 13/16, 16e 2/3A Z=96, Eb/N0 uniform in [1, 4] dB,
     int8 offset min-sum, per-code max iterations 8/12/16/20/10, autotuned tile shapes,
     one HIP graph per loop (encode + WGN/LLR + decode + count)
  4  trainable decoder: 11n 1/2 z81, T = 10, batch 4096, forward + adjoint per step; then 400
     training steps and the learned tables against the plain float and int8 decoders
  5  5G NR base graph 1, Z = 384, rate 1/3 (2Z punctured), int8 offset min-sum, max 20
     iterations, 16384-codeword batches, Eb/N0 0..1.5 dB; throughput at CWER 1e-3

Every function returns one JSON-serialisable dict.  Data are synthetic (the library's
counter-based codewords and noise).
"""
from __future__ import annotations

import math
import sys
import time

import numpy as np


def _log(s):
    print(s, file=sys.stderr, flush=True)


def _rate_db(cs):
    return 10 * math.log10((cs.enc.mcol - cs.enc.mrow) / cs.enc.mcol)


# VALU roofs of one MI355X: 256 CUs x 64 lanes x 2.4 GHz 32-bit lane-ops per second; the
# float and int32-lane kernels issue no packed (two-lane) instructions
VALU32_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
# algorithmic work per edge-update (one edge of one check row in one iteration):
# offset min-sum 13 int ops (residual with saturation, |.|, -beta, max 0, sign xor; exclusive
# min 3; sign select, negate, cap, saturating add; syndrome xor); sum-product 15 float ops and
# 2 transcendentals (tanh, atanh), counted as one op each
OPS_OMS, OPS_SP = 13, 17


def _timed(fn, stream, reps=1):
    """average milliseconds of fn() launched reps times on stream (HIP events on that stream)"""
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def _roofline(kernel, edge_updates, ms, ops, note, launches=1):
    """VALU roofline object: algorithmic ops over the measured kernel time"""
    tops = edge_updates * ops / (ms * 1e-3) / 1e12
    return {"bound": "valu", "kernel": kernel, "achieved": tops, "peak": VALU32_PEAK_TOPS, "unit": "Tops/s",
            "frac": tops / VALU32_PEAK_TOPS, "ops_per_edge_update": ops,
            "edge_updates_per_launch": edge_updates / launches, "avg_launch_ms": ms / launches,
            "traffic": None, "note": note}


def config2(args):
    import torch
    from . import codes
    from .decoder import Decoder, Hdec
    from .sim import Henc, Sim
    ebn0 = [0.5, 1.0, 1.5, 2.0, 2.5]
    ncw = 1 << 18
    per_z, tot_cw, tot_t = [], 0, 0.0
    roof_eu, roof_ms = 0.0, 0.0
    zs = [int(x) for x in args.z.split(",")] if getattr(args, "z", None) else list(range(24, 97, 4))
    for z in zs:
        cs = codes.gen_structs("16e 1/2", simd_size=512, elem_size=32, z_over=z)
        henc = Henc.from_tables(cs.enc)
        dec = Decoder(Hdec.from_tables(cs.dec), "Vec16f_prod", 10, 6e-8, 0.0)
        snrs = [e + _rate_db(cs) for e in ebn0]
        sim = Sim(henc, dec, ncw, snrs, llr_scale=1.0, max_abs_llr=1e9, seed=100 + z, use_graph=True)
        sim.run(1)
        torch.cuda.synchronize()
        sim.reset()
        t0 = time.perf_counter()
        sim.run(1)
        st = sim.stats()
        dt = time.perf_counter() - t0
        k = (cs.enc.mcol - cs.enc.mrow) * cs.z
        per_z.append({"z": z, "seconds": dt, "cw_per_s": ncw * len(snrs) / dt,
                      "ber": [x["bit_err"] / (x["cw"] * k) for x in st],
                      "cwer": [x["cw_err"] / x["cw"] for x in st],
                      "mean_iterations": [x["iters"] / x["cw"] for x in st]})
        tot_cw += ncw * len(snrs)
        tot_t += dt
        # decode-only launch at Eb/N0 1.5 dB (the middle point) for the roofline
        from . import lib as _L
        bits = henc.new_bits(ncw)
        henc.gen_codewords(bits, seed=7)
        llr = torch.empty((ncw, henc.n), dtype=torch.float32, device="cuda")
        henc.gen_llr(bits, llr, _L.F32, 1.5 + _rate_db(cs), seed=7, llr_scale=1.0, max_abs_llr=1e9)
        out = torch.empty_like(llr)
        it = torch.zeros(ncw, dtype=torch.int32, device="cuda")
        st_ = torch.cuda.current_stream()
        dec.decode(llr, out, iters=it, stream=st_)
        ms = _timed(lambda: dec.decode(llr, out, iters=it, stream=st_), st_)
        eu = float(it.double().sum().item()) * cs.dec.total_hc_ones * cs.z
        roof_eu += eu
        roof_ms += ms
        del bits, llr, out
        _log("config 2: z=%d %.3f s cwer %s; decode at 1.5 dB %.2f ms" % (z, dt, ["%.2e" % c for c in per_z[-1]["cwer"]],
                                                                          ms))
    return {"config": 2, "roofline": _roofline(
                "k_float_reg<FloatPolicy<float,true>,PairShape8> (16e 1/2, C2V in registers)", roof_eu, roof_ms, OPS_SP,
                "decode-only launches of 2^18 cw at Eb/N0 1.5 dB, one per Z, summed", launches=len(zs)), "workload": "16e 1/2 Z=24..96 step 4, Vec16f_prod (float32 sum-product), max 10 iterations, "
                                     "Eb/N0 0.5..2.5 dB step 0.5, 2^18 cw per point, simulation loop per Z",
            "ebn0_db": ebn0, "cw_per_s": tot_cw / tot_t, "codewords": tot_cw, "seconds": tot_t, "per_z": per_z,
            "dtype": "f32", "data": "synthetic"}


CODES3 = [("11n 1/2 z27", None, 8), ("11n 3/4 z54", None, 12), ("11n 5/6 z81", None, 16),
          ("11ad 13/16", None, 20), ("16e 2/3A", 96, 10)]


def config3(args):
    import torch
    from . import codes
    from .decoder import Decoder, Hdec
    from .multi import MixedSim, MultiDecoder
    from .sim import Henc
    n = 65536
    decs, hencs, css = [], [], []
    for name, zo, mi in CODES3:
        cs = codes.gen_structs(name, simd_size=512, elem_size=8, z_over=zo)
        css.append(cs)
        decs.append(Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", mi, 31.0, 1.0))
        hencs.append(Henc.from_tables(cs.enc))
    m = MultiDecoder(decs, n)
    shapes0 = [m.shape(c) for c in range(len(CODES3))]
    # autotune on one loop's queue of fresh channel LLRs (copied aside: every decode reads
    # them and writes the posteriors elsewhere)
    tune = MixedSim(hencs, m, n, ebn0=(1.0, 4.0), seed=5, use_graph=False)
    tune.fill()
    fresh = tune.llr().clone()
    torch.cuda.synchronize()
    import ctypes as C
    from . import lib as _L
    code, off = C.c_void_p(), C.c_void_p()
    _L.check(_L.lib().fecsim_msim_queue(tune._s, C.byref(code), C.byref(off), None, None), "queue")
    llr = C.c_void_p(fresh.data_ptr())
    scratch = torch.empty_like(fresh)
    _L.check(_L.lib().fecsim_multi_autotune(m._m, n, code, off, llr, C.c_void_p(scratch.data_ptr()), 3,
                                            C.c_void_p(torch.cuda.current_stream().cuda_stream)), "autotune")
    shapes1 = [m.shape(c) for c in range(len(CODES3))]
    kernel = m.kernel   # the autotuner's choice: the compiled-code kernel or the tuned generic one
    # decode-only launches of the tuned queue for the roofline (grouping sort included)
    qcode, qoff, _, _ = tune.queue()
    it = torch.zeros(n, dtype=torch.int32, device="cuda")
    st_ = torch.cuda.current_stream()
    sp = C.c_void_p(st_.cuda_stream)

    def _dec():
        _L.check(_L.lib().fecsim_multi_decode(m._m, n, code, off, llr, C.c_void_p(scratch.data_ptr()),
                                              C.c_void_p(it.data_ptr()), None, sp), "multi decode")
    _dec()
    ms = _timed(_dec, st_, reps=3)
    e_code = torch.tensor([cs.dec.total_hc_ones * cs.z for cs in css], dtype=torch.float64, device="cuda")
    eu = float((it.double() * e_code[qcode.long()]).sum().item())
    code_kernel = kernel == _L.KERNEL_CODE
    roof = _roofline("k_multi_code (codes compiled in)" if code_kernel else "k_multi_decode<IntPolicy<8>,24>", eu, ms,
                     OPS_OMS, "decode-only launches of the autotuned 65536-entry queue of fresh channel LLRs "
                     "(grouping sort included; %.2f mean iterations)" % (float(it.double().mean().item())))
    # the other kernel on the same queue, for comparison; then the tuned choice again
    try:
        m.select_kernel(_L.KERNEL_GENERIC if code_kernel else _L.KERNEL_CODE)
        _dec()
        other_ms = _timed(_dec, st_, reps=3)
    except _L.FecsimError:   # a code not compiled in: the generic kernel only
        other_ms = None
    m.select_kernel(_L.KERNEL_CODE if code_kernel else _L.KERNEL_GENERIC)
    del tune
    sim = MixedSim(hencs, m, n, ebn0=(1.0, 4.0), seed=100, use_graph=True)
    sim.run(1)
    torch.cuda.synchronize()
    sim.reset()
    loops = max(20, args.steps)   # ~50 ms of loops: a stable rate
    t0 = time.perf_counter()
    sim.run(loops)
    st, iters = sim.stats()
    dt = time.perf_counter() - t0
    info_bits = sum(s["cw"] * (cs.enc.mcol - cs.enc.mrow) * cs.z for s, cs in zip(st, css))
    return {"config": 3, "roofline": roof, "kernel": "k_multi_code" if code_kernel else "k_multi_decode",
            "decode_only_ms": ms, ("decode_only_ms_generic" if code_kernel else "decode_only_ms_code"): other_ms,
            "workload": "mixed queue of 65536 cw per loop over 5 codes (11n 1/2 z27, 11n 3/4 z54, "
                                     "11n 5/6 z81, 11ad 13/16, 16e 2/3A Z=96), Eb/N0 U[1,4] dB, int8 min-sum, max "
                                     "iterations 8/12/16/20/10, HIP graph per loop (encode+WGN+decode+count)",
            "loops": loops, "seconds": dt, "cw_per_s": n * loops / dt, "info_gbit_s": info_bits / dt / 1e9,
            "mean_iterations": iters / (n * loops),
            "tile_shapes_default": shapes0, "tile_shapes_autotuned": shapes1,
            "per_code": [{"code": c[0], "cw": s["cw"], "cwer": s["cw_err"] / max(1, s["cw"])}
                         for c, s in zip(CODES3, st)],
            "dtype": "int8", "data": "synthetic"}


def config4(args):
    import torch
    from . import codes
    from .decoder import Hdec
    from .sim import Henc
    from .train import Trainer, make_batch, min_sum_function, rate_db
    cs = codes.gen_structs("11n 1/2 z81", simd_size=512, elem_size=8)
    henc = Henc.from_tables(cs.enc)
    T, B = 10, 4096
    tr = Trainer(Hdec.from_tables(cs.dec), T, B)
    alpha = torch.full((T, cs.dec.mrow), 0.8, device="cuda", requires_grad=True)
    beta = torch.full((T, cs.dec.mrow), 0.5, device="cuda", requires_grad=True)
    fn = min_sum_function()
    rng = np.random.default_rng(1)
    rate = rate_db(cs.dec.mcol * cs.z - cs.enc.mrow * cs.z, cs.dec.mcol * cs.z)
    llr, b01 = make_batch(henc, B, 1.0, 3.0, rate, 7, 0, rng)
    steps = max(3, args.steps)
    for _ in range(2):
        post = fn.apply(llr, alpha, beta, tr)
        torch.nn.functional.binary_cross_entropy_with_logits(-post, b01).backward()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    fwd_ms = bwd_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        ev[0].record()
        post = fn.apply(llr, alpha, beta, tr)
        loss = torch.nn.functional.binary_cross_entropy_with_logits(-post, b01)
        ev[1].record()
        loss.backward()
        ev[2].record()
        torch.cuda.synchronize()
        fwd_ms += ev[0].elapsed_time(ev[1])
        bwd_ms += ev[1].elapsed_time(ev[2])
    dt = time.perf_counter() - t0
    # the decoder's two launches alone, back to back on one stream (no loss or optimiser
    # kernels between the events)
    x, al, be = llr.contiguous(), alpha.detach().contiguous(), beta.detach().contiguous()
    gpost = torch.randn_like(x)
    kf = kb = 0.0
    for _ in range(steps):
        ev[0].record()
        tr.forward(x, al, be)
        ev[1].record()
        tr.backward(gpost)
        ev[2].record()
        torch.cuda.synchronize()
        kf += ev[0].elapsed_time(ev[1])
        kb += ev[1].elapsed_time(ev[2])
    E = cs.dec.total_hc_ones * cs.z
    procedure = config4_procedure(cs, henc) if getattr(args, "procedure", True) else None
    return {"config": 4, "procedure": procedure,
            "workload": "11n 1/2 z81 trainable offset/normalized min-sum, T=10, batch 4096, "
                                     "forward + adjoint (checkpoints at segment starts, segment length of least memory, "
                                     "segments recomputed)",
            "steps": steps, "step_ms": dt / steps * 1e3, "forward_ms": fwd_ms / steps, "backward_ms": bwd_ms / steps,
            "decoder_forward_ms": kf / steps, "decoder_adjoint_ms": kb / steps, "adjoint_over_forward": kb / kf,
            "cw_per_s_train": B * steps / dt, "edge_updates_per_s_forward": B * T * E / (kf / steps * 1e-3),
            "dtype": "f32", "data": "synthetic"}


def config4_procedure(cs, henc, steps=400, ncw=1 << 18):
    """The training procedure of the config: Adam (lr 1e-3) on the BCE of the posteriors
    after T = 10 iterations, scales and offsets per (iteration, layer) from 0.8 and 0.5,
    batches of 4096 codewords at Eb/N0 U[1, 3] dB; then the learned float decoder against
    the plain one (offset 0.5, scale 1), and the learned offsets quantised into the int8
    decoder (LLR scale 2) against its plain offset 1, both at 10 iterations, CWER 1e-3
    crossings on the same codewords and noise."""
    from .decoder import Decoder, Hdec
    from .train import cwer_curve, eval_float, quantize_offsets, snr_at, train
    T = 10
    t0 = time.perf_counter()
    alpha, beta, losses = train(cs, henc, T=T, batch=4096, steps=steps, ebn0=(1.0, 3.0), lr=1e-3)
    t_train = time.perf_counter() - t0
    snrs = [-1.5 + 0.125 * i for i in range(17)]
    ones = np.ones_like(alpha)
    fp = eval_float(cs, henc, ones, 0.5 * ones, snrs, ncw, T)
    fl = eval_float(cs, henc, alpha, beta, snrs, ncw, T)
    q = quantize_offsets(beta, 2.0, max_iter=T)
    hdec = Hdec.from_tables(cs.dec)
    plain = Decoder(hdec, "Vec64c_min", T, 31.0, 1.0)
    learned = Decoder(hdec, "Vec64c_min", T, 31.0, 1.0)
    learned.set_offsets(q)
    ip = cwer_curve(cs, henc, plain, snrs, ncw)
    il = cwer_curve(cs, henc, learned, snrs, ncw)

    def gain(a, b):
        sa, sb = snr_at(a), snr_at(b)
        return (sa - sb) if sa is not None and sb is not None else None

    return {"steps": steps, "train_seconds": t_train, "loss_first": losses[0], "loss_last": losses[-1],
            "float_gain_db": gain(fp, fl), "int8_gain_db": gain(ip, il),
            "int8_table_entries": sorted({int(x) for x in q.ravel()}), "cw_per_point": ncw, "es_n0_db": snrs,
            "cwer_float_plain": [c for _, c, *_ in fp], "cwer_float_learned": [c for _, c, *_ in fl],
            "cwer_int8_plain": [c for _, c, *_ in ip], "cwer_int8_learned": [c for _, c, *_ in il]}


def bg1_learned_gain(cs=None, henc=None, plain=None):
    """The committed BG1 offset table (fecsim/learned/5GNR_n25344_k8448_T20_int8.json)
    against the plain int8 decoder (beta 1, cap 31, max 20 iterations), both on the same
    codewords and noise, on a fine Eb/N0 grid around CWER 1e-3 with 2^20 codewords per
    point: the SNR gain at CWER 1e-3 by log-linear interpolation."""
    from . import codes
    from .decoder import Decoder, Hdec
    from .sim import Henc
    from .train import cwer_curve, load_offsets, snr_at
    if cs is None:
        cs = codes.gen_structs("5GNR_n25344_k8448", simd_size=512, elem_size=8)
        henc = Henc.from_tables(cs.enc)
    if plain is None:
        plain = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 20, 31.0, 1.0)
    rate = 10 * math.log10(8448 / 25344)
    tab, rec = load_offsets("5GNR_n25344_k8448_T20_int8")
    ldec = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 20, 31.0, 1.0)
    ldec.set_offsets(tab)
    fine = [0.55, 0.6, 0.65, 0.7, 0.75, 0.8]
    fsnr = [e + rate for e in fine]
    cp = cwer_curve(cs, henc, plain, fsnr, 262144, seed=300, loops=4)
    cl = cwer_curve(cs, henc, ldec, fsnr, 262144, seed=300, loops=4)
    curve_p = [(e, c) for e, (_, c, *_) in zip(fine, cp)]
    curve_l = [(e, c) for e, (_, c, *_) in zip(fine, cl)]
    e3_p, e3_l = snr_at(curve_p), snr_at(curve_l)
    return {"table": "simd-fecfun_amd/fecsim/learned/5GNR_n25344_k8448_T20_int8.json", "made_by": rec["command"],
            "code": "5G NR BG1 Z=384 (K=8448, N=25344 sent), int8, max 20 iterations", "cw_per_point": 1 << 20,
            "ebn0_at_cwer_1e-3_plain": e3_p, "ebn0_at_cwer_1e-3_learned": e3_l,
            "gain_db": (e3_p - e3_l) if e3_p is not None and e3_l is not None else None,
            "curve_ebn0_cwer_plain": curve_p, "curve_ebn0_cwer_learned": curve_l, "_decoder": ldec}


def config5(args):
    import torch
    from . import codes
    from .decoder import Decoder, Hdec
    from .sim import Henc, Sim
    from .train import snr_at
    cs = codes.gen_structs("5GNR_n25344_k8448", simd_size=512, elem_size=8)
    henc = Henc.from_tables(cs.enc)
    dec = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 20, 31.0, 1.0)
    rate = 10 * math.log10(8448 / 25344)          # transmitted rate 1/3 (2Z columns punctured)
    ebn0 = [0.0, 0.25, 0.5, 0.75, 1.0, 1.25, 1.5]
    snrs = [e + rate for e in ebn0]
    ncw = 16384
    loops = max(1, args.steps)
    sim = Sim(henc, dec, ncw, snrs, seed=100, use_graph=True)
    sim.run(1)
    torch.cuda.synchronize()
    sim.reset()
    t0 = time.perf_counter()
    sim.run(loops)
    st = sim.stats()
    dt = time.perf_counter() - t0
    k = 8448
    curve = [(e, x["cw_err"] / x["cw"], x["bit_err"] / (x["cw"] * k), x["iters"] / x["cw"]) for e, x in zip(ebn0, st)]
    e3 = snr_at(curve)
    lg = bg1_learned_gain(cs, henc, dec)
    gain = lg["gain_db"]
    if gain is not None and gain > 0:
        dec = lg.pop("_decoder")
    lg.pop("_decoder", None)
    # decode-only throughput at the Eb/N0 closest to CWER 1e-3
    pick = min(range(len(ebn0)), key=lambda i: abs(math.log10(max(curve[i][1], 1e-9)) + 3))
    bits = henc.new_bits(ncw)
    henc.gen_codewords(bits, seed=7)
    from . import lib as _L
    llr = torch.empty((ncw, henc.n), dtype=torch.int8, device="cuda")
    henc.gen_llr(bits, llr, _L.I8, snrs[pick], seed=7)
    out = torch.empty_like(llr)
    it = torch.zeros(ncw, dtype=torch.int32, device="cuda")
    dec.decode(llr, out, iters=it)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 3
    e0.record()
    for _ in range(reps):
        dec.decode(llr, out, iters=it)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    roof = _roofline("k_code_i8<CodeTab0>" if dec.kernel == _L.KERNEL_CODE else "k_decode_wg<IntPolicy<8>,...>",
                     float(it.double().sum().item()) * cs.dec.total_hc_ones * cs.z, ms, OPS_OMS,
                     "decode-only launch of 16384 cw at Eb/N0 %.2f dB" % ebn0[pick])
    return {"config": 5, "roofline": roof, "workload": "5G NR BG1 Z=384 (K=8448, N=25344 sent, 2Z punctured), int8 offset min-sum "
                                     "beta 1 cap 31, max 20 iterations, 16384 cw per batch, Eb/N0 0..1.5 dB",
            "sim_seconds": dt, "sim_cw_per_s": ncw * len(snrs) * loops / dt,
            "curve_ebn0_cwer_ber_iters": curve, "ebn0_at_cwer_1e-3": e3,
            "learned_offsets": dict(lg, decoder_used="learned" if gain is not None and gain > 0
                                    else "plain (the learned table does not gain)"),
            "decode_at_ebn0": ebn0[pick], "decode_cw_per_s": ncw / (ms * 1e-3),
            "decode_info_gbit_s": ncw * k / (ms * 1e-3) / 1e9, "decode_mean_iterations": float(it.float().mean()),
            "dtype": "int8", "data": "synthetic"}


def config6(args):
    """802.3ca (Z = 256, 12 x 69 blocks, K = 14592 of N = 17664), int8 offset min-sum beta 1
    cap 31, max 10 iterations: simulation-loop CWER curves of the mother code and of the
    code with its last two parity block columns punctured, and decode-only throughput."""
    import torch
    from . import codes
    from . import lib as _L
    from .decoder import Decoder, Hdec
    from .sim import Henc, Sim
    from .train import snr_at
    ncw = 32768
    loops = max(1, args.steps)
    k = 14592
    out = {"config": 6, "workload": "802.3ca LDPC (N=17664, K=14592, Z=256) int8 offset min-sum beta 1 cap 31, "
                                    "max 10 iterations, 32768 cw per batch; mother code and last 512 parity bits "
                                    "punctured", "dtype": "int8", "data": "synthetic"}
    for punct in (False, True):
        cs = codes.gen_structs("802.3ca", simd_size=512, elem_size=8, puncture_8023ca=punct)
        henc = Henc.from_tables(cs.enc)
        dec = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 10, 31.0, 1.0)
        # Es/N0 dB around the waterfall (about 0.1 dB wide at 65536 cw per point)
        snrs = [round((2.15 if not punct else 2.9) + 0.05 * i, 2) for i in range(8)]
        sim = Sim(henc, dec, ncw, snrs, seed=100, use_graph=True)
        sim.run(1)
        torch.cuda.synchronize()
        sim.reset()
        t0 = time.perf_counter()
        sim.run(loops)
        st = sim.stats()
        dt = time.perf_counter() - t0
        curve = [(e, x["cw_err"] / x["cw"], x["bit_err"] / (x["cw"] * k), x["iters"] / x["cw"])
                 for e, x in zip(snrs, st)]
        tag = "punctured" if punct else "mother"
        out[tag] = {"curve_esn0_cwer_ber_iters": curve, "esn0_at_cwer_1e-3": snr_at(curve),
                    "sim_cw_per_s": ncw * len(snrs) * loops / dt}
        if punct:
            continue
        pick = min(range(len(snrs)), key=lambda i: abs(math.log10(max(curve[i][1], 1e-9)) + 3))
        bits = henc.new_bits(ncw)
        henc.gen_codewords(bits, seed=7)
        llr = torch.empty((ncw, henc.n), dtype=torch.int8, device="cuda")
        henc.gen_llr(bits, llr, _L.I8, snrs[pick], seed=7)
        post = torch.empty_like(llr)
        it = torch.zeros(ncw, dtype=torch.int32, device="cuda")
        dec.decode(llr, post, iters=it)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 3
        e0.record()
        for _ in range(reps):
            dec.decode(llr, post, iters=it)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        out.update({"decode_at_esn0": snrs[pick], "decode_ms_per_batch": ms, "decode_cw_per_s": ncw / (ms * 1e-3),
                    "decode_info_gbit_s": ncw * k / (ms * 1e-3) / 1e9,
                    "decode_mean_iterations": float(it.float().mean())})
    return out


def run_config(n, args):
    return {2: config2, 3: config3, 4: config4, 5: config5, 6: config6}[n](args)
