"""Benchmark: decoded info Gbit/s and codewords/s of the int8 offset min-sum decoder,
802.11ad rate 1/2 (N=672, K=336, Z=42), on one MI355X.

Workload (the reference's default run, build/run_LDPCsim.pl:8-39 / README.md:41-47):
beta 1, row cap 31, at most 10 iterations with syndrome early exit, channel LLRs
quantized to int8 (scale 2, |LLR| <= 24) for SNR -1.00 .. 0.00 dB in 0.25 dB steps,
1e6 codewords per SNR point, random codewords (seed 100) from the library's encoder
and noise generator.  One step = decoding all five SNR batches (5e6 cw), decode
only, inputs resident in HBM (the reference times LDPCdecode_set alone,
src/Top/LDPCtop.cpp:107-110).  The end-to-end simulation loop (new codewords,
LLRs, decode, error count per SNR point; LDPCtop.cpp:86-121) is timed after it
and reported as sim_cw_per_s.

Prints one JSON line.  The CPU baseline (the oracle's SIMD restatement of
LDPCtop_Vec64c_min on the host cores) is timed in the same run on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(REPO, "simd-fecfun_amd"), os.path.join(REPO, "oracle"), os.path.join(REPO, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "decoded info Gbit/s + CW/s, int8 min-sum 11ad 1/2; dB gain of learned offsets"
SNRS = (-1.0, -0.75, -0.5, -0.25, 0.0)
K_INFO = 336
SEED = 100                                  # run_LDPCsim.pl:36
HBM_PEAK_GBS = 8000.0
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12   # 256 CUs x 64 lanes x 2.4 GHz, 32-bit lane-ops
# the int8 decoders compute in packed 16-bit halves (v_pk_* ops: two 16-bit lane-ops per
# lane and cycle), so their roof is twice the 32-bit one
VALU_PK16_PEAK_TOPS = 2 * VALU_PEAK_TOPS
OPS_PER_EDGE_UPDATE = 13                    # algorithmic int ops per edge-update (offset min-sum)
E_11AD = 2184                               # edges per iteration, 11ad 1/2
# HBM bytes per codeword moved by one k_pair_i8c launch, from rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs scaled by a calibration copy (profiles/k_pair_i8c_11ad12_1e6cw_hbm_traffic.md)
PAIR_TRAFFIC_BYTES_PER_CW = 1370.6
LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12     # 256 CUs x 128 B/clk x 2.4 GHz
LDS_BYTES_PER_EDGE_UPDATE = 8


def make_inputs(henc, ncw, seed, torch, llr_scale=2.0, max_abs=24.0):
    """Random codewords (counter-based data bits + QC encoder), BPSK over AWGN,
    quantized like gen_metrics.cpp:76-98, all generated on the GPU by the library.
    As the reference does (LDPCtop.cpp:90-97), the codewords and the noise are the
    same at every SNR point."""
    from fecsim import lib as _L
    bits = henc.new_bits(ncw)
    henc.gen_codewords(bits, seed=seed, cw0=0)
    llrs = []
    for snr in SNRS:
        x = torch.empty((ncw, henc.n), dtype=torch.int8, device="cuda")
        henc.gen_llr(bits, x, _L.I8, snr, seed=seed, cw0=0, llr_scale=llr_scale, max_abs_llr=max_abs)
        llrs.append(x)
    return bits, llrs


def host_cpus():
    """The host cores this process may run on, and what else describes them."""
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count() or 1
    quota = None
    try:   # cgroup v2 CPU bandwidth limit, "max" or "<quota> <period>"
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, per = fh.read().split()[:2]
        if q != "max":
            quota = float(q) / float(per)
    except (OSError, ValueError):
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"affinity_cores": affinity, "os_cpu_count": os.cpu_count(), "cgroup_cpu_quota": quota,
            "cpu_model": model}


def cpu_baseline(cs, llr_samples, max_seconds=12.0):
    """Oracle SIMD decoder (LDPCtop_Vec64c_min restated) on every host core this process
    may use (its CPU affinity set; a cgroup CPU quota below that caps the threads)."""
    from refsim import O
    info = host_cpus()
    nthreads = info["affinity_cores"]
    if info["cgroup_cpu_quota"] is not None:
        nthreads = max(1, min(nthreads, int(round(info["cgroup_cpu_quota"]))))
    bufs = [O.dense_to_ref(x, cs.z, cs.dec.mcol, 64) for x in llr_samples]
    done_cw, t_dec, isa = 0, 0.0, 0
    while t_dec < max_seconds:
        for b in bufs:
            work = b.copy()
            t0 = time.perf_counter()
            _, isa = O.decode_vec64c(cs.dec, work, nthreads=nthreads)
            t_dec += time.perf_counter() - t0
            done_cw += b.shape[0]
    # one-thread figure on one SNR batch slice, for README.md:43-47
    one = bufs[0][:20000].copy()
    t0 = time.perf_counter()
    O.decode_vec64c(cs.dec, one, nthreads=1)
    t1 = time.perf_counter() - t0
    res = {
        "value": done_cw / t_dec, "unit": "cw/s", "cores": nthreads, "kind": "port",
        "sample": "%d cw decoded (%d-cw batches at SNR %s dB, repeated) by oracle/baseline_vec.c "
                  "(AVX%d int8 x64 lanes, %d threads), %.1f s of decode" % (
                      done_cw, bufs[0].shape[0], "/".join("%g" % s for s in SNRS), isa, nthreads, t_dec),
        "one_thread_cw_per_s": 20000 / t1,
        "info_gbit_s": done_cw / t_dec * K_INFO / 1e9,
    }
    res.update(info)
    return res


def learned_gain(cs, henc, ncw):
    """Es/N0 gain at CWER 1e-3 of the committed learned offset table over the plain
    offset decoder (beta 1), both int8 with at most 10 iterations, measured on the same
    codewords and noise (device simulation loop, seed 100) over -1.0 .. 0.5 dB."""
    from fecsim.decoder import Decoder, Hdec
    from fecsim.train import cwer_curve, load_offsets, snr_at
    tab, rec = load_offsets("11ad12_T10_int8")
    hdec = Hdec.from_tables(cs.dec)
    plain = Decoder(hdec, "Vec64c_min", 10, 31.0, 1.0)
    learned = Decoder(hdec, "Vec64c_min", 10, 31.0, 1.0)
    learned.set_offsets(tab)
    snrs = [-1.0 + 0.125 * i for i in range(13)]
    cp = cwer_curve(cs, henc, plain, snrs, ncw)
    cl = cwer_curve(cs, henc, learned, snrs, ncw)
    sp, sl = snr_at(cp), snr_at(cl)
    return {"gain_db": (sp - sl) if sp is not None and sl is not None else None,
            "table": "simd-fecfun_amd/fecsim/learned/11ad12_T10_int8.json", "made_by": rec["command"],
            "es_n0_at_cwer_1e-3_plain": sp, "es_n0_at_cwer_1e-3_learned": sl, "cw_per_point": ncw,
            "cwer_plain": [c for _, c, *_ in cp], "cwer_learned": [c for _, c, *_ in cl],
            "es_n0_db": snrs}


def probe(args):
    import torch
    from fecsim import codes, lib as _L
    from fecsim.decoder import Decoder, Hdec
    from fecsim.sim import Henc, Sim
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    henc = Henc.from_tables(cs.enc)
    dec = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 10, 31.0, 1.0)
    ncw = args.ncw
    if args.probe == "traffic":
        bits = henc.new_bits(ncw)
        henc.gen_codewords(bits, seed=SEED)
        x = torch.empty((ncw, cs.n), dtype=torch.int8, device="cuda")
        henc.gen_llr(bits, x, _L.I8, -0.5, seed=SEED)
        o = torch.empty_like(x)
        it = torch.zeros(ncw, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        for _ in range(2):
            dec.decode(x, o, iters=it)
            torch.cuda.synchronize()
        for _ in range(2):
            o.copy_(x)
            torch.cuda.synchronize()
        print(json.dumps({"probe": "traffic", "ncw": ncw, "copy_bytes": x.numel(),
                          "mean_iterations": float(it.float().mean())}))
        return
    if args.probe == "host":
        # fecsim_decode_set_host: LLRs in pageable host memory, decoded in place through
        # the GPU (copy in, decode, copy out), as a caller holding host buffers sees it
        bits = henc.new_bits(ncw)
        henc.gen_codewords(bits, seed=SEED)
        x = torch.empty((ncw, cs.n), dtype=torch.int8, device="cuda")
        henc.gen_llr(bits, x, _L.I8, -0.5, seed=SEED)
        host = x.cpu().numpy()
        work = host.copy()
        dec.decode_set_host(work)
        times = []
        for _ in range(max(1, args.steps)):
            work[...] = host
            t0 = time.perf_counter()
            dec.decode_set_host(work)
            times.append(time.perf_counter() - t0)
        best = min(times)
        print(json.dumps({"probe": "host", "ncw": ncw, "bytes_each_way": int(host.nbytes),
                          "seconds_best": best, "cw_per_s": ncw / best,
                          "info_gbit_s": ncw * K_INFO / best / 1e9}))
        return
    sim = Sim(henc, dec, ncw, list(SNRS), seed=SEED, use_graph=False)
    sim.run(1)
    torch.cuda.synchronize()
    sim.reset()
    t0 = time.perf_counter()
    sim.run(args.steps)
    st = sim.stats()
    dt = time.perf_counter() - t0
    print(json.dumps({"probe": "sim", "ncw": ncw, "loops": args.steps,
                      "sim_cw_per_s": ncw * len(SNRS) * args.steps / dt, "stats": st}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ncw", type=int, default=1_000_000, help="codewords per SNR point")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-offset CWER comparison")
    ap.add_argument("--learned-ncw", type=int, default=2_000_000, help="codewords per SNR point of that comparison")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--kernel", choices=["auto", "generic", "pair", "pair_rt"], default="auto",
                    help="decoder kernel (results are identical); pair_rt: the row-pair kernel's instance "
                         "that reads the edge table at run time")
    ap.add_argument("--probe", choices=["traffic", "sim", "host"], default=None,
                    help="short fixed workloads for profiler runs instead of the benchmark: 'traffic' = "
                         "one decode of --ncw codewords at -0.5 dB twice, then a plain copy of the LLR "
                         "tensor twice (counter calibration); 'sim' = --steps simulation loops; 'host' = "
                         "decode of --ncw codewords held in host memory (PCIe copies included), best of --steps")
    ap.add_argument("--config", type=int, choices=[1, 2, 3, 4, 5, 6], default=1,
                    help="1: 11ad 1/2 int8 decode (the headline metric); 2: 16e 1/2 float sum-product Z sweep; "
                         "3: mixed-code queue; 4: trainable decoder step; 5: 5G NR BG1 Z=384 int8; 6: 802.3ca int8")
    ap.add_argument("--z", default=None, help="config 2: comma-separated lifting sizes instead of 24..96")
    args = ap.parse_args()
    if args.probe:
        return probe(args)
    if args.config != 1:
        if args.gpus != 1:
            print("bench.py measures one GPU only", file=sys.stderr)
            sys.exit(2)
        from fecsim.benchmarks import run_config
        print(json.dumps(run_config(args.config, args)))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (codewords are "
              "independent, so a batch would shard by codeword range with a final sum of "
              "the error counters)", file=sys.stderr)
        sys.exit(2)

    import torch
    import numpy as np
    from fecsim import codes
    from fecsim.decoder import Decoder, Hdec

    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    dec = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 10, 31.0, 1.0)
    from fecsim import lib as _L
    dec.select_kernel({"auto": _L.KERNEL_AUTO, "generic": _L.KERNEL_GENERIC, "pair": _L.KERNEL_PAIR,
                       "pair_rt": _L.KERNEL_PAIR_RT}[args.kernel])
    kname = {_L.KERNEL_PAIR: "k_pair_i8c<PairShape0> (row-pair, edge table compiled in, persistent)",
             _L.KERNEL_PAIR_RT: "k_pair_i8<PairShape0> (row-pair, run-time edge table, persistent)"}.get(
                 dec.kernel, "k_decode<IntPolicy<8>,8>")
    ncw = args.ncw
    print("bench: generating %d x %d codeword LLRs on the GPU" % (len(SNRS), ncw), file=sys.stderr, flush=True)
    from fecsim.sim import Henc, Sim
    henc = Henc.from_tables(cs.enc)
    bits, llrs = make_inputs(henc, ncw, SEED, torch)
    outs = [torch.empty_like(x) for x in llrs]
    iters = [torch.zeros(ncw, dtype=torch.int32, device="cuda") for _ in SNRS]
    torch.cuda.synchronize()

    stream = torch.cuda.current_stream()

    def step():
        for x, o, it in zip(llrs, outs, iters):
            dec.decode(x, o, iters=it, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print("bench: timing %d steps" % args.steps, file=sys.stderr, flush=True)
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps * len(SNRS))]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    k = 0
    for _ in range(args.steps):
        for x, o, it in zip(llrs, outs, iters):
            ev[k][0].record(stream)
            dec.decode(x, o, iters=it, stream=stream)
            ev[k][1].record(stream)
            k += 1
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kern_ms = [a.elapsed_time(b) for a, b in ev]

    tot_iter = int(sum(int(it.sum().item()) for it in iters))   # per step
    cw_per_step = ncw * len(SNRS)
    cw_s = cw_per_step * args.steps / wall
    gbit = cw_s * K_INFO / 1e9
    avg_launch_ms = float(np.mean(kern_ms))
    # algorithmic HBM bytes per launch: N bytes LLR in + N bytes posterior out + 4 B iteration count per cw
    hbm_bytes = ncw * (cs.n + cs.n + 4)
    hbm_gbs = hbm_bytes / (avg_launch_ms * 1e-3) / 1e9
    edge_updates_per_launch = tot_iter / len(SNRS) * E_11AD
    valu_tops = edge_updates_per_launch * OPS_PER_EDGE_UPDATE / (avg_launch_ms * 1e-3) / 1e12
    lds_tbs = edge_updates_per_launch * LDS_BYTES_PER_EDGE_UPDATE / (avg_launch_ms * 1e-3) / 1e12
    mean_iters = {"%g" % s: float(it.float().mean().item()) for s, it in zip(SNRS, iters)}

    # error rates of the decoded batch (data columns vs transmitted bits)
    from fecsim import lib as _L
    rates = {}
    for s, o in zip(SNRS, outs):
        cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
        henc.count_errors(bits, o, _L.I8, cnt)
        be, ce = (int(v) for v in cnt.tolist())
        rates["%g" % s] = {"ber": be / (ncw * K_INFO), "cwer": ce / ncw}

    # end-to-end simulation loop: new codewords + LLRs + decode + count for all SNR
    # points, one graph replay per loop
    print("bench: simulation loop", file=sys.stderr, flush=True)
    del outs
    sim = Sim(henc, dec, ncw, list(SNRS), seed=SEED, use_graph=True)
    sim.run(1)
    torch.cuda.synchronize()
    sim.reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(args.steps)
    torch.cuda.synchronize()
    sim_wall = time.perf_counter() - t0
    sim_cw_s = ncw * len(SNRS) * args.steps / sim_wall

    res = {
        "metric": METRIC, "value": gbit, "unit": "info Gbit/s", "cw_per_s": cw_s,
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": gbit / 30.2,
        "vs_baseline_ref": "README.md:5 '90 million codewords/s' dual-socket Ice Lake = 30.2 info Gbit/s",
        "dtype": "int8", "data": "synthetic",
        "config": {"workload": "11ad 1/2 (N=672,K=336,Z=42) int8 offset min-sum decode, beta 1, cap 31, max 10 "
                               "iterations with syndrome exit, SNR -1.00..0.00 dB step 0.25, %d cw per SNR" % ncw,
                   "codewords_per_step": cw_per_step,
                   "input": "random codewords (seed %d, library encoder) + AWGN, int8 LLRs (scale 2, max 24)" % SEED},
        "mean_iterations": mean_iters,
        "error_rates": rates,
        "sim_cw_per_s": sim_cw_s,
        "sim_info_gbit_s": sim_cw_s * K_INFO / 1e9,
        "sim_note": "encode + WGN/LLR + decode + count per codeword and SNR point, HIP graph per loop",
        # the binding roof: VALU, algorithmic int ops at the packed 16-bit rate
        "roofline": {"bound": "valu", "achieved": valu_tops, "peak": VALU_PK16_PEAK_TOPS, "unit": "Tops/s",
                     "frac": valu_tops / VALU_PK16_PEAK_TOPS,
                     "traffic": PAIR_TRAFFIC_BYTES_PER_CW * ncw if dec.kernel == _L.KERNEL_PAIR else None,
                     "kernel": kname, "avg_launch_ms": avg_launch_ms,
                     "ops_per_edge_update": OPS_PER_EDGE_UPDATE, "edge_updates_per_launch": edge_updates_per_launch,
                     "peak_note": "256 CU x 64 lanes x 2.4 GHz x 2 (packed 16-bit v_pk_* ops)",
                     "frac_of_32bit_peak": valu_tops / VALU_PEAK_TOPS},
        "roofline_hbm": {"bound": "hbm", "achieved": hbm_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": hbm_gbs / HBM_PEAK_GBS,
                         "traffic": PAIR_TRAFFIC_BYTES_PER_CW * ncw if dec.kernel == _L.KERNEL_PAIR else None,
                         "kernel": kname, "avg_launch_ms": avg_launch_ms,
                         "algorithmic_bytes_per_launch": hbm_bytes},
        # LDS bytes the row-pair kernel moves per edge-update (one row of one edge, one iteration):
        # its pair read, pair + twin write and syndrome read are 16 bytes per lane for two rows
        "roofline_lds": {"bound": "lds", "achieved": lds_tbs, "peak": LDS_PEAK_TBS, "unit": "TB/s",
                         "frac": lds_tbs / LDS_PEAK_TBS, "bytes_per_edge_update": LDS_BYTES_PER_EDGE_UPDATE,
                         "note": "algorithmic bytes; 46% of LDS-active cycles are bank conflicts (32-bank passes)"},
        "learned_offset_gain_db": None,
    }
    if not args.no_learned:
        # decode throughput with the committed learned offset table on the same batches
        from fecsim.train import load_offsets
        ltab, _ = load_offsets("11ad12_T10_int8")
        ldec = Decoder(Hdec.from_tables(cs.dec), "Vec64c_min", 10, 31.0, 1.0)
        ldec.set_offsets(ltab)
        lo = torch.empty_like(llrs[0])
        lit = torch.zeros(ncw, dtype=torch.int32, device="cuda")
        for x in llrs:
            ldec.decode(x, lo, iters=lit, stream=stream)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for x in llrs:
            ldec.decode(x, lo, iters=lit, stream=stream)
        torch.cuda.synchronize()
        lt = time.perf_counter() - t0
        res["learned_table_decode"] = {
            "cw_per_s": ncw * len(SNRS) / lt, "info_gbit_s": ncw * len(SNRS) / lt * K_INFO / 1e9,
            "kernel": "k_pair_i8c<PairShape0> with the offset table" if ldec.kernel == _L.KERNEL_PAIR else "generic",
            "note": "one pass over the five SNR batches with the learned per-(iteration, layer) offsets"}
        del lo
        print("bench: learned offsets", file=sys.stderr, flush=True)
        lg = learned_gain(cs, henc, args.learned_ncw)
        res["learned_offset_gain_db"] = lg["gain_db"]
        res["learned_offsets"] = lg
        # the same comparison on 5G NR base graph 1 (the large code of the workloads, 20
        # iterations), whose rows range from weight 3 to 19
        print("bench: learned offsets, 5G NR BG1", file=sys.stderr, flush=True)
        from fecsim.benchmarks import bg1_learned_gain
        lb = bg1_learned_gain()
        lb.pop("_decoder", None)
        res["learned_offset_gain_db_bg1"] = lb["gain_db"]
        res["learned_offsets_bg1"] = lb
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        sample = [llrs[i][:100_000].cpu().numpy() for i in range(len(SNRS))]
        cb = cpu_baseline(cs, sample, args.cpu_seconds)
        res["cpu_baseline"] = cb
        res["gpu_over_cpu"] = cw_s / cb["value"]
    print(json.dumps(res))


if __name__ == "__main__":
    main()
