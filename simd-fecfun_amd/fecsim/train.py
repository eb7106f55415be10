"""Training harness for the differentiable layered min-sum decoder.

The decoder itself (forward and hand-written adjoint) runs in the HIP library
through the C ABI (fecsim_train_fwd / fecsim_train_bwd); this module wraps it in a
torch.autograd.Function and holds the training loop, the quantization of learned
offsets into an integer decoder's offset table, and the CWER-vs-SNR comparison that
measures the gain of the learned offsets.  No counterpart in the reference.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import lib as _L
from .decoder import Decoder, Hdec, _ptr


class TrainParams(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("max_batch", C.c_uint32), ("cap", C.c_float), ("elem", C.c_int),
                ("flags", C.c_uint32)]


GLOBAL_C2V = 1   # keep the C2V messages in a global array (the kernels for any shape)


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Trainer:
    """Device state of the unrolled decoder: checkpoints for up to max_batch codewords.
    int8=True: float32 emulation of the int8 decoder (integer-valued LLRs in the
    decoder's units, saturation to [-128, 127], offsets rounded with a straight-through
    gradient).  global_c2v=True keeps the C2V messages in a global array even where the
    code's shape has the register kernels (same results bit for bit, slower)."""

    def __init__(self, hdec: Hdec, iters: int, max_batch: int, cap: float = 1e30, double: bool = False,
                 int8: bool = False, global_c2v: bool = False):
        self.hdec, self.iters, self.max_batch, self.cap = hdec, iters, max_batch, cap
        self.elem = _L.I8 if int8 else (_L.F64 if double else _L.F32)
        p = TrainParams(iters, max_batch, cap, self.elem, GLOBAL_C2V if global_c2v else 0)
        t = C.c_void_p()
        _L.check(_L.lib().fecsim_trainer_create(hdec._h, C.byref(p), C.byref(t)), "fecsim_trainer_create")
        self._t = t

    @property
    def torch_dtype(self):
        import torch
        return torch.float64 if self.elem == _L.F64 else torch.float32

    def _check(self, *ts):
        for t in ts:
            if t.dtype != self.torch_dtype or not t.is_cuda or not t.is_contiguous():
                raise TypeError("tensors must be contiguous CUDA %s" % self.torch_dtype)

    def forward(self, llr, alpha, beta):
        import torch
        self._check(llr, alpha, beta)
        post = torch.empty_like(llr)
        _L.check(_L.lib().fecsim_train_fwd(self._t, llr.shape[0], _ptr(llr), _ptr(alpha), _ptr(beta), _ptr(post),
                                           _stream()), "fecsim_train_fwd")
        return post

    def backward(self, grad_post, want_llr=True):
        import torch
        grad_post = grad_post.contiguous()
        self._check(grad_post)
        ga = torch.empty((self.iters, self.hdec.info["mrow"]), dtype=self.torch_dtype, device="cuda")
        gb = torch.empty_like(ga)
        gl = torch.empty_like(grad_post) if want_llr else None
        _L.check(_L.lib().fecsim_train_bwd(self._t, grad_post.shape[0], _ptr(grad_post), _ptr(ga), _ptr(gb), _ptr(gl),
                                           _stream()), "fecsim_train_bwd")
        return gl, ga, gb

    def __del__(self):
        if getattr(self, "_t", None) and _L._lib is not None:
            _L.lib().fecsim_trainer_free(self._t)
            self._t = None


def min_sum_function():
    """torch.autograd.Function: post = decoder(llr; alpha, beta) on the HIP library."""
    import torch

    class MinSum(torch.autograd.Function):
        @staticmethod
        def forward(ctx, llr, alpha, beta, trainer):
            ctx.trainer = trainer
            ctx.want_llr = llr.requires_grad
            return trainer.forward(llr.contiguous(), alpha.contiguous(), beta.contiguous())

        @staticmethod
        def backward(ctx, grad_post):
            gl, ga, gb = ctx.trainer.backward(grad_post, want_llr=ctx.want_llr)
            return gl, ga, gb, None

    return MinSum


def rate_db(k, n):
    """10 log10(K/N): Es/N0 = Eb/N0 + rate_db for BPSK."""
    return 10.0 * math.log10(k / n)


def make_batch(henc, batch, ebn0_lo, ebn0_hi, rate, seed, cw0, rng, groups=8, scale=1.0, max_abs=1e9):
    """Random codewords and float LLRs; the batch is split into `groups` slices, each at
    an Eb/N0 drawn uniformly from [ebn0_lo, ebn0_hi].  Returns (llr, bits01)."""
    import torch
    bits = henc.new_bits(batch)
    henc.gen_codewords(bits, seed=seed, cw0=cw0)
    llr = torch.empty((batch, henc.n), dtype=torch.float32, device="cuda")
    edges = np.linspace(0, batch, groups + 1).astype(int)
    for g in range(groups):
        a, b = edges[g], edges[g + 1]
        if b <= a:
            continue
        snr = float(rng.uniform(ebn0_lo, ebn0_hi) + rate)
        henc.gen_llr(bits[a:b], llr[a:b], _L.F32, snr, seed=seed, cw0=cw0 + a, llr_scale=scale, max_abs_llr=max_abs)
    z, mcol = henc.z, henc.mcol
    shifts = torch.arange(32, device="cuda", dtype=torch.int32)
    b01 = ((bits.unsqueeze(-1) >> shifts) & 1).reshape(batch, mcol, -1)[:, :, :z].reshape(batch, mcol * z)
    return llr, b01.to(torch.float32)


def train(cs, henc, T=10, batch=4096, steps=400, ebn0=(1.0, 3.0), lr=1e-3, init_offset=0.5, init_scale=0.8,
          learn_scale=True, cap=1e30, seed=7, log_every=50, log=None):
    """Adam on the BCE of the posteriors (after T iterations) against the transmitted
    bits.  Returns (alpha [T, mrow], beta [T, mrow], losses)."""
    import torch
    import torch.nn.functional as Fn
    hdec = Hdec.from_tables(cs.dec)
    tr = Trainer(hdec, T, batch, cap=cap)
    Lr = cs.dec.mrow
    alpha = torch.full((T, Lr), init_scale if learn_scale else 1.0, device="cuda", requires_grad=learn_scale)
    beta = torch.full((T, Lr), init_offset, device="cuda", requires_grad=True)
    opt = torch.optim.Adam([p for p in (alpha, beta) if p.requires_grad], lr=lr)
    fn = min_sum_function()
    rng = np.random.default_rng(seed)
    rate = rate_db(cs.dec.mcol * cs.z - cs.enc.mrow * cs.z, cs.dec.mcol * cs.z)
    losses = []
    for step in range(steps):
        llr, b01 = make_batch(henc, batch, ebn0[0], ebn0[1], rate, seed, step * batch, rng)
        post = fn.apply(llr, alpha if learn_scale else alpha.detach(), beta, tr)
        loss = Fn.binary_cross_entropy_with_logits(-post, b01)
        opt.zero_grad()
        loss.backward()
        opt.step()
        with torch.no_grad():
            beta.clamp_(min=0.0)
            if learn_scale:
                alpha.clamp_(0.05, 1.5)
        losses.append(float(loss.item()))
        if log and (step % log_every == 0 or step == steps - 1):
            log("step %d loss %.5f beta[0] %.3f..%.3f alpha[0] %.3f" % (
                step, losses[-1], float(beta.detach()[0].min()), float(beta.detach()[0].max()),
                float(alpha.detach()[0].mean())))
    return alpha.detach().cpu().numpy(), beta.detach().cpu().numpy(), losses


def _int8_batch(henc, batch, ebn0, rate, seed, cw0, rng, llr_scale, max_abs):
    """Random codewords and the int8 decoder's LLRs (as float32 integers); the batch is
    split into 8 slices, each at an Eb/N0 drawn uniformly from ebn0.  Returns (llr, bits01)."""
    import torch
    bits = henc.new_bits(batch)
    henc.gen_codewords(bits, seed=seed, cw0=cw0)
    q = torch.empty((batch, henc.n), dtype=torch.int8, device="cuda")
    edges = np.linspace(0, batch, 9).astype(int)
    for g in range(8):
        a, b = edges[g], edges[g + 1]
        if b <= a:
            continue
        snr = float(rng.uniform(ebn0[0], ebn0[1]) + rate)
        henc.gen_llr(bits[a:b], q[a:b], _L.I8, snr, seed=seed, cw0=cw0 + a, llr_scale=llr_scale, max_abs_llr=max_abs)
    z, mcol = henc.z, henc.mcol
    shifts = torch.arange(32, device="cuda", dtype=torch.int32)
    b01 = ((bits.unsqueeze(-1) >> shifts) & 1).reshape(batch, mcol, -1)[:, :, :z].reshape(batch, mcol * z)
    return q.to(torch.float32), b01.to(torch.float32)


def train_int8(cs, henc, T=10, batch=512, steps=400, ebn0=(0.5, 1.0), lr=2e-2, init=None, seed=7, log_every=50,
               log=None, llr_scale=2.0, max_abs=24.0, cap=31.0, rate=None, greedy=False, loss_iters=None):
    """Offsets for the int8 decoder (Vec64c_min, LLR scale 2, |LLR| <= 24, cap 31) trained
    through the int8-emulating forward: the LLRs are the int8 decoder's own (integers),
    the decoder saturates as the int8 one does and rounds the offsets (straight-through
    gradient), so the table that training sees is the table the bit-exact decoder runs.
    Adam on the BCE of the posteriors (in LLR units, divided by llr_scale) against the
    transmitted bits, offsets clamped to [0, 20].

    greedy=False: all T iterations' offsets trained together on the loss after T
    iterations.  greedy=True: iteration by iteration -- for t = 1..T, row t-1 of the
    table (started from row t-2) is trained for steps // T steps on the loss after t
    iterations, the earlier rows held fixed; every iteration's offsets then see a loss
    whose posteriors are not yet saturated.  loss_iters (joint training only): the loss is
    the mean of the BCE after each of these iteration counts (default [T]), so early
    iterations' offsets also get gradient from posteriors that are not yet saturated.
    Returns (integer offsets [T, mrow], losses)."""
    import torch
    import torch.nn.functional as Fn
    hdec = Hdec.from_tables(cs.dec)
    Lr = cs.dec.mrow
    tab = np.ones((T, Lr), np.float32) if init is None else np.array(init, np.float32)
    fn = min_sum_function()
    rng = np.random.default_rng(seed)
    if rate is None:   # Es/N0 = Eb/N0 + rate (dB); pass the transmitted rate for punctured codes
        rate = rate_db(cs.dec.mcol * cs.z - cs.enc.mrow * cs.z, cs.dec.mcol * cs.z)
    losses = []
    stages = [(t, t - 1, max(1, steps // T)) for t in range(1, T + 1)] if greedy else [(T, None, steps)]
    step = 0
    for t, row, n in stages:
        ks = [t] if greedy or not loss_iters else sorted(set(int(k) for k in loss_iters) | {T})
        if ks[0] < 1 or ks[-1] > T:
            raise ValueError("loss_iters must lie in 1..T")
        trs = [(k, Trainer(hdec, k, batch, cap=cap, int8=True)) for k in ks]
        if row is not None and row > 0 and init is None:
            tab[row] = tab[row - 1]
        lo = 0 if row is None else row
        fixed = torch.tensor(tab[:lo], device="cuda")
        beta = torch.tensor(tab[lo:t], device="cuda", requires_grad=True)
        alpha = torch.ones((t, Lr), device="cuda")
        opt = torch.optim.Adam([beta], lr=lr)
        for _ in range(n):
            q, b01 = _int8_batch(henc, batch, ebn0, rate, seed, step * batch, rng, llr_scale, max_abs)
            full = torch.cat([fixed, beta])
            loss = sum(Fn.binary_cross_entropy_with_logits(-fn.apply(q, alpha[:k], full[:k], tr) / llr_scale, b01)
                       for k, tr in trs) / len(trs)
            opt.zero_grad()
            loss.backward()
            opt.step()
            with torch.no_grad():
                beta.clamp_(min=0.0, max=20.0)
            losses.append(float(loss.item()))
            if log and (step % log_every == 0 or step == steps - 1):
                r = np.rint(beta.detach().cpu().numpy()).astype(int)
                log("step %d iters %d loss %.5f offsets %s" % (step, t, losses[-1], np.bincount(r.ravel())))
            step += 1
        tab[lo:t] = beta.detach().cpu().numpy()
        del trs
    return np.rint(tab).astype(np.float32), losses


def quantize_offsets(beta, llr_scale=2.0, max_iter=None):
    """Learned offsets (true LLR units) -> integer offsets of a decoder fed LLRs scaled by
    llr_scale: round(llr_scale * beta), repeated for iterations past T."""
    q = np.rint(np.asarray(beta, np.float64) * llr_scale).astype(np.float32)
    if max_iter is not None and q.shape[0] < max_iter:
        q = np.concatenate([q, np.repeat(q[-1:], max_iter - q.shape[0], 0)], 0)
    return q


def refine_offsets(cs, henc, table, snr, ncw=2_000_000, sweeps=2, seed=4242, max_iter=10, log=None):
    """Greedy +-1 coordinate descent on an integer offset table of the int8 decoder
    (LLR scale 2, |LLR| <= 24, cap 31): every candidate is scored by its codeword
    errors on the same pre-generated batch at `snr` (paired comparison), a change is
    kept when it removes errors.  Returns (table, errors before, errors after)."""
    import torch
    hdec = Hdec.from_tables(cs.dec)
    dec = Decoder(hdec, "Vec64c_min", max_iter, 31.0, 1.0)
    bits = henc.new_bits(ncw)
    henc.gen_codewords(bits, seed=seed, cw0=0)
    llr = torch.empty((ncw, henc.n), dtype=torch.int8, device="cuda")
    henc.gen_llr(bits, llr, _L.I8, float(snr), seed=seed, llr_scale=2.0, max_abs_llr=24.0)
    post = torch.empty_like(llr)
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")

    def errors(tab):
        dec.set_offsets(tab)
        dec.decode(llr, post)
        cnt.zero_()
        henc.count_errors(bits, post, _L.I8, cnt)
        return int(cnt[1].item())

    tab = np.array(table, np.float32).copy()
    first = best = errors(tab)
    for sw in range(sweeps):
        changed = 0
        for t in range(tab.shape[0]):
            for l in range(tab.shape[1]):
                for d in (-1.0, 1.0):
                    v = tab[t, l] + d
                    if v < 0 or v > 6:
                        continue
                    old = tab[t, l]
                    tab[t, l] = v
                    e = errors(tab)
                    if e < best:
                        best = e
                        changed += 1
                        break
                    tab[t, l] = old
        if log:
            log("refine sweep %d: %d codeword errors (start %d), %d entries changed" % (sw, best, first, changed))
        if not changed:
            break
    return tab, first, best


def search_offsets(cs, henc, table, snr, groups, ncw=262144, sweeps=2, seed=4242, max_iter=20, vmin=1, vmax=6,
                   accept=0.98, log=None):
    """Derivative-free search for the int8 decoder's offset table (LLR scale 2, |LLR| <= 24,
    cap 31): the coordinates are (iteration group, layer) -- `groups` are iteration
    boundaries [0, g1, ..., max_iter] and every iteration of a group shares the layer's
    offset.  Each coordinate tries +1, -1, +2 in turn (within [vmin, vmax]) and keeps
    the first change that leaves fewer than `accept` times the codeword errors of the
    current table on one pre-generated batch at `snr` (paired comparison).  vmin = 1
    keeps tables that the code-specialized kernels take.  Returns (table, errors before,
    errors after, the same two counts on a second batch of other codewords and noise)."""
    import torch
    hdec = Hdec.from_tables(cs.dec)
    dec = Decoder(hdec, "Vec64c_min", max_iter, 31.0, 1.0)

    def batch(sd):
        bits = henc.new_bits(ncw)
        henc.gen_codewords(bits, seed=sd, cw0=0)
        llr = torch.empty((ncw, henc.n), dtype=torch.int8, device="cuda")
        henc.gen_llr(bits, llr, _L.I8, float(snr), seed=sd, llr_scale=2.0, max_abs_llr=24.0)
        return bits, llr

    b_fit, b_val = batch(seed), batch(seed + 1)
    post = torch.empty_like(b_fit[1])
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")

    def errors(tab, b):
        dec.set_offsets(tab)
        dec.decode(b[1], post)
        cnt.zero_()
        henc.count_errors(b[0], post, _L.I8, cnt)
        return int(cnt[1].item())

    bounds = list(groups)
    if bounds[0] != 0 or bounds[-1] != max_iter or any(a >= b for a, b in zip(bounds, bounds[1:])):
        raise ValueError("groups must rise from 0 to max_iter: %r" % (groups,))
    tab = np.array(table, np.float32).copy()
    gt = np.stack([tab[a] for a in bounds[:-1]])    # one row of offsets per group

    def expand(g):
        t = np.empty((max_iter, g.shape[1]), np.float32)
        for i in range(len(bounds) - 1):
            t[bounds[i]:bounds[i + 1]] = g[i]
        return t

    first = best = errors(expand(gt), b_fit)
    val0 = errors(expand(gt), b_val)
    if log:
        log("search start: %d codeword errors (validation batch %d)" % (first, val0))
    for sw in range(sweeps):
        changed = 0
        for g in range(gt.shape[0]):
            if log:
                log("search sweep %d, iterations %d..%d: %d errors" % (sw, bounds[g], bounds[g + 1] - 1, best))
            for l in range(gt.shape[1]):
                for d in (1.0, -1.0, 2.0):
                    v = gt[g, l] + d
                    if v < vmin or v > vmax:
                        continue
                    old = gt[g, l]
                    gt[g, l] = v
                    e = errors(expand(gt), b_fit)
                    if e < best * accept:
                        best = e
                        changed += 1
                        break
                    gt[g, l] = old
        if log:
            log("search sweep %d: %d codeword errors (start %d), %d coordinates changed" % (sw, best, first, changed))
        if not changed:
            break
    return expand(gt), first, best, val0, errors(expand(gt), b_val)


def cwer_curve(cs, henc, dec, snrs, ncw, seed=100, loops=1):
    """CWER / BER per SNR point of the device simulation loop."""
    from .sim import Sim
    sim = Sim(henc, dec, ncw, list(snrs), seed=seed, use_graph=False)
    sim.run(loops)
    st = sim.stats()
    k = (cs.dec.mcol - cs.enc.mrow) * cs.z
    return [(s, x["cw_err"] / x["cw"], x["bit_err"] / (x["cw"] * k), x["iters"] / x["cw"]) for s, x in zip(snrs, st)]


def load_offsets(name: str):
    """A committed integer offset table, fecsim/learned/<name>.json: ([max_iter][mrow]
    float32 array for Decoder.set_offsets, the full record with its provenance)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "learned", name + ".json")) as f:
        rec = json.load(f)
    return np.asarray(rec["beta_int"], np.float32), rec


def snr_at(curve, target=1e-3):
    """SNR where CWER crosses `target`, log-linear interpolation between the bracketing
    points; None when the curve does not bracket it."""
    pts = [(s, c) for s, c, *_ in curve]
    for (s0, c0), (s1, c1) in zip(pts, pts[1:]):
        if c0 >= target >= c1 and c0 > 0 and c1 > 0:
            t = (math.log(c0) - math.log(target)) / (math.log(c0) - math.log(c1))
            return s0 + t * (s1 - s0)
    return None


def eval_float(cs, henc, alpha, beta, snrs, ncw, T, cap=1e30, batch=4096, seed=1234):
    """CWER / BER of the unrolled float decoder (T iterations, given alpha/beta) per
    SNR point (Es/N0 dB), on LLRs of scale 1."""
    import torch
    hdec = Hdec.from_tables(cs.dec)
    tr = Trainer(hdec, T, batch, cap=cap)
    a = torch.tensor(np.asarray(alpha, np.float32), device="cuda")
    b = torch.tensor(np.asarray(beta, np.float32), device="cuda")
    k = (cs.dec.mcol - cs.enc.mrow) * cs.z
    out = []
    for snr in snrs:
        cwe = be = 0
        done = 0
        while done < ncw:
            n = min(batch, ncw - done)
            bits = henc.new_bits(n)
            henc.gen_codewords(bits, seed=seed, cw0=done)
            llr = torch.empty((n, henc.n), dtype=torch.float32, device="cuda")
            henc.gen_llr(bits, llr, _L.F32, float(snr), seed=seed, cw0=done, llr_scale=1.0, max_abs_llr=1e9)
            post = tr.forward(llr, a, b)
            cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
            henc.count_errors(bits, post, _L.F32, cnt)
            e = cnt.tolist()
            be += e[0]
            cwe += e[1]
            done += n
        out.append((float(snr), cwe / ncw, be / (ncw * k), float(T)))
    return out


def _main():
    import argparse
    import json
    import sys
    import time
    from . import codes
    from .sim import Henc
    ap = argparse.ArgumentParser(description="learn per-(iteration, layer) offsets/scales of the layered "
                                             "min-sum decoder and measure their SNR gain")
    ap.add_argument("--code", default="11ad 1/2")
    ap.add_argument("--T", type=int, default=10)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--steps", type=int, default=1500)
    ap.add_argument("--ebn0", type=float, nargs=2, default=(1.0, 3.0))
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-offset", type=float, default=0.5)
    ap.add_argument("--init-scale", type=float, default=0.8)
    ap.add_argument("--offset-only", action="store_true", help="keep the scales at 1 (for the integer decoder)")
    ap.add_argument("--cap", type=float, default=1e30, help="row cap in LLR units of scale 1")
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--eval-snr", type=float, nargs=3, default=None, metavar=("LO", "HI", "STEP"),
                    help="Es/N0 grid for the CWER comparison")
    ap.add_argument("--eval-ncw", type=int, default=2_000_000)
    ap.add_argument("--eval-loops", type=int, default=1, help="simulation loops of eval-ncw codewords per point")
    ap.add_argument("--refine-snr", type=float, default=None,
                    help="Es/N0 at which the quantized offset table is refined by coordinate descent")
    ap.add_argument("--refine-sweeps", type=int, default=2)
    ap.add_argument("--search-snr", type=float, default=None,
                    help="Es/N0 at which the int8 offset table is searched per (iteration group, layer)")
    ap.add_argument("--search-groups", default=None, help="iteration boundaries, e.g. 0,1,2,4,6,10,20")
    ap.add_argument("--search-ncw", type=int, default=262144)
    ap.add_argument("--search-sweeps", type=int, default=2)
    ap.add_argument("--search-from", default="trained",
                    help="start of the search: 'trained' (the quantized trained table), 'plain' (beta = 1 "
                         "everywhere) or the path of an earlier --out file (its integer table)")
    ap.add_argument("--int8-aware", action="store_true",
                    help="train the int8 decoder's integer offsets through its float emulation (implies --offset-only)")
    ap.add_argument("--greedy", action="store_true", help="int8-aware: train iteration by iteration")
    ap.add_argument("--loss-iters", default=None, help="int8-aware: iteration counts whose losses are averaged")
    ap.add_argument("--init-from", default=None,
                    help="int8-aware: start table, a committed table's name (fecsim/learned/<name>.json) or the "
                         "path of an earlier --out file (default: beta = 1 everywhere)")
    ap.add_argument("--rate-db", type=float, default=None,
                    help="int8-aware: 10 log10(K/N) of the transmitted code (Es/N0 = Eb/N0 + rate-db)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args()
    if args.int8_aware:
        args.offset_only = True
    log = lambda s: print(s, file=sys.stderr, flush=True)   # noqa: E731
    cs = codes.gen_structs(args.code, simd_size=512, elem_size=8)
    henc = Henc.from_tables(cs.enc)
    t0 = time.time()
    if args.steps > 0 and args.int8_aware:
        li = [int(x) for x in args.loss_iters.split(",")] if args.loss_iters else None
        init = None
        if args.init_from:
            if args.init_from.endswith(".json"):
                with open(args.init_from) as f:
                    init = np.asarray(json.load(f)["beta_int"], np.float32)
            else:
                init = load_offsets(args.init_from)[0]
            init = init[:args.T]
        tab, losses = train_int8(cs, henc, T=args.T, batch=args.batch, steps=args.steps, ebn0=tuple(args.ebn0),
                                 lr=args.lr, init=init, seed=args.seed, log=log, log_every=max(1, args.steps // 20),
                                 rate=args.rate_db, greedy=args.greedy, loss_iters=li)
        alpha = np.ones_like(tab)
        beta = tab / 2.0   # LLR units of scale 1: quantize_offsets gives back the trained integers
    elif args.steps > 0:
        alpha, beta, losses = train(cs, henc, T=args.T, batch=args.batch, steps=args.steps, ebn0=tuple(args.ebn0),
                                    lr=args.lr, init_offset=args.init_offset, init_scale=args.init_scale,
                                    learn_scale=not args.offset_only, cap=args.cap, seed=args.seed, log=log,
                                    log_every=max(1, args.steps // 20))
    else:   # no gradient training: the plain offset 0.5 (int8 beta 1) and scale 1
        alpha = np.ones((args.T, cs.dec.mrow), np.float32)
        beta = np.full((args.T, cs.dec.mrow), 0.5, np.float32)
        losses = []
    res = {"code": args.code, "T": args.T, "batch": args.batch, "steps": args.steps, "ebn0": list(args.ebn0),
           "lr": args.lr, "init_offset": args.init_offset, "init_scale": args.init_scale,
           "offset_only": args.offset_only, "int8_aware": args.int8_aware, "greedy": args.greedy,
           "loss_iters": args.loss_iters, "init_from": args.init_from, "rate_db": args.rate_db, "cap": args.cap, "seed": args.seed, "train_seconds": time.time() - t0,
           "loss_first": losses[:5], "loss_last": losses[-5:], "alpha": alpha.tolist(), "beta": beta.tolist()}
    if args.eval_snr:
        lo, hi, st = args.eval_snr
        snrs = [float(x) for x in np.arange(lo, hi + 1e-9, st)]
        if args.offset_only:
            # integer decoder (LLR scale 2, |LLR| <= 24, cap 31): plain beta = 1 vs learned table
            q = quantize_offsets(beta, 2.0, max_iter=args.T)
            res["beta_int_rounded"] = q.tolist()
            if args.search_snr is not None:
                if args.search_from == "plain":
                    start = np.ones_like(q)
                elif args.search_from == "trained":
                    start = np.maximum(q, 1.0)
                else:
                    with open(args.search_from) as f:
                        start = np.asarray(json.load(f)["beta_int"], np.float32)
                groups = [int(x) for x in args.search_groups.split(",")] if args.search_groups else [0, args.T]
                q, e0, e1, v0, v1 = search_offsets(cs, henc, start, args.search_snr, groups, ncw=args.search_ncw,
                                                   sweeps=args.search_sweeps, max_iter=args.T, log=log)
                res.update({"search_snr": args.search_snr, "search_groups": groups, "search_from": args.search_from,
                            "search_errors_before": e0, "search_errors_after": e1,
                            "search_validation_errors_before": v0, "search_validation_errors_after": v1})
            if args.refine_snr is not None:
                q, e0, e1 = refine_offsets(cs, henc, q, args.refine_snr, sweeps=args.refine_sweeps, max_iter=args.T,
                                           log=log)
                res.update({"refine_snr": args.refine_snr, "refine_errors_before": e0, "refine_errors_after": e1})
            res["beta_int"] = q.tolist()
            hdec = Hdec.from_tables(cs.dec)
            plain = Decoder(hdec, "Vec64c_min", args.T, 31.0, 1.0)
            learned = Decoder(hdec, "Vec64c_min", args.T, 31.0, 1.0)
            learned.set_offsets(q)
            cp = cwer_curve(cs, henc, plain, snrs, args.eval_ncw, loops=args.eval_loops)
            cl = cwer_curve(cs, henc, learned, snrs, args.eval_ncw, loops=args.eval_loops)
            kind = "int8 Vec64c_min, max %d iterations with syndrome exit" % args.T
        else:
            # float unrolled decoder: plain offset 0.5 / scale 1 vs learned
            ones = np.ones_like(alpha)
            cp = eval_float(cs, henc, ones, 0.5 * ones, snrs, args.eval_ncw, args.T, cap=args.cap)
            cl = eval_float(cs, henc, alpha, beta, snrs, args.eval_ncw, args.T, cap=args.cap)
            kind = "float32 unrolled, %d iterations" % args.T
        sp, sl = snr_at(cp), snr_at(cl)
        res.update({"eval_decoder": kind, "eval_ncw": args.eval_ncw * args.eval_loops, "curve_plain": cp,
                    "curve_learned": cl,
                    "snr_at_1e-3_plain": sp, "snr_at_1e-3_learned": sl,
                    "gain_db": (sp - sl) if sp is not None and sl is not None else None})
        log("gain at CWER 1e-3: %s dB" % res["gain_db"])
    with open(args.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps({k: res[k] for k in res if k not in ("alpha", "beta", "beta_int", "curve_plain",
                                                            "curve_learned")}))


if __name__ == "__main__":
    _main()
