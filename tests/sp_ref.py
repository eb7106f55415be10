"""Independent layered sum-product decoder in numpy (test infrastructure).

A second restatement of the reference's float tanh/atanh decoder, written apart from
oracle/oracle.c so that the two can check each other:

  * tanh / atanh come from the platform libm through numpy, evaluated in float64; for
    the float32 decoder the result is rounded to float32 (a correctly rounded float
    tanh/atanh for all but a vanishing set of arguments), standing in for vectorclass's
    tanh/atanh (src/LDPCDec/min_find_wrap.pl:12-16), whose bits are not available here;
  * every other operation is done in the decoder's element type in the reference's
    order: the exclusive products follow the schedule that src/LDPCDec/min_find.pl
    prints (tests/golden/minprod_tree.json holds its output for weights 2..53), with
    the row cap applied where the generator applies it (min_find.pl:100-107);
  * the cap is decode_functions.cpp:28-32's: 1 - (CTYPE)max_row_metric when
    max_row_metric < 1e-6, converted to the element type;
  * one layer is decode_functions.cpp:39-102 (residual, |.| - beta clipped at 0, sign
    XOR, exclusive product, sign_combine, posterior update), with circulant rows
    addressed modulo Z; the syndrome is the XOR of posterior sign bits per check
    (calc_syndrome_row, decode_functions.cpp:108-137; -0.0 counts as negative), and a
    codeword stops after the first iteration whose syndrome is zero or at max_iter
    (LDPCdecode.cpp:172-197).

The decoder is vectorised over codewords; it is meant for batches of a few hundred
codewords at the sizes of the 802.16e rate-1/2 sweep.
"""
from __future__ import annotations

import json
import os

import numpy as np

TREE_FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "minprod_tree.json")
_TREES = None


def trees():
    global _TREES
    if _TREES is None:
        with open(TREE_FIXTURE) as fh:
            _TREES = {int(w): (s["ops"], s["out"]) for w, s in json.load(fh).items()}
    return _TREES


def libm_tanh(x, dtype):
    """tanh via float64 libm, rounded to dtype"""
    return np.tanh(np.asarray(x, np.float64)).astype(dtype)


def libm_atanh(x, dtype):
    with np.errstate(divide="ignore"):
        return np.arctanh(np.asarray(x, np.float64)).astype(dtype)


def row_cap(max_row_metric, dtype):
    """decode_functions.cpp:28-32 with CTYPE = dtype (max_row_metric is a float parameter)"""
    m = float(np.float32(max_row_metric))
    if m < 1e-6:
        return dtype(1.0 - float(dtype(m)))
    return dtype(m)


def exclusive_products(t, cap, dtype):
    """t: list of w arrays tanh(A_k / 2); returns the w exclusive products in the order
    min_find.pl's 'prod' functions evaluate them (cap applied at the marked nodes)."""
    w = len(t)
    ops, out = trees()[w]
    v = {1 << i: t[i] for i in range(w)}
    for d, a, b in ops:
        if a < 0:
            x = v[d]
            v[d] = np.where(cap < x, cap, x).astype(dtype)  # min(Max_val, m): _mm_min_ps(cap, m)
        else:
            v[d] = (v[a] * v[b]).astype(dtype)
    return [v[o] for o in out]


def _signbit(x):
    return np.signbit(x)


def decode_sp(dec, llr, max_iter=10, max_row_metric=6e-8, beta=0.0, dtype=np.float32):
    """Decode llr [ncw, mcol*z] (dense layout).  Returns (posteriors, iterations, peak),
    peak[i] = the largest |check-to-variable message| codeword i produced."""
    dtype = np.dtype(dtype).type
    Z, mrow = int(dec.z), int(dec.mrow)
    rw = np.asarray(dec.row_weights, np.int64)
    cols = np.asarray(dec.row_active_cols, np.int64)
    offs = np.asarray(dec.row_offset, np.int64)
    start = np.concatenate([[0], np.cumsum(rw)])
    cap = row_cap(max_row_metric, dtype)
    b = dtype(beta)
    half, two, zero = dtype(0.5), dtype(2.0), dtype(0.0)
    r = np.arange(Z)
    idx = [[cols[e] * Z + (offs[e] + r) % Z for e in range(start[l], start[l + 1])] for l in range(mrow)]

    L = np.array(llr, dtype=dtype, copy=True)
    ncw = L.shape[0]
    R = np.zeros((ncw, int(dec.total_hc_ones), Z), dtype)
    iters = np.zeros(ncw, np.uint32)
    peak = np.zeros(ncw, np.float64)
    act = np.arange(ncw)
    it = 0
    while act.size:
        it += 1
        La, Ra = L[act], R[act]
        pk = np.zeros(act.size, np.float64)
        for l in range(mrow):
            es = range(start[l], start[l + 1])
            Q = [(La[:, idx[l][k]] - Ra[:, e]).astype(dtype) for k, e in enumerate(es)]
            S = np.zeros(Q[0].shape, bool)
            for q in Q:
                S ^= _signbit(q)
            A = [np.maximum((np.abs(q) - b).astype(dtype), zero) for q in Q]
            t = [libm_tanh((a * half).astype(dtype), dtype) for a in A]
            p = exclusive_products(t, cap, dtype)
            for k, e in enumerate(es):
                m = (two * libm_atanh(p[k], dtype)).astype(dtype)
                pk = np.maximum(pk, m.max(axis=1))
                neg = S ^ _signbit(Q[k])
                rn = np.where(neg, -m, m).astype(dtype)  # sign_combine: flips the sign bit (0 -> -0.0)
                La[:, idx[l][k]] = (Q[k] + rn).astype(dtype)
                Ra[:, e] = rn
        bad = np.zeros(act.size, bool)
        for l in range(mrow):
            par = np.zeros((act.size, Z), bool)
            for k in range(rw[l]):
                par ^= _signbit(La[:, idx[l][k]])
            bad |= par.any(axis=1)
        L[act], R[act] = La, Ra
        peak[act] = np.maximum(peak[act], pk)
        stop = ~bad | (it >= max_iter)
        iters[act[stop]] = it
        act = act[~stop]
    return L, iters, peak


# messages at most this large come from products p <= tanh(4) = 0.99933, where a change
# of one unit in the last place of tanh moves 2 atanh(p) by less than 1e-5 relative
WELL_CONDITIONED_PEAK = 8.0


def sp_parity(got, got_iters, want, want_iters, peak, rtol=1e-4, hard_floor=1e-3,
              peak_max=WELL_CONDITIONED_PEAK):
    """The float sum-product tolerance.

    * every codeword: the same iteration count, and the same hard decision (sign bit)
      wherever |want| >= hard_floor;
    * codewords whose messages all stayed <= peak_max in magnitude (well conditioned):
      |got - want| <= rtol * max(|want|, 1) for every posterior.

    Codewords with larger messages are not held to rtol: near the cap 1 - 6e-8 a
    product p sits within a few float32 units of 1, and 2 atanh(p) then moves by up to
    ln 2 for one unit of difference in a tanh -- two float32 implementations whose
    tanh differ in the last bit cannot agree to rtol there (the reference's own
    vectorclass tanh included).

    Returns a dict of counts and the worst relative error over well-conditioned
    codewords; 'ok' tells whether every condition holds."""
    g = np.asarray(got, np.float64)
    w = np.asarray(want, np.float64)
    it_ok = np.array_equal(np.asarray(got_iters, np.uint32), np.asarray(want_iters, np.uint32))
    big = np.abs(w) >= hard_floor
    flips = int(np.count_nonzero((np.signbit(g) != np.signbit(w)) & big))
    wc = np.asarray(peak) <= peak_max
    rel = np.abs(g - w) / np.maximum(np.abs(w), 1.0)
    worst = float(rel[wc].max()) if wc.any() else 0.0
    return dict(ok=bool(it_ok and flips == 0 and worst <= rtol), iters_equal=bool(it_ok), flips=flips,
                well_conditioned=int(wc.sum()), codewords=int(wc.size), worst_rel=worst)
