"""Exclusive min / product tree schedules.

Restates the schedule construction of the reference's code generator
src/LDPCDec/min_find.pl:24-164: inputs are paired up level by level (levels sorted by
how many inputs a node covers), the two top nodes are clamped to the row cap
(min(Max_val, .), :99-107), then nodes are combined back down the levels until every
"all inputs but one" node exists (:111-163).  Output j is the node covering every
input except j.  For the min-sum decoder the result equals min(cap, min_{k!=j} A_k)
whatever the order; for the sum-product decoder the multiplication order and the cap
position decide the float rounding, so the HIP kernels evaluate exactly this schedule.

`emit_hip_header` writes the straight-line device functions the sum-product kernels
call (one per row weight), as the reference's Makefile:45-49 generates C++ from Perl.
"""
from __future__ import annotations

CAP = -1


def tree_schedule(n: int):
    """(ops, out) for row weight n: ops = [(dst, a, b)] with a == CAP meaning
    dst = min(cap, dst); out[j] = mask of the node that is output j."""
    if n < 2:
        raise ValueError("row weight must be >= 2")
    mask = [[1 << i for i in range(n)]]
    cnt = [[1] * n]
    ops = []
    lvl = 1
    while True:
        pm, pc = mask[lvl - 1], cnt[lvl - 1]
        n_prev = len(pm)
        half = n_prev >> 1
        cm, cc = [], []
        for i in range(half):
            m = pm[2 * i] | pm[2 * i + 1]
            cm.append(m)
            cc.append(pc[2 * i] + pc[2 * i + 1])
            if n != 2:
                ops.append((m, pm[2 * i], pm[2 * i + 1]))
        if n_prev & 1:
            half += 1
            cm.append(pm[-1])
            cc.append(pc[-1])
        order = sorted(range(len(cm)), key=lambda k: cc[k])  # stable, like Perl's mergesort
        mask.append([cm[k] for k in order])
        cnt.append([cc[k] for k in order])
        lvl += 1
        if not half > 2:
            break
    lvl -= 1
    comb = lvl - 1
    curm, curc = list(mask[lvl]), list(cnt[lvl])
    for i in range(2):
        t = (i + 1) if n == 2 else curm[i]
        ops.append((t, CAP, CAP))
    end_cnt = 0
    allkey = set()
    while True:
        nextm, nextc = [], []
        this_pass = set()
        prevm, prevc = mask[comb], cnt[comb]   # comb == -1 reads the last level, as Perl does
        for i in range(len(curm)):
            for j in range(len(prevm)):
                t1, t2 = prevc[j], curc[i]
                m1, m2 = prevm[j], curm[i]
                m = m1 | m2
                if m1 & m2 or t1 + t2 == n or m in allkey:
                    continue
                if t1 + t2 == n - 1:
                    end_cnt += 1
                ops.append((m, m1, m2))
                allkey.add(m)
                this_pass.add(m2)
                if t1 + t2 != n - 1:
                    nextm.append(m)
                    nextc.append(t1 + t2)
        for i in range(len(curm)):
            if curm[i] not in this_pass:
                nextm.append(curm[i])
                nextc.append(curc[i])
        comb -= 1
        curm, curc = nextm, nextc
        if end_cnt >= n or comb < -1:
            break
    full = (1 << n) - 1
    out = [full ^ (1 << j) for j in range(n)]
    return ops, out


def check_schedule(n: int, ops, out) -> None:
    """Every output covers all inputs but one and passes through a capped node."""
    defined = {1 << i: False for i in range(n)}  # mask -> carries a cap
    for d, a, b in ops:
        if a == CAP:
            assert d in defined
            defined[d] = True
        else:
            assert a in defined and b in defined and not (a & b) and d == a | b
            defined[d] = defined[a] or defined[b]
    full = (1 << n) - 1
    for j, m in enumerate(out):
        assert m == full ^ (1 << j) and defined[m], (n, j)


def emit_hip_header(max_w: int = 32) -> str:
    """Device functions `prod_tree_<w>(const F *t, F cap, F *o)` (F = float or double):
    o[j] = product over k != j of t[k] in the reference's order, with the cap clamps
    min(cap, m) = (cap < m ? cap : m), the operand order of vectorclass min()."""
    lines = ["// Generated by fecsim/minprod_tree.py (restating src/LDPCDec/min_find.pl). Do not edit.",
             "#pragma once", "",
             "template <typename F> __device__ __forceinline__ F vmin_ref(F a, F b) { return a < b ? a : b; }", ""]
    for n in range(2, max_w + 1):
        ops, out = tree_schedule(n)
        lines.append("template <typename F> __device__ __forceinline__ void prod_tree_%d(const F* t, F cap, F* o) {" % n)
        for i in range(n):
            lines.append("  F m%x = t[%d];" % (1 << i, i))
        seen = {1 << i for i in range(n)}
        for d, a, b in ops:
            if a == CAP:
                lines.append("  m%x = vmin_ref(cap, m%x);" % (d, d))
            else:
                decl = "" if d in seen else "F "
                seen.add(d)
                lines.append("  %sm%x = m%x * m%x;" % (decl, d, a, b))
        for j, m in enumerate(out):
            lines.append("  o[%d] = m%x;" % (j, m))
        lines.append("}")
        lines.append("")
    lines.append("// run-time row weight -> straight-line schedule")
    lines.append("template <typename F> __device__ __forceinline__ void prod_tree_dyn(int w, const F* t, F cap, F* o) {")
    lines.append("  switch (w) {")
    for n in range(2, max_w + 1):
        lines.append("  case %d: prod_tree_%d<F>(t, cap, o); break;" % (n, n))
    lines.append("  default: break;")
    lines.append("  }")
    lines.append("}")
    return "\n".join(lines) + "\n"
