"""LDS bank-conflict layouts of the row-pair int8 decoder with compiled edge tables
(k_pair_i8c in csrc/pair_i8.hip).

A wave decodes 64 // L codewords ("slots"), L = Z/2.  Lane (slot, lr) handles rows lr
and lr + L of every layer; for an edge of shift s it reads the pair at position
(s + lr) mod L of its slot's column, and it writes that pair and its twin (256 bytes
away, in the same bank).  Per iteration every edge is read twice (layer and syndrome)
and written twice, all with the same bank pattern.

Bank model (on MI355X the run-time-table kernel's LDS-active cycles per LDS instruction,
3.93 with its refill and write-out included, are within 3% of the model's 3.81 for its
decode accesses): a wave's 32-bit LDS access is served in two
passes of 32 lanes (lanes 0-31, then 32-63) over 32 banks of 4 bytes; a pass takes as
many cycles as the most distinct dwords one bank is asked for (equal addresses are a
broadcast).  With the slots in lane order and the pair positions in slot order (the
run-time-table kernel's layout) 802.11ad 1/2 costs 3.81 cycles per access; 2 is the
floor.

Both the lane of each (slot, lr) and the dword of each (slot, position) are free: the
kernel takes its class addresses from a per-lane table.  `cost` scores a layout for a
code's shift classes; `search` anneals one (lds_anneal.c, compiled with gcc); the
layouts found are committed in pair_lds_layouts.json and compiled into the shape
header by gen_pair_shapes.py.

    python -m fecsim.lds_layout [--steps N] [--seeds a,b,...]   (from simd-fecfun_amd/)
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
LAYOUTS = os.path.join(HERE, "pair_lds_layouts.json")


def class_multiplicity(shifts, z):
    """{s mod Z/2: number of edges} of an edge table"""
    half = z // 2
    m = {}
    for s in shifts:
        m[s % half] = m.get(s % half, 0) + 1
    return m


def identity_layout(half):
    """slots in lane order, positions in slot order (the run-time-table kernel's layout):
    (lane_item[64], dword[items]); lane_item -1 marks a spare lane"""
    ni = (64 // half) * half
    return [x if x < ni else -1 for x in range(64)], list(range(ni))


def cost(lane_item, dword, mult, half):
    """weighted LDS cycles of one access per edge: sum over shift classes of
    multiplicity x (cycles of the first half-wave + cycles of the second)"""
    ni = (64 // half) * half
    items = sorted(x for x in lane_item if x >= 0)
    assert items == list(range(ni)), "every lane item once"
    assert len(dword) == ni and len(set(dword)) == ni and all(0 <= d < 64 for d in dword)
    total = 0
    for c, m in mult.items():
        for lo in (0, 32):
            cnt = {}
            for lane in range(lo, lo + 32):
                it = lane_item[lane]
                if it < 0:
                    continue   # spare lane: mirrors a lane of its own half (broadcast)
                sl, lr = divmod(it, half)
                b = dword[sl * half + (c + lr) % half] % 32
                cnt[b] = cnt.get(b, 0) + 1
            total += m * max(cnt.values())
    return total


def _annealer():
    exe = os.path.join(tempfile.gettempdir(), "fecsim_lds_anneal")
    src = os.path.join(HERE, "lds_anneal.c")
    if not os.path.exists(exe) or os.path.getmtime(exe) < os.path.getmtime(src):
        subprocess.run(["gcc", "-O2", "-o", exe, src, "-lm"], check=True)
    return exe


def search(mult, half, seeds=(1, 2, 3, 4), steps=4_000_000):
    """best layout over annealing runs: (weighted cycles, lane_item[64], dword[items])"""
    exe = _annealer()
    inp = "%d\n%s\n" % (half, " ".join("%d:%d" % kv for kv in sorted(mult.items())))
    procs = [subprocess.Popen([exe, str(sd), str(steps)], stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
             for sd in seeds]
    for p in procs:   # every run gets its input before any is waited for: they run in parallel
        p.stdin.write(inp)
        p.stdin.close()
    best = None
    for p in procs:
        out = p.stdout.read()
        p.wait()
        lines = dict(ln.split(" ", 1) for ln in out.strip().splitlines())
        second = [int(v) for v in lines["pass"].split()]
        dword = [int(v) for v in lines["dword"].split()]
        first_items = [i for i, h in enumerate(second) if h == 0]
        second_items = [i for i, h in enumerate(second) if h == 1]
        lane_item = first_items + [-1] * (32 - len(first_items)) + second_items
        lane_item += [-1] * (64 - len(lane_item))
        c = cost(lane_item, dword, mult, half)
        assert c == int(lines["cost"])
        if best is None or c < best[0]:
            best = (c, lane_item, dword)
    return best


def load():
    """{matrix set: {"lane_item", "dword", "cycles_per_access", ...}} of the committed file"""
    if not os.path.exists(LAYOUTS):
        return {}
    with open(LAYOUTS) as f:
        return json.load(f)


def main(argv=None):
    import argparse
    from . import gen_pair_shapes as g
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--steps", type=int, default=12_000_000)
    ap.add_argument("--seeds", default="11,12,13,14")
    args = ap.parse_args(argv)
    seeds = [int(x) for x in args.seeds.split(",")]
    out = load()
    for w, ms, (zc, mc) in g.shapes():
        if not zc:
            continue
        col, sft = g.edge_table(ms)
        half = zc // 2
        mult = class_multiplicity(sft, zc)
        li0, dw0 = identity_layout(half)
        base = cost(li0, dw0, mult, half)
        c, lane_item, dword = search(mult, half, seeds, args.steps)
        prev = out.get(ms)
        if prev and prev["weighted_cycles"] <= c:
            print("%s: kept %.3f cycles per access (search %.3f)" % (ms, prev["cycles_per_access"], c / len(sft)))
            continue
        out[ms] = {"lane_item": lane_item, "dword": dword, "weighted_cycles": c,
                   "cycles_per_access": round(c / len(sft), 4), "identity_cycles_per_access": round(base / len(sft), 4),
                   "search": {"seeds": seeds, "steps": args.steps}}
        print("%s: %.3f cycles per access (identity layout %.3f, floor 2)" % (ms, c / len(sft), base / len(sft)))
    with open(LAYOUTS, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main(sys.argv[1:])
