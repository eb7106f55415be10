"""Writes tests/golden/minprod_tree.json: the exclusive min / product tree schedules
that the reference's code generator src/LDPCDec/min_find.pl emits for row weights
2..53 (min_find_wrap.pl:6).  Runs the generator in this container (perl is here);
the fixture holds only the parsed schedule (data), not the generated code.

Schedule format per weight w: "ops" = [[dst, a, b], ...] in emission order, where
dst/a/b are bit masks of the inputs a node covers and a == -1 marks the cap
"dst = min(Max_val, dst)"; "out"[j] = mask whose value becomes output j.
"""
import json
import os
import re
import subprocess
import sys

GEN = "/root/reference/src/LDPCDec/min_find.pl"


def parse(text, w):
    ops, out, leaves = [], [None] * w, [None] * w
    for line in text.splitlines():
        line = line.strip()
        m = re.match(r"SIMD_CLASS m([0-9a-f]+) = (?:tanh\( )?grp_min\[(\d+)\*SIMD_PER_Z \+ i\]", line)
        if m:
            leaves[int(m.group(2))] = int(m.group(1), 16)
            continue
        m = re.match(r"SIMD_CLASS m([0-9a-f]+) = (?:m([0-9a-f]+) \* m([0-9a-f]+)|min\( m([0-9a-f]+), m([0-9a-f]+) \));", line)
        if m:
            a = m.group(2) or m.group(4)
            b = m.group(3) or m.group(5)
            ops.append([int(m.group(1), 16), int(a, 16), int(b, 16)])
            continue
        m = re.match(r"m([0-9a-f]+) = min\( Max_val, m([0-9a-f]+) \);", line)
        if m:
            ops.append([int(m.group(1), 16), -1, -1])
            continue
        m = re.match(r"fin_min\[(\d+)\*SIMD_PER_Z \+ i\] = (?:2\.0\*atanh\( )?m([0-9a-f]+)", line)
        if m:
            out[int(m.group(1))] = int(m.group(2), 16)
    assert leaves == [1 << i for i in range(w)], (w, leaves)
    assert None not in out
    return {"ops": ops, "out": out}


def main():
    res = {}
    for w in range(2, 54):
        sched = {}
        for mode in ("min", "prod"):
            txt = subprocess.run(["perl", GEN, str(w), mode], check=True, capture_output=True, text=True).stdout
            sched[mode] = parse(txt, w)
        assert sched["min"] == sched["prod"], w
        res[str(w)] = sched["min"]
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "minprod_tree.json")
    with open(path, "w") as f:
        json.dump(res, f, separators=(",", ":"))
    print("wrote", path, file=sys.stderr)


if __name__ == "__main__":
    main()
