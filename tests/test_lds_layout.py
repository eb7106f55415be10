"""LDS layouts of the row-pair decoder with compiled edge tables (fecsim/lds_layout.py,
csrc/pair_i8.hip k_pair_i8c): the committed layouts are valid placements, their recorded
costs follow from the bank model, and the model itself is checked by hand on a small case
and on the lane-order layout whose cost was measured on the GPU."""
import pytest

from fecsim import gen_pair_shapes, lds_layout


def _shape_tables():
    out = []
    for w, ms, (zc, mc) in gen_pair_shapes.shapes():
        if zc:
            col, sft = gen_pair_shapes.edge_table(ms)
            out.append((ms, zc, col, sft))
    return out


def test_cost_model_by_hand():
    """L = 16 (Z = 32): four slots of 16 lanes fill the wave; lanes 0-31 are slots 0 and
    1.  With slots in lane order and positions in slot order, slot s position u sits at
    dword 16 s + u, bank (16 s + u) mod 32: slots 0 and 1 use banks 0-15 and 16-31 in
    the first half-wave, slots 2 and 3 the same banks in the second, for every shift:
    one cycle per half, 2 per access.  Putting every slot's positions at dwords 2 u + x
    (x = 0, 1 for slots 0/1 ... ) folds two slots onto 16 banks: 2 cycles per half."""
    li, dw = lds_layout.identity_layout(16)
    assert lds_layout.cost(li, dw, {0: 1, 5: 2}, 16) == 2 * 3
    # slots 0 and 1 share banks: slot s position u at dword 32 * (s % 2) + 2 u + s // 2
    dw2 = [32 * (s % 2) + 2 * u + s // 2 for s in range(4) for u in range(16)]
    # in the first half, slot 0 lanes use banks 2u (u = 0..15) and slot 1 lanes the same
    # banks 2u (dwords 32 + 2u): every bank twice -> 2 cycles, same in the second half
    assert lds_layout.cost(li, dw2, {0: 1}, 16) == 4


def test_cost_model_brute_force_small():
    """The cost of a random layout equals an independent count of distinct dwords per
    bank and half-wave."""
    import random
    rnd = random.Random(3)
    half = 21
    for _ in range(5):
        items = list(range(63))
        rnd.shuffle(items)
        li = items[:32] + items[32:] + [-1]
        dw = rnd.sample(range(64), 63)
        mult = {c: rnd.randint(1, 4) for c in rnd.sample(range(half), 6)}
        want = 0
        for c, m in mult.items():
            for lanes in (range(0, 32), range(32, 64)):
                per_bank = {}
                for ln in lanes:
                    if li[ln] < 0:
                        continue
                    s, lr = divmod(li[ln], half)
                    d = dw[s * half + (c + lr) % half]
                    per_bank.setdefault(d % 32, set()).add(d)
                want += m * max(len(v) for v in per_bank.values())
        assert lds_layout.cost(li, dw, mult, half) == want


def test_lane_order_layout_matches_measurement():
    """802.11ad 1/2 with slots in lane order: 3.81 LDS cycles per access.  The run-time-table
    kernel, which uses that layout, measured 3.93 LDS-active cycles per LDS instruction
    with rocprofv3 (SQ_LDS_IDX_ACTIVE 1.058e9 / SQ_INSTS_LDS 2.693e8 per launch, refill
    and write-out accesses included)"""
    ms, zc, col, sft = _shape_tables()[0]
    assert ms == "11ad 1/2"
    li, dw = lds_layout.identity_layout(zc // 2)
    c = lds_layout.cost(li, dw, lds_layout.class_multiplicity(sft, zc), zc // 2)
    assert c == 198 and abs(c / len(sft) - 3.808) < 1e-3


@pytest.mark.parametrize("tab", _shape_tables(), ids=lambda t: t[0])
def test_committed_layouts(tab):
    """Every fixed-Z shape has a committed layout; it places each (slot, lr) lane item once
    and each (slot, position) item at its own dword, its recorded cost follows from the
    bank model, and it beats the lane-order layout."""
    ms, zc, col, sft = tab
    lay = lds_layout.load()
    assert ms in lay, "no LDS layout committed for %s" % ms
    half = zc // 2
    li, dw = lay[ms]["lane_item"], lay[ms]["dword"]
    assert len(li) == 64 and sorted(x for x in li if x >= 0) == list(range((64 // half) * half))
    assert len(set(dw)) == len(dw) == (64 // half) * half and all(0 <= d < 64 for d in dw)
    mult = lds_layout.class_multiplicity(sft, zc)
    c = lds_layout.cost(li, dw, mult, half)
    assert c == lay[ms]["weighted_cycles"]
    li0, dw0 = lds_layout.identity_layout(half)
    assert c < lds_layout.cost(li0, dw0, mult, half)
    assert c >= 2 * len(sft)   # two half-waves of at least one cycle per access


def test_shape_header_carries_tables_and_layouts():
    """The generated shape header holds, per fixed-Z shape, the edge table of the 64-lane
    int8 decoder file and the committed layout."""
    import os
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    hdr = open(os.path.join(here, "simd-fecfun_amd", "csrc", "pair_shapes.h")).read()
    lay = lds_layout.load()
    for ms, zc, col, sft in _shape_tables():
        assert "static constexpr int esft[%d] = {%s};" % (len(sft), ", ".join(map(str, sft))) in hdr
        assert "static const int32_t li[64] = {%s};" % ", ".join(map(str, lay[ms]["lane_item"])) in hdr
    # shift classes: each edge's class index points at its shift mod Z/2, flip = shift >= Z/2
    for ms, zc, col, sft in _shape_tables():
        cls, ecls, eflip = gen_pair_shapes.shift_classes(sft, zc)
        for s, k, f in zip(sft, ecls, eflip):
            assert cls[k] == s % (zc // 2) and f == int(s >= zc // 2)
        assert len(cls) == len(set(cls)) > 0
