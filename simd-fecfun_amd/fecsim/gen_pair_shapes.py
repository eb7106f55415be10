"""Generates csrc/pair_shapes.h, csrc/pair_shapes_dispatch.h,
csrc/float_shapes_dispatch.h, csrc/train_shapes.h and csrc/train_shapes_dispatch.h: the
code shapes (layer count and row weights) the row-pair int8 decoder (csrc/pair_i8.hip),
the register-C2V float decoder (csrc/float_reg.hip) and the register-C2V differentiable
decoder (csrc/train_reg.hip) are compiled for.

A shape is the row-weight sequence of a base matrix; the shifts and columns stay run-time
data, so one shape serves every lifting size, num_par and row relabeling of a code.
The fixed-Z shapes also carry their code's edge table (column and shift of every edge,
as gen_structs writes it for P = 64 lanes of int8, the LDPCtop_Vec64c_min layout):
a decoder whose tables equal it runs the kernel instance with the table compiled in
(edge addresses from one register per shift class and orientation, columns as
immediate offsets); other tables of the same shape take the run-time-table instance.
Run by simd-fecfun_amd/build.py before compiling.
"""
from __future__ import annotations

import os

from . import codes

# codes whose lifting size is fixed by their standard: their shape also gets a kernel with
# Z and the column count compiled in (loads and stores of a codeword at immediate offsets);
# other lifting sizes (num_par > 1, 802.16e's sweep) take the run-time-Z kernel
FIXED_Z_SETS = ["11ad 1/2", "11ad 13/16", "11ad 3/4", "11ad 5/8"]
# codes whose lifting size can be even and <= 64 (pair kernel: Z/2 <= 32 banks)
PAIR_SETS = ["11ad 1/2", "11ad 13/16", "11ad 3/4", "11ad 5/8", "11n 1/2 z54", "11n 2/3 z54", "11n 3/4 z54",
             "11n 1/2 z27", "16e 1/2", "16e 2/3A", "ITU G.hn 1/2 z14"]
MAX_EDGES = 90
# shapes the float32 decoder with C2V in registers is compiled for (the float workloads:
# rate-1/2 codes with row weights <= 8; one C2V register per edge of a row's layers)
FLOAT_SETS = ["11ad 1/2", "11n 1/2 z54", "11n 1/2 z27", "16e 1/2"]
# shapes the differentiable decoder keeps its C2V messages in registers for (csrc/train_reg.hip:
# rate-1/2 codes, Z <= 128; two rows per lane of 86 edges at most: 172 registers)
TRAIN_SETS = ["11n 1/2 z81", "11n 1/2 z54", "11n 1/2 z27", "11ad 1/2", "16e 1/2"]


def train_shapes():
    """[(row weights, matrix set)] of the register-C2V trainer, one per distinct shape"""
    out = []
    for ms in TRAIN_SETS:
        hc, _, _ = codes.get_hmatrix(ms)
        w = tuple(int((r >= 0).sum()) for r in hc)
        if w not in [x for x, _ in out]:
            out.append((w, ms))
    return out


def float_shapes():
    """indices into shapes() of the shapes compiled for the register-C2V float decoder"""
    sh = shapes()
    out = []
    for ms in FLOAT_SETS:
        hc, _, _ = codes.get_hmatrix(ms)
        w = tuple(int((r >= 0).sum()) for r in hc)
        i = [s for s, _, _ in sh].index(w)
        if i not in out:
            out.append(i)
    return out


def edge_table(ms: str):
    """(columns, shifts) of every edge of matrix set ms in the decoder-table order that
    gen_structs writes for 64 int8 lanes (build/gen_structs.m with simd 512 / elem 8)"""
    d = codes.gen_structs(ms, simd_size=512, elem_size=8).dec
    return [int(c) for c in d.row_active_cols], [int(x) for x in d.row_offset]


def shift_classes(shifts, z):
    """the distinct shift classes s mod Z/2 in order of first use, and per edge its class
    index and whether s >= Z/2 (the pair it reads is the twin of the class's pair)"""
    half = z // 2
    cls = []
    for x in shifts:
        if x % half not in cls:
            cls.append(x % half)
    return cls, [cls.index(x % half) for x in shifts], [int(x >= half) for x in shifts]


def shapes():
    """[(row weights, matrix set, (Z, mcol) compiled in or (0, 0))]"""
    out = []
    for ms in PAIR_SETS:
        hc, z, _ = codes.get_hmatrix(ms)
        w = tuple(int((r >= 0).sum()) for r in hc)
        if sum(w) <= MAX_EDGES and w not in [s for s, _, _ in out]:
            zc = (z, hc.shape[1]) if ms in FIXED_Z_SETS and z % 2 == 0 else (0, 0)
            out.append((w, ms, zc))
    return out


def emit(csrc_dir: str) -> None:
    sh = shapes()
    h = ["// Generated by fecsim/gen_pair_shapes.py. Do not edit.", "#pragma once", "#include <stdint.h>", "",
         "namespace fecsim {", ""]
    for i, (w, ms, (zc, mc)) in enumerate(sh):
        h.append("struct PairShape%d {  // %s" % (i, ms))
        h.append("  static constexpr int mrow = %d, ne = %d, maxw = %d;" % (len(w), sum(w), max(w)))
        h.append("  static constexpr int zc = %d, mc = %d;  // Z and columns of the fixed-Z kernel (0: none)" % (zc, mc))
        h.append("  static constexpr int w[%d] = {%s};" % (len(w), ", ".join(map(str, w))))
        if zc:
            col, sft = edge_table(ms)
            cls, ecls, eflip = shift_classes(sft, zc)
            h.append("  // edge table compiled in (P = 64 int8 layout): column, shift, shift class "
                     "(shift mod Z/2) and shift >= Z/2 per edge")
            h.append("  static constexpr bool ct = true;")
            h.append("  static constexpr int ncls = %d;" % len(cls))
            for nm, arr in (("ecol", col), ("esft", sft), ("ecls", ecls), ("eflip", eflip)):
                h.append("  static constexpr int %s[%d] = {%s};" % (nm, len(arr), ", ".join(map(str, arr))))
            h.append("  static constexpr int cls[%d] = {%s};" % (len(cls), ", ".join(map(str, cls))))
        else:
            h.append("  static constexpr bool ct = false;")
            h.append("  static constexpr int ncls = 0;")
            h.append("  static constexpr int ecol[1] = {0}, esft[1] = {0}, ecls[1] = {0}, eflip[1] = {0}, cls[1] = {0};")
        h.append("};")
    h += ["", "inline int pair_shape_index(int mrow, const int32_t* w) {"]
    for i, (w, _, _) in enumerate(sh):
        cond = " && ".join("w[%d] == %d" % (k, x) for k, x in enumerate(w))
        h.append("  if (mrow == %d && %s) return %d;" % (len(w), cond, i))
    h += ["  return -1;", "}", "",
          "// the shape has a kernel with lifting size z and mcol columns compiled in",
          "inline bool pair_shape_fixed_z(int shape, int z, int mcol) {", "  switch (shape) {"]
    for i, (_, _, (zc, mc)) in enumerate(sh):
        if zc:
            h.append("  case %d: return z == %d && mcol == %d;" % (i, zc, mc))
    h += ["  default: return false;", "  }", "}", "",
          "// the shape's compiled-in edge table equals (cols, shifts) of a decoder",
          "inline bool pair_shape_table(int shape, int ne, const int32_t* cols, const int32_t* shifts) {",
          "  switch (shape) {"]
    for i, (w, ms, (zc, mc)) in enumerate(sh):
        if zc:
            col, sft = edge_table(ms)
            h.append("  case %d: {" % i)
            h.append("    static const int32_t c[%d] = {%s};" % (len(col), ", ".join(map(str, col))))
            h.append("    static const int32_t s[%d] = {%s};" % (len(sft), ", ".join(map(str, sft))))
            h.append("    if (ne != %d) return false;" % len(col))
            h.append("    for (int e = 0; e < ne; ++e)")
            h.append("      if (cols[e] != c[e] || shifts[e] != s[e]) return false;")
            h.append("    return true;")
            h.append("  }")
    # LDS layouts of k_pair_i8c (bank conflicts), searched by lds_layout.py
    from . import lds_layout
    lay = lds_layout.load()
    h += ["  default: return false;", "  }", "}", "",
          "// k_pair_i8c's LDS layout of the shape (lds_layout.py): per hardware lane the",
          "// (slot, lr) item slot * Z/2 + lr (-1: spare lane), per (slot, position) item its",
          "// dword in a column row; false: slots in lane order, positions in slot order",
          "inline bool pair_shape_lds_layout(int shape, int32_t* lane_item, int32_t* dword) {",
          "  switch (shape) {"]
    for i, (w, ms, (zc, mc)) in enumerate(sh):
        if zc and ms in lay:
            li, dw = lay[ms]["lane_item"], lay[ms]["dword"]
            h.append("  case %d: {  // %s: %.3f LDS cycles per access (%.3f in lane order)" % (
                i, ms, lay[ms]["cycles_per_access"], lay[ms]["identity_cycles_per_access"]))
            h.append("    static const int32_t li[64] = {%s};" % ", ".join(map(str, li)))
            h.append("    static const int32_t dw[%d] = {%s};" % (len(dw), ", ".join(map(str, dw))))
            h.append("    for (int x = 0; x < 64; ++x) lane_item[x] = li[x];")
            h.append("    for (int x = 0; x < %d; ++x) dword[x] = dw[x];" % len(dw))
            h.append("    return true;")
            h.append("  }")
    fl = float_shapes()
    h += ["  default: return false;", "  }", "}", "",
          "// the shape has a register-C2V float32 decoder",
          "inline bool float_shape_ok(int shape) {", "  switch (shape) {"]
    h += ["  case %d:" % i for i in fl]
    h += ["    return true;" if fl else "", "  default: return false;", "  }", "}", "", "struct PairArgs;",
          "hipError_t pair_dispatch(int shape, const PairArgs& a, hipStream_t st);", "", "}  // namespace fecsim", ""]
    d = ["// Generated by fecsim/gen_pair_shapes.py. Do not edit.",
         "// included inside namespace fecsim by pair_i8.hip after FECSIM_PAIR_LAUNCH is defined",
         "hipError_t pair_dispatch(int shape, const PairArgs& a, hipStream_t st) {", "  switch (shape) {"]
    for i in range(len(sh)):
        d.append("  case %d: FECSIM_PAIR_LAUNCH(PairShape%d)" % (i, i))
    d += ["  default: return hipErrorInvalidValue;", "  }", "}", ""]
    _write(os.path.join(csrc_dir, "pair_shapes.h"), "\n".join(h))
    _write(os.path.join(csrc_dir, "pair_shapes_dispatch.h"), "\n".join(d))
    ts = train_shapes()
    t = ["// Generated by fecsim/gen_pair_shapes.py. Do not edit.", "#pragma once", "#include <stdint.h>", "",
         "namespace fecsim {", "",
         "// code shapes (layer count and row weights) of the register-C2V differentiable decoder",
         "// (csrc/train_reg.hip); shifts and columns stay run-time data"]
    for i, (w, ms) in enumerate(ts):
        t.append("struct TrainShape%d {  // %s" % (i, ms))
        t.append("  static constexpr int mrow = %d, ne = %d, maxw = %d;" % (len(w), sum(w), max(w)))
        t.append("  static constexpr int w[%d] = {%s};" % (len(w), ", ".join(map(str, w))))
        t.append("};")
    t += ["", "inline int train_shape_index(int mrow, const int32_t* w) {"]
    for i, (w, _) in enumerate(ts):
        cond = " && ".join("w[%d] == %d" % (k, x) for k, x in enumerate(w))
        t.append("  if (mrow == %d && %s) return %d;" % (len(w), cond, i))
    t += ["  return -1;", "}", "", "}  // namespace fecsim", ""]
    _write(os.path.join(csrc_dir, "train_shapes.h"), "\n".join(t))
    td = ["// Generated by fecsim/gen_pair_shapes.py. Do not edit.",
          "// included inside namespace fecsim by train_reg.hip after FECSIM_TRAIN_LAUNCH is defined",
          "hipError_t train_reg_dispatch(int shape, const TrainArgs& a, int bwd, hipStream_t st) {", "  switch (shape) {"]
    for i in range(len(ts)):
        td.append("  case %d: FECSIM_TRAIN_LAUNCH(TrainShape%d)" % (i, i))
    td += ["  default: return hipErrorInvalidValue;", "  }", "}", ""]
    _write(os.path.join(csrc_dir, "train_shapes_dispatch.h"), "\n".join(td))
    f = ["// Generated by fecsim/gen_pair_shapes.py. Do not edit.",
         "// included inside namespace fecsim by float_reg.hip after FECSIM_FREG_LAUNCH is defined",
         "template <class P>", "hipError_t float_dispatch(int shape, const KernelArgs& a, hipStream_t st) {",
         "  switch (shape) {"]
    for i in fl:
        f.append("  case %d: FECSIM_FREG_LAUNCH(P, PairShape%d)" % (i, i))
    f += ["  default: return hipErrorInvalidValue;", "  }", "}", ""]
    _write(os.path.join(csrc_dir, "float_shapes_dispatch.h"), "\n".join(f))


def _write(path, text):
    old = open(path).read() if os.path.exists(path) else None
    if old != text:
        with open(path, "w") as f:
            f.write(text)
