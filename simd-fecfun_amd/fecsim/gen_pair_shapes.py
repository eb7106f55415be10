"""Generates csrc/pair_shapes.h, csrc/pair_shapes_dispatch.h and
csrc/float_shapes_dispatch.h: the code shapes (layer count and row weights) the
row-pair int8 decoder (csrc/pair_i8.hip) and the register-C2V float decoder
(csrc/float_reg.hip) are compiled for.

A shape is the row-weight sequence of a base matrix; the shifts and columns stay run-time
data, so one shape serves every lifting size, num_par and row relabeling of a code.
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
