"""Generates csrc/code_tables.h and csrc/code_tables_dispatch.h: the codes compiled into the
code-specialized int8 decoder (csrc/code_i8.hip), each with its lifting size, columns and
shifts, so that every LDS address in that kernel is a register plus an immediate offset.

A code is taken as it appears in the decoder's tables (row relabelling of decode_method.m
applied); the decoder uses the kernel only when a loaded table matches one of these
exactly.  Run by simd-fecfun_amd/build.py before compiling.
"""
from __future__ import annotations

import os

from . import codes
from .gen_pair_shapes import _write

# large-Z codes: one codeword per workgroup of Z threads
CODE_SETS = ["5GNR_n25344_k8448", "802.3ca"]
LDS_LIMIT = 160 * 1024
# codes of the mixed-code queue kernel (csrc/multi_code.hip): several codewords per
# 256-thread workgroup, one row of each per thread; (code, lifting size if not the
# standard's own)
MULTI_SETS = [("11n 1/2 z27", None), ("11n 3/4 z54", None), ("11n 5/6 z81", None), ("11ad 13/16", None),
              ("16e 2/3A", 96),
              # the rest of the 802.11n and 802.11ad families
              ("11n 1/2 z54", None), ("11n 1/2 z81", None), ("11n 2/3 z27", None), ("11n 2/3 z54", None),
              ("11n 2/3 z81", None), ("11n 3/4 z27", None), ("11n 3/4 z81", None), ("11n 5/6 z27", None),
              ("11n 5/6 z54", None), ("11ad 1/2", None), ("11ad 5/8", None), ("11ad 3/4", None)]
MULTI_THREADS = 256


def tables():
    out = []
    for ms in CODE_SETS:
        d = codes.gen_structs(ms, simd_size=512, elem_size=8).dec
        z, mcol = int(d.z), int(d.mcol)
        w = [int(x) for x in d.row_weights]
        assert z % 16 == 0 and z <= 1024 and min(w) >= 2, ms
        assert (3 * mcol + 1) * z <= LDS_LIMIT, ms
        out.append((ms, z, mcol, w, [int(x) for x in d.row_active_cols], [int(x) for x in d.row_offset]))
    return out


def multi_tables():
    out = []
    for ms, zo in MULTI_SETS:
        d = codes.gen_structs(ms, simd_size=512, elem_size=8, z_over=zo).dec
        z, mcol = int(d.z), int(d.mcol)
        w = [int(x) for x in d.row_weights]
        assert 2 <= z <= MULTI_THREADS and min(w) >= 2 and mcol <= 128, ms
        name = ms if zo is None else "%s Z=%d" % (ms, zo)
        out.append((name, z, mcol, w, [int(x) for x in d.row_active_cols], [int(x) for x in d.row_offset]))
    return out


def _struct(h, prefix, i, ms, z, mcol, w, col, sh):
    h.append("struct %s%d {  // %s" % (prefix, i, ms))
    h.append("  static constexpr int z = %d, mcol = %d, mrow = %d, ne = %d, maxw = %d;"
             % (z, mcol, len(w), len(col), max(w)))
    h.append("  static constexpr int w[%d] = {%s};" % (len(w), _arr(w)))
    h.append("  static constexpr int col[%d] = {%s};" % (len(col), _arr(col)))
    h.append("  static constexpr int sh[%d] = {%s};" % (len(sh), _arr(sh)))
    h.append("};")


def _arr(v):
    return ", ".join(map(str, v))


def emit(csrc_dir: str) -> None:
    tb = tables()
    mt = multi_tables()
    h = ["// Generated by fecsim/gen_code_tables.py. Do not edit.", "#pragma once", "#include <stdint.h>", "",
         "namespace fecsim {", ""]
    for i, t in enumerate(tb):
        _struct(h, "CodeTab", i, *t)
    h.append("")
    h.append("// codes of the mixed-code queue kernel")
    for i, t in enumerate(mt):
        _struct(h, "MultiTab", i, *t)
    h.append("constexpr int kMultiTabs = %d;" % len(mt))
    h.append("template <int I>")
    h.append("struct MultiTabAt;")
    for i in range(len(mt)):
        h.append("template <>")
        h.append("struct MultiTabAt<%d> {" % i)
        h.append("  using T = MultiTab%d;" % i)
        h.append("};")
    h += ["", "template <class CT>",
          "inline bool code_tab_match(int z, int mcol, int mrow, const int32_t* w, const int32_t* col,",
          "                           const int32_t* sh) {",
          "  if (z != CT::z || mcol != CT::mcol || mrow != CT::mrow) return false;",
          "  int ne = 0;",
          "  for (int l = 0; l < mrow; ++l) {",
          "    if (w[l] != CT::w[l]) return false;",
          "    ne += w[l];",
          "  }",
          "  for (int e = 0; e < ne; ++e)",
          "    if (col[e] != CT::col[e] || sh[e] != CT::sh[e]) return false;",
          "  return true;",
          "}", "",
          "inline int code_tab_index(int z, int mcol, int mrow, const int32_t* w, const int32_t* col, const int32_t* sh) {"]
    for i in range(len(tb)):
        h.append("  if (code_tab_match<CodeTab%d>(z, mcol, mrow, w, col, sh)) return %d;" % (i, i))
    h += ["  return -1;", "}", "",
          "inline int multi_tab_index(int z, int mcol, int mrow, const int32_t* w, const int32_t* col, const int32_t* sh) {"]
    for i in range(len(mt)):
        h.append("  if (code_tab_match<MultiTab%d>(z, mcol, mrow, w, col, sh)) return %d;" % (i, i))
    h += ["  return -1;", "}", "", "struct CodeArgs;",
          "hipError_t code_dispatch(int code, const CodeArgs& a, hipStream_t st);", "", "}  // namespace fecsim", ""]
    d = ["// Generated by fecsim/gen_code_tables.py. Do not edit.",
         "// included inside namespace fecsim by code_i8.hip after FECSIM_CODE_LAUNCH is defined",
         "hipError_t code_dispatch(int code, const CodeArgs& a, hipStream_t st) {", "  switch (code) {"]
    for i in range(len(tb)):
        d.append("  case %d: FECSIM_CODE_LAUNCH(CodeTab%d)" % (i, i))
    d += ["  default: return hipErrorInvalidValue;", "  }", "}", ""]
    _write(os.path.join(csrc_dir, "code_tables.h"), "\n".join(h))
    _write(os.path.join(csrc_dir, "code_tables_dispatch.h"), "\n".join(d))
