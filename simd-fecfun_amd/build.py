"""Builds libfecsim.so (HIP for gfx950) in-tree: simd-fecfun_amd/fecsim/libfecsim.so.

hipcc flags:
  --offload-arch=gfx950           MI355X only
  -ffp-contract=off               float kernels must not depend on FMA contraction
  -fhip-fp32-correctly-rounded-divide-sqrt   IEEE division and sqrt in float kernels
  -Xclang -target-feature -Xclang -packed-fp32-ops   keep packed-FP32 instructions out
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "fecsim")
LIB = os.path.join(OUT_DIR, "libfecsim.so")
OBJ_DIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["decode.hip", "float_reg.hip", "pair_i8.hip", "code_i8.hip", "sim.hip", "refrng.hip", "train.hip", "train_reg.hip",
           "multi.hip", "api.cpp", "sim_api.cpp", "msim_api.cpp",
           "train_api.cpp", "multi_api.cpp"]

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + FLAGS + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, r.stdout, r.stderr))
    return obj


def generate() -> None:
    """Code generation steps (shape list of the row-pair kernel, compiled-in codes of the
    code-specialized kernel, product-tree schedules)."""
    sys.path.insert(0, HERE)
    from fecsim import gen_code_tables, gen_pair_shapes, minprod_tree
    gen_pair_shapes.emit(CSRC)
    gen_code_tables.emit(CSRC)
    gen_pair_shapes._write(os.path.join(CSRC, "prodtree_gen.h"), minprod_tree.emit_hip_header(32))


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    generate()
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "--offload-arch=gfx950", "-fPIC"] + objs + ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    build_drop_in()
    if verbose:
        print("built", LIB, file=sys.stderr)
    return LIB


BIN_DIR = os.path.join(HERE, "bin")
# executable names of the reference's decoder variants (build/Makefile:9-14), all served
# by one binary that reads its own name
DROP_IN_KINDS = ["Vec64c_min", "Vec32s_min", "Vec16w_min", "Vec16f_min", "Vec16f_prod", "Vec8d_min", "Vec8d_prod"]


def build_drop_in() -> str:
    """bin/LDPCtop_fecsim (the reference worker's stdin/stdout protocol on the GPU) and
    LDPCtop_<kind> links to it for run_LDPCsim.pl."""
    os.makedirs(BIN_DIR, exist_ok=True)
    exe = os.path.join(BIN_DIR, "LDPCtop_fecsim")
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-x", "hip", os.path.join(CSRC, "ldpctop_main.cpp"),
           "-o", exe + ".tmp", "-L" + os.path.dirname(LIB), "-lfecsim", "-Wl,-rpath,$ORIGIN/../fecsim"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("drop-in executable failed to build:\n%s\n%s" % (r.stdout, r.stderr))
    os.replace(exe + ".tmp", exe)
    for k in DROP_IN_KINDS:
        link = os.path.join(BIN_DIR, "LDPCtop_" + k)
        if not os.path.islink(link):
            if os.path.exists(link):
                os.remove(link)
            os.symlink("LDPCtop_fecsim", link)
    return exe


if __name__ == "__main__":
    build(verbose=True)
