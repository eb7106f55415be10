"""The C-ABI library loads and exports every symbol include/fecsim.h declares; host-only
entry points behave (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fecsim import codes
from fecsim import lib as L

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    syms = set()
    for fn in os.listdir(os.path.join(REPO, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(REPO, "include", fn)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            syms |= set(re.findall(r"\b(fecsim_\w+)\s*\(", txt))
    return syms


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in sorted(syms):
        assert hasattr(lib, s), s
    assert {n for n, _, _ in L.SIGNATURES} == syms


def test_abi_version_and_errors():
    lib = L.lib()
    assert lib.fecsim_abi_version() == 1
    assert lib.fecsim_strerror(-3) == b"malformed table file"


def test_hdec_parse_and_info():
    cs = codes.gen_structs("11n 1/2 z27", num_par=2, simd_size=512, elem_size=8)
    words = np.frombuffer(cs.dec.to_bytes(), "<i4").copy()
    lib = L.lib()
    h = C.c_void_p()
    assert lib.fecsim_hdec_from_words(words.ctypes.data, words.size, C.byref(h)) == 0
    info = (C.c_int32 * 7)()
    assert lib.fecsim_hdec_info(h, info) == 0
    assert list(info) == [12, 24, 54, cs.dec.total_hc_ones, cs.dec.max_row_weight, len(cs.dec.f2b_cols),
                          len(cs.dec.b2f_cols)]
    lib.fecsim_hdec_free(h)


def test_hdec_errors(tmp_path):
    """A truncated or missing table file is an error code, not exit(1)
    (decode_matrix_allocate.cpp:37-40)."""
    cs = codes.gen_structs("11ad 1/2")
    words = np.frombuffer(cs.dec.to_bytes(), "<i4").copy()
    lib = L.lib()
    h = C.c_void_p()
    assert lib.fecsim_hdec_from_words(words.ctypes.data, words.size - 5, C.byref(h)) == -3
    assert b"truncated" in lib.fecsim_last_error()
    bad = words.copy()
    bad[7] = 99   # row weight 99 > 53
    assert lib.fecsim_hdec_from_words(bad.ctypes.data, bad.size, C.byref(h)) == -3
    assert lib.fecsim_hdec_load(str(tmp_path / "missing.dat").encode(), C.byref(h)) == -2
    p = tmp_path / "d.dat"
    p.write_bytes(cs.dec.to_bytes())
    assert lib.fecsim_hdec_load(str(p).encode(), C.byref(h)) == 0
    lib.fecsim_hdec_free(h)


def test_decoder_kind_validation():
    """Kinds follow the executable names; prod needs a float element (no GPU is
    touched before the kind is validated)."""
    cs = codes.gen_structs("11ad 1/2")
    words = np.frombuffer(cs.dec.to_bytes(), "<i4").copy()
    lib = L.lib()
    h = C.c_void_p()
    assert lib.fecsim_hdec_from_words(words.ctypes.data, words.size, C.byref(h)) == 0
    prm = (C.c_uint32 * 3)(10, 0, 0)
    d = C.c_void_p()
    assert lib.fecsim_decoder_create(h, b"Vec64x_min", prm, C.byref(d)) == -1
    assert lib.fecsim_decoder_create(h, b"Vec64c_prod", prm, C.byref(d)) == -6
    assert lib.fecsim_decoder_create(h, b"garbage", prm, C.byref(d)) == -1
    zero = (C.c_uint32 * 3)(0, 0, 0)
    assert lib.fecsim_decoder_create(h, b"Vec64c_min", zero, C.byref(d)) == -1   # max_iter = 0
    lib.fecsim_hdec_free(h)


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError):
        L.lib()
