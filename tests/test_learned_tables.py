"""Committed learned offset tables (fecsim/learned/*.json): they load, match their
code's layer count, hold integer offsets the int8 decoder accepts, and say how they
were made (CPU)."""
import glob
import json
import os

import numpy as np
import pytest

from fecsim import codes
from fecsim.train import load_offsets

LEARNED = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "simd-fecfun_amd", "fecsim", "learned")
NAMES = sorted(os.path.splitext(os.path.basename(p))[0] for p in glob.glob(os.path.join(LEARNED, "*.json")))


def test_tables_present():
    assert "11ad12_T10_int8" in NAMES and "5GNR_n25344_k8448_T20_int8" in NAMES


@pytest.mark.parametrize("name", NAMES)
def test_learned_table_shape_and_values(name):
    tab, rec = load_offsets(name)
    cs = codes.gen_structs(rec["code"], simd_size=512, elem_size=8)
    assert tab.dtype == np.float32
    assert tab.shape == (rec["max_iter"], cs.dec.mrow)
    assert np.array_equal(tab, np.rint(tab)) and tab.min() >= 0 and tab.max() <= 6   # refine_offsets' range
    assert rec["command"].startswith("python -m fecsim.train")
    assert "gain_db" in rec["training_eval"]
    with open(os.path.join(LEARNED, name + ".json")) as f:
        assert json.load(f)["beta_int"] == tab.tolist()
