"""Writes inv_m2_8023ca.json from the reference's matlab/LDPC/invM2_size256.mat: the
GF(2) inverse of the 802.3ca parity part M2 (3072 x 3072) that gen_encode.m:19-20
loads for that code (encode_method.m:78-79).  The fixture keeps a SHA-256 of the
whole inverse (rows packed little-endian, 384 bytes each) and, in full, the three rows
that the encoder's directly computed parity columns use (direct_cols 8, 11, 12:
rows 7*256, 10*256, 11*256).  Run from the repository root, where the reference is
at /root/reference; the .mat is read with scipy.io.loadmat, which executes nothing."""
import hashlib
import json
import os

import numpy as np
import scipy.io

SRC = "/root/reference/matlab/LDPC/invM2_size256.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "inv_m2_8023ca.json")

inv = (np.asarray(scipy.io.loadmat(SRC)["invM2"]) != 0).astype(np.uint8)
packed = np.packbits(inv, axis=1, bitorder="little")
rows = {str(r): packed[r].tobytes().hex() for r in (7 * 256, 10 * 256, 11 * 256)}
with open(OUT, "w") as f:
    json.dump({"source": "matlab/LDPC/invM2_size256.mat", "shape": list(inv.shape),
               "sha256_rows_packed_little": hashlib.sha256(packed.tobytes()).hexdigest(),
               "ones": int(inv.sum()), "rows": rows}, f, indent=1)
print(OUT, inv.shape, int(inv.sum()))
