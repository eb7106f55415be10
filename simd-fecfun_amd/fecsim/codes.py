"""Quasi-cyclic LDPC code tables and the two `.dat` files the simulator reads.

This is the host-side restatement of the reference's offline table generator
(build/gen_structs.m and matlab/LDPC/*.m, which need MATLAB/Octave).  It writes
byte-compatible `LDPCencode.raw.dat` / `LDPCdecode.raw.dat`:

* `get_hmatrix`      <- matlab/LDPC/GetHMatrix.m:1-436 (+ lift_matrix.m, Hbg_matrix.m)
* `encode_method`    <- matlab/LDPC/encode_method.m:1-194
* `decode_method`    <- matlab/LDPC/decode_method.m:1-52   (per-layer cyclic row relabeling)
* `gen_decode`       <- matlab/LDPC/gen_decode.m:1-68      (decoder tables + f2b/b2f copy lists)
* `gen_encode`       <- matlab/LDPC/gen_encode.m:1-85      (encoder schedule, par_offset byte addresses)
* `gen_structs`      <- build/gen_structs.m:28-64          (num_par lifting, z_over)
* `hbg_fill_puncture`: the function GetHMatrix.m:403 calls but the reference never defines;
  written here from 3GPP TS 38.212 sec. 5.2.2 / 5.3.2 (base-graph choice, lifting size).

Conventions kept from the reference: a block with shift s connects check row r of
its block row to variable (r + s) mod Z of its block column (ExpandH.m:14,
`circshift(eye(Z),-s)`); data columns come first, parity columns last.
"""
from __future__ import annotations

import re
from dataclasses import dataclass, field

import numpy as np

from .basegraphs import BASE, LIFT_11AY, BG1, BG2, BG1_SHAPE, BG2_SHAPE

# ---------------------------------------------------------------------------
# base matrices
# ---------------------------------------------------------------------------

LIFTING_SETS = {  # TS 38.212 Table 5.3.2-1: iLS -> lifting sizes
    0: (2, 4, 8, 16, 32, 64, 128, 256),
    1: (3, 6, 12, 24, 48, 96, 192, 384),
    2: (5, 10, 20, 40, 80, 160, 320),
    3: (7, 14, 28, 56, 112, 224),
    4: (9, 18, 36, 72, 144, 288),
    5: (11, 22, 44, 88, 176, 352),
    6: (13, 26, 52, 104, 208),
    7: (15, 30, 60, 120, 240),
}

MATRIX_SETS = tuple(sorted(BASE)) + tuple(sorted(LIFT_11AY))


def _parse(rows):
    return np.array([[int(x) for x in r.split()] for r in rows], dtype=np.int64)


def lift_matrix(hc: np.ndarray, liftr, liftc) -> np.ndarray:
    """2x lifting of 11ay from 11ad (lift_matrix.m:1-20): every block becomes the
    2x2 diagonal [[h,-1],[-1,h]]; the listed (1-based) blocks become anti-diagonal."""
    r, c = hc.shape
    out = -np.ones((2 * r, 2 * c), dtype=np.int64)
    out[0::2, 0::2] = hc
    out[1::2, 1::2] = hc
    for lr, lc in zip(liftr, liftc):
        row, col = 2 * (lr - 1), 2 * (lc - 1)
        tmp = out[row, col:col + 2].copy()
        out[row, col:col + 2] = out[row + 1, col:col + 2]
        out[row + 1, col:col + 2] = tmp
    return out


def hbg_matrix(bg: int, ils: int) -> np.ndarray:
    """Base graph `bg` (1 or 2) shift table for lifting-size set `ils` (Hbg_matrix.m)."""
    shape, tab = (BG1_SHAPE, BG1) if bg == 1 else (BG2_SHAPE, BG2)
    h = -np.ones(shape, dtype=np.int64)
    for (r, c), v in tab.items():
        h[r, c] = v[ils]
    return h


def hbg_fill_puncture(n: int, k: int):
    """(bg, Z, iLS, k_mcol, start_punc_mcol) for a 5G NR code of n sent bits and k info bits.

    Base graph choice: TS 38.212 sec. 6.2.2 / 7.2.2 (BG2 if k<=292, or k<=3824 and
    R<=0.67, or R<=0.25).  Kb: 22 for BG1; 10/9/8/6 for BG2 by k (sec. 5.3.2).  Z is the
    smallest lifting size with Kb*Z >= k.  The first 2Z systematic bits are punctured,
    so the base graph's info columns are k_mcol + 2 (22 or 10), as GetHMatrix.m:405 uses.
    """
    rate = k / n
    bg = 2 if (k <= 292 or (k <= 3824 and rate <= 0.67) or rate <= 0.25) else 1
    if bg == 1:
        kb = 22
    else:
        kb = 10 if k > 640 else 9 if k > 560 else 8 if k > 192 else 6
    best = None
    for ils, zs in LIFTING_SETS.items():
        for z in zs:
            if kb * z >= k and (best is None or z < best[0]):
                best = (z, ils)
    if best is None:
        raise ValueError("k=%d too large for 5G NR base graph %d" % (k, bg))
    z, ils = best
    info_cols = 22 if bg == 1 else 10
    return bg, z, ils, info_cols - 2, 2


def get_hmatrix(matrix_set: str):
    """(Hc, Z, extra_5g) as GetHMatrix.m returns them."""
    if matrix_set in BASE:
        z, rows = BASE[matrix_set]
        return _parse(rows), z, None
    if matrix_set in LIFT_11AY:
        src, lr, lc = LIFT_11AY[matrix_set]
        hc, z, _ = get_hmatrix(src)
        return lift_matrix(hc, lr, lc), z, None
    m = re.fullmatch(r"5GNR_n(\d+)_k(\d+)", matrix_set)
    if m:
        n, k = int(m.group(1)), int(m.group(2))
        bg, z, ils, k_mcol, start_punc = hbg_fill_puncture(n, k)
        h = hbg_matrix(bg, ils)
        p_mcol = -(-(n - k) // z) + start_punc
        info_mcol = k_mcol + start_punc
        h = h[:p_mcol, :p_mcol + info_mcol]
        punc_parity = z * (p_mcol - start_punc) - (n - k)
        punc_idx = [1, start_punc * z] if start_punc > 0 else []
        if punc_parity > 0:
            tot = z * (p_mcol + info_mcol)
            punc_idx += [tot - punc_parity + 1, tot]
        fill_cnt = z * info_mcol - k
        fill_idx = [z * info_mcol - fill_cnt + 1, z * info_mcol] if fill_cnt > 0 else []
        extra = [p_mcol, len(punc_idx), len(fill_idx)] + punc_idx + fill_idx
        return h, z, extra
    raise ValueError("invalid MatrixSet argument: %r" % matrix_set)


def encode_method(matrix_set: str):
    """(direct_cols, subst_row, subst_col, z_value), 1-based as in encode_method.m."""
    def rng(a, b):
        return list(range(a, b + 1))
    if matrix_set.startswith("11ad") or (matrix_set.startswith("11ay") and matrix_set != "11ay 7/8"):
        mrow = get_hmatrix(matrix_set)[0].shape[0]
        return [], rng(1, mrow), rng(1, mrow), 42
    if matrix_set == "11ay 7/8":
        return [], rng(1, 4), [2, 1, 3, 4], 42
    if matrix_set.startswith("16e") or matrix_set.startswith("11n") or matrix_set.startswith("ITU G.hn"):
        mrow = get_hmatrix(matrix_set)[0].shape[0]
        z = {"16e": 24, "11n": None, "ITU": 14}[matrix_set[:3]]
        if z is None:
            z = int(matrix_set.split("z")[-1])
        return [1], rng(1, mrow - 1), rng(2, mrow), z
    if matrix_set.startswith("5GNR_"):
        return [1, 3], [2, 4] + rng(5, 99), [2, 4] + rng(5, 99), 0
    if matrix_set == "802.3ca":
        # encode_method.m:74-80; its inverse file invM2_size256 is inv(M2) over GF(2),
        # which gen_encode solves for directly
        return [8, 11, 12], [2, 6, 12, 1, 4, 5, 7, 11, 8], [4, 2, 3, 10, 7, 1, 9, 6, 5], 256
    raise ValueError("invalid MatrixSet argument: %r" % matrix_set)


# ---------------------------------------------------------------------------
# dense helpers
# ---------------------------------------------------------------------------

def expand_h(hc: np.ndarray, z: int) -> np.ndarray:
    """Dense parity-check matrix (ExpandH.m:1-17): block row i, col j with shift s has
    H[i*Z + r, j*Z + (r + s) mod Z] = 1."""
    mrow, mcol = hc.shape
    h = np.zeros((mrow * z, mcol * z), dtype=np.uint8)
    r = np.arange(z)
    for i in range(mrow):
        for j in range(mcol):
            s = hc[i, j]
            if s >= 0:
                h[i * z + r, j * z + (r + s) % z] = 1
    return h


def gf2_solve_rows(a: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    """Solve A x = rhs over GF(2) for each column of rhs (A square, invertible).

    Gauss-Jordan elimination on rows of [A | rhs] packed 64 columns to a word, so a
    3072 x 3072 system (802.3ca's parity part) takes about a second."""
    a = np.asarray(a)
    rhs = np.asarray(rhs)
    n, m = a.shape[0], rhs.shape[1]
    nbits = n + m
    nbytes = -(-nbits // 64) * 8
    aug8 = np.zeros((n, nbytes), np.uint8)
    aug8[:, :-(-nbits // 8)] = np.packbits(np.concatenate([a.astype(np.uint8) & 1, rhs.astype(np.uint8) & 1], axis=1),
                                           axis=1, bitorder="little")
    aug = aug8.view(np.uint64).copy()
    one = np.uint64(1)
    for col in range(n):
        w, b = divmod(col, 64)
        bits = (aug[:, w] >> np.uint64(b)) & one
        piv = np.nonzero(bits[col:])[0]
        if piv.size == 0:
            raise ValueError("matrix is singular over GF(2)")
        p = col + piv[0]
        if p != col:
            aug[[col, p]] = aug[[p, col]]
            bits[[col, p]] = bits[[p, col]]
        rows = np.nonzero(bits)[0]
        rows = rows[rows != col]
        if rows.size:
            aug[rows] ^= aug[col]
    out = np.unpackbits(aug.view(np.uint8), axis=1, bitorder="little", count=nbits)
    return out[:, n:]


# ---------------------------------------------------------------------------
# decoder tables
# ---------------------------------------------------------------------------

def _phase(off: np.ndarray, z: int, p: int) -> np.ndarray:
    """Offset of a shift inside the P-element SIMD word layout (decode_method.m:38,
    gen_decode.m:85-90): shifts >= Z mod P are mapped to mod(shift - Z, P)."""
    zrem = z % p
    off = off.copy()
    m = off >= zrem
    off[m] = np.mod(off[m] - z, p)
    return off


def decode_method(matrix_set: str, hc: np.ndarray, z: int, p: int, bypass: bool = False) -> np.ndarray:
    """Cyclically relabel the check rows of every layer so that as many active columns
    as possible keep the previous layer's phase (decode_method.m:1-52).  Ties keep the
    first (smallest) rotation, as MATLAB's strict `>` does."""
    hc = hc.copy()
    if matrix_set == "802.3ca":
        hc[[0, 1]] = hc[[1, 0]]
    if bypass:
        return hc
    mrow, mcol = hc.shape
    prev_all = np.zeros(mcol, dtype=np.int64)
    for row in range(mrow):
        cols = np.nonzero(hc[row] >= 0)[0]
        prev = prev_all[cols]
        best_cnt, best_off, best_row = -1, None, None
        for i in range(z):
            new_row = np.mod(hc[row, cols] + i, z)
            off = _phase(new_row, z, p)
            cnt = int(np.sum(off == prev))
            if cnt > best_cnt:
                best_cnt, best_off, best_row = cnt, off, new_row
        hc[row, cols] = best_row
        prev_all[cols] = best_off
    return hc


@dataclass
class DecTables:
    """Contents of LDPCdecode.raw.dat (decode_matrix_allocate.cpp:22-50)."""
    mrow: int
    mcol: int
    z: int
    total_hc_ones: int
    max_row_weight: int
    row_weights: np.ndarray
    row_active_cols: np.ndarray
    row_offset: np.ndarray
    f2b_copy_cnt: np.ndarray
    f2b_cols: np.ndarray
    f2b_offsets: np.ndarray
    b2f_copy_cnt: np.ndarray
    b2f_cols: np.ndarray
    b2f_offsets: np.ndarray

    def to_bytes(self) -> bytes:
        hdr = [self.mrow, self.mcol, self.z, self.total_hc_ones, self.max_row_weight,
               len(self.f2b_cols), len(self.b2f_cols)]
        body = np.concatenate([np.asarray(hdr), self.row_weights, self.row_active_cols, self.row_offset,
                               self.f2b_copy_cnt, self.f2b_cols, self.f2b_offsets,
                               self.b2f_copy_cnt, self.b2f_cols, self.b2f_offsets])
        return body.astype("<i4").tobytes()

    @classmethod
    def from_bytes(cls, b: bytes) -> "DecTables":
        a = np.frombuffer(b, dtype="<i4").astype(np.int64)
        mrow, mcol, z, tot, maxrow, lf, lb = (int(x) for x in a[:7])
        pos = 7
        def take(n):
            nonlocal pos
            v = a[pos:pos + n]
            if v.size != n:
                raise ValueError("truncated LDPCdecode.raw.dat")
            pos += n
            return v.copy()
        rw = take(mrow); rac = take(tot); ro = take(tot)
        fc = take(mrow); fcol = take(lf); foff = take(lf)
        bc = take(mrow + 1); bcol = take(lb); boff = take(lb)
        return cls(mrow, mcol, z, tot, maxrow, rw, rac, ro, fc, fcol, foff, bc, bcol, boff)

    def hc(self) -> np.ndarray:
        """The (row-relabeled) base matrix these tables describe."""
        h = -np.ones((self.mrow, self.mcol), dtype=np.int64)
        e = 0
        for r in range(self.mrow):
            for _ in range(self.row_weights[r]):
                h[r, self.row_active_cols[e]] = self.row_offset[e]
                e += 1
        return h


def gen_decode(hc: np.ndarray, p: int, z: int) -> DecTables:
    """Decoder tables and f2b/b2f copy lists (gen_decode.m:1-68)."""
    mrow, mcol = hc.shape
    zrem = z % p
    perfect = zrem == 0
    rac, ro, rw = [], [], []
    for r in range(mrow):
        cols = np.nonzero(hc[r] >= 0)[0]
        rac += list(cols)
        rw.append(len(cols))
        ro += list(hc[r, cols])
    f2b_cnt, b2f_cnt, f2b_off, b2f_off, f2b_cols, b2f_cols = [], [], [], [], [], []
    prev_all = np.zeros(mcol, dtype=np.int64)
    for r in range(mrow):
        cols = np.nonzero(hc[r] >= 0)[0]
        off = hc[r, cols].astype(np.int64)
        prev = prev_all[cols]
        off = np.mod(off, p) if perfect else _phase(off, z, p)
        d = off - prev
        pos, neg = d > 0, d < 0
        f2b_cnt.append(int(pos.sum())); b2f_cnt.append(int(neg.sum()))
        f2b_off += list(prev[pos]); b2f_off += list(prev[neg])
        f2b_cols += list(cols[pos]); b2f_cols += list(cols[neg])
        prev_all[cols] = off
    neg = prev_all > 0
    b2f_cnt.append(int(neg.sum()))
    b2f_off += list(prev_all[neg])
    b2f_cols += list(np.arange(mcol)[neg])
    i64 = lambda v: np.asarray(v, dtype=np.int64)
    return DecTables(mrow, mcol, z, int(sum(rw)), int(max(rw)), i64(rw), i64(rac), i64(ro),
                     i64(f2b_cnt), i64(f2b_cols), i64(f2b_off), i64(b2f_cnt), i64(b2f_cols), i64(b2f_off))


# ---------------------------------------------------------------------------
# encoder tables
# ---------------------------------------------------------------------------

ENC_BYTES_PER_WORD = 64  # gen_encode.m:6 (512-bit columns)


@dataclass
class EncTables:
    """Contents of LDPCencode.raw.dat (encode_matrix_allocate.cpp:51-84)."""
    z: int
    num_par: int
    mrow: int
    mcol: int
    par_cnts: np.ndarray
    par_colnum: np.ndarray
    par_offset: np.ndarray
    punc_list: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))
    fill_list: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int64))

    def to_bytes(self) -> bytes:
        hdr = [self.z, len(self.punc_list), len(self.fill_list), self.num_par, self.mrow, self.mcol]
        body = np.concatenate([np.asarray(hdr), self.par_cnts, self.par_colnum, self.par_offset,
                               self.punc_list, self.fill_list])
        return body.astype("<i4").tobytes()

    @classmethod
    def from_bytes(cls, b: bytes) -> "EncTables":
        a = np.frombuffer(b, dtype="<i4").astype(np.int64)
        z, pc, fc, npar, mrow, mcol = (int(x) for x in a[:6])
        pos = 6
        cnts = a[pos:pos + mrow]; pos += mrow
        colnum = a[pos:pos + mrow]; pos += mrow
        s = int(cnts.sum())
        off = a[pos:pos + s]; pos += s
        punc = a[pos:pos + pc]; pos += pc
        fill = a[pos:pos + fc]; pos += fc
        if fill.size != fc:
            raise ValueError("truncated LDPCencode.raw.dat")
        return cls(z, npar, mrow, mcol, cnts.copy(), colnum.copy(), off.copy(), punc.copy(), fill.copy())

    def schedule(self):
        """Per parity column (in computation order): (parity col index among the parity
        columns, [(source column, shift)]).  Decodes the rotation-area byte address
        col*16*64 + (shift&7)*2*64 + (shift>>3) of gen_encode.m:38,57."""
        out, idx = [], 0
        for i in range(self.mrow):
            srcs = []
            for _ in range(self.par_cnts[i]):
                a = int(self.par_offset[idx]); idx += 1
                col = a // (16 * ENC_BYTES_PER_WORD)
                rem = a - col * 16 * ENC_BYTES_PER_WORD
                mod8 = rem // (2 * ENC_BYTES_PER_WORD)
                byte = rem - mod8 * 2 * ENC_BYTES_PER_WORD
                srcs.append((col, byte * 8 + mod8))
            out.append((int(self.par_colnum[i]), srcs))
        return out

    def masks(self):
        """(puncture, filler) boolean masks of shape [mcol, Z] (enc_fill_mask,
        encode_matrix_allocate.cpp:13-36: inclusive 0-based [start, end] pairs)."""
        res = []
        for lst in (self.punc_list, self.fill_list):
            m = np.zeros(self.mcol * self.z, dtype=bool)
            for i in range(0, len(lst), 2):
                m[lst[i]:lst[i + 1] + 1] ^= True
            res.append(m.reshape(self.mcol, self.z))
        return tuple(res)


def _wa_addr(col, shift):
    shift = np.asarray(shift, dtype=np.int64)
    return np.asarray(col, dtype=np.int64) * 16 * ENC_BYTES_PER_WORD + (shift & 7) * 2 * ENC_BYTES_PER_WORD + (shift >> 3)


def gen_encode(hc, z, num_par, direct_cols, subst_row, subst_col, punc_fill_list) -> EncTables:
    """Encoder schedule (gen_encode.m:1-85).

    Direct parity columns come from the first row of each block row of inv(M2)*M1 over
    GF(2) (gen_encode.m:223-242).  The reference takes inv(M2) with a real-valued
    inverse that "only works for simple M2"; here it is solved over GF(2) on the
    smallest leading block of rows whose parity columns are closed (all rows for
    802.11n/16e, the 4-row core for 5G NR), which gives the same unique codeword.
    """
    mrow, mcol = hc.shape
    if z > 8 * ENC_BYTES_PER_WORD:
        raise ValueError("Z = %d does not fit the encoder's %d-bit column word" % (z, 8 * ENC_BYTES_PER_WORD))
    hlog = hc >= 0
    ndata = mcol - mrow
    par_offset, col_cnts = [], []
    if direct_cols:
        # smallest leading set of block rows that only touches the first `core` parity cols
        core = mrow
        for k in range(1, mrow + 1):
            used = np.nonzero(hlog[:k, ndata:].any(axis=0))[0]
            if used.size and used.max() < k and used.size == k:
                core = k
                break
        h = expand_h(hc[:core, :ndata + core], z)
        m1 = h[:, :ndata * z]
        m2 = h[:, ndata * z:]
        # row (dc-1)*z of inv(M2)*M1: e^T inv(M2) M1 = y^T M1 with M2^T y = e
        e = np.zeros((core * z, len(direct_cols)), dtype=np.uint8)
        for i, dc in enumerate(direct_cols):
            e[(dc - 1) * z, i] = 1
        ys = gf2_solve_rows(m2.T.copy(), e)
        for i in range(len(direct_cols)):
            row = np.bitwise_xor.reduce(m1[ys[:, i].astype(bool)], axis=0)
            idx = np.nonzero(row)[0]
            cols = idx // z
            par_offset += list(_wa_addr(cols, idx - cols * z))
            col_cnts.append(len(idx))
    for sr, sc in zip(subst_row, subst_col):
        tgt = sc + ndata - 1
        cols = np.nonzero(hlog[sr - 1])[0]
        cols = cols[cols != tgt]
        col_cnts.append(len(cols))
        rot = hc[sr - 1, tgt]
        off = np.mod(z - rot + hc[sr - 1, cols], z)
        par_offset += list(_wa_addr(cols, off))
    par_colnum = [c - 1 for c in list(direct_cols) + list(subst_col)]
    if punc_fill_list is not None and len(punc_fill_list) > 0:
        pc, fc = int(punc_fill_list[0]), int(punc_fill_list[1])
        punc = list(punc_fill_list[2:2 + pc])
        fill = list(punc_fill_list[2 + pc:])
    else:
        punc, fill = [], []
    i64 = lambda v: np.asarray(v, dtype=np.int64)
    return EncTables(z, num_par, mrow, mcol, i64(col_cnts), i64(par_colnum), i64(par_offset), i64(punc), i64(fill))


# ---------------------------------------------------------------------------
# gen_structs
# ---------------------------------------------------------------------------

@dataclass
class CodeSet:
    matrix_set: str
    num_par: int
    z: int
    hc: np.ndarray          # lifted, before row relabeling (what the encoder sees)
    enc: EncTables
    dec: DecTables

    @property
    def n(self):
        return self.hc.shape[1] * self.z

    @property
    def k(self):
        return (self.hc.shape[1] - self.hc.shape[0]) * self.z


def gen_structs(matrix_set: str = "11ad 1/2", num_par: int = 1, simd_size: int = 256,
                elem_size: int = 32, z_over: int | None = None,
                enc_path: str | None = None, dec_path: str | None = None,
                puncture_8023ca: bool = False) -> CodeSet:
    """build/gen_structs.m:28-64.  simd_size/elem_size give P, the lane count the
    row relabeling is tuned for (run_LDPCsim.pl:65-70 derives them from the
    executable name, e.g. Vec64c -> 512/8).

    puncture_8023ca: puncture the last two parity block columns of 802.3ca (their
    LLRs are 0), as sim_struct_LDPC.h:9 describes `puncture_8023ca_flag`; the reference
    reads that flag but its masking is commented out (gen_metrics.cpp:47-49,88), so
    here the puncture travels in the encoder file's puncture list instead."""
    hc, z_geth, extra = get_hmatrix(matrix_set)
    direct, srow, scol, z = encode_method(matrix_set)
    if matrix_set.startswith("5GNR_"):
        par_mcol = extra[0]
        lists = list(extra[1:])
        pc = lists[0]
        for i in range(2, len(lists), 2):
            lists[i] = (lists[i] - 1) * num_par + 1
        for i in range(3, len(lists), 2):
            lists[i] = lists[i] * num_par
        for i in range(2, len(lists)):
            lists[i] -= 1
        srow = srow[:par_mcol - 2]
        scol = scol[:par_mcol - 2]
        z = z_geth
        punc_fill = lists
        del pc
    else:
        punc_fill = [0, 0]
    if puncture_8023ca:
        if matrix_set != "802.3ca":
            raise ValueError("puncture_8023ca applies to the 802.3ca matrix set only")
        zz = (int(z_over) if z_over is not None else z * num_par)
        mcol = hc.shape[1]
        punc_fill = [2, 0, zz * (mcol - 2), zz * mcol - 1]   # 0-based, inclusive
    z = z * num_par
    hc = np.where(hc < 0, -1, hc * num_par)
    if z_over is not None:
        z = int(z_over)
    enc = gen_encode(hc, z, num_par, direct, srow, scol, punc_fill)
    p = simd_size // elem_size
    hre = decode_method(matrix_set, hc, z, p)
    dec = gen_decode(hre, p, z)
    if enc_path:
        with open(enc_path, "wb") as f:
            f.write(enc.to_bytes())
    if dec_path:
        with open(dec_path, "wb") as f:
            f.write(dec.to_bytes())
    return CodeSet(matrix_set, num_par, z, hc, enc, dec)


EXEC_ELEM = {"c": 8, "s": 16, "w": 32, "q": 64, "f": 32, "d": 64}  # run_LDPCsim.pl:65


def parse_exec_name(name: str):
    """'LDPCtop_Vec64c_min' -> (P, elem char, 'min'|'prod', simd_size, elem_size)
    (run_LDPCsim.pl:65-68; Makefile:9-14)."""
    m = re.search(r"Vec(\d+)([cswqfd])_(min|prod)", name)
    if not m:
        raise ValueError("bad decoder name %r" % name)
    p, ch, mode = int(m.group(1)), m.group(2), m.group(3)
    es = EXEC_ELEM[ch]
    return p, ch, mode, p * es, es


def _main(argv=None):
    """Write LDPCencode.raw.dat and LDPCdecode.raw.dat, as build/gen_structs.m does."""
    import argparse
    import os
    ap = argparse.ArgumentParser(description="write the simulator's LDPCencode.raw.dat / LDPCdecode.raw.dat")
    ap.add_argument("--matrix-set", default="11ad 1/2", help="e.g. '11ad 1/2', '802.3ca', 5GNR_n25344_k8448")
    ap.add_argument("--num-par", type=int, default=1)
    ap.add_argument("--simd-size", type=int, default=256)
    ap.add_argument("--elem-size", type=int, default=32)
    ap.add_argument("--z-over", type=int, default=None)
    ap.add_argument("--puncture-8023ca", action="store_true")
    ap.add_argument("--out-dir", default=".")
    a = ap.parse_args(argv)
    cs = gen_structs(a.matrix_set, a.num_par, a.simd_size, a.elem_size, a.z_over,
                     os.path.join(a.out_dir, "LDPCencode.raw.dat"), os.path.join(a.out_dir, "LDPCdecode.raw.dat"),
                     puncture_8023ca=a.puncture_8023ca)
    print("%s: Z=%d, %dx%d blocks, n=%d, k=%d" % (a.matrix_set, cs.z, cs.hc.shape[0], cs.hc.shape[1], cs.n, cs.k))


if __name__ == "__main__":
    _main()
