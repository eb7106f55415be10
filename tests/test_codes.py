"""Table generator (fecsim.codes) and tree schedule (fecsim.minprod_tree) checks (CPU)."""
import json
import os

import numpy as np
import pytest

from fecsim import codes
from fecsim.minprod_tree import check_schedule, tree_schedule
import refsim

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _encode(cs, data, oracle):
    """Encode data bits [ncw, ndata*z] with the oracle encoder; returns dense [ncw, n]."""
    ncw = data.shape[0]
    z, mcol, mrow = cs.z, cs.enc.mcol, cs.enc.mrow
    words = np.zeros((ncw, mcol, 64), np.uint8)
    for c in range(mcol - mrow):
        bits = np.zeros((ncw, 512), np.uint8)
        bits[:, :z] = data[:, c * z:(c + 1) * z]
        words[:, c] = np.packbits(bits, axis=1, bitorder="little")
    cnt, colnum, sc, ss = refsim.enc_arrays(cs.enc)
    oracle.lib().orc_encode(ncw, z, mrow, mcol, cnt.ctypes.data, colnum.ctypes.data, sc.ctypes.data, ss.ctypes.data,
                            words.ctypes.data)
    return refsim.unpack_bits(words, z, mcol)


SETS = list(codes.MATRIX_SETS)


@pytest.mark.parametrize("ms", SETS)
@pytest.mark.parametrize("npar", [1, 3])
def test_encoder_satisfies_parity_checks(ms, npar, oracle):
    """H c = 0 (mod 2) for random data: the encoder schedule and the dense H agree
    (LDPCenc_test.m:39-43 invariant)."""
    if ms == "802.3ca" and npar == 3:
        npar = 2   # Z = 768 would not fit the 512-bit encoder column word
    cs = codes.gen_structs(ms, num_par=npar, simd_size=512, elem_size=8)
    rng = np.random.default_rng(7)
    data = rng.integers(0, 2, size=(4, cs.k), dtype=np.uint8)
    c = _encode(cs, data, oracle)
    assert np.array_equal(c[:, :cs.k], data)
    assert not _syndrome(cs.hc, cs.z, c).any()


def _syndrome(hc, z, c):
    """H c^T mod 2 from the block shifts, without the dense H (ExpandH.m:14 convention)."""
    s = np.zeros((c.shape[0], hc.shape[0] * z), np.uint8)
    r = np.arange(z)
    for i in range(hc.shape[0]):
        for j in range(hc.shape[1]):
            if hc[i, j] >= 0:
                s[:, i * z:(i + 1) * z] ^= c[:, j * z + (r + hc[i, j]) % z]
    return s


def test_8023ca_inverse_matches_reference_file():
    """The GF(2) inverse of 802.3ca's parity part equals the reference's
    invM2_size256.mat (which gen_encode.m:19-20 loads) bit for bit: SHA-256 of the
    whole 3072 x 3072 matrix and the three rows the direct parity columns use."""
    with open(os.path.join(GOLDEN, "inv_m2_8023ca.json")) as f:
        ref = json.load(f)
    import hashlib
    hc, z, _ = codes.get_hmatrix("802.3ca")
    ndata = hc.shape[1] - hc.shape[0]
    m2 = codes.expand_h(hc, z)[:, ndata * z:]
    inv = codes.gf2_solve_rows(m2, np.eye(m2.shape[0], dtype=np.uint8))
    assert list(inv.shape) == ref["shape"] and int(inv.sum()) == ref["ones"]
    packed = np.packbits(inv, axis=1, bitorder="little")
    for r, hx in ref["rows"].items():
        assert packed[int(r)].tobytes().hex() == hx, r
    assert hashlib.sha256(packed.tobytes()).hexdigest() == ref["sha256_rows_packed_little"]
    # a plausible slip (one flipped bit) changes the digest
    packed[5, 7] ^= 1
    assert hashlib.sha256(packed.tobytes()).hexdigest() != ref["sha256_rows_packed_little"]


def test_gf2_solver_small_cases():
    """Packed Gauss-Jordan against brute force on 8 x 8 systems, and singular input."""
    rng = np.random.default_rng(11)
    done = 0
    while done < 20:
        a = rng.integers(0, 2, size=(8, 8), dtype=np.uint8)
        x = rng.integers(0, 2, size=(8, 3), dtype=np.uint8)
        b = (a.astype(np.int64) @ x) & 1
        try:
            got = codes.gf2_solve_rows(a, b)
        except ValueError:
            # singular: brute force over all 256 inputs confirms rank < 8
            vecs = {tuple((a.astype(np.int64) @ np.array([(v >> i) & 1 for i in range(8)])) & 1) for v in range(256)}
            assert len(vecs) < 256
            continue
        assert np.array_equal(got, x)
        done += 1
    with pytest.raises(ValueError):
        codes.gf2_solve_rows(np.zeros((70, 70), np.uint8), np.zeros((70, 1), np.uint8))


def test_8023ca_structs_and_puncture(tmp_path):
    """802.3ca: encode_method.m:74-80 schedule, decode_method.m:7-13 row swap, the
    512-bit encoder word limit on Z, and the optional puncture of the last two parity
    block columns (sim_struct_LDPC.h:9) written to the encoder file's puncture list."""
    cs = codes.gen_structs("802.3ca", simd_size=512, elem_size=8)
    assert (cs.z, cs.hc.shape, cs.n, cs.k) == (256, (12, 69), 17664, 14592)
    assert list(cs.enc.par_colnum) == [7, 10, 11, 3, 1, 2, 9, 6, 0, 8, 5, 4]
    assert len(cs.enc.punc_list) == 0
    hre = cs.dec.hc()
    assert np.array_equal(hre[0] >= 0, cs.hc[1] >= 0) and np.array_equal(hre[1] >= 0, cs.hc[0] >= 0)
    with pytest.raises(ValueError):
        codes.gen_structs("802.3ca", num_par=3)
    with pytest.raises(ValueError):
        codes.gen_structs("11ad 1/2", puncture_8023ca=True)
    codes._main(["--matrix-set", "802.3ca", "--puncture-8023ca", "--simd-size", "512", "--elem-size", "8",
                 "--out-dir", str(tmp_path)])
    e = codes.EncTables.from_bytes((tmp_path / "LDPCencode.raw.dat").read_bytes())
    assert list(e.punc_list) == [67 * 256, 69 * 256 - 1] and len(e.fill_list) == 0
    pm, fm = e.masks()
    assert pm[67:].all() and not pm[:67].any() and not fm.any()
    d = codes.DecTables.from_bytes((tmp_path / "LDPCdecode.raw.dat").read_bytes())
    assert d.to_bytes() == cs.dec.to_bytes()


def test_5g_bg1_z384_encoder(oracle):
    cs = codes.gen_structs("5GNR_n25344_k8448", simd_size=512, elem_size=8)
    assert (cs.z, cs.hc.shape, cs.dec.total_hc_ones * cs.z) == (384, (46, 68), 121344)
    pm, fm = cs.enc.masks()
    assert pm[:2].all() and not pm[2:].any() and not fm.any()   # first 2Z systematic bits punctured
    h = codes.expand_h(cs.hc, cs.z)
    data = np.random.default_rng(1).integers(0, 2, size=(1, cs.k), dtype=np.uint8)
    c = _encode(cs, data, oracle)
    syn = (h.astype(np.int32) @ c[0].astype(np.int32)) & 1
    assert not syn.any()


@pytest.mark.parametrize("ms", ["11ad 1/2", "16e 1/2", "11n 5/6 z81", "11ay 1/2"])
@pytest.mark.parametrize("p", [8, 16, 64])
def test_decoder_tables_are_layerwise_rotations(ms, p):
    """decode_method only relabels the rows of each layer cyclically: every active
    block's shift moves by the same amount mod Z (decode_method.m:34-47)."""
    cs = codes.gen_structs(ms, simd_size=p * 8, elem_size=8)
    hre = cs.dec.hc()
    for r in range(hre.shape[0]):
        m = hre[r] >= 0
        assert np.array_equal(m, cs.hc[r] >= 0)
        d = np.mod(hre[r][m] - np.mod(cs.hc[r][m], cs.z), cs.z)
        assert len(set(d.tolist())) == 1


def test_11ad_copy_list_sizes():
    """27 front-to-back and 31 back-to-front copies per iteration for 11ad 1/2 at
    P = 64, as counted from gen_decode.m."""
    cs = codes.gen_structs("11ad 1/2", simd_size=512, elem_size=8)
    assert len(cs.dec.f2b_cols) == 27 and len(cs.dec.b2f_cols) == 31
    assert cs.dec.f2b_copy_cnt.sum() == 27 and cs.dec.b2f_copy_cnt.sum() == 31


@pytest.mark.parametrize("ms", ["11ad 1/2", "11n 1/2 z27", "16e 2/3A"])
def test_dat_round_trip(ms, tmp_path):
    cs = codes.gen_structs(ms, num_par=2, simd_size=256, elem_size=32, enc_path=str(tmp_path / "e.dat"),
                           dec_path=str(tmp_path / "d.dat"))
    d2 = codes.DecTables.from_bytes((tmp_path / "d.dat").read_bytes())
    e2 = codes.EncTables.from_bytes((tmp_path / "e.dat").read_bytes())
    assert d2.to_bytes() == cs.dec.to_bytes() and e2.to_bytes() == cs.enc.to_bytes()
    # header words as gen_decode.m:114 / gen_encode.m:278 write them
    hd = np.frombuffer((tmp_path / "d.dat").read_bytes(), "<i4")[:7]
    assert list(hd[:5]) == [cs.dec.mrow, cs.dec.mcol, cs.z, cs.dec.total_hc_ones, cs.dec.max_row_weight]
    he = np.frombuffer((tmp_path / "e.dat").read_bytes(), "<i4")[:6]
    assert list(he) == [cs.z, 0, 0, 2, cs.enc.mrow, cs.enc.mcol]


def test_num_par_lifting_and_z_over():
    """gen_structs.m:45-51: z = z*num_par, shifts*num_par, z_over overrides z; 16e
    shifts are then taken mod Z (not floor(p*Z/96))."""
    cs = codes.gen_structs("16e 1/2", num_par=2, z_over=96)
    raw, _, _ = codes.get_hmatrix("16e 1/2")
    assert cs.z == 96 and np.array_equal(cs.hc, np.where(raw < 0, -1, raw * 2))
    cs1 = codes.gen_structs("16e 1/2", z_over=48)
    h = codes.expand_h(cs1.hc, 48)
    r, c = 0, 1  # block (0,1) has shift 94 -> 94 mod 48 = 46
    assert h[r * 48 + 0, c * 48 + (94 % 48)] == 1


def test_tree_schedule_matches_reference_generator():
    """The Python restatement emits exactly the schedules min_find.pl emits (golden
    fixture parsed from the generator's output, w = 2..53)."""
    with open(os.path.join(GOLDEN, "minprod_tree.json")) as f:
        ref = json.load(f)
    for w in range(2, 54):
        ops, out = tree_schedule(w)
        assert [list(o) for o in ops] == ref[str(w)]["ops"], w
        assert out == ref[str(w)]["out"], w
        check_schedule(w, ops, out)


def test_perl_seeds():
    """srand(100); 2**56*rand as run_LDPCsim.pl:74-77 packs them (values printed by
    perl 5.34 here)."""
    s = refsim.perl_seeds(100)
    assert list(s[:4]) == [18090700487852288, 15056319331661824, 67801001354732288, 30447614900236800]


def test_5g_lifting_choice():
    assert codes.hbg_fill_puncture(25344, 8448)[:3] == (1, 384, 1)
    bg, z, ils, kc, sp = codes.hbg_fill_puncture(600, 200)   # small K -> BG2, Kb = 8
    assert bg == 2 and z == 26 and ils == 6 and kc == 8 and sp == 2


def test_exec_name_parse():
    assert codes.parse_exec_name("LDPCtop_Vec64c_min") == (64, "c", "min", 512, 8)
    assert codes.parse_exec_name("LDPCtop_Vec16f_prod") == (16, "f", "prod", 512, 32)
    with pytest.raises(ValueError):
        codes.parse_exec_name("LDPCtop_Vec64x_min")


def test_code_tables_compiled_in():
    """The codes compiled into the code-specialized int8 kernel are exactly the decoder
    tables gen_structs writes (row relabelling included): BG1 Z=384 and 802.3ca, with
    the LDS budget of three column copies checked."""
    from fecsim import gen_code_tables
    tb = {ms: rest for ms, *rest in gen_code_tables.tables()}
    assert set(tb) == {"5GNR_n25344_k8448", "802.3ca"}
    for ms, (z, mcol, w, col, sh) in tb.items():
        d = codes.gen_structs(ms, simd_size=512, elem_size=8).dec
        assert (z, mcol) == (d.z, d.mcol)
        assert w == list(d.row_weights) and col == list(d.row_active_cols) and sh == list(d.row_offset)
        assert (3 * mcol + 1) * z <= gen_code_tables.LDS_LIMIT
    assert tb["802.3ca"][:2] == [256, 69] and len(tb["802.3ca"][3]) == 275


def test_multi_code_tables_compiled_in():
    """The codes compiled into the mixed-queue kernel are exactly the decoder tables
    gen_structs writes, every 802.11n and 802.11ad code and 16e 2/3A at Z = 96, and each
    fits the kernel's layout: at least one codeword slot per 256 threads, a slot's LDS
    (two copies and a gap per column) within the instruction's 15-bit immediate offsets,
    and the generated header's match list in the same order."""
    import os
    from fecsim import gen_code_tables
    mt = gen_code_tables.multi_tables()
    names = [t[0] for t in mt]
    assert len(set(names)) == len(names) == 17
    assert {"11n %s z%d" % (r, z) for r in ("1/2", "2/3", "3/4", "5/6") for z in (27, 54, 81)} <= set(names)
    assert {"11ad 1/2", "11ad 5/8", "11ad 3/4", "11ad 13/16", "16e 2/3A Z=96"} <= set(names)
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "simd-fecfun_amd", "csrc", "code_tables.h")).read()
    for i, (name, z, mcol, w, col, sh) in enumerate(mt):
        ms, zo = (name[:-len(" Z=96")], 96) if name.endswith(" Z=96") else (name, None)
        d = codes.gen_structs(ms, simd_size=512, elem_size=8, z_over=zo).dec
        assert (z, mcol) == (d.z, d.mcol)
        assert w == list(d.row_weights) and col == list(d.row_active_cols) and sh == list(d.row_offset)
        assert 1 <= gen_code_tables.MULTI_THREADS // z
        assert ((3 * mcol + 1) * z + 15) // 16 * 16 < 32768
        assert min(w) >= 2
        assert "struct MultiTab%d {  // %s" % (i, name) in hdr
        assert "if (code_tab_match<MultiTab%d>(z, mcol, mrow, w, col, sh)) return %d;" % (i, i) in hdr
