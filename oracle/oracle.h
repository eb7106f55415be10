/*
 * CPU oracle for the LDPC simulator hot path.  TEST INFRASTRUCTURE ONLY: it is
 * linked by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
 * checker, never by the product library.
 *
 * Restates, in plain C, the reference (tomt-1/simd-fecfun) algorithm:
 *   orc_decode_ideal    src/LDPCDec/LDPCdecode.cpp:79-201 + decode_functions.cpp:11-137
 *                       with modular circulant addressing (one check row at a time)
 *   orc_decode_literal  the same file, restated literally: P-lane words, guard words,
 *                       f2b/b2f copy lists from LDPCdecode.raw.dat, the exclusive
 *                       min/product tree of src/LDPCDec/min_find.pl
 *   orc_decode_vec64c   LDPCtop_Vec64c_min as a SIMD CPU decoder (AVX-512BW / AVX2),
 *                       multithreaded: the CPU baseline
 *   orc_sfc64 / orc_xoshiro / orc_normdist2   src/RNGfunc/prng_functions.cpp:22-186
 *   orc_gen_rand_databits   src/Top/gen_rand_databits.cpp:11-40
 *   orc_encode              src/LDPCEnc/LDPCencode.cpp:40-140 (as XORs of rotations)
 *   orc_gen_metrics         src/Top/gen_metrics.cpp:18-121
 *   orc_count_errors        src/Top/LDPC_count_errors.cpp:40-136
 *   orc_noise_llr           the product's counter-based noise scheme (splitmix64 +
 *                           the reference's Box-Muller transform), for bit-exact tests
 *
 * Parity pins: tests/test_oracle.py and tests/test_codes.py.  Functions with no pin beyond their own
 * restatement are marked "parity unpinned" where they are declared.
 */
#ifndef FECSIM_ORACLE_H
#define FECSIM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_I8 = 0, ORC_I16 = 1, ORC_I32 = 2, ORC_F32 = 3, ORC_F64 = 4 };
enum { ORC_MIN = 0, ORC_PROD = 1 };

/* LDPCdecode.raw.dat contents (dec_struct_LDPC.h:3-24) */
typedef struct {
    int32_t mrow, mcol, z, total_hc_ones, max_row_weight;
    const int32_t *row_weights, *row_active_cols, *row_offset;
    const int32_t *f2b_copy_cnt, *f2b_cols, *f2b_offsets;
    const int32_t *b2f_copy_cnt, *b2f_cols, *b2f_offsets;
} orc_hdec;

/* Exclusive min / product tree schedule for one row weight w, as emitted by
 * min_find.pl: ops[i] = {dst, a, b} (bit masks of the inputs a node covers);
 * a == -1 marks "dst = min(cap, dst)".  Leaves m[1<<i] = input i; output
 * j = m[all ^ (1<<j)]. */
int orc_set_tree(int w, int nops, const int64_t *ops /* [nops][3] */);

/* Ideal layered decoder.  llr: [num_cw][mcol*z] elements of `elem` (in: channel
 * LLRs, out: posteriors).  c2v: [num_cw][total_hc_ones*z] (out, may be NULL), row r of
 * edge e at e*z + r.  iters: [num_cw] (out).  use_tree: evaluate the exclusive
 * min with the tree (1) or min1/min2 (0); prod mode always uses the tree.
 * group_exit: number of interleaved codewords sharing one early-termination
 * decision (1 = per codeword; n > 1 emulates num_par=n with per-slot exits).
 * Returns the summed iterations, or a negative error code. */
int64_t orc_decode_ideal(const orc_hdec *h, int elem, int mode, uint32_t num_cw, uint32_t max_iter,
                         float max_row_metric, float beta_offset, void *llr, void *c2v, uint32_t *iters,
                         int use_tree);

/* Per-(iteration, layer) offsets [max_iter][mrow] used by orc_decode_ideal in place of
 * beta_offset (converted the same way), or NULL for the constant.  The table must
 * cover every iteration of the next calls; it is read, not copied. */
void orc_set_beta_table(const float *tab);

/* Literal decoder on the reference layout: metrics [num_cw][mcol][(SPZ+2)*P] with
 * column c's data at element P + c*(SPZ+2)*P.  c2v out as in orc_decode_ideal. */
int64_t orc_decode_literal(const orc_hdec *h, int elem, int mode, int P, uint32_t num_cw, uint32_t max_iter,
                           float max_row_metric, float beta_offset, void *metrics, void *c2v, uint32_t *iters);

/* CPU baseline: LDPCtop_Vec64c_min (P = 64, int8) with SIMD, nthreads threads.
 * metrics in the reference layout (P = 64).  Returns summed iterations, negative on
 * error.  isa_out: 512 (AVX-512BW) or 256 (AVX2). */
int64_t orc_decode_vec64c(const orc_hdec *h, uint32_t num_cw, uint32_t max_iter, float max_row_metric,
                          float beta_offset, int8_t *metrics, uint32_t *iters, int nthreads, int *isa_out);

/* ---- RNG / WGN (prng_functions.cpp) ---- */
void orc_init_randstate(const uint64_t *inp /* [32] */);
void orc_sfc64_N(int sel, uint64_t out[4]);
void orc_xoshiro_N(int sel, uint64_t out[4]);
float orc_fastm2log(float x);
void orc_sincos_norm(float x, float *s, float *c);
void orc_normdist2(float *result /* [simd_count*8] */, int simd_count);

/* ---- bits: per cw, mcol words of 64 bytes, bit i of a column at byte i>>3 bit i&7 ---- */
void orc_gen_rand_databits(uint32_t num_cw, int z, int mrow, int mcol, const uint8_t *filler_mask /* [mcol][64] */,
                           uint8_t *cw_bits /* [num_cw][mcol][64] */);
/* schedule: for each parity column in order, cnt[i] sources (col, shift) */
void orc_encode(uint32_t num_cw, int z, int mrow, int mcol, const int32_t *par_cnts, const int32_t *par_colnum,
                const int32_t *src_col, const int32_t *src_shift, uint8_t *cw_bits);

/* gen_metrics.cpp:18-121.  layout: P > 0 -> reference layout with P lanes,
 * P == 0 -> dense [num_cw][mcol*z].  Returns the uncoded error count. */
uint32_t orc_gen_metrics(uint32_t num_cw, float snr, const uint8_t *cw_bits, int z, int mcol, const float *wgn,
                         void *metrics, int elem, int P, float llr_scale, float max_abs_llr,
                         const uint8_t *punct_mask, const uint8_t *filler_mask);

/* LDPC_count_errors.cpp:40-136 on dense posteriors [num_cw][mcol*z] of `elem` */
void orc_count_errors(uint32_t num_cw, int z, int mrow, int mcol, int num_par, const uint8_t *cw_bits,
                      const void *post, int elem, uint32_t *tot_cw_err, uint32_t *tot_bit_err);

/* sigma and LLR factor of gen_metrics.cpp:28-30 */
void orc_snr_params(float snr, float llr_scale, float *sigma, float *factor);

/* ---- the product's counter-based noise scheme (restated for bit-exact tests) ---- */
uint64_t orc_splitmix64(uint64_t x);
/* noise sample for (cw, col, row): pair p = (cw*ceil(mcol/2) + col/2)*z + row,
 * half = col&1 (0: sin*g, 1: cos*g) */
float orc_noise_sample(uint64_t seed, uint64_t cw, int col, int row, int z, int mcol);
/* Dense int8/float LLRs for given cw bits (data+parity, [num_cw][mcol*z] 0/1 bytes). */
uint32_t orc_noise_llr(uint64_t seed, uint64_t cw0, uint32_t num_cw, const uint8_t *bits, int z, int mcol,
                       float sigma, float factor, float max_abs_llr, const uint8_t *punct, const uint8_t *fill,
                       int elem, void *llr);

/* the product's counter-based data bits (words layout as orc_gen_rand_databits) */
void orc_gpu_databits(uint64_t seed, uint64_t cw0, uint32_t num_cw, int z, int mrow, int mcol,
                      const uint8_t *filler_mask, uint8_t *cw_bits);

/* ---- differentiable offset/normalized min-sum (oracle/train.c) ---- */
/* T iterations, no early exit; alpha/beta [T][mrow]; llr [num_cw][N] in: channel, out:
 * posterior; c2v_out [num_cw][E] or NULL.  Returns 0, negative on error. */
int orc_train_fwd_f32(const orc_hdec *h, uint32_t num_cw, uint32_t T, const float *alpha, const float *beta,
                      float cap, float *llr, float *c2v_out);
int orc_train_fwd_f64(const orc_hdec *h, uint32_t num_cw, uint32_t T, const double *alpha, const double *beta,
                      double cap, double *llr, double *c2v_out);
/* reverse mode of orc_train_fwd_f64: gpost = dL/d(posterior) [num_cw][N] -> summed
 * galpha/gbeta [T][mrow] and gllr = dL/d(channel LLR) [num_cw][N] (may be NULL) */
int orc_train_grad_f64(const orc_hdec *h, uint32_t num_cw, uint32_t T, const double *alpha, const double *beta,
                       double cap, const double *llr, const double *gpost, double *galpha, double *gbeta,
                       double *gllr);

/* own float/double math used by the sum-product decoder (parity unpinned against
 * vectorclass's tanh/atanh; see the header comment of oracle.c) */
float orc_tanhf(float x);
float orc_atanhf(float x);
double orc_tanh(double x);
double orc_atanh(double x);

#ifdef __cplusplus
}
#endif
#endif
