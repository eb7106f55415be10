/*
 * fecsim: MI355X (gfx950) LDPC simulator hot path -- C ABI.
 *
 * Drop-in boundary for the reference simulator tomt-1/simd-fecfun.  Every entry
 * point names the reference interface it replaces (file:line in that repository).
 * Conventions:
 *   - plain C types and pointers only; opaque handles; every call returns an int
 *     status (FECSIM_OK = 0, negative on error; nothing calls exit());
 *   - "device" pointers are HIP device memory; "stream" is a hipStream_t passed as
 *     void* (NULL = the default stream); device calls are asynchronous on it;
 *   - a handle may be used by one host thread at a time; distinct handles are
 *     independent.
 *
 * Element kinds follow the reference's executable names LDPCtop_Vec<P><t>_<mode>
 * (include/simd_defs.h:34-54, build/Makefile:11-14, build/run_LDPCsim.pl:65-68):
 * t = c (int8), s (int16), w (int32), f (float), d (double); mode = min (offset
 * min-sum) or prod (tanh/atanh sum-product, f and d only).  P only sets the
 * reference memory layout and the row relabeling already stored in the .dat file.
 */
#ifndef FECSIM_H
#define FECSIM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FECSIM_ABI_VERSION 1

enum {
    FECSIM_OK = 0,
    FECSIM_E_INVAL = -1,       /* bad argument or parameter */
    FECSIM_E_IO = -2,          /* file cannot be opened or read */
    FECSIM_E_FORMAT = -3,      /* malformed or truncated table file */
    FECSIM_E_NOMEM = -4,       /* host or device allocation failed */
    FECSIM_E_HIP = -5,         /* HIP runtime error (see fecsim_last_error) */
    FECSIM_E_UNSUPPORTED = -6  /* valid request this build does not handle */
};

enum { FECSIM_I8 = 0, FECSIM_I16 = 1, FECSIM_I32 = 2, FECSIM_F32 = 3, FECSIM_F64 = 4 };
enum { FECSIM_MIN = 0, FECSIM_PROD = 1 };

int fecsim_abi_version(void);
const char *fecsim_strerror(int status);
/* message of the last failing call on this host thread ("" if none) */
const char *fecsim_last_error(void);

/* ------------------------------------------------------------------------- */
/* Decoder tables: LDPCdecode.raw.dat                                         */
/* ------------------------------------------------------------------------- */

typedef struct fecsim_hdec fecsim_hdec;

typedef struct {
    int32_t mrow, mcol, z, total_hc_ones, max_row_weight, len_f2b, len_b2f;
} fecsim_hdec_info_t;

/* Replaces decode_matrix_allocate() (src/LDPCDec/decode_matrix_allocate.cpp:12-61):
 * same file format (dec_struct_LDPC.h:3-24; gen_decode.m:113-116).  The f2b/b2f
 * copy lists are validated and kept but not needed by the GPU decoder, which
 * addresses circulants modulo Z.  Buffers are not tied to a codeword count. */
int fecsim_hdec_load(const char *path, fecsim_hdec **out);
/* same, from the file's int32 words already in memory */
int fecsim_hdec_from_words(const int32_t *words, size_t nwords, fecsim_hdec **out);
int fecsim_hdec_info(const fecsim_hdec *h, fecsim_hdec_info_t *info);
void fecsim_hdec_free(fecsim_hdec *h);

/* ------------------------------------------------------------------------- */
/* Decoder                                                                    */
/* ------------------------------------------------------------------------- */

typedef struct fecsim_decoder fecsim_decoder;

typedef struct {
    uint32_t max_iter;     /* sim_struct.max_iter (sim_struct_LDPC.h:7) */
    float max_row_metric;  /* sim_struct.max_row_metric: row cap, or 1-cap for prod if < 1e-6 */
    float beta_offset;     /* sim_struct.beta_offset */
} fecsim_dec_params;

/* Builds a decoder for the tables `h` and the kind named like the reference
 * executables ("Vec64c_min", "LDPCtop_Vec16f_prod", ...; decoder selection of
 * Makefile:11-14 + simd_defs.h).  The tables are copied to the device. */
int fecsim_decoder_create(const fecsim_hdec *h, const char *kind, const fecsim_dec_params *params,
                          fecsim_decoder **out);
/* Kernel selection.  AUTO picks the row-pair int8 kernel when the code shape was
 * compiled for it (even Z <= 64, int8 min-sum, 1 <= beta, 0 <= cap, dense layout), the
 * code-specialized int8 kernel when the whole code was compiled in (5G NR base graph 1
 * at Z = 384; int8 min-sum, 1 <= beta, 0 <= cap, dense 16-byte aligned buffers, any
 * offset table with entries 1..127), the register-C2V float kernel for float32 decoders
 * (min-sum or sum-product) whose code shape was compiled for it (802.11ad, 802.11n and
 * 802.16e rate 1/2, Z <= 256, dense layout, any offset table), and the generic kernel
 * otherwise; all give identical results.  No reference counterpart (the reference
 * selects its decoder at compile time only). */
enum {
    FECSIM_KERNEL_AUTO = 0,
    FECSIM_KERNEL_GENERIC = 1,
    FECSIM_KERNEL_PAIR = 2,
    FECSIM_KERNEL_CODE = 3,
    FECSIM_KERNEL_FLOAT = 4,
    /* the row-pair kernel's instance that reads the edge table at run time, also for a
     * code whose table is compiled into the row-pair kernel (the 802.11ad codes in the
     * 64-lane int8 table order, which FECSIM_KERNEL_PAIR runs with the table compiled
     * in); the same results */
    FECSIM_KERNEL_PAIR_RT = 5
};
/* FECSIM_E_INVAL while a queue decoder (fecsim_multi_create) or a simulation loop
 * (fecsim_sim_create) holds d, as fecsim_decoder_set_offsets */
int fecsim_decoder_select_kernel(fecsim_decoder *d, int kernel);
/* kernel a dense decode_set call on aligned buffers would run now
 * (FECSIM_KERNEL_GENERIC, _PAIR, _CODE, _FLOAT, or _PAIR_RT when selected) */
int fecsim_decoder_kernel(const fecsim_decoder *d);

/* element kind (FECSIM_I8..F64), mode (FECSIM_MIN/PROD), bytes per element */
int fecsim_decoder_info(const fecsim_decoder *d, int *elem, int *mode, int *elem_bytes);

/* Per-(iteration, layer) offsets: beta[max_iter][mrow] host floats, each converted as
 * sim_struct.beta_offset is (float -> int truncation -> element width for integer
 * kinds, decode_functions.cpp:60-61), used by iteration t, layer l in place of
 * beta_offset.  NULL restores the constant offset.  No reference counterpart: this is
 * how offsets learned with the trainable decoder below are run in the bit-exact
 * decoders; with every entry equal to beta_offset the results are the plain
 * decoder's.  Tables run on the row-pair kernel (802.11ad shapes at their own Z, at
 * most 4096 entries, int8 entries 0..127) and on the code-specialized kernel (entries
 * 1..127); other tables run on the generic kernel.  Returns FECSIM_E_INVAL
 * while a queue decoder (fecsim_multi_create) or a simulation loop (fecsim_sim_create)
 * holds d: they copy its tables and kernel choice when they are made, so set offsets
 * before creating them (or after freeing them). */
int fecsim_decoder_set_offsets(fecsim_decoder *d, const float *beta);
void fecsim_decoder_free(fecsim_decoder *d);

/* Replaces LDPCdecode_set() (src/LDPCDec/LDPCdecode.cpp:203-245): decodes num_cw
 * codewords in place; every codeword stops after the iteration whose syndrome is
 * zero or at max_iter (LDPCdecode.cpp:172-197).
 *   metrics  device, in: channel LLRs, out: posterior LLRs, of the decoder's element
 *            type.  layout_p = 0: dense [num_cw][mcol*z]; layout_p = P > 0: the
 *            reference layout [num_cw][mcol][(ceil(z/P)+2)*P] with column c's data at
 *            element P + c*(ceil(z/P)+2)*P (LDPCdecode.cpp:100,225,242).
 *   iters    device [num_cw] uint32 iteration counts, or NULL (the reference
 *            allocates Hdec->iterations but never fills it, decode_matrix_allocate.cpp:60)
 *   c2v      device [num_cw][total_hc_ones*z] final check-to-variable messages (row r
 *            of edge e at e*z + r, rows labeled as the .dat's row_offset), or NULL
 *   tot_iter device uint64 to which the summed iterations are added, or NULL
 *            (the return value of LDPCdecode_set)
 * At most 8 decodes of one decoder may be in flight on the device at once (on
 * different streams): the row-pair kernel's work queue takes one of the decoder's 8
 * counter sets per call, in turn.  A captured graph keeps the set of its capture. */
int fecsim_decode_set(fecsim_decoder *d, uint32_t num_cw, int layout_p, void *metrics, uint32_t *iters,
                      void *c2v, unsigned long long *tot_iter, void *stream);

/* Out-of-place form of fecsim_decode_set: reads channel LLRs from llr_in and writes
 * posteriors to post_out (same layout; the two may be equal).  No reference
 * counterpart beyond LDPCdecode_set; it lets a caller keep its LLRs for a rerun. */
int fecsim_decode(fecsim_decoder *d, uint32_t num_cw, int layout_p, const void *llr_in, void *post_out,
                  uint32_t *iters, void *c2v, unsigned long long *tot_iter, void *stream);

/* Synchronous form over host memory, the literal counterpart of LDPCdecode_set on
 * Hdec->decoded_metrics: copies in, decodes, copies back, returns the summed
 * iterations in *tot_iter.  iters_host may be NULL. */
int fecsim_decode_set_host(fecsim_decoder *d, uint32_t num_cw, int layout_p, void *metrics_host,
                           uint32_t *iters_host, unsigned long long *tot_iter);

/* The float32 sum-product decoder's own device functions, elementwise over device
 * arrays, for verifying their accuracy: fn 0 gives tanh(x[i] / 2) (x >= 0), fn 1 gives
 * 2 atanh(x[i]) (0 <= x < 1).  They stand where the reference's generated product
 * functions call vectorclass tanh / atanh (src/LDPCDec/min_find.pl:39,45). */
int fecsim_sp_math(int fn, const float *x, float *y, uint32_t n, void *stream);

/* ------------------------------------------------------------------------- */
/* Mixed-code queue decoder (no reference counterpart)                        */
/* ------------------------------------------------------------------------- */

/* One persistent launch decodes a queue whose codewords belong to different codes:
 * decoder c (its tables, max_iter, beta, cap) decodes every queue entry with code id
 * c, with results identical to fecsim_decode on that decoder.  All decoders need the
 * same element kind and mode, and their messages must fit one wave's LDS slice.
 * Two kernels serve a queue: the generic one (C2V messages in LDS, any codes) and,
 * when every decoder's code is compiled into it (every 802.11n code, rates 1/2 to 5/6 at
 * Z = 27, 54 and 81, every 802.11ad code, and 802.16e 2/3A at Z = 96; int8 min-sum,
 * 1 <= beta, 0 <= cap, offset tables with entries 1..127), one with the codes compiled in (C2V messages in
 * registers); the latter is the default then, and both give identical results. */
typedef struct fecsim_multi fecsim_multi;

int fecsim_multi_create(fecsim_decoder *const *decs, int ncode, uint32_t max_cw, fecsim_multi **out);
/* code: device int32 [num_cw] code ids; off: device uint64 [num_cw] element offset of
 * each codeword's dense LLRs in llr_in (and of its posteriors in post_out); iters:
 * device [num_cw] or NULL; tot_iter: device uint64 added to, or NULL.  The queue is
 * grouped by code on the device (no host round trip; graph-capturable). */
int fecsim_multi_decode(fecsim_multi *m, uint32_t num_cw, const int32_t *code, const unsigned long long *off,
                        const void *llr_in, void *post_out, uint32_t *iters, unsigned long long *tot_iter,
                        void *stream);
/* Autotuner: times the queue (decoded into post_out, best of `rounds`) for candidate
 * tile shapes of the generic kernel -- rows per lane, hence codewords per wave, for each
 * code in turn, then waves per block -- and keeps the fastest; then, when the codes are
 * compiled in and the kernel choice is FECSIM_KERNEL_AUTO, times that kernel and runs
 * the queue on whichever of the two is faster.  With FECSIM_KERNEL_CODE selected it
 * does nothing.  Synchronises the stream. */
int fecsim_multi_autotune(fecsim_multi *m, uint32_t num_cw, const int32_t *code, const unsigned long long *off,
                          const void *llr_in, void *post_out, int rounds, void *stream);
int fecsim_multi_shape(const fecsim_multi *m, int code, int *rows_per_lane, int *cw_per_wave,
                       int *waves_per_block);
/* kernel: FECSIM_KERNEL_AUTO (the compiled-code kernel when every code is compiled in),
 * FECSIM_KERNEL_GENERIC, or FECSIM_KERNEL_CODE (FECSIM_E_UNSUPPORTED when a code is not
 * compiled in).  fecsim_multi_kernel: FECSIM_KERNEL_GENERIC or FECSIM_KERNEL_CODE, the
 * kernel fecsim_multi_decode runs now. */
int fecsim_multi_select_kernel(fecsim_multi *m, int kernel);
int fecsim_multi_kernel(const fecsim_multi *m);
/* fecsim_multi_decode without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc): the queue
 * workspace (grouping permutation, per-code counters, work counter) is a static device
 * array of FECSIM_MULTI_STATIC_MAX_CW entries instead of the handle's, so num_cw may
 * exceed the handle's max_cw up to that bound; larger queues return FECSIM_E_INVAL.
 * Results equal fecsim_multi_decode's.  Every call on one GPU shares that GPU's array
 * (each GPU holds its own copy, 4 MiB, until the process ends): calls that overlap on
 * one GPU -- concurrent streams or host threads -- race and must be serialised by the
 * caller.  No counterpart in the reference. */
#define FECSIM_MULTI_STATIC_MAX_CW 1048576u
int fecsim_multi_decode_static(fecsim_multi *m, uint32_t num_cw, const int32_t *code, const unsigned long long *off,
                               const void *llr_in, void *post_out, uint32_t *iters, unsigned long long *tot_iter,
                               void *stream);
void fecsim_multi_free(fecsim_multi *m);


/* ------------------------------------------------------------------------- */
/* Encoder tables: LDPCencode.raw.dat                                         */
/* ------------------------------------------------------------------------- */

typedef struct fecsim_henc fecsim_henc;

typedef struct {
    int32_t z, mrow, mcol, num_par;  /* header of the file (enc_struct_LDPC.h:12-28) */
    int32_t punc_cnt, fill_cnt;      /* lengths of the puncture / filler range lists */
    int32_t punc_bits, fill_bits;    /* bits those lists mark (fill_cnt of "Matrix Params") */
    int32_t col_words;               /* 32-bit words per column of the dense bit layout, ceil(z/32) */
    int32_t num_src;                 /* total rotated sources of the encoding schedule */
} fecsim_henc_info_t;

/* Replaces encode_matrix_allocate() (src/LDPCEnc/encode_matrix_allocate.cpp:38-92):
 * same file format (gen_encode.m:83-84), puncture/filler lists turned into masks
 * as enc_fill_mask() does (:13-36, ranges toggle bits).  Host-only; device copies
 * of the tables are made on first use.  Needs z <= 2048. */
int fecsim_henc_load(const char *path, fecsim_henc **out);
int fecsim_henc_from_words(const int32_t *words, size_t nwords, fecsim_henc **out);
int fecsim_henc_info(const fecsim_henc *h, fecsim_henc_info_t *info);
/* masks as [mcol][col_words] uint32 (row r of column c at word c*col_words + r/32,
 * bit r%32); either pointer may be NULL */
int fecsim_henc_masks(const fecsim_henc *h, uint32_t *punct, uint32_t *fill);
void fecsim_henc_free(fecsim_henc *h);

/* Codeword bits in device memory are [num_cw][mcol][col_words] uint32 words, row r of
 * a column at word r/32, bit r%32, data columns first (col_words >= ceil(z/32)).
 * col_words = 16 is the reference's cw_bits layout (one 512-bit ENC_CLASS word per
 * column, enc_defs.h). */

/* Replaces gen_rand_databits() (src/Top/gen_rand_databits.cpp:11-40) + LDPCencode()
 * (src/LDPCEnc/LDPCencode.cpp:40-100): fills the data columns of codewords
 * cw0 .. cw0+num_cw-1 with counter-based random bits (seed, codeword index; masked to
 * z bits and to non-filler bits) and computes the parity columns.  The stream is
 * this library's own (splitmix64 per 64-bit chunk), not the reference's xoshiro
 * stream, whose state depends on every earlier draw. */
int fecsim_gen_codewords(fecsim_henc *h, uint64_t seed, uint64_t cw0, uint32_t num_cw, int col_words,
                         uint32_t *bits, void *stream);
/* Replaces LDPCencode() alone: reads the data columns of bits (masked to z bits) and
 * writes the parity columns. */
int fecsim_encode(fecsim_henc *h, uint32_t num_cw, int col_words, uint32_t *bits, void *stream);

/* ------------------------------------------------------------------------- */
/* Channel: BPSK + AWGN -> quantized LLRs                                     */
/* ------------------------------------------------------------------------- */

typedef struct {
    float snr_db;       /* Es/N0 in dB: sigma = 10^(-(snr + 10 log10 2)/20) (gen_metrics.cpp:28-30) */
    float llr_scale;    /* LLR_scaling_factor: LLR = 2*scale/sigma^2 * y */
    float max_abs_llr;  /* clamp; filler bits get +max_abs_llr, punctured bits 0 */
} fecsim_chan_params;

/* Replaces normdist2() + gen_metrics() (src/RNGfunc/prng_functions.cpp:148-186,
 * src/Top/gen_metrics.cpp:18-121): for codewords cw0.. of `bits`, noise from the
 * counter-based generator keyed by (seed, codeword, column pair, row), the
 * reference's Box-Muller transform (fastm2log, sincos_norm), y = fma(n, sigma, +-1),
 * LLR clamped, punctured -> 0, filler -> +max, integers rounded to nearest even and
 * truncated to the element width.  Writes `elem` LLRs in layout_p's layout (as
 * fecsim_decode_set) and adds the count of sign flips of y (the uncoded bit errors)
 * to *uncoded_err (device uint64, may be NULL).  The same seed and codeword index
 * give the same noise at every SNR, as the reference reuses its noise buffer
 * across SNR points (LDPCtop.cpp:90-97). */
int fecsim_gen_llr(fecsim_henc *h, const fecsim_chan_params *p, uint64_t seed, uint64_t cw0, uint32_t num_cw,
                   int col_words, const uint32_t *bits, int elem, int layout_p, void *llr,
                   unsigned long long *uncoded_err, void *stream);

/* sigma and LLR factor the channel uses for p (host computation, gen_metrics.cpp:28-30) */
int fecsim_chan_sigma(const fecsim_chan_params *p, float *sigma, float *factor);

/* fecsim_gen_llr with a channel per codeword: sigma_factor is device [num_cw][2] floats
 * (sigma, LLR factor) as fecsim_chan_sigma returns them; everything else as
 * fecsim_gen_llr.  No reference counterpart (its SNR is per batch); used for batches
 * whose codewords see different SNRs. */
int fecsim_gen_llr_var(fecsim_henc *h, const float *sigma_factor, float max_abs_llr, uint64_t seed, uint64_t cw0,
                       uint32_t num_cw, int col_words, const uint32_t *bits, int elem, int layout_p, void *llr,
                       unsigned long long *uncoded_err, void *stream);

/* Replaces LDPC_count_errors() (src/Top/LDPC_count_errors.cpp:40-136): hard decisions
 * (value < 0) of the data columns of `post` against `bits`; adds the bit errors to
 * counts[0] and the codeword errors (one per interleaved codeword slot r % num_par
 * with an error) to counts[1] (device uint64[2]).  cw_bit_err (device [num_cw] or
 * NULL) receives each codeword's bit errors. */
int fecsim_count_errors(fecsim_henc *h, uint32_t num_cw, int col_words, const uint32_t *bits, int elem,
                        int layout_p, const void *post, unsigned long long *counts, uint32_t *cw_bit_err,
                        void *stream);

/* Hard decisions (value < 0 -> 1, as LDPC_count_errors.cpp:84) of every column of
 * num_cw decoded codewords (`post`, element `elem`, layout as fecsim_decode_set),
 * packed like the codeword bits of fecsim_gen_codewords: bits_out[num_cw][mcol][col_words],
 * bit r of column c at word r >> 5, bit r & 31.  The reference packs only the data
 * columns, in place, to compare them (LDPC_count_errors.cpp:67-96). */
int fecsim_hard_decisions(fecsim_henc *h, uint32_t num_cw, int col_words, int elem, int layout_p,
                          const void *post, uint32_t *bits_out, void *stream);

/* Histogram of per-codeword iteration counts (device iters[num_cw] as the decode calls
 * write them): hist[min(iters[i], nbins - 1)] += 1 (device uint64[nbins], added to),
 * nbins 1..1024.  The reference keeps only the sum (LDPCdecode.cpp:236,244). */
int fecsim_iter_histogram(const uint32_t *iters, uint32_t num_cw, uint32_t nbins, unsigned long long *hist,
                          void *stream);

/* ------------------------------------------------------------------------- */
/* The reference's own sequential streams (replay of one of its runs)        */
/* ------------------------------------------------------------------------- */

/* State of src/RNGfunc/prng_functions.cpp: s[8] four-lane 64-bit words seeded from the
 * 32 rand_state values (init_randstate :22-26, the values run_LDPCsim.pl passes or
 * read_sim_params.cpp:33-43 draws): sfc64 in s[0..3] for the noise, xoshiro256++ in
 * s[4..7] for the data bits.  The streams are sequential (normdist2 takes an extra draw
 * of all lanes when a value is tiny), so one wave replays them: for replaying a
 * reference run bit for bit, not a throughput path (fecsim_gen_codewords /
 * fecsim_gen_llr are). */
typedef struct fecsim_refrng fecsim_refrng;
int fecsim_refrng_create(const uint64_t *rand_state /* [32] */, fecsim_refrng **out);
void fecsim_refrng_free(fecsim_refrng *g);
/* normdist2(wgn, simd_count) (prng_functions.cpp:148-186): simd_count * 8 floats into
 * device wgn; simd_count even; the state advances as the reference's does */
int fecsim_refrng_normdist2(fecsim_refrng *g, float *wgn, uint32_t simd_count, void *stream);
/* gen_rand_databits(num_cw, Henc) (src/Top/gen_rand_databits.cpp:11-40): the data
 * columns of bits [num_cw][mcol][col_words] (Z <= 512), masked to Z rows and to the
 * non-filler rows; the parity columns are left to fecsim_encode */
int fecsim_refrng_databits(fecsim_refrng *g, fecsim_henc *h, uint32_t num_cw, int col_words, uint32_t *bits,
                           void *stream);
/* gen_metrics (src/Top/gen_metrics.cpp:18-121) on a given noise buffer in normdist2's
 * order (per codeword, column and group of eight rows, ceil(Z/8) groups; rows past Z
 * are cut off and not stored); otherwise as fecsim_gen_llr */
int fecsim_gen_llr_from_noise(fecsim_henc *h, const fecsim_chan_params *p, const float *wgn, uint32_t num_cw,
                              int col_words, const uint32_t *bits, int elem, int layout_p, void *llr,
                              unsigned long long *uncoded_err, void *stream);

/* ------------------------------------------------------------------------- */
/* Simulation loop                                                            */
/* ------------------------------------------------------------------------- */

typedef struct fecsim_sim fecsim_sim;

typedef struct {
    uint32_t num_cw;        /* codewords per inner loop (sim_struct.num_codeword) */
    uint32_t num_snr;       /* SNR points */
    const float *snr_db;    /* [num_snr], Es/N0 dB as gen_metrics */
    float llr_scale, max_abs_llr;
    uint64_t seed;
    int use_graph;          /* 1: capture one inner loop in a HIP graph and replay it */
} fecsim_sim_params;

/* per SNR point, accumulated since the last reset: the fields of the reference's
 * "loop" line (LDPCtop.cpp:127-129) */
typedef struct {
    unsigned long long cw, uncoded_bit_err, iters, bit_err, cw_err;
} fecsim_sim_stats;

/* Replaces the driver loop of src/Top/LDPCtop.cpp:86-121 for one decoder: every inner
 * loop generates num_cw new codewords (data bits + encoder), then for every SNR
 * point generates their LLRs (same noise at every point), decodes them in place
 * and counts errors, all on the device.  The decoder's element type sets the LLR
 * type; the layout is dense.  Owns its device buffers. */
int fecsim_sim_create(fecsim_henc *h, fecsim_decoder *d, const fecsim_sim_params *p, fecsim_sim **out);
/* enqueues `loops` inner loops on stream (asynchronous) */
int fecsim_sim_run(fecsim_sim *s, uint32_t loops, void *stream);
/* synchronises the device and copies the counters ([num_snr]) */
int fecsim_sim_stats_get(fecsim_sim *s, fecsim_sim_stats *out);
/* zeroes the counters (stream-ordered); the codeword index keeps advancing */
int fecsim_sim_reset(fecsim_sim *s, void *stream);
/* synchronises and copies the per-SNR iteration histograms accumulated since the last
 * reset: hist_out[num_snr][nbins] (host), bin i = codewords that stopped after i
 * iterations; nbins must be the decoder's max_iter + 1 (at most 1024, the last bin
 * then holds every longer count) */
int fecsim_sim_iter_hist(fecsim_sim *s, unsigned long long *hist_out, uint32_t nbins);
void fecsim_sim_free(fecsim_sim *s);

/* Mixed-code simulation loop: a fixed queue of num_cw entries, entry i drawn uniformly
 * from the queue decoder's codes and given an Eb/N0 drawn uniformly from [ebn0_lo,
 * ebn0_hi] dB (Es/N0 = Eb/N0 + 10 log10(K/N) of its code).  Every loop generates new
 * codewords and noise for all entries (encoder + channel of each code, LLRs of the
 * decoders' element type), decodes the interleaved queue with fecsim_multi_decode and
 * counts errors per code; with use_graph the loop is one replayed HIP graph.  The
 * per-code launches run on up to 8 side streams of the handle, forked from and joined
 * back into the caller's stream.  No reference counterpart (each reference process
 * simulates one code). */
typedef struct fecsim_msim fecsim_msim;
typedef struct {
    uint32_t num_cw;
    float ebn0_lo, ebn0_hi;
    float llr_scale, max_abs_llr;
    uint64_t seed;
    int use_graph;
} fecsim_msim_params;
/* hencs[c]: encoder tables of the queue decoder's code c */
int fecsim_msim_create(fecsim_henc *const *hencs, fecsim_multi *m, const fecsim_msim_params *p, fecsim_msim **out);
int fecsim_msim_run(fecsim_msim *s, uint32_t loops, void *stream);
/* synchronises; per_code[c] (iters field 0) and the summed iterations of all codes */
int fecsim_msim_stats(fecsim_msim *s, fecsim_sim_stats *per_code, unsigned long long *iters);
/* device views of the queue (code ids, LLR offsets), the packed LLR / posterior buffer and
 * the per-entry (sigma, factor) pairs in code order, for inspection */
int fecsim_msim_queue(const fecsim_msim *s, const int32_t **code, const unsigned long long **off, void **llr,
                      const float **sigma_factor);
/* bytes of the packed LLR / posterior buffer that fecsim_msim_queue returns */
int fecsim_msim_llr_bytes(const fecsim_msim *s, unsigned long long *bytes);
/* Writes the channel LLRs of the next loop (the codewords and noise the next
 * fecsim_msim_run decodes, which it generates again identically) into that buffer,
 * without decoding or counting: a queue of fresh channel outputs for timing or tuning
 * the queue decoder on (copy it aside first: the loop decodes in place). */
int fecsim_msim_fill(fecsim_msim *s, void *stream);
int fecsim_msim_reset(fecsim_msim *s, void *stream);
void fecsim_msim_free(fecsim_msim *s);

/* ------------------------------------------------------------------------- */
/* Differentiable decoder (no reference counterpart)                          */
/* ------------------------------------------------------------------------- */

/* Layered offset / normalized min-sum in float or double, unrolled for T iterations
 * (no syndrome exit), with an offset beta[t][l] and a scale alpha[t][l] per
 * (iteration, layer): C2V_j = sign * alpha * min(cap, min_{k != j} max(|q_k| - beta, 0)).
 * With alpha = 1 and beta = beta_offset it is the Vec16f_min / Vec8d_min layer update
 * (src/LDPCDec/decode_functions.cpp:11-103).  The forward keeps checkpoints
 * (posteriors and C2V) only at the starts of ceil(T/K) segments, the last segment K
 * iterations long, and the per-iteration state the adjoint needs for that last segment;
 * the adjoint recomputes each earlier segment from its checkpoint, keeping that
 * segment's state, and sweeps it backwards.  The state of an iteration is S values: its
 * V2C messages (S = E), or, for the kernels that hold the C2V messages in registers
 * (float32 or FECSIM_I8, Z <= 128, the compiled rate-1/2 shapes of 802.11n, 802.11ad and
 * 802.16e), each row's two smallest magnitudes and a word of indices and per-edge flags
 * (S = 3 mrow z, 4 mrow z with FECSIM_I8).  K is the segment length in 1..T of least
 * memory ceil(T/K) (N + E) + K S (the larger on a tie), of order sqrt(T (N + E) / S).
 * Device memory per codeword: ceil(T/K) (N + E) + K S + 2 E values (E instead of 2 E with the
 * register kernels), plus 2 N when the posteriors sit in global memory (N = mcol z variables, E =
 * total_hc_ones z edges). */
typedef struct fecsim_trainer fecsim_trainer;

typedef struct {
    uint32_t iters;      /* T */
    uint32_t max_batch;  /* codewords per call at most (sizes the checkpoints) */
    float cap;           /* row cap (max_row_metric) */
    int elem;            /* FECSIM_F32 or FECSIM_F64; or FECSIM_I8: float32 tensors emulating the
                            int8 decoder (Vec64c_min) -- LLRs given as integer-valued floats in
                            the decoder's units, V2C values and posteriors saturated to
                            [-128, 127], offsets rounded to integers (the gradient passes the
                            rounding unchanged, a saturated value passes none), an offset of 0
                            giving A = 0 at -128 as abs(-128) does; alpha should be 1 */
    uint32_t flags;      /* 0, or FECSIM_TRAIN_GLOBAL_C2V */
} fecsim_train_params;

/* fecsim_train_params.flags: keep the C2V messages in a global array even where the code's
 * shape has kernels that hold them in registers (float, Z <= 128, the compiled rate-1/2
 * shapes of 802.11n, 802.11ad and 802.16e).  Both give the same results bit for bit; the
 * register kernels are faster. */
#define FECSIM_TRAIN_GLOBAL_C2V 1u

int fecsim_trainer_create(const fecsim_hdec *h, const fecsim_train_params *p, fecsim_trainer **out);
/* llr, post: device [batch][mcol*z]; alpha, beta: device [T][mrow]; all of elem's type
 * (float32 for FECSIM_I8).  The trainer copies alpha and beta on `stream` for the
 * backward: the caller's buffers may be freed or changed once the call is ordered. */
int fecsim_train_fwd(fecsim_trainer *t, uint32_t batch, const void *llr, const void *alpha, const void *beta,
                     void *post, void *stream);
/* Gradients of a loss L given dL/d(post) of the last forward (same batch; one backward
 * per forward, FECSIM_E_INVAL otherwise): grad_alpha, grad_beta [T][mrow] (overwritten,
 * summed over the batch), grad_llr [batch][mcol*z] or NULL. */
int fecsim_train_bwd(fecsim_trainer *t, uint32_t batch, const void *grad_post, void *grad_alpha, void *grad_beta,
                     void *grad_llr, void *stream);
void fecsim_trainer_free(fecsim_trainer *t);

#ifdef __cplusplus
}
#endif
#endif /* FECSIM_H */
