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
 * compiled for it (even Z <= 64, int8 min-sum, 1 <= beta, 0 <= cap, dense layout) and
 * the generic kernel otherwise; both give identical results.  No reference
 * counterpart (the reference selects its decoder at compile time only). */
enum { FECSIM_KERNEL_AUTO = 0, FECSIM_KERNEL_GENERIC = 1, FECSIM_KERNEL_PAIR = 2 };
int fecsim_decoder_select_kernel(fecsim_decoder *d, int kernel);
/* kernel a dense decode_set call would run now (FECSIM_KERNEL_GENERIC or _PAIR) */
int fecsim_decoder_kernel(const fecsim_decoder *d);

/* element kind (FECSIM_I8..F64), mode (FECSIM_MIN/PROD), bytes per element */
int fecsim_decoder_info(const fecsim_decoder *d, int *elem, int *mode, int *elem_bytes);
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
 *            (the return value of LDPCdecode_set) */
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

#ifdef __cplusplus
}
#endif
#endif /* FECSIM_H */
