/*
 * SIMD CPU decoder LDPCtop_Vec64c_min for the CPU baseline -- TEST INFRASTRUCTURE.
 * One 64-lane int8 word = one __m512i (AVX-512BW) or two __m256i (AVX2); the
 * saturating/abs/min/max ops are the x86 instructions vectorclass maps Vec64c to
 * (include/simd_defs.h:34-37, decode_functions.cpp:48-97).
 */
#include <immintrin.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---------------- AVX-512BW ---------------- */
#pragma GCC push_options
#pragma GCC target("avx512f,avx512bw")
#define VEC __m512i
#define FN(x) avx512_##x
#define V_LOADU(p) _mm512_loadu_si512((const void *)(p))
#define V_STOREU(p, v) _mm512_storeu_si512((void *)(p), (v))
#define V_STOREN(p, v, n) _mm512_mask_storeu_epi8((void *)(p), (n) >= 64 ? ~0ull : ((1ull << (n)) - 1), (v))
#define V_SET1(x) _mm512_set1_epi8((char)(x))
#define V_ZERO() _mm512_setzero_si512()
#define V_SUBS(a, b) _mm512_subs_epi8((a), (b))
#define V_ADDS(a, b) _mm512_adds_epi8((a), (b))
#define V_SUB(a, b) _mm512_sub_epi8((a), (b))
#define V_ABS(a) _mm512_abs_epi8(a)
#define V_MAX(a, b) _mm512_max_epi8((a), (b))
#define V_MIN(a, b) _mm512_min_epi8((a), (b))
#define V_XOR(a, b) _mm512_xor_si512((a), (b))
#define V_OR(a, b) _mm512_or_si512((a), (b))
#define V_SELEQ(a, b, x, y) _mm512_mask_blend_epi8(_mm512_cmpeq_epi8_mask((a), (b)), (y), (x))
#define V_NEGIF(s, x) _mm512_mask_sub_epi8((x), _mm512_movepi8_mask(s), _mm512_setzero_si512(), (x))
#define V_KEEPN(x, n) _mm512_maskz_mov_epi8((n) >= 64 ? ~0ull : ((1ull << (n)) - 1), (x))
#define V_ANYNEG(x) (_mm512_movepi8_mask(x) != 0)
#include "baseline_vec_body.h"
#undef VEC
#undef FN
#undef V_LOADU
#undef V_STOREU
#undef V_STOREN
#undef V_SET1
#undef V_ZERO
#undef V_SUBS
#undef V_ADDS
#undef V_SUB
#undef V_ABS
#undef V_MAX
#undef V_MIN
#undef V_XOR
#undef V_OR
#undef V_SELEQ
#undef V_NEGIF
#undef V_KEEPN
#undef V_ANYNEG
#pragma GCC pop_options

/* ---------------- AVX2: one 64-lane word = two __m256i ---------------- */
#pragma GCC push_options
#pragma GCC target("avx2")
typedef struct { __m256i lo, hi; } v64x2;
#define VEC v64x2
#define FN(x) avx2_##x
#define BIN(f, a, b) ({ v64x2 _a = (a), _b = (b), _r; _r.lo = f(_a.lo, _b.lo); _r.hi = f(_a.hi, _b.hi); _r; })
static inline v64x2 a2_load(const void *p) {
    v64x2 r;
    r.lo = _mm256_loadu_si256((const __m256i *)p);
    r.hi = _mm256_loadu_si256((const __m256i *)((const char *)p + 32));
    return r;
}
static inline void a2_store(void *p, v64x2 v) {
    _mm256_storeu_si256((__m256i *)p, v.lo);
    _mm256_storeu_si256((__m256i *)((char *)p + 32), v.hi);
}
static inline void a2_storen(void *p, v64x2 v, int n) {
    char tmp[64];
    a2_store(tmp, v);
    memcpy(p, tmp, (size_t)n);
}
static inline v64x2 a2_set1(char x) { v64x2 r; r.lo = r.hi = _mm256_set1_epi8(x); return r; }
static inline v64x2 a2_abs(v64x2 a) { a.lo = _mm256_abs_epi8(a.lo); a.hi = _mm256_abs_epi8(a.hi); return a; }
static inline v64x2 a2_seleq(v64x2 a, v64x2 b, v64x2 x, v64x2 y) {
    v64x2 r;
    r.lo = _mm256_blendv_epi8(y.lo, x.lo, _mm256_cmpeq_epi8(a.lo, b.lo));
    r.hi = _mm256_blendv_epi8(y.hi, x.hi, _mm256_cmpeq_epi8(a.hi, b.hi));
    return r;
}
static inline v64x2 a2_negif(v64x2 s, v64x2 x) {
    v64x2 r, z;
    z.lo = z.hi = _mm256_setzero_si256();
    r.lo = _mm256_blendv_epi8(x.lo, _mm256_sub_epi8(z.lo, x.lo), s.lo);
    r.hi = _mm256_blendv_epi8(x.hi, _mm256_sub_epi8(z.hi, x.hi), s.hi);
    return r;
}
static inline v64x2 a2_keepn(v64x2 x, int n) {
    char tmp[64];
    a2_store(tmp, x);
    if (n < 64) memset(tmp + n, 0, (size_t)(64 - n));
    return a2_load(tmp);
}
static inline int a2_anyneg(v64x2 x) { return (_mm256_movemask_epi8(x.lo) | _mm256_movemask_epi8(x.hi)) != 0; }
#define V_LOADU(p) a2_load((const void *)(p))
#define V_STOREU(p, v) a2_store((void *)(p), (v))
#define V_STOREN(p, v, n) a2_storen((void *)(p), (v), (n))
#define V_SET1(x) a2_set1((char)(x))
#define V_ZERO() a2_set1(0)
#define V_SUBS(a, b) BIN(_mm256_subs_epi8, a, b)
#define V_ADDS(a, b) BIN(_mm256_adds_epi8, a, b)
#define V_SUB(a, b) BIN(_mm256_sub_epi8, a, b)
#define V_ABS(a) a2_abs(a)
#define V_MAX(a, b) BIN(_mm256_max_epi8, a, b)
#define V_MIN(a, b) BIN(_mm256_min_epi8, a, b)
#define V_XOR(a, b) BIN(_mm256_xor_si256, a, b)
#define V_OR(a, b) BIN(_mm256_or_si256, a, b)
#define V_SELEQ(a, b, x, y) a2_seleq((a), (b), (x), (y))
#define V_NEGIF(s, x) a2_negif((s), (x))
#define V_KEEPN(x, n) a2_keepn((x), (n))
#define V_ANYNEG(x) a2_anyneg(x)
#include "baseline_vec_body.h"
#pragma GCC pop_options

int64_t orc_decode_vec64c(const orc_hdec *h, uint32_t num_cw, uint32_t max_iter, float max_row_metric,
                          float beta_offset, int8_t *metrics, uint32_t *iters, int nthreads, int *isa_out) {
    if (!h || h->z <= 0 || max_iter == 0) return -1;
    for (int l = 0; l < h->mrow; ++l)
        if (h->row_weights[l] > 64) return -2;
    /* (int8) truncation of the float parameters, decode_functions.cpp:28-32,60 */
    int8_t beta8 = (int8_t)(int)beta_offset;
    int8_t cap8 = max_row_metric < 1e-6f ? (int8_t)(int)(1.0 - (double)(int8_t)(int)max_row_metric)
                                         : (int8_t)(int)max_row_metric;
    __builtin_cpu_init();
    if (__builtin_cpu_supports("avx512bw")) {
        if (isa_out) *isa_out = 512;
        return avx512_run(h, num_cw, max_iter, beta8, cap8, metrics, iters, nthreads);
    }
    if (isa_out) *isa_out = 256;
    return avx2_run(h, num_cw, max_iter, beta8, cap8, metrics, iters, nthreads);
}
