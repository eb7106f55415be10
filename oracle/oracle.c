/*
 * CPU oracle for the LDPC simulator hot path -- TEST INFRASTRUCTURE ONLY.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * as the checker.  Every function cites the reference file:line it restates.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off, so float results do not
 * depend on FMA contraction; explicit fmaf() is used where the reference's
 * vectorclass mul_add is a fused multiply-add on its AVX2/AVX-512 targets).
 *
 * Own float math (orc_tanhf / orc_atanhf / log / expm1): the reference links
 * vectorclass's tanh/atanh (min_find_wrap.pl:12-16), which is not available here;
 * these are written from the mathematics (Cody-Waite range reduction of exp with a
 * Taylor polynomial; log from the atanh series of (m-1)/(m+1); atanh from its Taylor
 * series below 0.25 and 0.5*log((1+x)/(1-x)) above).  The product's HIP kernels
 * evaluate the same expressions in the same order.  Their bit patterns against
 * vectorclass are "parity unpinned".
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* own math                                                                  */
/* ------------------------------------------------------------------------ */

#define LN2_HI_F 0.693145751953125f
#define LN2_LO_F 1.42860677e-06f
#define LN2_HI_D 0.6931471803691238
#define LN2_LO_D 1.9082149292705877e-10

static float expm1_poly_f(float r) {
    /* r + r^2/2! + ... + r^9/9! (Horner) */
    float p = 1.0f / 362880.0f;
    p = p * r + 1.0f / 40320.0f;
    p = p * r + 1.0f / 5040.0f;
    p = p * r + 1.0f / 720.0f;
    p = p * r + 1.0f / 120.0f;
    p = p * r + 1.0f / 24.0f;
    p = p * r + 1.0f / 6.0f;
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    return p * r;
}

static float pow2i_f(int k) { /* 2^k for -126 <= k <= 127 */
    union { uint32_t u; float f; } v;
    v.u = (uint32_t)(k + 127) << 23;
    return v.f;
}

static float own_expm1f(float y) {
    if (y < -88.0f) return -1.0f;
    if (y > 88.0f) return INFINITY;
    if (fabsf(y) < 0.34657359f) return expm1_poly_f(y);
    float k = rintf(y * 1.44269504f);
    float r = (y - k * LN2_HI_F) - k * LN2_LO_F;
    float p = expm1_poly_f(r);
    float s = pow2i_f((int)k);
    return s * p + (s - 1.0f);
}

static float own_logf(float v) { /* v > 0, normal */
    union { float f; uint32_t u; } b;
    b.f = v;
    int e = (int)((b.u >> 23) & 0xff) - 127;
    b.u = (b.u & 0x007fffffu) | 0x3f800000u; /* m in [1,2) */
    float m = b.f;
    if (m > 1.41421356f) { m = m * 0.5f; e += 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 1.0f / 13.0f;
    p = p * s2 + 1.0f / 11.0f;
    p = p * s2 + 1.0f / 9.0f;
    p = p * s2 + 1.0f / 7.0f;
    p = p * s2 + 1.0f / 5.0f;
    p = p * s2 + 1.0f / 3.0f;
    p = p * s2 + 1.0f;
    float lm = 2.0f * s * p;
    float fe = (float)e;
    return fe * LN2_HI_F + (fe * LN2_LO_F + lm);
}

float orc_tanhf(float x) {
    float a = fabsf(x);
    float e = own_expm1f(-2.0f * a);
    float t = (0.0f - e) / (2.0f + e);
    return x < 0.0f ? -t : t;
}

float orc_atanhf(float x) {
    float a = fabsf(x), r;
    if (a < 0.25f) {
        float s2 = a * a;
        float p = 1.0f / 13.0f;
        p = p * s2 + 1.0f / 11.0f;
        p = p * s2 + 1.0f / 9.0f;
        p = p * s2 + 1.0f / 7.0f;
        p = p * s2 + 1.0f / 5.0f;
        p = p * s2 + 1.0f / 3.0f;
        r = a + a * s2 * p;
    } else {
        r = 0.5f * own_logf((1.0f + a) / (1.0f - a));
    }
    return x < 0.0f ? -r : r;
}

static double expm1_poly_d(double r) {
    double p = 1.0 / 6227020800.0; /* 1/13! */
    static const double c[] = {1.0 / 479001600.0, 1.0 / 39916800.0, 1.0 / 3628800.0, 1.0 / 362880.0,
                               1.0 / 40320.0,     1.0 / 5040.0,     1.0 / 720.0,     1.0 / 120.0,
                               1.0 / 24.0,        1.0 / 6.0,        0.5,             1.0};
    for (int i = 0; i < 12; ++i) p = p * r + c[i];
    return p * r;
}

static double own_expm1(double y) {
    if (y < -708.0) return -1.0;
    if (y > 709.0) return INFINITY;
    if (fabs(y) < 0.34657359027997264) return expm1_poly_d(y);
    double k = rint(y * 1.4426950408889634);
    double r = (y - k * LN2_HI_D) - k * LN2_LO_D;
    double p = expm1_poly_d(r);
    union { uint64_t u; double d; } s;
    s.u = (uint64_t)((int)k + 1023) << 52;
    return s.d * p + (s.d - 1.0);
}

static double own_log(double v) {
    union { double d; uint64_t u; } b;
    b.d = v;
    int e = (int)((b.u >> 52) & 0x7ff) - 1023;
    b.u = (b.u & 0x000fffffffffffffull) | 0x3ff0000000000000ull;
    double m = b.d;
    if (m > 1.4142135623730951) { m = m * 0.5; e += 1; }
    double s = (m - 1.0) / (m + 1.0), s2 = s * s;
    double p = 1.0 / 21.0;
    for (int k = 19; k >= 1; k -= 2) p = p * s2 + 1.0 / (double)k;
    double fe = (double)e;
    return fe * LN2_HI_D + (fe * LN2_LO_D + 2.0 * s * p);
}

double orc_tanh(double x) {
    double a = fabs(x);
    double e = own_expm1(-2.0 * a);
    double t = (0.0 - e) / (2.0 + e);
    return x < 0.0 ? -t : t;
}

double orc_atanh(double x) {
    double a = fabs(x), r;
    if (a < 0.25) {
        double s2 = a * a, p = 1.0 / 29.0;
        for (int k = 27; k >= 3; k -= 2) p = p * s2 + 1.0 / (double)k;
        r = a + a * s2 * p;
    } else {
        r = 0.5 * own_log((1.0 + a) / (1.0 - a));
    }
    return x < 0.0 ? -r : r;
}

/* ------------------------------------------------------------------------ */
/* exclusive min / product tree (min_find.pl:29-164)                         */
/* ------------------------------------------------------------------------ */

#define MAXW 64
typedef struct {
    int nops, nslots;
    int (*op)[3];   /* dst slot, a slot (-1: cap), b slot */
    int out[MAXW];  /* slot of output j */
} tree_t;
static tree_t g_tree[MAXW];

int orc_set_tree(int w, int nops, const int64_t *ops) {
    if (w < 2 || w >= MAXW || nops <= 0) return -1;
    int64_t *masks = (int64_t *)malloc(sizeof(int64_t) * (size_t)(w + nops));
    int (*op)[3] = (int (*)[3])malloc(sizeof(int[3]) * (size_t)nops);
    int ns = 0;
    for (int i = 0; i < w; ++i) masks[ns++] = (int64_t)1 << i;
#define FIND(m) ({ int _s = -1; for (int _k = 0; _k < ns; ++_k) if (masks[_k] == (m)) { _s = _k; break; } _s; })
    for (int i = 0; i < nops; ++i) {
        int64_t d = ops[3 * i], a = ops[3 * i + 1], b = ops[3 * i + 2];
        if (a == -1) {
            int s = FIND(d);
            if (s < 0) { free(masks); free(op); return -2; }
            op[i][0] = s; op[i][1] = -1; op[i][2] = -1;
        } else {
            int sa = FIND(a), sb = FIND(b);
            if (sa < 0 || sb < 0) { free(masks); free(op); return -3; }
            int sd = FIND(d);
            if (sd < 0) { masks[ns] = d; sd = ns++; }
            op[i][0] = sd; op[i][1] = sa; op[i][2] = sb;
        }
    }
    int64_t all = ((int64_t)1 << w) - 1;
    tree_t *t = &g_tree[w];
    for (int j = 0; j < w; ++j) {
        int s = FIND(all ^ ((int64_t)1 << j));
        if (s < 0) { free(masks); free(op); return -4; }
        t->out[j] = s;
    }
#undef FIND
    free(t->op);
    t->op = op;
    t->nops = nops;
    t->nslots = ns;
    free(masks);
    return 0;
}

/* _mm_min_ps(a, b) semantics: a < b ? a : b (returns b on NaN/equal) */
#define VMIN(a, b) ((a) < (b) ? (a) : (b))

/* evaluate the min tree on doubles (exact for min on any element type's values) */
static int tree_min(int w, const double *in, double cap, double *out) {
    tree_t *t = &g_tree[w];
    if (!t->op) return -1;
    double v[4 * MAXW + 8];
    for (int i = 0; i < w; ++i) v[i] = in[i];
    for (int i = 0; i < t->nops; ++i) {
        int d = t->op[i][0], a = t->op[i][1], b = t->op[i][2];
        if (a < 0) v[d] = cap < v[d] ? cap : v[d];     /* min(Max_val, m) */
        else v[d] = VMIN(v[a], v[b]);                 /* min(ma, mb) */
    }
    for (int j = 0; j < w; ++j) out[j] = v[t->out[j]];
    return 0;
}

static int tree_prod_f(int w, const float *in, float cap, float *out) {
    tree_t *t = &g_tree[w];
    if (!t->op) return -1;
    float v[4 * MAXW + 8];
    for (int i = 0; i < w; ++i) v[i] = orc_tanhf(in[i] * 0.5f);
    for (int i = 0; i < t->nops; ++i) {
        int d = t->op[i][0], a = t->op[i][1], b = t->op[i][2];
        if (a < 0) v[d] = VMIN(cap, v[d]);
        else v[d] = v[a] * v[b];
    }
    for (int j = 0; j < w; ++j) out[j] = 2.0f * orc_atanhf(v[t->out[j]]);
    return 0;
}

static int tree_prod_d(int w, const double *in, double cap, double *out) {
    tree_t *t = &g_tree[w];
    if (!t->op) return -1;
    double v[4 * MAXW + 8];
    for (int i = 0; i < w; ++i) v[i] = orc_tanh(in[i] * 0.5);
    for (int i = 0; i < t->nops; ++i) {
        int d = t->op[i][0], a = t->op[i][1], b = t->op[i][2];
        if (a < 0) v[d] = VMIN(cap, v[d]);
        else v[d] = v[a] * v[b];
    }
    for (int j = 0; j < w; ++j) out[j] = 2.0 * orc_atanh(v[t->out[j]]);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* integer element semantics (x86 epi8/epi16/epi32 as vectorclass maps them) */
/* ------------------------------------------------------------------------ */

static inline int64_t ilo(int bits) { return -((int64_t)1 << (bits - 1)); }
static inline int64_t ihi(int bits) { return ((int64_t)1 << (bits - 1)) - 1; }
static inline int64_t isat(int64_t x, int bits) {           /* subs/adds_epiN */
    int64_t lo = ilo(bits), hi = ihi(bits);
    return x < lo ? lo : (x > hi ? hi : x);
}
static inline int64_t iwrap(int64_t x, int bits) {          /* sub_epiN */
    uint64_t m = (bits == 64) ? ~0ull : (((uint64_t)1 << bits) - 1);
    uint64_t u = (uint64_t)x & m;
    if (u >> (bits - 1)) u |= ~m;
    return (int64_t)u;
}
static inline int64_t iwabs(int64_t x, int bits) {          /* abs_epiN: abs(MIN) = MIN */
    return x == ilo(bits) ? x : (x < 0 ? -x : x);
}

static int elem_bits(int elem) { return elem == ORC_I8 ? 8 : elem == ORC_I16 ? 16 : 32; }
static size_t elem_size(int elem) {
    switch (elem) {
    case ORC_I8: return 1;
    case ORC_I16: return 2;
    case ORC_I32: return 4;
    case ORC_F32: return 4;
    default: return 8;
    }
}
static int64_t iget(const void *p, size_t i, int elem) {
    switch (elem) {
    case ORC_I8: return ((const int8_t *)p)[i];
    case ORC_I16: return ((const int16_t *)p)[i];
    default: return ((const int32_t *)p)[i];
    }
}
static void iput(void *p, size_t i, int elem, int64_t v) {
    switch (elem) {
    case ORC_I8: ((int8_t *)p)[i] = (int8_t)v; break;
    case ORC_I16: ((int16_t *)p)[i] = (int16_t)v; break;
    default: ((int32_t *)p)[i] = (int32_t)v; break;
    }
}

/* beta and cap as the reference's implicit conversions give them
 * (decode_functions.cpp:28-32,60-61: float -> int (truncation) -> element broadcast) */
static int64_t int_beta(float beta, int bits) { return iwrap((int64_t)(int)beta, bits); }
static int64_t int_cap(float mrm, int bits) {
    if (mrm < 1e-6f) {
        int64_t t = iwrap((int64_t)(int)mrm, bits);  /* (CTYPE)max_row_metric */
        return iwrap((int64_t)(int)(1.0 - (double)t), bits);
    }
    return iwrap((int64_t)(int)mrm, bits);
}
static float flt_cap(float mrm) { return mrm < 1e-6f ? (float)(1.0 - (double)(float)mrm) : mrm; }
static double dbl_cap(float mrm) { return mrm < 1e-6f ? 1.0 - (double)mrm : (double)mrm; }

/* ------------------------------------------------------------------------ */
/* ideal layered decoder                                                     */
/* ------------------------------------------------------------------------ */

/* per-(iteration, layer) offsets replacing beta (converted the same way), or NULL */
static const float *g_beta_tab = NULL;
void orc_set_beta_table(const float *tab) { g_beta_tab = tab; }

static int64_t ideal_int(const orc_hdec *h, int elem, uint32_t num_cw, uint32_t max_iter, float mrm, float beta,
                         void *llr, void *c2v, uint32_t *iters, int use_tree) {
    const int bits = elem_bits(elem);
    const int Z = h->z, N = h->mcol * h->z, E = h->total_hc_ones * h->z;
    const int64_t b = int_beta(beta, bits), cap = int_cap(mrm, bits);
    int64_t *L = (int64_t *)malloc(sizeof(int64_t) * (size_t)N);
    int64_t *R = (int64_t *)malloc(sizeof(int64_t) * (size_t)E);
    int64_t tot = 0;
    for (uint32_t cw = 0; cw < num_cw; ++cw) {
        for (int v = 0; v < N; ++v) L[v] = iget(llr, (size_t)cw * N + v, elem);
        memset(R, 0, sizeof(int64_t) * (size_t)E);               /* LDPCdecode.cpp:115-119 */
        uint32_t iter = 0;
        int done = 0;
        while (!done) {
            int e0 = 0;
            for (int l = 0; l < h->mrow; ++l) {
                const int w = h->row_weights[l];
                const int64_t bl = g_beta_tab ? int_beta(g_beta_tab[iter * h->mrow + l], bits) : b;
                for (int r = 0; r < Z; ++r) {
                    int64_t Q[MAXW];
                    double A[MAXW], M[MAXW];
                    int vidx[MAXW];
                    int sgn = 0;
                    for (int j = 0; j < w; ++j) {                   /* decode_functions.cpp:39-69 */
                        const int e = e0 + j;
                        vidx[j] = h->row_active_cols[e] * Z + (h->row_offset[e] + r) % Z;
                        Q[j] = isat(L[vidx[j]] - R[e * Z + r], bits);
                        int64_t a = iwrap(iwabs(Q[j], bits) - bl, bits);
                        A[j] = (double)(a > 0 ? a : 0);
                        sgn ^= (Q[j] < 0);
                    }
                    if (use_tree) {
                        tree_min(w, A, (double)cap, M);
                    } else {
                        double m1 = INFINITY, m2 = INFINITY;
                        int im = -1;
                        for (int j = 0; j < w; ++j) {
                            if (A[j] < m1) { m2 = m1; m1 = A[j]; im = j; }
                            else if (A[j] < m2) m2 = A[j];
                        }
                        for (int j = 0; j < w; ++j) {
                            double x = (j == im) ? m2 : m1;
                            M[j] = (double)cap < x ? (double)cap : x;
                        }
                    }
                    for (int j = 0; j < w; ++j) {                   /* decode_functions.cpp:75-102 */
                        const int e = e0 + j;
                        int64_t m = (int64_t)M[j];
                        int neg = sgn ^ (Q[j] < 0);
                        int64_t rn = neg ? iwrap(-m, bits) : m;
                        L[vidx[j]] = isat(Q[j] + rn, bits);
                        R[e * Z + r] = rn;
                    }
                }
                e0 += w;
            }
            int fail = 0;                                           /* LDPCdecode.cpp:172-195 */
            e0 = 0;
            for (int l = 0; l < h->mrow && !fail; ++l) {
                const int w = h->row_weights[l];
                for (int r = 0; r < Z && !fail; ++r) {
                    int par = 0;
                    for (int j = 0; j < w; ++j) {
                        const int e = e0 + j;
                        par ^= (L[h->row_active_cols[e] * Z + (h->row_offset[e] + r) % Z] < 0);
                    }
                    fail |= par;
                }
                e0 += w;
            }
            done = !fail;
            ++iter;
            done = done || (iter == max_iter);
        }
        for (int v = 0; v < N; ++v) iput(llr, (size_t)cw * N + v, elem, L[v]);
        if (c2v)
            for (int e = 0; e < E; ++e) iput(c2v, (size_t)cw * E + e, elem, R[e]);
        if (iters) iters[cw] = iter;
        tot += iter;
    }
    free(L);
    free(R);
    return tot;
}

/* float / double: no saturation; sign handling on the IEEE sign bit
 * (vectorclass sign_combine = XOR of sign bits; syndrome reads the sign bit, so
 * -0.0 counts as a 1 there, LDPCdecode.cpp:192-193) */
#define DEFINE_IDEAL_FP(NAME, FT, FABS, CAPF, TREEPROD)                                                      \
    static int64_t NAME(const orc_hdec *h, int mode, uint32_t num_cw, uint32_t max_iter, float mrm, float beta, \
                        FT *llr, FT *c2v, uint32_t *iters, int use_tree) {                                      \
        const int Z = h->z, N = h->mcol * h->z, E = h->total_hc_ones * h->z;                                    \
        const FT b = (FT)beta, cap = (FT)CAPF(mrm);                                                             \
        FT *L = (FT *)malloc(sizeof(FT) * (size_t)N);                                                           \
        FT *R = (FT *)malloc(sizeof(FT) * (size_t)E);                                                           \
        int64_t tot = 0;                                                                                        \
        for (uint32_t cw = 0; cw < num_cw; ++cw) {                                                              \
            memcpy(L, llr + (size_t)cw * N, sizeof(FT) * (size_t)N);                                            \
            for (int e = 0; e < E; ++e) R[e] = (FT)0;                                                           \
            uint32_t iter = 0;                                                                                  \
            int done = 0;                                                                                       \
            while (!done) {                                                                                     \
                int e0 = 0;                                                                                     \
                for (int l = 0; l < h->mrow; ++l) {                                                             \
                    const int w = h->row_weights[l];                                                            \
                    const FT bl = g_beta_tab ? (FT)g_beta_tab[iter * h->mrow + l] : b;                          \
                    for (int r = 0; r < Z; ++r) {                                                               \
                        FT Q[MAXW], A[MAXW], M[MAXW];                                                           \
                        int vidx[MAXW];                                                                         \
                        int sgn = 0;                                                                            \
                        for (int j = 0; j < w; ++j) {                                                           \
                            const int e = e0 + j;                                                               \
                            vidx[j] = h->row_active_cols[e] * Z + (h->row_offset[e] + r) % Z;                   \
                            Q[j] = L[vidx[j]] - R[e * Z + r];                                                   \
                            FT a = FABS(Q[j]) - bl;                                                             \
                            A[j] = a > (FT)0 ? a : (FT)0; /* _mm_max_ps(a, 0) */                                \
                            sgn ^= (signbit(Q[j]) != 0);                                                        \
                        }                                                                                       \
                        if (mode == ORC_PROD) {                                                                 \
                            TREEPROD(w, A, cap, M);                                                             \
                        } else if (use_tree) {                                                                  \
                            double Ad[MAXW], Md[MAXW];                                                          \
                            for (int j = 0; j < w; ++j) Ad[j] = A[j];                                           \
                            tree_min(w, Ad, (double)cap, Md);                                                   \
                            for (int j = 0; j < w; ++j) M[j] = (FT)Md[j];                                       \
                        } else {                                                                                \
                            FT m1 = (FT)INFINITY, m2 = (FT)INFINITY;                                            \
                            int im = -1;                                                                        \
                            for (int j = 0; j < w; ++j) {                                                       \
                                if (A[j] < m1) { m2 = m1; m1 = A[j]; im = j; }                                  \
                                else if (A[j] < m2) m2 = A[j];                                                  \
                            }                                                                                   \
                            for (int j = 0; j < w; ++j) {                                                       \
                                FT x = (j == im) ? m2 : m1;                                                     \
                                M[j] = cap < x ? cap : x;                                                       \
                            }                                                                                   \
                        }                                                                                       \
                        for (int j = 0; j < w; ++j) {                                                           \
                            const int e = e0 + j;                                                               \
                            int neg = sgn ^ (int)(signbit(Q[j]) != 0);                                          \
                            FT rn = neg ? -M[j] : M[j];                                                         \
                            L[vidx[j]] = Q[j] + rn;                                                             \
                            R[e * Z + r] = rn;                                                                  \
                        }                                                                                       \
                    }                                                                                           \
                    e0 += w;                                                                                    \
                }                                                                                               \
                int fail = 0;                                                                                   \
                e0 = 0;                                                                                         \
                for (int l = 0; l < h->mrow && !fail; ++l) {                                                    \
                    const int w = h->row_weights[l];                                                            \
                    for (int r = 0; r < Z && !fail; ++r) {                                                      \
                        int par = 0;                                                                            \
                        for (int j = 0; j < w; ++j) {                                                           \
                            const int e = e0 + j;                                                               \
                            par ^= (signbit(L[h->row_active_cols[e] * Z + (h->row_offset[e] + r) % Z]) != 0);   \
                        }                                                                                       \
                        fail |= par;                                                                            \
                    }                                                                                           \
                    e0 += w;                                                                                    \
                }                                                                                               \
                done = !fail;                                                                                   \
                ++iter;                                                                                         \
                done = done || (iter == max_iter);                                                              \
            }                                                                                                   \
            memcpy(llr + (size_t)cw * N, L, sizeof(FT) * (size_t)N);                                            \
            if (c2v) memcpy(c2v + (size_t)cw * E, R, sizeof(FT) * (size_t)E);                                   \
            if (iters) iters[cw] = iter;                                                                        \
            tot += iter;                                                                                        \
        }                                                                                                       \
        free(L);                                                                                                \
        free(R);                                                                                                \
        return tot;                                                                                             \
    }

DEFINE_IDEAL_FP(ideal_f32, float, fabsf, flt_cap, tree_prod_f)
DEFINE_IDEAL_FP(ideal_f64, double, fabs, dbl_cap, tree_prod_d)

static int check_tables(const orc_hdec *h) {
    if (!h || h->z <= 0 || h->mrow <= 0 || h->mcol <= h->mrow) return -1;
    int s = 0;
    for (int l = 0; l < h->mrow; ++l) {
        if (h->row_weights[l] < 2 || h->row_weights[l] >= MAXW) return -2;
        s += h->row_weights[l];
    }
    if (s != h->total_hc_ones) return -3;
    for (int e = 0; e < s; ++e)
        if (h->row_active_cols[e] < 0 || h->row_active_cols[e] >= h->mcol || h->row_offset[e] < 0) return -4;
    return 0;
}

int64_t orc_decode_ideal(const orc_hdec *h, int elem, int mode, uint32_t num_cw, uint32_t max_iter,
                         float max_row_metric, float beta_offset, void *llr, void *c2v, uint32_t *iters,
                         int use_tree) {
    int rc = check_tables(h);
    if (rc) return rc;
    if (max_iter == 0) return -10;
    if (elem == ORC_F32)
        return ideal_f32(h, mode, num_cw, max_iter, max_row_metric, beta_offset, (float *)llr, (float *)c2v, iters,
                         use_tree);
    if (elem == ORC_F64)
        return ideal_f64(h, mode, num_cw, max_iter, max_row_metric, beta_offset, (double *)llr, (double *)c2v,
                         iters, use_tree);
    if (mode != ORC_MIN) return -11;
    return ideal_int(h, elem, num_cw, max_iter, max_row_metric, beta_offset, llr, c2v, iters, use_tree);
}

/* ------------------------------------------------------------------------ */
/* literal restatement of LDPCdecode.cpp (P-lane words, guard words, copies)  */
/* ------------------------------------------------------------------------ */

/* lane values are held as double for float/double and int64 for ints; the buffer
 * keeps the element type so partial stores and guard copies are exact. */
typedef struct {
    int elem, bits, P;
} lctx;

static double lget(const lctx *c, const void *buf, long i) {
    if (c->elem == ORC_F32) return ((const float *)buf)[i];
    if (c->elem == ORC_F64) return ((const double *)buf)[i];
    return (double)iget(buf, (size_t)i, c->elem);
}
static void lput(const lctx *c, void *buf, long i, double v) {
    if (c->elem == ORC_F32) ((float *)buf)[i] = (float)v;
    else if (c->elem == ORC_F64) ((double *)buf)[i] = v;
    else iput(buf, (size_t)i, c->elem, (int64_t)v);
}

int64_t orc_decode_literal(const orc_hdec *h, int elem, int mode, int P, uint32_t num_cw, uint32_t max_iter,
                           float mrm, float beta, void *metrics, void *c2v_out, uint32_t *iters) {
    int rc = check_tables(h);
    if (rc) return rc;
    if (P <= 0 || max_iter == 0) return -10;
    if (mode == ORC_PROD && elem != ORC_F32 && elem != ORC_F64) return -11;
    const lctx C = {elem, (elem <= ORC_I32) ? elem_bits(elem) : 0, P};
    const int isint = elem <= ORC_I32;
    const int Z = h->z, SPZ = (Z + P - 1) / P, Zrem = Z - (SPZ - 1) * P;
    const size_t es = elem_size(elem);
    const long cw_elems = (long)h->mcol * (SPZ + 2) * P;
    const int maxw = h->max_row_weight;
    /* scratch (words of P lanes): local_metrics, group_min, fin_min, row_xor */
    void *local = calloc((size_t)h->total_hc_ones * SPZ * P, es);
    double *gmin = (double *)malloc(sizeof(double) * (size_t)maxw * SPZ * P);
    double *fmin = (double *)malloc(sizeof(double) * (size_t)maxw * SPZ * P);
    uint8_t *rxor = (uint8_t *)malloc((size_t)SPZ * P); /* sign parity per lane */
    int64_t bI = isint ? int_beta(beta, C.bits) : 0, capI = isint ? int_cap(mrm, C.bits) : 0;
    double bF = (elem == ORC_F32) ? (double)(float)beta : (double)beta;
    double capF = (elem == ORC_F32) ? (double)flt_cap(mrm) : dbl_cap(mrm);
    int64_t tot = 0;
    for (uint32_t cw = 0; cw < num_cw; ++cw) {
        char *all = (char *)metrics + (size_t)cw * cw_elems * es;
        char *dm = all + (size_t)P * es; /* decoded_metrics = all + 1 word (LDPCdecode.cpp:100) */
        memset(local, 0, (size_t)h->total_hc_ones * SPZ * P * es);
        uint32_t iter = 0;
        int done = 0;
        while (!done) {
            int e0 = 0, fi = 0, bi = 0;
            for (int l = 0; l < h->mrow; ++l) {
                for (int k = 0; k < h->f2b_copy_cnt[l]; ++k, ++fi) {   /* LDPCdecode.cpp:133-141 */
                    char *cb = dm + (size_t)(SPZ + 2) * P * h->f2b_cols[fi] * es;
                    long off = h->f2b_offsets[fi];
                    memmove(cb + (off + Z) * (long)es, cb + off * (long)es, (size_t)P * es);
                }
                for (int k = 0; k < h->b2f_copy_cnt[l]; ++k, ++bi) {   /* :142-150 */
                    char *cb = dm + (size_t)(SPZ + 2) * P * h->b2f_cols[bi] * es;
                    long off = h->b2f_offsets[bi];
                    memmove(cb + (off - P) * (long)es, cb + (off - P + Z) * (long)es, (size_t)P * es);
                }
                /* decode_row (decode_functions.cpp:11-103) */
                const int w = h->row_weights[l];
                memset(rxor, 0, (size_t)SPZ * P);
                for (int j = 0; j < w; ++j) {
                    const int e = e0 + j;
                    char *glb = dm + (size_t)(SPZ + 2) * P * h->row_active_cols[e] * es;
                    long li = h->row_offset[e];
                    for (int i = 0; i < SPZ; ++i) {
                        for (int p = 0; p < P; ++p) {
                            long lidx = ((long)e * SPZ + i) * P + p;
                            double g = lget(&C, glb, li + p), loc = lget(&C, local, lidx), q, a;
                            if (isint) {
                                int64_t Q = isat((int64_t)g - (int64_t)loc, C.bits);
                                int64_t A = iwrap(iwabs(Q, C.bits) - bI, C.bits);
                                q = (double)Q;
                                a = (double)(A > 0 ? A : 0);
                                rxor[i * P + p] ^= (Q < 0);
                            } else {
                                q = (elem == ORC_F32) ? (double)((float)g - (float)loc) : g - loc;
                                if (elem == ORC_F32) {
                                    float t = fabsf((float)q) - (float)bF;
                                    a = t > 0.0f ? t : 0.0f;
                                } else {
                                    double t = fabs(q) - bF;
                                    a = t > 0.0 ? t : 0.0;
                                }
                                rxor[i * P + p] ^= (signbit(q) != 0);
                            }
                            lput(&C, local, lidx, q);
                            gmin[((long)j * SPZ + i) * P + p] = a;
                        }
                        li += P;
                        if (li >= Z) li -= Z;
                    }
                }
                for (int i = 0; i < SPZ; ++i) {                         /* min_fn(...) per lane */
                    for (int p = 0; p < P; ++p) {
                        double in[MAXW], out[MAXW];
                        for (int j = 0; j < w; ++j) in[j] = gmin[((long)j * SPZ + i) * P + p];
                        if (mode == ORC_PROD && elem == ORC_F32) {
                            float fi_[MAXW], fo[MAXW];
                            for (int j = 0; j < w; ++j) fi_[j] = (float)in[j];
                            tree_prod_f(w, fi_, (float)capF, fo);
                            for (int j = 0; j < w; ++j) out[j] = fo[j];
                        } else if (mode == ORC_PROD) {
                            tree_prod_d(w, in, capF, out);
                        } else {
                            tree_min(w, in, isint ? (double)capI : capF, out);
                        }
                        for (int j = 0; j < w; ++j) fmin[((long)j * SPZ + i) * P + p] = out[j];
                    }
                }
                for (int j = 0; j < w; ++j) {
                    const int e = e0 + j;
                    char *glb = dm + (size_t)(SPZ + 2) * P * h->row_active_cols[e] * es;
                    long li = h->row_offset[e];
                    for (int i = 0; i < SPZ; ++i) {
                        const int nst = (i != SPZ - 1) ? P : Zrem;    /* store_partial on the last word */
                        for (int p = 0; p < P; ++p) {
                            long lidx = ((long)e * SPZ + i) * P + p;
                            double q = lget(&C, local, lidx), m = fmin[((long)j * SPZ + i) * P + p], rn, g;
                            int qneg = isint ? (q < 0) : (signbit(q) != 0);
                            int neg = rxor[i * P + p] ^ qneg;
                            if (isint) {
                                int64_t M = (int64_t)m;
                                int64_t Rn = neg ? iwrap(-M, C.bits) : M;
                                rn = (double)Rn;
                                g = (double)isat((int64_t)q + Rn, C.bits);
                            } else if (elem == ORC_F32) {
                                float Rn = neg ? -(float)m : (float)m;
                                rn = Rn;
                                g = (double)((float)q + Rn);
                            } else {
                                rn = neg ? -m : m;
                                g = q + rn;
                            }
                            lput(&C, local, lidx, rn);
                            if (p < nst) lput(&C, glb, li + p, g);
                        }
                        li += P;
                        if (li >= Z) li -= Z;
                    }
                }
                e0 += w;
            }
            for (int k = 0; k < h->b2f_copy_cnt[h->mrow]; ++k, ++bi) {  /* :161-169 */
                char *cb = dm + (size_t)(SPZ + 2) * P * h->b2f_cols[bi] * es;
                long off = h->b2f_offsets[bi];
                memmove(cb + (off - P) * (long)es, cb + (off - P + Z) * (long)es, (size_t)P * es);
            }
            for (int c = 0; c < h->mcol; ++c) {                           /* :174-179 */
                char *cb = dm + (size_t)(SPZ + 2) * P * c * es;
                memmove(cb + (size_t)Z * es, cb, (size_t)P * es);
            }
            int fail = 0;                                                 /* calc_syndrome_row :108-137 */
            e0 = 0;
            for (int l = 0; l < h->mrow; ++l) {
                const int w = h->row_weights[l];
                memset(rxor, 0, (size_t)SPZ * P);
                for (int j = 0; j < w; ++j) {
                    const int e = e0 + j;
                    char *glb = dm + (size_t)(SPZ + 2) * P * h->row_active_cols[e] * es;
                    long li = h->row_offset[e];
                    for (int i = 0; i < SPZ; ++i) {
                        for (int p = 0; p < P; ++p) {
                            double g = lget(&C, glb, li + p);
                            rxor[i * P + p] ^= isint ? (g < 0) : (signbit(g) != 0);
                        }
                        li += P;
                        if (li >= Z) li -= Z;
                    }
                }
                for (int i = 0; i < SPZ; ++i)
                    for (int p = 0; p < ((i == SPZ - 1) ? Zrem : P); ++p) fail |= rxor[i * P + p];
                e0 += w;
            }
            done = !fail;
            ++iter;
            done = done || (iter == max_iter);
        }
        if (c2v_out) {
            long E = (long)h->total_hc_ones * Z;
            for (int e = 0; e < h->total_hc_ones; ++e)
                for (int r = 0; r < Z; ++r) {
                    double v = lget(&C, local, ((long)e * SPZ + r / P) * P + r % P);
                    lput(&C, (char *)c2v_out + (size_t)cw * E * es, (long)e * Z + r, v);
                }
        }
        if (iters) iters[cw] = iter;
        tot += iter;
    }
    free(local);
    free(gmin);
    free(fmin);
    free(rxor);
    return tot;
}

/* ------------------------------------------------------------------------ */
/* RNG and WGN (prng_functions.cpp)                                          */
/* ------------------------------------------------------------------------ */

static uint64_t g_s[8][4]; /* s[4*PAR_GEN] Vec4uq, PAR_GEN = 2 (prng_functions.cpp:14-20) */

void orc_init_randstate(const uint64_t *inp) { /* :22-26 */
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k < 4; ++k) g_s[i][k] = inp[4 * i + k];
}

static inline uint64_t rotl64(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

void orc_sfc64_N(int sel, uint64_t out[4]) { /* :33-41 */
    uint64_t(*s)[4] = g_s + sel * 4;
    for (int k = 0; k < 4; ++k) {
        uint64_t t = s[0][k] + s[1][k] + s[3][k];
        s[3][k] = s[3][k] + 1;
        s[0][k] = s[1][k] ^ (s[1][k] >> 11);
        s[1][k] = s[2][k] + (s[2][k] << 3);
        s[2][k] = rotl64(s[2][k], 24) + t;
        out[k] = t;
    }
}

void orc_xoshiro_N(int sel, uint64_t out[4]) { /* :44-61 */
    uint64_t(*s)[4] = g_s + sel * 4;
    for (int k = 0; k < 4; ++k) {
        uint64_t a = s[0][k] + s[3][k];
        uint64_t result = rotl64(a, 23) + s[0][k];
        uint64_t t1 = s[1][k] << 17;
        s[2][k] ^= s[0][k];
        s[3][k] ^= s[1][k];
        s[1][k] ^= s[2][k];
        s[0][k] ^= s[3][k];
        s[2][k] ^= t1;
        s[3][k] = rotl64(s[3][k], 45);
        out[k] = result;
    }
}

static inline uint32_t fbits(float f) { union { float f; uint32_t u; } v; v.f = f; return v.u; }
static inline float bitsf(uint32_t u) { union { float f; uint32_t u; } v; v.u = u; return v.f; }

float orc_fastm2log(float x) { /* :81-103 */
    const float a = -0.6296735f, b = 1.466967f, m2log2 = -2.0f * 0.6931472f;
    int32_t xi = (int32_t)fbits(x);
    int32_t ex = xi >> 23;
    int32_t more1p5 = (xi & 0x00400000) ^ 0x00400000;
    int32_t div2 = more1p5 >> 22;
    int32_t expbits = (more1p5 << 1) | 0x3f000000;
    int32_t frac_i = (xi & 0x007fffff) | expbits;
    float frac = bitsf((uint32_t)frac_i) - 1.0f;
    ex = ex - (126 | div2);
    float exf = (float)ex;
    float r = fmaf(fmaf(a, frac, b), frac, exf);
    return r * m2log2;
}

void orc_sincos_norm(float x, float *s, float *c) { /* :109-141 */
    /* x in [-1, 1) is the angle / pi.  vectorclass andnot(a, b) is a & ~b, so the
     * flag is the sign bit exactly when |x| - 0.5 >= 0 (angle beyond +-pi/2). */
    const float ca = 0.99940307f, cb = -4.8911858f, cc = 3.5838442f;
    uint32_t xs = fbits(x) & 0x80000000u;
    float ax = fabsf(x);
    float axm = ax - 0.5f;
    uint32_t mag_gt_p5 = 0x80000000u & ~fbits(axm);
    float xcos = (mag_gt_p5 == 0x80000000u) ? ax - 1.0f : ax;   /* if_sub */
    xcos = xcos * xcos;
    float xsin = axm * axm;
    float lcos = fmaf(fmaf(xcos, cc, cb), xcos, ca);
    float lsin = fmaf(fmaf(xsin, cc, cb), xsin, ca);
    *c = bitsf(fbits(lcos) | mag_gt_p5);
    *s = bitsf(fbits(lsin) | xs);
}

void orc_normdist2(float *result, int simd_count) { /* :148-186 */
    for (int it = 0; it < simd_count / 2; ++it) {
        uint64_t r64[4], r2[4], rc[4];
        orc_sfc64_N(0, r64);
        int32_t rv[8];
        for (int k = 0; k < 4; ++k) { rv[2 * k] = (int32_t)(uint32_t)r64[k]; rv[2 * k + 1] = (int32_t)(r64[k] >> 32); }
        float u[8];
        int any0 = 0;
        for (int i = 0; i < 8; ++i) {
            int32_t v = (rv[i] & 0x7fffffff) ^ 0x7fffffff;
            u[i] = (float)v * 0x1.0p-31f;
            any0 |= ((v & 0x7ffffff8) == 0);
        }
        if (any0) {
            orc_sfc64_N(0, r2);
            for (int k = 0; k < 4; ++k) {
                int32_t lo = (int32_t)(uint32_t)r2[k], hi = (int32_t)(r2[k] >> 32);
                int32_t vv[2] = {lo, hi};
                for (int q = 0; q < 2; ++q) {
                    int32_t v = (vv[q] & 0x7fffffff) ^ 0x7fffffff;
                    float f2 = fmaf((float)v, 0x1.0p-62f, 0x1.0p-63f);
                    u[2 * k + q] = u[2 * k + q] + f2;
                }
            }
        }
        orc_sfc64_N(0, rc);
        for (int i = 0; i < 8; ++i) {
            int32_t v = (i & 1) ? (int32_t)(rc[i >> 1] >> 32) : (int32_t)(uint32_t)rc[i >> 1];
            float g = sqrtf(orc_fastm2log(u[i]));
            float ang = (float)v * 0x1.0p-31f;
            float s, c;
            orc_sincos_norm(ang, &s, &c);
            result[16 * it + i] = s * g;
            result[16 * it + 8 + i] = c * g;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* data bits, encoder, metrics, error counts                                 */
/* ------------------------------------------------------------------------ */

#define WB 64 /* bytes per column word (ENC_BW/8) */

static inline int getbit(const uint8_t *w, int i) { return (w[i >> 3] >> (i & 7)) & 1; }
static inline void setbit(uint8_t *w, int i, int b) {
    if (b) w[i >> 3] |= (uint8_t)(1u << (i & 7));
    else w[i >> 3] &= (uint8_t)~(1u << (i & 7));
}

void orc_gen_rand_databits(uint32_t num_cw, int z, int mrow, int mcol, const uint8_t *filler, uint8_t *bits) {
    /* gen_rand_databits.cpp:11-40: two xoshiro draws form the 512-bit word, masked
     * to Z bits and to non-filler bits */
    uint8_t *p = bits;
    for (uint32_t cw = 0; cw < num_cw; ++cw) {
        for (int dc = 0; dc < mcol - mrow; ++dc) {
            uint64_t a[4], b[4], w[8];
            orc_xoshiro_N(1, a);
            orc_xoshiro_N(1, b);
            for (int k = 0; k < 4; ++k) { w[k] = a[k]; w[4 + k] = b[k]; }
            int zz = z;
            for (int k = 0; k < 8; ++k) {
                uint64_t m = zz >= 64 ? ~0ull : (zz == 0 ? 0 : (~0ull >> (64 - zz)));
                zz = zz >= 64 ? zz - 64 : 0;
                uint64_t fm;
                memcpy(&fm, filler + (size_t)dc * WB + 8 * k, 8);
                w[k] &= m & ~fm;
            }
            memcpy(p, w, WB);
            p += WB;
        }
        p += (size_t)mrow * WB;
    }
}

void orc_encode(uint32_t num_cw, int z, int mrow, int mcol, const int32_t *cnt, const int32_t *colnum,
                const int32_t *src_col, const int32_t *src_shift, uint8_t *bits) {
    /* LDPCencode.cpp:40-100: parity column = XOR over the schedule of rotated columns;
     * the rotation area holds the doubled column, so source bit for row r is
     * col[(r + shift) mod Z] (copy_rotations :102-140) */
    const int nd = mcol - mrow;
    for (uint32_t cw = 0; cw < num_cw; ++cw) {
        uint8_t *c = bits + (size_t)cw * mcol * WB;
        int idx = 0;
        for (int i = 0; i < mrow; ++i) {
            uint8_t acc[WB];
            memset(acc, 0, WB);
            for (int k = 0; k < cnt[i]; ++k, ++idx) {
                const uint8_t *src = c + (size_t)src_col[idx] * WB;
                const int s = src_shift[idx];
                for (int r = 0; r < z; ++r) {
                    int rr = r + s;
                    while (rr >= z) rr -= z;
                    if (getbit(src, rr)) acc[r >> 3] ^= (uint8_t)(1u << (r & 7));
                }
            }
            memcpy(c + (size_t)(nd + colnum[i]) * WB, acc, WB);
        }
    }
}

void orc_snr_params(float snr, float llr_scale, float *sigma, float *factor) {
    /* gen_metrics.cpp:28-30 */
    float three_db = 10 * log10f(2.0);
    float s = powf(10.0, ((snr + three_db) / -20.0));
    *sigma = s;
    *factor = 2 * llr_scale / (s * s);
}

/* roundi (cvtps2dq, nearest-even; out of range -> INT_MIN) then the element store */
static int32_t roundi_x86(float x) {
    if (!(fabsf(x) < 2147483648.0f)) return INT32_MIN;
    return (int32_t)rintf(x);
}

uint32_t orc_gen_metrics(uint32_t num_cw, float snr, const uint8_t *bits, int z, int mcol, const float *wgn,
                         void *metrics, int elem, int P, float llr_scale, float max_abs, const uint8_t *punct,
                         const uint8_t *fill) {
    float sigma, f;
    orc_snr_params(snr, llr_scale, &sigma, &f);
    const int nb = (z + 7) / 8;
    const int spz = P > 0 ? (z + P - 1) / P : 0;
    const long col_elems = P > 0 ? (long)(spz + 2) * P : z;
    const long base = P > 0 ? P : 0;
    const size_t es = elem_size(elem);
    uint32_t err = 0;
    const float *wp = wgn;
    for (uint32_t cw = 0; cw < num_cw; ++cw) {
        for (int col = 0; col < mcol; ++col) {
            const uint8_t *cb = bits + ((size_t)cw * mcol + col) * WB;
            for (int byte = 0; byte < nb; ++byte) {
                for (int k = 0; k < 8; ++k) {
                    const int r = byte * 8 + k;
                    float orig = ((cb[byte] >> k) & 1) ? -1.0f : 1.0f;
                    float rx = fmaf(*wp++, sigma, orig);
                    if (r >= z) rx = 0.0f;            /* cutoff(last_count) */
                    err += (fbits(orig) ^ fbits(rx)) >> 31;
                    float llr = rx * f;
                    llr = max_abs < llr ? max_abs : llr;  /* min(max_abs, LLR) */
                    llr = -max_abs > llr ? -max_abs : llr; /* max(-max_abs, LLR) */
                    if (r < z && getbit(punct + (size_t)col * WB, r)) llr = 0.0f;
                    if (r < z && getbit(fill + (size_t)col * WB, r)) llr = max_abs;
                    if (P == 0 && r >= z) continue;
                    long at = (long)cw * mcol * col_elems + (long)col * col_elems + base + r;
                    if (P > 0 && r >= spz * P) continue;  /* would land in the guard word */
                    char *m = (char *)metrics;
                    switch (elem) {
                    case ORC_F32: ((float *)m)[at] = llr; break;
                    case ORC_F64: ((double *)m)[at] = (double)llr; break;
                    default: iput(m, (size_t)at, elem, (int64_t)roundi_x86(llr)); (void)es; break;
                    }
                }
            }
        }
    }
    return err;
}

void orc_count_errors(uint32_t num_cw, int z, int mrow, int mcol, int num_par, const uint8_t *bits,
                      const void *post, int elem, uint32_t *tot_cw_err, uint32_t *tot_bit_err) {
    /* LDPC_count_errors.cpp:40-136: data columns, hard decision (x < 0), interleaved
     * slot of bit i is i mod num_par; a codeword error per slot with any error */
    uint32_t cwe = 0, be = 0;
    const long N = (long)mcol * z;
    for (uint32_t cw = 0; cw < num_cw; ++cw) {
        uint32_t slot[64] = {0}, cb = 0;
        for (int col = 0; col < mcol - mrow; ++col) {
            const uint8_t *w = bits + ((size_t)cw * mcol + col) * WB;
            for (int r = 0; r < z; ++r) {
                long at = (long)cw * N + (long)col * z + r;
                int hard;
                if (elem == ORC_F32) hard = ((const float *)post)[at] < 0.0f;
                else if (elem == ORC_F64) hard = ((const double *)post)[at] < 0.0;
                else hard = iget(post, (size_t)at, elem) < 0;
                if (hard != getbit(w, r)) { ++cb; ++slot[r % num_par]; }
            }
        }
        if (cb) {
            for (int p = 0; p < num_par; ++p) cwe += slot[p] > 0;
            be += cb;
        }
    }
    *tot_cw_err = cwe;
    *tot_bit_err = be;
}

/* ------------------------------------------------------------------------ */
/* the product's counter-based noise scheme                                  */
/* ------------------------------------------------------------------------ */

uint64_t orc_splitmix64(uint64_t x) {
    /* SplitMix64 output function (Steele, Lea, Flood, OOPSLA 2014) */
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

#define GOLDEN_GAMMA 0x9e3779b97f4a7c15ull

static void noise_pair(uint64_t seed, uint64_t p, float *s_out, float *c_out) {
    uint64_t r = orc_splitmix64(seed + (p + 1) * GOLDEN_GAMMA);
    int32_t lo = (int32_t)(uint32_t)r, hi = (int32_t)(r >> 32);
    int32_t v = (lo & 0x7fffffff) ^ 0x7fffffff;
    float u = (float)v * 0x1.0p-31f;
    if ((v & 0x7ffffff8) == 0) {
        uint64_t r2 = orc_splitmix64((seed ^ 0x5851f42d4c957f2dull) + (p + 1) * GOLDEN_GAMMA);
        int32_t v2 = ((int32_t)(uint32_t)r2 & 0x7fffffff) ^ 0x7fffffff;
        u = u + fmaf((float)v2, 0x1.0p-62f, 0x1.0p-63f);
    }
    float g = sqrtf(orc_fastm2log(u));
    float ang = (float)hi * 0x1.0p-31f;
    float s, c;
    orc_sincos_norm(ang, &s, &c);
    *s_out = s * g;
    *c_out = c * g;
}

float orc_noise_sample(uint64_t seed, uint64_t cw, int col, int row, int z, int mcol) {
    uint64_t p = (cw * (uint64_t)((mcol + 1) / 2) + (uint64_t)(col / 2)) * (uint64_t)z + (uint64_t)row;
    float s, c;
    noise_pair(seed, p, &s, &c);
    return (col & 1) ? c : s;
}

uint32_t orc_noise_llr(uint64_t seed, uint64_t cw0, uint32_t num_cw, const uint8_t *bits, int z, int mcol,
                       float sigma, float f, float max_abs, const uint8_t *punct, const uint8_t *fill, int elem,
                       void *llr) {
    const long N = (long)mcol * z;
    uint32_t err = 0;
    for (uint32_t cw = 0; cw < num_cw; ++cw) {
        for (int col = 0; col < mcol; ++col) {
            for (int r = 0; r < z; ++r) {
                long v = (long)col * z + r;
                float n = orc_noise_sample(seed, cw0 + cw, col, r, z, mcol);
                float orig = bits[(long)cw * N + v] ? -1.0f : 1.0f;
                float rx = fmaf(n, sigma, orig);
                err += (fbits(orig) ^ fbits(rx)) >> 31;
                float l = rx * f;
                l = max_abs < l ? max_abs : l;
                l = -max_abs > l ? -max_abs : l;
                if (punct && punct[v]) l = 0.0f;
                if (fill && fill[v]) l = max_abs;
                long at = (long)cw * N + v;
                if (elem == ORC_F32) ((float *)llr)[at] = l;
                else iput(llr, (size_t)at, elem, (int64_t)roundi_x86(l));
            }
        }
    }
    return err;
}

/* the product's counter-based data bits: 64-bit chunk k of data column c of codeword
 * cw is splitmix64((seed ^ DATA_KEY) + (q + 1) * gamma), q = (cw*mcol + c)*8 + k,
 * masked to Z bits and to non-filler bits (the role of gen_rand_databits.cpp:11-40) */
#define DATA_KEY 0xd1b54a32d192ed03ull
void orc_gpu_databits(uint64_t seed, uint64_t cw0, uint32_t num_cw, int z, int mrow, int mcol,
                      const uint8_t *filler, uint8_t *bits) {
    for (uint32_t cw = 0; cw < num_cw; ++cw) {
        uint8_t *base = bits + (size_t)cw * mcol * WB;
        memset(base, 0, (size_t)mcol * WB);
        for (int c = 0; c < mcol - mrow; ++c) {
            uint64_t w[8];
            int zz = z;
            for (int k = 0; k < 8; ++k) {
                uint64_t q = (((uint64_t)(cw0 + cw) * (uint64_t)mcol + (uint64_t)c) * 8u + (uint64_t)k);
                uint64_t v = orc_splitmix64((seed ^ DATA_KEY) + (q + 1) * GOLDEN_GAMMA);
                uint64_t m = zz >= 64 ? ~0ull : (zz <= 0 ? 0 : (~0ull >> (64 - zz)));
                zz = zz >= 64 ? zz - 64 : 0;
                uint64_t fm;
                memcpy(&fm, filler + (size_t)c * WB + 8 * k, 8);
                w[k] = v & m & ~fm;
            }
            memcpy(base + (size_t)c * WB, w, WB);
        }
    }
}
