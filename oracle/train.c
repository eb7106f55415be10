/*
 * Oracle of the differentiable offset / normalized min-sum decoder.  TEST
 * INFRASTRUCTURE ONLY (linked into liboracle.so, used by tests/ as the checker).
 *
 * The reference has no trainable decoder.  The forward here is the float min-sum
 * path of orc_decode_ideal (src/LDPCDec/decode_functions.cpp:11-103 semantics:
 * A_j = max(|q_j| - beta, 0), exclusive min capped at max_row_metric, sign of the
 * product of the other signs) with three changes: a fixed number T of iterations
 * (no syndrome exit), an offset beta[t][l] and a scale alpha[t][l] per (iteration,
 * layer), the scale multiplying the capped minimum: M_j = alpha * min(cap, min A).
 * With alpha = 1 and beta[t][l] = beta it is orc_decode_ideal's float min-sum for
 * every codeword that needs all T iterations (pinned in tests/test_oracle_train.py).
 *
 * orc_train_grad_f64 is a plain reverse-mode restatement of the same forward in
 * double (per-iteration checkpoints, recomputation within an iteration); it is
 * pinned against central finite differences of orc_train_fwd_f64.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define TMAXW 64

#define DEFINE_TRAIN_FWD(NAME, FT, FABS)                                                                        \
    int NAME(const orc_hdec *h, uint32_t num_cw, uint32_t T, const FT *alpha, const FT *beta, FT cap, FT *llr,  \
             FT *c2v_out) {                                                                                     \
        const int Z = h->z, N = h->mcol * h->z, E = h->total_hc_ones * h->z;                                    \
        FT *R = (FT *)malloc(sizeof(FT) * (size_t)E);                                                           \
        if (!R) return -1;                                                                                      \
        for (uint32_t cw = 0; cw < num_cw; ++cw) {                                                              \
            FT *L = llr + (size_t)cw * N;                                                                       \
            for (int e = 0; e < E; ++e) R[e] = (FT)0;                                                           \
            for (uint32_t t = 0; t < T; ++t) {                                                                  \
                int e0 = 0;                                                                                     \
                for (int l = 0; l < h->mrow; ++l) {                                                             \
                    const int w = h->row_weights[l];                                                            \
                    const FT b = beta[t * h->mrow + l], a = alpha[t * h->mrow + l];                             \
                    for (int r = 0; r < Z; ++r) {                                                               \
                        FT Q[TMAXW], A[TMAXW];                                                                  \
                        int vidx[TMAXW];                                                                        \
                        int sgn = 0;                                                                            \
                        for (int j = 0; j < w; ++j) {                                                           \
                            const int e = e0 + j;                                                               \
                            vidx[j] = h->row_active_cols[e] * Z + (h->row_offset[e] + r) % Z;                   \
                            Q[j] = L[vidx[j]] - R[e * Z + r];                                                   \
                            FT x = FABS(Q[j]) - b;                                                              \
                            A[j] = x > (FT)0 ? x : (FT)0;                                                       \
                            sgn ^= (signbit(Q[j]) != 0);                                                        \
                        }                                                                                       \
                        FT m1 = (FT)INFINITY, m2 = (FT)INFINITY;                                                \
                        int im = -1;                                                                            \
                        for (int j = 0; j < w; ++j) {                                                           \
                            if (A[j] < m1) { m2 = m1; m1 = A[j]; im = j; }                                      \
                            else if (A[j] < m2) m2 = A[j];                                                      \
                        }                                                                                       \
                        for (int j = 0; j < w; ++j) {                                                           \
                            const int e = e0 + j;                                                               \
                            FT x = (j == im) ? m2 : m1;                                                         \
                            FT m = a * (cap < x ? cap : x);                                                     \
                            int neg = sgn ^ (int)(signbit(Q[j]) != 0);                                          \
                            FT rn = neg ? -m : m;                                                               \
                            L[vidx[j]] = Q[j] + rn;                                                             \
                            R[e * Z + r] = rn;                                                                  \
                        }                                                                                       \
                    }                                                                                           \
                    e0 += w;                                                                                    \
                }                                                                                               \
            }                                                                                                   \
            if (c2v_out) memcpy(c2v_out + (size_t)cw * E, R, sizeof(FT) * (size_t)E);                           \
        }                                                                                                       \
        free(R);                                                                                                \
        return 0;                                                                                               \
    }

DEFINE_TRAIN_FWD(orc_train_fwd_f32, float, fabsf)
DEFINE_TRAIN_FWD(orc_train_fwd_f64, double, fabs)

int orc_train_grad_f64(const orc_hdec *h, uint32_t num_cw, uint32_t T, const double *alpha, const double *beta,
                       double cap, const double *llr, const double *gpost, double *galpha, double *gbeta,
                       double *gllr) {
    const int Z = h->z, N = h->mcol * h->z, E = h->total_hc_ones * h->z, Lr = h->mrow;
    double *ckP = (double *)malloc(sizeof(double) * (size_t)T * N);
    double *ckR = (double *)malloc(sizeof(double) * (size_t)T * E);
    double *P = (double *)malloc(sizeof(double) * (size_t)N);
    double *R = (double *)malloc(sizeof(double) * (size_t)E);
    double *Qs = (double *)malloc(sizeof(double) * (size_t)E);
    double *gP = (double *)malloc(sizeof(double) * (size_t)N);
    double *gRn = (double *)malloc(sizeof(double) * (size_t)E);   /* dL / d(C2V out of iteration t) */
    double *gRo = (double *)malloc(sizeof(double) * (size_t)E);   /* dL / d(C2V into iteration t) */
    if (!ckP || !ckR || !P || !R || !Qs || !gP || !gRn || !gRo) return -1;
    memset(galpha, 0, sizeof(double) * (size_t)T * Lr);
    memset(gbeta, 0, sizeof(double) * (size_t)T * Lr);
    for (uint32_t cw = 0; cw < num_cw; ++cw) {
        memcpy(P, llr + (size_t)cw * N, sizeof(double) * (size_t)N);
        for (int e = 0; e < E; ++e) R[e] = 0.0;
        /* forward with a checkpoint at the start of every iteration */
        for (uint32_t t = 0; t < T; ++t) {
            memcpy(ckP + (size_t)t * N, P, sizeof(double) * (size_t)N);
            memcpy(ckR + (size_t)t * E, R, sizeof(double) * (size_t)E);
            int e0 = 0;
            for (int l = 0; l < Lr; ++l) {
                const int w = h->row_weights[l];
                const double b = beta[t * Lr + l], a = alpha[t * Lr + l];
                for (int r = 0; r < Z; ++r) {
                    double Q[TMAXW], A[TMAXW];
                    int vidx[TMAXW], sgn = 0;
                    for (int j = 0; j < w; ++j) {
                        const int e = e0 + j;
                        vidx[j] = h->row_active_cols[e] * Z + (h->row_offset[e] + r) % Z;
                        Q[j] = P[vidx[j]] - R[e * Z + r];
                        double x = fabs(Q[j]) - b;
                        A[j] = x > 0 ? x : 0;
                        sgn ^= (signbit(Q[j]) != 0);
                    }
                    double m1 = INFINITY, m2 = INFINITY;
                    int im = -1;
                    for (int j = 0; j < w; ++j) {
                        if (A[j] < m1) { m2 = m1; m1 = A[j]; im = j; }
                        else if (A[j] < m2) m2 = A[j];
                    }
                    for (int j = 0; j < w; ++j) {
                        const int e = e0 + j;
                        double x = (j == im) ? m2 : m1;
                        double m = a * (cap < x ? cap : x);
                        double rn = (sgn ^ (signbit(Q[j]) != 0)) ? -m : m;
                        P[vidx[j]] = Q[j] + rn;
                        R[e * Z + r] = rn;
                    }
                }
                e0 += w;
            }
        }
        /* reverse sweep */
        memcpy(gP, gpost + (size_t)cw * N, sizeof(double) * (size_t)N);
        for (int e = 0; e < E; ++e) gRn[e] = 0.0;
        for (int t = (int)T - 1; t >= 0; --t) {
            memcpy(P, ckP + (size_t)t * N, sizeof(double) * (size_t)N);
            const double *Ro = ckR + (size_t)t * E;
            /* recompute this iteration, keeping every V2C message */
            int e0 = 0;
            for (int l = 0; l < Lr; ++l) {
                const int w = h->row_weights[l];
                const double b = beta[t * Lr + l], a = alpha[t * Lr + l];
                for (int r = 0; r < Z; ++r) {
                    double Q[TMAXW], A[TMAXW];
                    int vidx[TMAXW], sgn = 0;
                    for (int j = 0; j < w; ++j) {
                        const int e = e0 + j;
                        vidx[j] = h->row_active_cols[e] * Z + (h->row_offset[e] + r) % Z;
                        Q[j] = P[vidx[j]] - Ro[e * Z + r];
                        Qs[e * Z + r] = Q[j];
                        double x = fabs(Q[j]) - b;
                        A[j] = x > 0 ? x : 0;
                        sgn ^= (signbit(Q[j]) != 0);
                    }
                    double m1 = INFINITY, m2 = INFINITY;
                    int im = -1;
                    for (int j = 0; j < w; ++j) {
                        if (A[j] < m1) { m2 = m1; m1 = A[j]; im = j; }
                        else if (A[j] < m2) m2 = A[j];
                    }
                    for (int j = 0; j < w; ++j) {
                        double x = (j == im) ? m2 : m1;
                        double m = a * (cap < x ? cap : x);
                        double rn = (sgn ^ (signbit(Q[j]) != 0)) ? -m : m;
                        P[vidx[j]] = Q[j] + rn;
                    }
                }
                e0 += w;
            }
            /* layers in reverse */
            int ends[4096];
            e0 = 0;
            for (int l = 0; l < Lr; ++l) { ends[l] = e0; e0 += h->row_weights[l]; }
            for (int l = Lr - 1; l >= 0; --l) {
                const int w = h->row_weights[l];
                const int eb = ends[l];
                const double b = beta[t * Lr + l], a = alpha[t * Lr + l];
                for (int r = 0; r < Z; ++r) {
                    double Q[TMAXW], A[TMAXW], gA[TMAXW], gq[TMAXW];
                    int vidx[TMAXW], sgn = 0;
                    for (int j = 0; j < w; ++j) {
                        const int e = eb + j;
                        vidx[j] = h->row_active_cols[e] * Z + (h->row_offset[e] + r) % Z;
                        Q[j] = Qs[e * Z + r];
                        double x = fabs(Q[j]) - b;
                        A[j] = x > 0 ? x : 0;
                        sgn ^= (signbit(Q[j]) != 0);
                        gA[j] = 0.0;
                    }
                    double m1 = INFINITY, m2 = INFINITY;
                    int im = -1, im2 = -1;
                    for (int j = 0; j < w; ++j) {
                        if (A[j] < m1) { m2 = m1; im2 = im; m1 = A[j]; im = j; }
                        else if (A[j] < m2) { m2 = A[j]; im2 = j; }
                    }
                    for (int j = 0; j < w; ++j) {
                        const int e = eb + j;
                        const double x = (j == im) ? m2 : m1;
                        const int src = (j == im) ? im2 : im;
                        const double capped = cap < x ? cap : x;
                        const double s = (sgn ^ (signbit(Q[j]) != 0)) ? -1.0 : 1.0;
                        const double gout = gP[vidx[j]] + gRn[e * Z + r];   /* d/d C2V_j (new) */
                        gq[j] = gP[vidx[j]];                                 /* P' = q + C2V */
                        const double gm = s * gout;
                        galpha[t * Lr + l] += gm * capped;
                        if (!(cap < x)) gA[src] += a * gm;
                    }
                    for (int j = 0; j < w; ++j) {
                        if (fabs(Q[j]) - b > 0) {
                            gq[j] += gA[j] * (signbit(Q[j]) ? -1.0 : 1.0);
                            gbeta[t * Lr + l] -= gA[j];
                        }
                    }
                    for (int j = 0; j < w; ++j) {
                        const int e = eb + j;
                        gP[vidx[j]] = gq[j];
                        gRo[e * Z + r] = -gq[j];
                    }
                }
            }
            memcpy(gRn, gRo, sizeof(double) * (size_t)E);
        }
        if (gllr) memcpy(gllr + (size_t)cw * N, gP, sizeof(double) * (size_t)N);
    }
    free(ckP); free(ckR); free(P); free(R); free(Qs); free(gP); free(gRn); free(gRo);
    return 0;
}
