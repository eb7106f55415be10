/*
 * Body of the SIMD CPU decoder LDPCtop_Vec64c_min (P = 64 int8 lanes), included by
 * baseline_vec.c once per instruction set with V_* helpers defined for it.
 * Restates src/LDPCDec/LDPCdecode.cpp:79-245 and decode_functions.cpp:11-137
 * literally (guard words, f2b/b2f copies, partial stores), with the exclusive
 * minimum taken as min1/min2 (the reference's min_find tree computes the same
 * exact values: fin_j = min(cap, min_{k != j} A_k)).  TEST INFRASTRUCTURE ONLY.
 */

static uint32_t FN(decode_one)(const orc_hdec *h, uint32_t max_iter, int8_t beta8, int8_t cap8,
                               int8_t *all, VEC *local, VEC *amin, VEC *qv) {
    const int Z = h->z, P = 64, SPZ = (Z + P - 1) / P, Zrem = Z - (SPZ - 1) * P;
    int8_t *dm = all + P;
    const long cstride = (long)(SPZ + 2) * P;
    const VEC vbeta = V_SET1(beta8), vcap = V_SET1(cap8), vzero = V_ZERO(), vmax = V_SET1(127);
    for (int i = 0; i < h->total_hc_ones * SPZ; ++i) local[i] = vzero;
    uint32_t iter = 0;
    int done = 0;
    while (!done) {
        int e0 = 0, fi = 0, bi = 0;
        for (int l = 0; l < h->mrow; ++l) {
            for (int k = 0; k < h->f2b_copy_cnt[l]; ++k, ++fi) {
                int8_t *cb = dm + cstride * h->f2b_cols[fi];
                const int off = h->f2b_offsets[fi];
                V_STOREU(cb + off + Z, V_LOADU(cb + off));
            }
            for (int k = 0; k < h->b2f_copy_cnt[l]; ++k, ++bi) {
                int8_t *cb = dm + cstride * h->b2f_cols[bi];
                const int off = h->b2f_offsets[bi];
                V_STOREU(cb - P + off, V_LOADU(cb - P + off + Z));
            }
            const int w = h->row_weights[l];
            for (int i = 0; i < SPZ; ++i) {
                VEC sx = vzero, m1 = vmax, m2 = vmax;
                /* pass 1 */
                for (int j = 0; j < w; ++j) {
                    const int e = e0 + j;
                    int8_t *glb = dm + cstride * h->row_active_cols[e];
                    int li = h->row_offset[e] + i * P;
                    while (li >= Z) li -= Z;
                    VEC g = V_LOADU(glb + li);
                    VEC q = V_SUBS(g, local[e * SPZ + i]);
                    VEC a = V_MAX(V_SUB(V_ABS(q), vbeta), vzero);
                    qv[j] = q;
                    amin[j] = a;
                    sx = V_XOR(sx, q);
                    m2 = V_MIN(m2, V_MAX(m1, a));
                    m1 = V_MIN(m1, a);
                }
                m1 = V_MIN(vcap, m1);
                m2 = V_MIN(vcap, m2);
                /* pass 2 */
                for (int j = 0; j < w; ++j) {
                    const int e = e0 + j;
                    int8_t *glb = dm + cstride * h->row_active_cols[e];
                    int li = h->row_offset[e] + i * P;
                    while (li >= Z) li -= Z;
                    VEC fin = V_SELEQ(amin[j], m1, m2, m1);
                    VEC rn = V_NEGIF(V_XOR(sx, qv[j]), fin);
                    VEC g = V_ADDS(qv[j], rn);
                    local[e * SPZ + i] = rn;
                    if (i != SPZ - 1) V_STOREU(glb + li, g);
                    else V_STOREN(glb + li, g, Zrem);
                }
            }
            e0 += w;
        }
        for (int k = 0; k < h->b2f_copy_cnt[h->mrow]; ++k, ++bi) {
            int8_t *cb = dm + cstride * h->b2f_cols[bi];
            const int off = h->b2f_offsets[bi];
            V_STOREU(cb - P + off, V_LOADU(cb - P + off + Z));
        }
        for (int c = 0; c < h->mcol; ++c) {
            int8_t *cb = dm + cstride * c;
            V_STOREU(cb + Z, V_LOADU(cb));
        }
        VEC syn = vzero;
        e0 = 0;
        for (int l = 0; l < h->mrow; ++l) {
            const int w = h->row_weights[l];
            for (int i = 0; i < SPZ; ++i) {
                VEC x = vzero;
                for (int j = 0; j < w; ++j) {
                    const int e = e0 + j;
                    int li = h->row_offset[e] + i * P;
                    while (li >= Z) li -= Z;
                    x = V_XOR(x, V_LOADU(dm + cstride * h->row_active_cols[e] + li));
                }
                if (i == SPZ - 1) x = V_KEEPN(x, Zrem);
                syn = V_OR(syn, x);
            }
            e0 += w;
        }
        done = !V_ANYNEG(syn);
        ++iter;
        done = done || (iter == max_iter);
    }
    return iter;
}

typedef struct {
    const orc_hdec *h;
    uint32_t max_iter, c0, c1;
    int8_t beta8, cap8;
    int8_t *metrics;
    uint32_t *iters;
    int64_t tot;
} FN(job_t);

static void *FN(worker)(void *arg) {
    FN(job_t) *jb = (FN(job_t) *)arg;
    const orc_hdec *h = jb->h;
    const int SPZ = (h->z + 63) / 64;
    const long cw_elems = (long)h->mcol * (SPZ + 2) * 64;
    VEC *local = (VEC *)aligned_alloc(64, sizeof(VEC) * (size_t)(h->total_hc_ones * SPZ + 1));
    VEC *amin = (VEC *)aligned_alloc(64, sizeof(VEC) * 64);
    VEC *qv = (VEC *)aligned_alloc(64, sizeof(VEC) * 64);
    int64_t tot = 0;
    for (uint32_t cw = jb->c0; cw < jb->c1; ++cw) {
        uint32_t it = FN(decode_one)(h, jb->max_iter, jb->beta8, jb->cap8, jb->metrics + cw * cw_elems, local,
                                     amin, qv);
        if (jb->iters) jb->iters[cw] = it;
        tot += it;
    }
    free(local);
    free(amin);
    free(qv);
    jb->tot = tot;
    return NULL;
}

static int64_t FN(run)(const orc_hdec *h, uint32_t num_cw, uint32_t max_iter, int8_t beta8, int8_t cap8,
                       int8_t *metrics, uint32_t *iters, int nthreads) {
    if (nthreads < 1) nthreads = 1;
    if ((uint32_t)nthreads > num_cw) nthreads = num_cw ? (int)num_cw : 1;
    pthread_t th[256];
    FN(job_t) jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].h = h;
        jobs[t].max_iter = max_iter;
        jobs[t].c0 = (uint32_t)((uint64_t)num_cw * t / nthreads);
        jobs[t].c1 = (uint32_t)((uint64_t)num_cw * (t + 1) / nthreads);
        jobs[t].beta8 = beta8;
        jobs[t].cap8 = cap8;
        jobs[t].metrics = metrics;
        jobs[t].iters = iters;
        jobs[t].tot = 0;
    }
    if (nthreads == 1) {
        FN(worker)(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t) pthread_create(&th[t], NULL, FN(worker), &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    int64_t tot = 0;
    for (int t = 0; t < nthreads; ++t) tot += jobs[t].tot;
    return tot;
}
