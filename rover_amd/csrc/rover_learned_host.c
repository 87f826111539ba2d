/*
 * rover_learned_host.c -- learned early termination: trace generation,
 * training and installation of the stop predictor.
 *
 * Traces come from the plain search itself (the search kernel in trace mode
 * records the 16 stop features every `stop_every` expansions and, per query,
 * the steps at which the members of its final top-K were scored).  A
 * record's label is 1 when all but `label_slack` of the final top-K had been
 * scored by that step (slack 0: the top-K is already final; a final top-K
 * member stays in the top-K from the step it is scored, since candidates
 * better than it only accumulate).  The MLP is trained by the library's own HIP
 * forward/backward/Adam kernels and installed on the session; searches with
 * learned_stop set stop once the predicted probability reaches the
 * threshold, and otherwise run to the plain fixed-L termination.
 */
#define _GNU_SOURCE
#include "rover_ctx.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NF RV_MLP_IN
#define NH RV_MLP_HIDDEN
#define NBF (NH * NF + NH * NH + NH) /* bf16 weights w1|w2|w3 */
#define NFL (2 * NH + 1 + 2 * NF)    /* b1|b2|b3|mean|inv_std */

int rover_mlp_param_count(void) { return rv_mlp_param_count(); }

AEEResult rover_mlp_forward_backward(const float *d_params, const float *d_x,
                                     const float *d_y, const float *d_weight,
                                     int n, float *d_grad, float *d_loss,
                                     float *d_out) {
    float *d_slab = NULL;
    AEEResult rc = ROVER_SUCCESS;
    if (!d_params || !d_x || !d_y || !d_weight || n < 0) return ROVER_EBADPARM;
    /* a slab of its own per call, so concurrent callers never share one */
    if (rv_malloc((void **)&d_slab, rv_mlp_slab_floats() * sizeof(float))) {
        rover_set_err("rover_mlp_forward_backward: slab");
        return ROVER_ENOMEMORY;
    }
    if (rv_mlp_fwd_bwd_sampled(d_params, d_x, d_y, d_weight, NULL, n, d_grad, d_loss,
                               d_out, d_slab, NULL)) {
        rover_set_errf(rv_learned_error());
        rc = ROVER_EFAILED;
    }
    if (rv_synchronize() && rc == ROVER_SUCCESS) {
        rover_set_err("rover_mlp_forward_backward");
        rc = ROVER_EFAILED;
    }
    rover_dfree(d_slab);
    return rc;
}

AEEResult rover_mlp_adam_step(float *d_params, const float *d_grad, float *d_m1,
                              float *d_m2, int step, float lr, float grad_scale) {
    if (!d_params || !d_grad || !d_m1 || !d_m2 || step < 1) return ROVER_EBADPARM;
    if (rv_mlp_adam(d_params, d_grad, d_m1, d_m2, step, lr, 0.9f, 0.999f, 1e-8f,
                    grad_scale, NULL)) {
        rover_set_errf(rv_learned_error());
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

/* Installs FP32 parameters (rounded to BF16 on the device) and the feature
 * normalization on the session. */
static AEEResult install(RoverCtx *c, const float *d_params, const float *params,
                         const float *mean, const float *inv_std) {
    RoverLearned *l = &c->learned;
    int P = rv_mlp_param_count();
    float fl[NFL];
    float *d_tmp = NULL;
    if (!l->d_w && rv_malloc((void **)&l->d_w, NBF * sizeof(uint16_t))) goto fail;
    if (!l->d_f && rv_malloc((void **)&l->d_f, NFL * sizeof(float))) goto fail;
    if (!l->h_params && !(l->h_params = malloc((size_t)P * sizeof(float)))) goto fail;
    if (!d_params) {
        if (rv_malloc((void **)&d_tmp, (size_t)P * sizeof(float)) ||
            rv_memcpy_h2d(d_tmp, params, (size_t)P * sizeof(float)))
            goto fail;
        d_params = d_tmp;
        memcpy(l->h_params, params, (size_t)P * sizeof(float));
    } else if (rv_memcpy_d2h(l->h_params, d_params, (size_t)P * sizeof(float))) {
        goto fail;
    }
    if (rv_mlp_to_bf16(d_params, l->d_w, NULL)) goto fail;
    /* b1 | b2 | b3 from the flat layout w1|b1|w2|b2|w3|b3 */
    memcpy(fl, l->h_params + NH * NF, NH * sizeof(float));
    memcpy(fl + NH, l->h_params + NH * NF + NH + NH * NH, NH * sizeof(float));
    fl[2 * NH] = l->h_params[P - 1];
    memcpy(fl + 2 * NH + 1, mean, NF * sizeof(float));
    memcpy(fl + 2 * NH + 1 + NF, inv_std, NF * sizeof(float));
    memcpy(l->h_mean, mean, sizeof(l->h_mean));
    memcpy(l->h_inv_std, inv_std, sizeof(l->h_inv_std));
    if (rv_memcpy_h2d(l->d_f, fl, sizeof(fl)) || rv_synchronize()) goto fail;
    rover_dfree(d_tmp);
    l->ready = 1;
    return ROVER_SUCCESS;
fail:
    rover_set_err("learned stop install");
    rover_dfree(d_tmp);
    return ROVER_EFAILED;
}

AEEResult rover_learned_stop_set(remote_handle64 h, const float *params,
                                 int paramsLen, const float *mean,
                                 const float *inv_std, int stop_every,
                                 int label_k, int min_steps) {
    RoverCtx *c = rover_ctx_of(h);
    AEEResult rc;
    if (!c || !params || !mean || !inv_std || paramsLen != rv_mlp_param_count() ||
        stop_every <= 0 || label_k <= 0 || min_steps < 0)
        return ROVER_EBADPARM;
    pthread_mutex_lock(&c->lock);
    rc = install(c, NULL, params, mean, inv_std);
    if (rc == ROVER_SUCCESS) {
        c->learned.stop_every = stop_every;
        c->learned.label_k = label_k;
        c->learned.min_steps = min_steps;
    }
    pthread_mutex_unlock(&c->lock);
    return rc;
}

AEEResult rover_learned_stop_get(remote_handle64 h, float *params, int paramsLen,
                                 float *mean, float *inv_std) {
    RoverCtx *c = rover_ctx_of(h);
    if (!c || !params || !mean || !inv_std || paramsLen != rv_mlp_param_count())
        return ROVER_EBADPARM;
    if (!c->learned.ready) return ROVER_EFAILED;
    memcpy(params, c->learned.h_params, (size_t)paramsLen * sizeof(float));
    memcpy(mean, c->learned.h_mean, sizeof(c->learned.h_mean));
    memcpy(inv_std, c->learned.h_inv_std, sizeof(c->learned.h_inv_std));
    return ROVER_SUCCESS;
}

/* Deterministic He-style initialization from a seed (xorshift + Box-Muller). */
static void init_params(float *p, int P, uint32_t seed) {
    uint64_t s = 0x9e3779b97f4a7c15ull ^ seed;
    const double two_pi = 6.283185307179586;
    for (int i = 0; i < P; i++) {
        double u1, u2, g;
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        u1 = ((s >> 11) + 1.0) / 9007199254740993.0;
        s ^= s << 13; s ^= s >> 7; s ^= s << 17;
        u2 = (s >> 11) / 9007199254740992.0;
        g = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
        p[i] = (float)g;
    }
    /* scale: w1 by sqrt(2/16), w2 by sqrt(2/64), w3 by sqrt(1/64); biases 0 */
    for (int i = 0; i < NH * NF; i++) p[i] *= (float)sqrt(2.0 / NF);
    for (int i = 0; i < NH; i++) p[NH * NF + i] = 0.0f;
    for (int i = 0; i < NH * NH; i++) p[NH * NF + NH + i] *= (float)sqrt(2.0 / NH);
    for (int i = 0; i < NH; i++) p[NH * NF + NH + NH * NH + i] = 0.0f;
    for (int i = 0; i < NH; i++) p[NH * NF + 2 * NH + NH * NH + i] *= (float)sqrt(1.0 / NH);
    p[P - 1] = 0.0f;
}

/* The plain search of a device batch with its stop features recorded: every
 * stop_every expansions the 16 features the learned stop would evaluate
 * (d_feats [nq][trace_cap][16], raw) and the expansion count
 * (d_steps [nq][trace_cap]); per query d_meta [nq][RV_TRACE_META]: records
 * kept, last top-K change, then the expansion at which each final top-K
 * member was scored, ascending. */
AEEResult rover_search_trace_device(remote_handle64 h, const RoverSearchBatch *b,
                                    int stop_every, int label_k, int trace_cap,
                                    float *d_feats, int32_t *d_steps, int32_t *d_meta,
                                    void *stream) {
    RoverCtx *c = rover_ctx_of(h);
    RvIndex ix;
    RvBatch rb;
    AEEResult rc;
    if (!c || !b || !d_feats || !d_steps || !d_meta || stop_every <= 0 || trace_cap <= 0 ||
        label_k <= 0 || label_k > RV_TRACE_TOPK || b->learned_stop || b->route_entries > 0 ||
        (b->method == ROVER_METHOD_SQ && b->traversal != ROVER_TRAVERSAL_BEST_FIRST))
        return ROVER_EBADPARM;
    pthread_mutex_lock(&c->lock);
    rc = rover_prepare_launch(c, b, b->pool_size, 0, &ix, &rb);
    if (rc == ROVER_SUCCESS) {
        rb.stop_mode = 2;
        rb.stop_every = stop_every;
        rb.stop_label_k = label_k;
        rb.trace_feats = d_feats;
        rb.trace_steps = d_steps;
        rb.trace_meta = d_meta;
        rb.trace_cap = trace_cap;
        if (rover_stream_turn(c, stream) || rover_launch(c, &ix, &rb, b->grid_waves, stream) ||
            rover_stream_done(c, stream)) {
            rover_set_err("trace search");
            rc = ROVER_EFAILED;
        }
    }
    pthread_mutex_unlock(&c->lock);
    return rc;
}

AEEResult rover_learned_stop_train(remote_handle64 h, RoverLearnedStopTrain *t) {
    RoverCtx *c = rover_ctx_of(h);
    RoverSearchBatch b;
    RvIndex ix;
    RvBatch rb;
    AEEResult rc = ROVER_EFAILED;
    const int P = rv_mlp_param_count();
    size_t nrec;
    float *d_q = NULL, *d_feat = NULL, *d_y = NULL, *d_w = NULL, *d_params = NULL;
    float *d_grad = NULL, *d_m1 = NULL, *d_m2 = NULL, *d_loss = NULL;
    float *d_mean = NULL, *d_istd = NULL, *d_od = NULL;
    int32_t *d_steps = NULL, *d_meta = NULL, *d_oid = NULL, *d_cnt = NULL;
    int32_t *d_idx = NULL, *d_ent = NULL;
    float *d_slab = NULL;
    float *h_feat = NULL, *h_y = NULL, *h_w = NULL, *params = NULL;
    int32_t *h_meta = NULL;
    double sum[NF], sq[NF], cnt = 0.0, pos = 0.0;
    float mean[NF], istd[NF];
    if (!c || !t || !t->queries || t->num_queries <= 0 || t->pool_size <= 0 ||
        t->pool_size > ROVER_MAX_POOL || t->stop_every <= 0 || t->trace_cap <= 0 ||
        t->label_k <= 0 || t->label_k > t->pool_size || t->label_k > RV_TRACE_TOPK ||
        t->label_slack < 0 || t->label_slack >= t->label_k || t->steps <= 0 ||
        t->batch <= 0 || !(t->lr > 0.0f))
        return ROVER_EBADPARM;
    pthread_mutex_lock(&c->lock);
    nrec = (size_t)t->num_queries * t->trace_cap;
    /* ---- traces from the plain search (trace mode) */
    if (rv_malloc((void **)&d_q, (size_t)t->num_queries * c->dim * sizeof(float)) ||
        rv_memcpy_h2d(d_q, t->queries, (size_t)t->num_queries * c->dim * sizeof(float)) ||
        rv_malloc((void **)&d_feat, nrec * NF * sizeof(float)) ||
        rv_malloc((void **)&d_steps, nrec * sizeof(int32_t)) ||
        rv_malloc((void **)&d_meta, (size_t)t->num_queries * RV_TRACE_META * sizeof(int32_t)) ||
        rv_malloc((void **)&d_oid, (size_t)t->num_queries * sizeof(int32_t)) ||
        rv_malloc((void **)&d_od, (size_t)t->num_queries * sizeof(float)) ||
        rv_malloc((void **)&d_cnt, (size_t)t->num_queries * sizeof(int32_t)) ||
        /* the trace search writes only the records it keeps; the others are
           sampled with weight 0 and must hold finite values (0 * NaN would
           poison the gradient), so start from zeros */
        rv_fill_zero(d_feat, nrec * NF * sizeof(float), NULL) ||
        rv_fill_zero(d_steps, nrec * sizeof(int32_t), NULL)) {
        rover_set_err("learned stop: trace buffers");
        goto out;
    }
    memset(&b, 0, sizeof(b));
    b.queries = d_q;
    b.num_queries = t->num_queries;
    b.pool_size = t->pool_size;
    b.num_out = 1;
    b.method = ROVER_METHOD_PQ;
    b.correction = t->correction ? 1 : 0;
    b.entry_points = t->entry_points > 0 ? t->entry_points : 1;
    if (t->route_entries > 0) {
        int S = t->route_entries;
        if (rv_malloc((void **)&d_ent, (size_t)t->num_queries * S * sizeof(int32_t)) ||
            rover_route_into(c, d_q, t->num_queries, S, d_ent, NULL) != ROVER_SUCCESS) {
            rover_set_err("learned stop: routing");
            goto out;
        }
        b.entry_ids = d_ent;
        b.entry_count = S;
        b.entry_stride = S;
    } else if (t->entry_ids) {
        /* one shared list, or one row of entry_stride ids per query */
        size_t cnt = t->entry_stride > 0 ? (size_t)t->entry_stride * t->num_queries
                                         : (size_t)(t->entry_count > 0 ? t->entry_count : 0);
        if (cnt == 0 || rv_malloc((void **)&d_ent, cnt * sizeof(int32_t)) ||
            rv_memcpy_h2d(d_ent, t->entry_ids, cnt * sizeof(int32_t))) {
            rover_set_err("learned stop: entry ids");
            goto out;
        }
        b.entry_ids = d_ent;
        b.entry_count = t->entry_stride > 0 ? t->entry_stride : t->entry_count;
        b.entry_stride = t->entry_stride > 0 ? t->entry_stride : 0;
    }
    b.expand_count = t->expand_count > 0 ? t->expand_count : c->degree;
    b.out_ids = d_oid;
    b.out_dists = d_od;
    b.out_counts = d_cnt;
    b.out_stride = 1;
    rc = rover_prepare_launch(c, &b, t->pool_size, 0, &ix, &rb);
    if (rc != ROVER_SUCCESS) goto out;
    rc = ROVER_EFAILED;
    rb.stop_mode = 2;
    rb.stop_every = t->stop_every;
    rb.stop_label_k = t->label_k;
    rb.trace_feats = d_feat;
    rb.trace_steps = d_steps;
    rb.trace_meta = d_meta;
    rb.trace_cap = t->trace_cap;
    if (rover_stream_turn(c, NULL) || rover_launch(c, &ix, &rb, 0, NULL) ||
        rover_stream_done(c, NULL) || rv_synchronize()) {
        rover_set_err("learned stop: trace search");
        goto out;
    }
    /* ---- normalization statistics from a strided sample of queries */
    {
        int stride = t->num_queries > 4096 ? t->num_queries / 4096 : 1;
        int nsamp = (t->num_queries + stride - 1) / stride;
        const int need = 2 + t->label_k - 1 - t->label_slack;
        h_meta = malloc((size_t)t->num_queries * RV_TRACE_META * sizeof(int32_t));
        h_feat = malloc((size_t)t->trace_cap * NF * sizeof(float));
        if (!h_meta || !h_feat ||
            rv_memcpy_d2h(h_meta, d_meta,
                          (size_t)t->num_queries * RV_TRACE_META * sizeof(int32_t)))
            goto out;
        for (int f = 0; f < NF; f++) sum[f] = sq[f] = 0.0;
        for (int s = 0; s < nsamp; s++) {
            int q = s * stride, recs = h_meta[(size_t)q * RV_TRACE_META];
            if (recs <= 0) continue;
            if (rv_memcpy_d2h(h_feat, d_feat + (size_t)q * t->trace_cap * NF,
                              (size_t)recs * NF * sizeof(float)))
                goto out;
            int32_t steps_q[4096];
            int take = recs < 4096 ? recs : 4096;
            if (rv_memcpy_d2h(steps_q, d_steps + (size_t)q * t->trace_cap,
                              (size_t)take * sizeof(int32_t)))
                goto out;
            for (int r = 0; r < take; r++) {
                for (int f = 0; f < NF; f++) {
                    double v = h_feat[(size_t)r * NF + f];
                    sum[f] += v;
                    sq[f] += v * v;
                }
                cnt += 1.0;
                pos += steps_q[r] >= h_meta[(size_t)q * RV_TRACE_META + need] ? 1.0 : 0.0;
            }
        }
        if (cnt < 1.0) {
            rover_set_errf("learned stop: traces are empty");
            goto out;
        }
        for (int f = 0; f < NF; f++) {
            double m = sum[f] / cnt, var = sq[f] / cnt - m * m;
            mean[f] = (float)m;
            istd[f] = (float)(1.0 / sqrt(var > 1e-12 ? var : 1e-12));
        }
        t->samples = 0;
        for (int q = 0; q < t->num_queries; q++)
            t->samples += h_meta[(size_t)q * RV_TRACE_META];
        t->positive_frac = (float)(pos / cnt);
    }
    /* ---- labels, weights, normalized features */
    {
        float pf = t->positive_frac;
        float pos_w = t->pos_weight > 0.0f ? t->pos_weight : 1.0f;
        float neg_w = 1.0f;
        (void)pf;
        if (rv_malloc((void **)&d_y, nrec * sizeof(float)) ||
            rv_malloc((void **)&d_w, nrec * sizeof(float)) ||
            rv_malloc((void **)&d_mean, NF * sizeof(float)) ||
            rv_malloc((void **)&d_istd, NF * sizeof(float)) ||
            rv_memcpy_h2d(d_mean, mean, sizeof(mean)) ||
            rv_memcpy_h2d(d_istd, istd, sizeof(istd)) ||
            rv_trace_labels(d_steps, d_meta, t->num_queries, t->trace_cap, t->label_k,
                            t->label_slack, pos_w, neg_w, d_y, d_w, NULL) ||
            rv_normalize(d_feat, nrec, d_mean, d_istd, NULL)) {
            rover_set_err("learned stop: labels");
            goto out;
        }
    }
    /* ---- training: Adam on sampled mini-batches */
    params = malloc((size_t)P * sizeof(float));
    if (!params) goto out;
    init_params(params, P, t->seed ? t->seed : 1u);
    if (rv_malloc((void **)&d_params, (size_t)P * sizeof(float)) ||
        rv_memcpy_h2d(d_params, params, (size_t)P * sizeof(float)) ||
        rv_malloc((void **)&d_grad, (size_t)P * sizeof(float)) ||
        rv_malloc((void **)&d_m1, (size_t)P * sizeof(float)) ||
        rv_malloc((void **)&d_m2, (size_t)P * sizeof(float)) ||
        rv_malloc((void **)&d_loss, 2 * sizeof(float)) ||
        rv_malloc((void **)&d_idx, (size_t)t->batch * sizeof(int32_t)) ||
        rv_malloc((void **)&d_slab, rv_mlp_slab_floats() * sizeof(float)) ||
        rv_fill_zero(d_m1, (size_t)P * sizeof(float), NULL) ||
        rv_fill_zero(d_m2, (size_t)P * sizeof(float), NULL) ||
        rv_fill_zero(d_loss, 2 * sizeof(float), NULL)) {
        rover_set_err("learned stop: training buffers");
        goto out;
    }
    {
        const int tail = t->steps > 50 ? 50 : t->steps;
        for (int step = 1; step <= t->steps; step++) {
            if (step == t->steps - tail + 1 && rv_fill_zero(d_loss, sizeof(float), NULL))
                goto out;
            if (rv_fill_zero(d_grad, (size_t)P * sizeof(float), NULL) ||
                rv_sample_indices(d_idx, t->batch, nrec, t->seed, (uint32_t)step, NULL) ||
                rv_mlp_fwd_bwd_sampled(d_params, d_feat, d_y, d_w, d_idx, t->batch,
                                       d_grad, d_loss, NULL, d_slab, NULL) ||
                rv_mlp_adam(d_params, d_grad, d_m1, d_m2, step, t->lr, 0.9f, 0.999f,
                            1e-8f, 1.0f / (float)t->batch, NULL)) {
                rover_set_errf(rv_learned_error());
                goto out;
            }
        }
        float l = 0.0f;
        if (rv_synchronize() || rv_memcpy_d2h(&l, d_loss, sizeof(float))) goto out;
        t->final_loss = l / (float)(tail * t->batch);
    }
    rc = install(c, d_params, NULL, mean, istd);
    if (rc == ROVER_SUCCESS) {
        c->learned.stop_every = t->stop_every;
        c->learned.label_k = t->label_k;
        c->learned.min_steps = t->min_steps;
    }
out:
    pthread_mutex_unlock(&c->lock);
    rover_dfree(d_q);
    rover_dfree(d_feat);
    rover_dfree(d_steps);
    rover_dfree(d_meta);
    rover_dfree(d_oid);
    rover_dfree(d_od);
    rover_dfree(d_cnt);
    rover_dfree(d_y);
    rover_dfree(d_w);
    rover_dfree(d_mean);
    rover_dfree(d_istd);
    rover_dfree(d_params);
    rover_dfree(d_grad);
    rover_dfree(d_m1);
    rover_dfree(d_m2);
    rover_dfree(d_loss);
    rover_dfree(d_idx);
    rover_dfree(d_slab);
    rover_dfree(d_ent);
    free(h_feat);
    free(h_y);
    free(h_w);
    free(h_meta);
    free(params);
    return rc;
}
