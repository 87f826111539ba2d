/*
 * ref_harness.c -- ctypes-callable driver around the reference's own scalar C
 * (compiled from the Rover sources in place by oracle/Makefile into
 * oracle/_ref/librover_ref.so).
 *
 * TEST INFRASTRUCTURE ONLY.  It contains no reference code: it builds the
 * reference's input structs from flat arrays and calls the reference
 * functions, so that tests/golden/make_golden.py can record the reference's
 * outputs as fixtures and the tests can cross-check the C restatement
 * (rover_oracle.c) against the reference itself.
 */
#define _GNU_SOURCE
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include "dsp_demo/retrieval_final_refine.h"
#include "dsp_demo/retrieval_l1_int8_core.h"
#include "dsp_demo/retrieval_pq_core.h"
#include "dsp_demo/retrieval_pq_residual.h"
#include "dsp_demo/retrieval_sqmd_pure_pq.h"
#include "dsp_demo/retrieval_sqmd_seed.h"
#include "rover_config.h"

int ref_build_lut(const float *q, const float *cb, int m, int ks, int sub,
                  float *lut) {
    return retrieval_pq_build_lut(q, cb, m, ks, sub, lut);
}

/* codes are split into chunks of chunk_vectors rows to exercise the
 * reference's chunked addressing. */
int ref_adc_ids(const uint8_t *codes, int n, int code_bytes, int chunk_vectors,
                const int32_t *ids, int count, const float *lut, int m, int ks,
                float *out) {
    const uint8_t *chunks[PURE_PQ_MAX_CODE_CHUNKS];
    int counts[PURE_PQ_MAX_CODE_CHUNKS];
    int nchunks = (n + chunk_vectors - 1) / chunk_vectors;
    if (nchunks > PURE_PQ_MAX_CODE_CHUNKS) return -1;
    for (int c = 0; c < nchunks; c++) {
        chunks[c] = codes + (size_t)c * chunk_vectors * code_bytes;
        counts[c] = n - c * chunk_vectors < chunk_vectors ? n - c * chunk_vectors
                                                          : chunk_vectors;
    }
    return retrieval_pq_adc_ids_batch_scalar(chunks, counts, nchunks,
                                             chunk_vectors, ids, count, lut, m,
                                             ks, out);
}

int ref_pq_train(const float *x, int n, int dim, int m, int ks, int iters,
                 float *codebook) {
    RetrievalPqConfig cfg = {m, ks, dim / m, dim, n, iters};
    return retrieval_pq_train(x, n, &cfg, codebook);
}

int ref_pq_encode(const float *x, int n, int dim, int m, int ks,
                  const float *codebook, uint8_t *codes) {
    RetrievalPqConfig cfg = {m, ks, dim / m, dim, n, 1};
    return retrieval_pq_encode_batch(x, n, codebook, &cfg, codes);
}

int ref_sqmd_seed_ids(int n, int entry_points, int graph_entry,
                      unsigned int seed, int *out, int cap) {
    return retrieval_sqmd_make_seed_ids(n, entry_points, graph_entry, seed, out,
                                        cap);
}

int ref_quantize_query(const float *q, int dim, int bits, float scale,
                       float zero, uint8_t *out) {
    return bits == 4 ? retrieval_quantize_query_u4(q, dim, scale, zero, out)
                     : retrieval_quantize_query_u8(q, dim, scale, zero, out);
}

float ref_sq_distance(const uint8_t *q, const uint8_t *code, int dim, int bits,
                      float scale) {
    return bits == 4
               ? retrieval_compute_int4_packed_distance_u8_scalar(q, code, dim,
                                                                  scale)
               : retrieval_compute_int8_native_distance_u8_scalar(q, code, dim,
                                                                  scale);
}

/* ---- residual artifacts, written and opened by the reference itself ---- */

typedef struct RefResidual {
    char prefix[256];
    char dir[128];
    RetrievalPqArtifactMetadata meta;
    RetrievalPqResidualIndex *index;
    RetrievalPqResidualQuery *query;
    const uint8_t *chunks[1];
    int counts[1];
} RefResidual;

/* Writes PQRES01 artifacts for (x, codes, codebook) with the reference
 * writer, reopens them with the reference reader, and copies the stored
 * residual codes and qparams out (flat [n][code_bytes] and [n][2]). */
void *ref_residual_open(const float *x, const uint8_t *codes,
                        const float *codebook, int n, int dim, int m, int ks,
                        int bits, uint8_t *out_codes, float *out_qparams) {
    RefResidual *r = calloc(1, sizeof(*r));
    RetrievalPqResidualWriter *w = NULL;
    RetrievalPqConfig cfg = {m, ks, dim / m, dim, n, 1};
    RetrievalPqResidualOpenRequest req;
    int cb = bits == 4 ? dim / 2 + dim % 2 : dim;
    int vpb = 4096 / cb;
    char path[300];
    FILE *f;
    if (!r) return NULL;
    snprintf(r->dir, sizeof(r->dir), "/tmp/rover_ref_XXXXXX");
    if (!mkdtemp(r->dir)) goto fail;
    snprintf(r->prefix, sizeof(r->prefix), "%s/index", r->dir);
    for (int i = 0; i < 16; i++) r->meta.index_uuid[i] = (uint8_t)(i + 1);
    r->meta.num_vectors = (uint64_t)n;
    r->meta.dim = dim;
    r->meta.m = m;
    r->meta.ks = ks;
    r->meta.subdim = dim / m;
    r->meta.code_bytes = retrieval_pq_code_bytes(m, ks);
    r->meta.train_count = (uint64_t)n;
    r->meta.kmeans_iters = 1;
    if (retrieval_pq_residual_writer_begin(r->prefix, &r->meta, bits, &w) ||
        retrieval_pq_residual_writer_append(w, x, codes, n, codebook, &cfg) ||
        retrieval_pq_residual_writer_commit(w))
        goto fail;
    snprintf(path, sizeof(path), "%s.pqres%d", r->prefix, bits);
    f = fopen(path, "rb");
    if (!f) goto fail;
    for (int i = 0; i < n; i++) {
        long off = 4096L + (long)(i / vpb) * 4096L + (long)(i % vpb) * cb;
        if (fseek(f, off, SEEK_SET) ||
            fread(out_codes + (size_t)i * cb, 1, cb, f) != (size_t)cb) {
            fclose(f);
            goto fail;
        }
    }
    fclose(f);
    snprintf(path, sizeof(path), "%s.pqres%d.qparams", r->prefix, bits);
    f = fopen(path, "rb");
    if (!f || fseek(f, 256, SEEK_SET) ||
        fread(out_qparams, 8, n, f) != (size_t)n) {
        if (f) fclose(f);
        goto fail;
    }
    fclose(f);
    r->chunks[0] = codes;
    r->counts[0] = n;
    memset(&req, 0, sizeof(req));
    req.artifact_prefix = r->prefix;
    req.residual_bits = bits;
    req.cache_slots = 64;
    req.pq_metadata = &r->meta;
    req.codebook = codebook;
    req.code_chunks = r->chunks;
    req.chunk_vector_counts = r->counts;
    req.chunk_count = 1;
    req.chunk_vectors = n;
    if (retrieval_pq_residual_index_open(&req, &r->index) ||
        retrieval_pq_residual_query_create(r->index, &r->query))
        goto fail;
    return r;
fail:
    free(r);
    return NULL;
}

void ref_residual_close(void *handle) {
    RefResidual *r = handle;
    char path[300];
    if (!r) return;
    retrieval_pq_residual_query_destroy(r->query);
    retrieval_pq_residual_index_close(r->index);
    snprintf(path, sizeof(path), "%s.pqres4", r->prefix); unlink(path);
    snprintf(path, sizeof(path), "%s.pqres8", r->prefix); unlink(path);
    snprintf(path, sizeof(path), "%s.pqres4.qparams", r->prefix); unlink(path);
    snprintf(path, sizeof(path), "%s.pqres8.qparams", r->prefix); unlink(path);
    snprintf(path, sizeof(path), "%s.pqres4.json", r->prefix); unlink(path);
    snprintf(path, sizeof(path), "%s.pqres8.json", r->prefix); unlink(path);
    rmdir(r->dir);
    free(r);
}

int ref_residual_evaluate(void *handle, const float *query, int id,
                          float *out) {
    RefResidual *r = handle;
    if (retrieval_pq_residual_query_bind(r->query, query)) return -1;
    return retrieval_pq_residual_query_evaluate(r->query, id, 0.0f, out);
}

int ref_residual_decode(void *handle, int id, float *out) {
    RefResidual *r = handle;
    return retrieval_pq_residual_query_decode(r->query, id, out);
}

/* ---- the scalar search itself ---- */

typedef struct {
    const uint8_t *qsq;
    const uint8_t *codes;
    int dim, bits;
    float scale;
} RefSqContext;

static int ref_sq_evaluator(void *ctx, int id, float coarse, float *out) {
    const RefSqContext *s = ctx;
    int cb = s->bits == 4 ? (s->dim + 1) / 2 : s->dim;
    (void)coarse;
    *out = ref_sq_distance(s->qsq, s->codes + (size_t)id * cb, s->dim, s->bits,
                           s->scale);
    return isfinite(*out) ? 0 : -1;
}

static int ref_online_evaluator(void *ctx, int id, float coarse, float *out) {
    RefResidual *r = ctx;
    return retrieval_pq_residual_query_evaluate(r->query, id, coarse, out);
}

/* mode 0 = correction OFF, 1 = ONLINE (needs residual handle),
 * 2 = SQ distance through the policy hook (needs sq_codes). */
int ref_search(const int32_t *graph, int degree, int n, const uint8_t *codes,
               int chunk_vectors, const float *codebook, int dim, int m, int ks,
               const float *query, const int *entries, int entry_count,
               int expand_count, int pool, int num_out, int mode,
               void *residual, const uint8_t *sq_codes, int sq_bits,
               float sq_scale, float sq_zero, int *out_ids, float *out_dists,
               int out_cap, uint64_t *out_stats) {
    const uint8_t *chunks[PURE_PQ_MAX_CODE_CHUNKS];
    int counts[PURE_PQ_MAX_CODE_CHUNKS];
    int nchunks = (n + chunk_vectors - 1) / chunk_vectors;
    int code_bytes = retrieval_pq_code_bytes(m, ks);
    PurePqSearchIndexView view;
    PurePqSearchRequest req;
    PurePqDistancePolicy policy;
    PurePqSearchStats st;
    uint8_t qsq[4096];
    RefSqContext sctx;
    int rc;
    if (nchunks > PURE_PQ_MAX_CODE_CHUNKS || dim > 4096) return -2;
    for (int c = 0; c < nchunks; c++) {
        chunks[c] = codes + (size_t)c * chunk_vectors * code_bytes;
        counts[c] = n - c * chunk_vectors < chunk_vectors ? n - c * chunk_vectors
                                                          : chunk_vectors;
    }
    memset(&view, 0, sizeof(view));
    view.graph = graph;
    view.graph_degree = degree;
    view.num_vectors = n;
    view.code_chunks = chunks;
    view.chunk_vector_counts = counts;
    view.chunk_count = nchunks;
    view.chunk_vectors = chunk_vectors;
    view.codebook = codebook;
    view.dim = dim;
    view.m = m;
    view.ks = ks;
    view.subdim = dim / m;
    memset(&req, 0, sizeof(req));
    req.query = query;
    req.entry_ids = entries;
    req.entry_count = entry_count;
    req.expand_count = expand_count;
    req.pool_capacity = pool;
    req.num_out = num_out;
    memset(&policy, 0, sizeof(policy));
    if (mode == 1) {
        RefResidual *r = residual;
        if (!r || retrieval_pq_residual_query_bind(r->query, query)) return -2;
        policy.mode = PURE_PQ_CORRECTION_ONLINE;
        policy.evaluator = ref_online_evaluator;
        policy.context = r;
    } else if (mode == 2) {
        if (ref_quantize_query(query, dim, sq_bits, sq_scale, sq_zero, qsq))
            return -2;
        sctx.qsq = qsq;
        sctx.codes = sq_codes;
        sctx.dim = dim;
        sctx.bits = sq_bits;
        sctx.scale = sq_scale;
        policy.mode = PURE_PQ_CORRECTION_ONLINE;
        policy.evaluator = ref_sq_evaluator;
        policy.context = &sctx;
    }
    rc = retrieval_pure_pq_search_scalar_with_policy(
        &view, &req, mode ? &policy : NULL, out_ids, out_dists, out_cap, &st);
    if (out_stats) {
        out_stats[0] = st.pq_scored_candidates;
        out_stats[1] = st.corrected_candidates;
        out_stats[2] = st.expanded_nodes;
        out_stats[3] = st.visited_dedup_skips;
    }
    return rc;
}

static int ref_refine_decode(void *provider, int id, float *out) {
    RefResidual *r = provider;
    return retrieval_pq_residual_query_decode(r->query, id, out);
}

/* retrieval_final_refine in CURRENT mode with the PQ residual decoder, the
 * composition rpc_demo_pure_pq.c wires up for --refine on. */
int ref_refine(void *residual, int n, int dim, const float *query,
               int32_t *ids, float *dists, int count, int output_k,
               const int32_t *graph_to_original) {
    RetrievalFinalRefineConfig c;
    RetrievalFinalRefineContext *ctx = NULL;
    int rc;
    memset(&c, 0, sizeof(c));
    c.mode = RETRIEVAL_REFINE_CURRENT;
    c.num_vectors = n;
    c.dim = dim;
    c.capacity = count > 0 ? count : 1;
    c.cache_slots = 1;
    c.decode = ref_refine_decode;
    c.provider = residual;
    c.graph_to_original = graph_to_original;
    if (retrieval_final_refine_open_pq(&c, "unused", &ctx)) return -100;
    rc = retrieval_final_refine(ctx, query, ids, dists, count, output_k, NULL);
    retrieval_final_refine_close(ctx);
    return rc;
}
