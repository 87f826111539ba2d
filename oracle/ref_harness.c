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

#include "dsp_demo/retrieval_build_artifacts.h"
#include "dsp_demo/retrieval_final_refine.h"
#include "dsp_demo/retrieval_nsg_loader.h"
#include "dsp_demo/retrieval_search_build_core.h"
#include "dsp_demo/retrieval_pq_artifact.h"
#include "dsp_demo/retrieval_raw_reader.h"
#include "dsp_demo/retrieval_l1_int8_core.h"
#include "dsp_demo/retrieval_pq_core.h"
#include "dsp_demo/retrieval_pq_residual.h"
#include "dsp_demo/retrieval_sqmd_common.h"
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
    const float *codebook;
    /* error4 artifact and the files it is bound to (ref_error4_build) */
    int res_bits;
    int e4_ready;
    char e4_path[300], e4_raw[300], e4_codes[300], e4_book[300];
    char e4_base[3][300];
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
    r->res_bits = bits;
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
    r->codebook = codebook;
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

/* A second reader of the artifacts `handle` wrote, with its own residual
 * index and query (the reference's reader keeps per-query state and a block
 * cache, so concurrent searches need one each).  Close forks before the
 * handle that owns the files. */
void *ref_residual_fork(void *handle) {
    RefResidual *r = handle, *f;
    RetrievalPqResidualOpenRequest req;
    if (!r || !(f = calloc(1, sizeof(*f)))) return NULL;
    memcpy(f->prefix, r->prefix, sizeof(f->prefix));
    f->dir[0] = 0; /* owns no files */
    f->meta = r->meta;
    f->res_bits = r->res_bits;
    f->chunks[0] = r->chunks[0];
    f->counts[0] = r->counts[0];
    memset(&req, 0, sizeof(req));
    req.artifact_prefix = f->prefix;
    req.residual_bits = f->res_bits;
    req.cache_slots = 64;
    req.pq_metadata = &f->meta;
    req.codebook = r->codebook;
    f->codebook = r->codebook;
    req.code_chunks = f->chunks;
    req.chunk_vector_counts = f->counts;
    req.chunk_count = 1;
    req.chunk_vectors = f->counts[0];
    if (retrieval_pq_residual_index_open(&req, &f->index) ||
        retrieval_pq_residual_query_create(f->index, &f->query)) {
        free(f);
        return NULL;
    }
    return f;
}

void ref_residual_close(void *handle) {
    RefResidual *r = handle;
    char path[300];
    if (!r) return;
    if (!r->dir[0]) {
        retrieval_pq_residual_query_destroy(r->query);
        retrieval_pq_residual_index_close(r->index);
        free(r);
        return;
    }
    if (r->e4_ready) {
        unlink(r->e4_path);
        unlink(r->e4_raw);
        unlink(r->e4_codes);
        unlink(r->e4_book);
    }
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

/* The reference's live pipeline for one query with `window` source turns
 * in flight: pure_pq_cpu_search_create/reset/seed, then
 * pure_pq_source_pipeline_step until it reports the search done, with the
 * DSP side (pure_pq_sqmd_process_buffer: ADC of each turn's work) run on
 * this thread between steps on every buffer the CPU side has filled -- the
 * demo's per-query driver (rpc_demo_pure_pq.c:283-321) with its seeds
 * scored by ADC first (:286-290).  mode 0 = correction OFF, 1 = ONLINE
 * (residual handle).  The pool is returned sorted (pure_pq_cpu_search_finish)
 * and cut to num_out; out_stats = (turns dispatched, work items, high water,
 * steps). */
int ref_pipeline_search(const int32_t *graph, int degree, int n, const uint8_t *codes,
                        int chunk_vectors, const float *codebook, int dim, int m,
                        int ks, const float *query, const int *entries,
                        int entry_count, int expand_count, int pool, int num_out,
                        int window, int lanes, int mode, void *residual,
                        int *out_ids, float *out_dists, int out_cap,
                        uint64_t *out_stats) {
    const uint8_t *chunks[PURE_PQ_MAX_CODE_CHUNKS];
    int counts[PURE_PQ_MAX_CODE_CHUNKS];
    int nchunks = (n + chunk_vectors - 1) / chunk_vectors;
    static const uint8_t uuid[16] = {0x52, 0x6f, 0x76, 0x65, 0x72, 0x2d, 0x47, 0x50,
                                     0x55, 0x2d, 0x70, 0x69, 0x70, 0x65, 0x2d, 0x31};
    PurePqQueryLutHeader hdr;
    PurePqSqmdSharedControl *control = NULL;
    PurePqSqmdAdcContext ctx;
    PurePqCpuSearchState *state = NULL;
    PurePqDistancePolicy policy;
    PurePqSourcePipeline pl;
    float *lut = NULL, *seed_d = NULL, *all_d = NULL;
    int *all_ids = NULL;
    uint64_t steps = 0;
    int rc = -2, count;
    if (nchunks > PURE_PQ_MAX_CODE_CHUNKS || num_out <= 0 || num_out > pool) return -2;
    for (int c = 0; c < nchunks; c++) {
        chunks[c] = codes + (size_t)c * chunk_vectors * retrieval_pq_code_bytes(m, ks);
        counts[c] = n - c * chunk_vectors < chunk_vectors ? n - c * chunk_vectors
                                                          : chunk_vectors;
    }
    lut = malloc((size_t)m * ks * sizeof(float));
    seed_d = malloc((size_t)entry_count * sizeof(float));
    all_ids = malloc((size_t)pool * sizeof(int));
    all_d = malloc((size_t)pool * sizeof(float));
    control = aligned_alloc(PURE_PQ_SQMD_CACHE_LINE_BYTES,
                            PURE_PQ_SQMD_ALIGN_UP(sizeof(*control),
                                                  PURE_PQ_SQMD_CACHE_LINE_BYTES));
    if (!lut || !seed_d || !all_ids || !all_d || !control) goto out;
    if (retrieval_pq_build_lut(query, codebook, m, ks, dim / m, lut)) goto out;
    memset(&hdr, 0, sizeof(hdr));
    hdr.m = m;
    hdr.ks = ks;
    hdr.lut_count = m * ks;
    memcpy(hdr.index_uuid, uuid, 16);
    hdr.lut_gen = 1;
    hdr.query_gen = 1;
    pure_pq_sqmd_control_init(control, lanes);
    control->query_gen = 1;
    memset(&ctx, 0, sizeof(ctx));
    ctx.code_chunks = chunks;
    ctx.chunk_vector_counts = counts;
    ctx.chunk_count = nchunks;
    ctx.chunk_vectors = chunk_vectors;
    ctx.num_vectors = n;
    ctx.m = m;
    ctx.ks = ks;
    ctx.index_uuid = uuid;
    ctx.lut_header = &hdr;
    ctx.lut = lut;
    if (retrieval_pq_adc_ids_batch_scalar(chunks, counts, nchunks, chunk_vectors,
                                          (const int32_t *)entries, entry_count, lut,
                                          m, ks, seed_d))
        goto out;
    memset(&policy, 0, sizeof(policy));
    if (mode == 1) {
        RefResidual *r = residual;
        if (!r || retrieval_pq_residual_query_bind(r->query, query)) goto out;
        policy.mode = PURE_PQ_CORRECTION_ONLINE;
        policy.evaluator = ref_online_evaluator;
        policy.context = r;
    }
    rc = -1;
    if (pure_pq_cpu_search_create(n, pool, expand_count, &state) ||
        pure_pq_cpu_search_reset(state, n, pool, expand_count) ||
        pure_pq_cpu_search_set_distance_policy(state, mode ? &policy : NULL) ||
        pure_pq_cpu_search_seed(state, entries, seed_d, entry_count) ||
        pure_pq_source_pipeline_init(&pl, window))
        goto out;
    for (;;) {
        int step = pure_pq_source_pipeline_step(&pl, state, graph, degree, control, 1);
        steps++;
        if (step < 0) goto out;
        if (step == 1) break;
        for (int lane = 0; lane < control->lane_count; lane++)
            for (int bi = 0; bi < PURE_PQ_SQMD_BUFFERS_PER_LANE; bi++) {
                PurePqSqmdLaneBuffer *buf = &control->lanes[lane].buffers[bi];
                if (buf->state == PURE_PQ_SQMD_BUFFER_CPU_READY &&
                    pure_pq_sqmd_process_buffer(buf, buf->turn_id, 1, &ctx))
                    goto out;
            }
        if (steps > 100000000ull) goto out;
    }
    count = pure_pq_cpu_search_finish(state, all_ids, all_d, pool);
    if (count < 0) goto out;
    rc = count < num_out ? count : num_out;
    if (rc > out_cap) rc = out_cap;
    memcpy(out_ids, all_ids, (size_t)rc * sizeof(int));
    memcpy(out_dists, all_d, (size_t)rc * sizeof(float));
    if (out_stats) {
        out_stats[0] = pl.jobs;
        out_stats[1] = pl.work_items;
        out_stats[2] = (uint64_t)pl.high_water;
        out_stats[3] = steps;
    }
out:
    pure_pq_cpu_search_destroy(state);
    free(lut);
    free(seed_d);
    free(all_ids);
    free(all_d);
    free(control);
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

/* ---- third-level refine (error4), written and read by the reference ---- */

static int ref_write_file(const char *path, const void *a, size_t an,
                          const void *b, size_t bn) {
    FILE *f = fopen(path, "wb");
    int ok;
    if (!f) return -1;
    ok = fwrite(a, 1, an, f) == an && (!bn || fwrite(b, 1, bn, f) == bn);
    return fclose(f) == 0 && ok ? 0 : -1;
}

/* Writes the base vectors as a raw fvecs file, the PQ codes and codebook as
 * files (only hashed), runs retrieval_build_error4 with the residual decoder
 * as the reconstruction, and copies the artifact's header [256], qparams
 * [n][2] and codes (flattened out of the 4096-byte blocks, [n][(dim+1)/2])
 * out.  The artifact stays on disk for ref_refine_mode(..., ERROR4). */
int ref_error4_build(void *handle, const float *x, const uint8_t *codes,
                     int code_bytes, const float *codebook, int book_floats,
                     int n, int dim, uint8_t *out_header, float *out_qparams,
                     uint8_t *out_codes) {
    RefResidual *r = handle;
    RetrievalRawReader *raw = NULL;
    const char *paths[3];
    int cb = (dim + 1) / 2, vpb = 4096 / cb, rc;
    FILE *f, *raw_f;
    if (!r || r->e4_ready) return -1;
    snprintf(r->e4_raw, sizeof(r->e4_raw), "%s/base.fvecs", r->dir);
    snprintf(r->e4_codes, sizeof(r->e4_codes), "%s/index.pqcodes", r->dir);
    snprintf(r->e4_book, sizeof(r->e4_book), "%s/index.pqbook", r->dir);
    snprintf(r->e4_path, sizeof(r->e4_path), "%s/error4.refine", r->dir);
    snprintf(r->e4_base[0], sizeof(r->e4_base[0]), "%s", r->e4_codes);
    snprintf(r->e4_base[1], sizeof(r->e4_base[1]), "%s.pqres%d", r->prefix,
             r->res_bits);
    snprintf(r->e4_base[2], sizeof(r->e4_base[2]), "%s.pqres%d.qparams",
             r->prefix, r->res_bits);
    raw_f = fopen(r->e4_raw, "wb");
    if (!raw_f) return -2;
    for (int i = 0; i < n; i++) {
        int32_t d32 = dim;
        if (fwrite(&d32, 4, 1, raw_f) != 1 ||
            fwrite(x + (size_t)i * dim, 4, (size_t)dim, raw_f) != (size_t)dim) {
            fclose(raw_f);
            return -2;
        }
    }
    if (fclose(raw_f) ||
        ref_write_file(r->e4_codes, codes, (size_t)n * code_bytes, NULL, 0) ||
        ref_write_file(r->e4_book, codebook, (size_t)book_floats * 4, NULL, 0))
        return -2;
    if (retrieval_raw_open(r->e4_raw, n, dim, (size_t)1 << 20, &raw)) return -3;
    for (int i = 0; i < 3; i++) paths[i] = r->e4_base[i];
    rc = retrieval_build_error4(r->e4_path, raw, n, dim, ref_refine_decode, r,
                                paths, r->e4_book, 1);
    retrieval_raw_close(raw);
    if (rc) return -4;
    r->e4_ready = 1;
    f = fopen(r->e4_path, "rb");
    if (!f) return -5;
    rc = fread(out_header, 1, 256, f) == 256 &&
         fread(out_qparams, 8, (size_t)n, f) == (size_t)n ? 0 : -5;
    for (int i = 0; !rc && i < n; i++) {
        long off = 256L + 8L * n + (long)(i / vpb) * 4096L + (long)(i % vpb) * cb;
        if (fseek(f, off, SEEK_SET) ||
            fread(out_codes + (size_t)i * cb, 1, cb, f) != (size_t)cb)
            rc = -5;
    }
    fclose(f);
    return rc;
}

/* Bytes of the artifact file ref_error4_build wrote (to copy it whole). */
long ref_error4_file(void *handle, uint8_t *out, long cap) {
    RefResidual *r = handle;
    FILE *f;
    long size;
    if (!r || !r->e4_ready || !(f = fopen(r->e4_path, "rb"))) return -1;
    fseek(f, 0, SEEK_END);
    size = ftell(f);
    if (out && size <= cap) {
        fseek(f, 0, SEEK_SET);
        if (fread(out, 1, (size_t)size, f) != (size_t)size) size = -1;
    }
    fclose(f);
    return size;
}

/* retrieval_final_refine with the PQ residual decoder in `mode`
 * (RETRIEVAL_REFINE_CURRENT or RETRIEVAL_REFINE_ERROR4).  ERROR4 opens the
 * artifact written by ref_error4_build through retrieval_final_refine_open_pq,
 * the path rpc_demo_pure_pq.c:518-533 takes: whole-file digest, payload
 * digest, the three base files and the codebook are all checked.  Returns the
 * output count, or the open error - 100. */
int ref_refine_mode(void *residual, int mode, int n, int dim,
                    const float *query, int32_t *ids, float *dists, int count,
                    int output_k, const int32_t *graph_to_original) {
    RefResidual *r = residual;
    RetrievalFinalRefineConfig c;
    RetrievalFinalRefineContext *ctx = NULL;
    unsigned char digest[32];
    const char *book = "unused";
    int rc;
    memset(&c, 0, sizeof(c));
    c.mode = mode;
    c.num_vectors = n;
    c.dim = dim;
    c.capacity = count > 0 ? count : 1;
    c.cache_slots = 4;
    c.decode = ref_refine_decode;
    c.provider = residual;
    c.graph_to_original = graph_to_original;
    if (mode == RETRIEVAL_REFINE_ERROR4) {
        if (!r->e4_ready || retrieval_build_hash(r->e4_path, digest)) return -200;
        c.artifact_path = r->e4_path;
        c.artifact_sha256 = digest;
        for (int i = 0; i < 3; i++) c.base_paths[i] = r->e4_base[i];
        book = r->e4_book;
    }
    rc = retrieval_final_refine_open_pq(&c, book, &ctx);
    if (rc) return rc - 100;
    rc = retrieval_final_refine(ctx, query, ids, dists, count, output_k, NULL);
    retrieval_final_refine_close(ctx);
    return rc;
}

/* ---- SQ final refine (SQERR4V1), written and read by the reference ---- */

typedef struct RefSqRefine {
    char dir[128], raw[300], l1[300], res[300], qp[300], art[300];
    const uint8_t *l1_codes, *res_codes;
    const float *res_qp;
    int n, dim, l1_bits, res_bits, l1_cb, res_cb, ready;
    float l1_scale, l1_zero;
} RefSqRefine;

/* retrieval_final_refine_decode_sq over the SQ codes and residuals: the
 * reconstruction the live SQ path refines from (pipeline.c:2725-2738). */
static int ref_sq_refine_decode(void *provider, int id, float *out) {
    const RefSqRefine *r = provider;
    if (id < 0 || id >= r->n) return RETRIEVAL_REFINE_INVALID;
    return retrieval_final_refine_decode_sq(
        r->l1_codes + (size_t)id * r->l1_cb, r->l1_bits, r->l1_scale, r->l1_zero,
        r->res_codes + (size_t)id * r->res_cb, r->res_bits, r->res_qp[2 * (size_t)id],
        r->res_qp[2 * (size_t)id + 1], r->dim, out);
}

/* Writes the base vectors (raw fvecs) and the SQ code, residual and qparams
 * arrays as files (only hashed), builds the SQERR4V1 artifact with
 * retrieval_build_error4 (pq = 0) over the SQ decoder, and returns a handle
 * for ref_sq_refine; the artifact's bytes via ref_sq_refine_artifact. */
void *ref_sq_refine_open(const float *x, const uint8_t *l1, int l1_bits, float l1_scale,
                         float l1_zero, const uint8_t *res, int res_bits,
                         const float *res_qp, int n, int dim) {
    RefSqRefine *r = calloc(1, sizeof(*r));
    RetrievalRawReader *raw = NULL;
    const char *paths[3];
    FILE *f;
    if (!r) return NULL;
    r->l1_codes = l1;
    r->res_codes = res;
    r->res_qp = res_qp;
    r->n = n;
    r->dim = dim;
    r->l1_bits = l1_bits;
    r->res_bits = res_bits;
    r->l1_cb = l1_bits == 4 ? (dim + 1) / 2 : dim;
    r->res_cb = res_bits == 4 ? (dim + 1) / 2 : dim;
    r->l1_scale = l1_scale;
    r->l1_zero = l1_zero;
    snprintf(r->dir, sizeof(r->dir), "/tmp/rover_sqref_XXXXXX");
    if (!mkdtemp(r->dir)) goto fail;
    snprintf(r->raw, sizeof(r->raw), "%s/base.fvecs", r->dir);
    snprintf(r->l1, sizeof(r->l1), "%s/l1.bin", r->dir);
    snprintf(r->res, sizeof(r->res), "%s/residual.bin", r->dir);
    snprintf(r->qp, sizeof(r->qp), "%s/qparams.bin", r->dir);
    snprintf(r->art, sizeof(r->art), "%s/error4.refine", r->dir);
    f = fopen(r->raw, "wb");
    if (!f) goto fail;
    for (int i = 0; i < n; i++) {
        int32_t d32 = dim;
        if (fwrite(&d32, 4, 1, f) != 1 ||
            fwrite(x + (size_t)i * dim, 4, (size_t)dim, f) != (size_t)dim) {
            fclose(f);
            goto fail;
        }
    }
    if (fclose(f) || ref_write_file(r->l1, l1, (size_t)n * r->l1_cb, NULL, 0) ||
        ref_write_file(r->res, res, (size_t)n * r->res_cb, NULL, 0) ||
        ref_write_file(r->qp, res_qp, (size_t)n * 8, NULL, 0))
        goto fail;
    if (retrieval_raw_open(r->raw, n, dim, (size_t)1 << 20, &raw)) goto fail;
    paths[0] = r->l1;
    paths[1] = r->res;
    paths[2] = r->qp;
    if (retrieval_build_error4(r->art, raw, n, dim, ref_sq_refine_decode, r, paths,
                               r->raw, 0)) {
        retrieval_raw_close(raw);
        goto fail;
    }
    retrieval_raw_close(raw);
    r->ready = 1;
    return r;
fail:
    free(r);
    return NULL;
}

long ref_sq_refine_artifact(void *handle, uint8_t *out, long cap) {
    RefSqRefine *r = handle;
    FILE *f;
    long size;
    if (!r || !r->ready || !(f = fopen(r->art, "rb"))) return -1;
    fseek(f, 0, SEEK_END);
    size = ftell(f);
    if (out && size <= cap) {
        fseek(f, 0, SEEK_SET);
        if (fread(out, 1, (size_t)size, f) != (size_t)size) size = -1;
    }
    fclose(f);
    return size;
}

/* retrieval_final_refine with the SQ decoder in `mode` (CURRENT or ERROR4,
 * the latter through retrieval_final_refine_open with no codebook: the
 * SQERR4V1 path of pipeline.c:3026-3045).  Returns the output count, or the
 * open error - 100. */
int ref_sq_refine(void *handle, int mode, const float *query, int32_t *ids, float *dists,
                  int count, int output_k, const int32_t *graph_to_original) {
    RefSqRefine *r = handle;
    RetrievalFinalRefineConfig c;
    RetrievalFinalRefineContext *ctx = NULL;
    unsigned char digest[32];
    int rc;
    memset(&c, 0, sizeof(c));
    c.mode = mode;
    c.num_vectors = r->n;
    c.dim = r->dim;
    c.capacity = count > 0 ? count : 1;
    c.cache_slots = 4;
    c.decode = ref_sq_refine_decode;
    c.provider = r;
    c.graph_to_original = graph_to_original;
    if (mode == RETRIEVAL_REFINE_ERROR4) {
        if (retrieval_build_hash(r->art, digest)) return -200;
        c.artifact_path = r->art;
        c.artifact_sha256 = digest;
        c.base_paths[0] = r->l1;
        c.base_paths[1] = r->res;
        c.base_paths[2] = r->qp;
    }
    rc = retrieval_final_refine_open(&c, &ctx);
    if (rc) return rc - 100;
    rc = retrieval_final_refine(ctx, query, ids, dists, count, output_k, NULL);
    retrieval_final_refine_close(ctx);
    return rc;
}

void ref_sq_refine_close(void *handle) {
    RefSqRefine *r = handle;
    if (!r) return;
    unlink(r->raw);
    unlink(r->l1);
    unlink(r->res);
    unlink(r->qp);
    unlink(r->art);
    rmdir(r->dir);
    free(r);
}

/* ---- the live SQ traversal: sqmd_cpu_state_run_fake_dsp ---- */

typedef struct {
    const uint8_t *codes, *qsq, *res;
    const float *query, *res_qp;
    int dim, bits, cb, res_bits, res_cb;
    float scale, zero;
} RefSqmdCtx;

static float ref_sqmd_coarse(const RefSqmdCtx *c, int id) {
    return ref_sq_distance(c->qsq, c->codes + (size_t)id * c->cb, c->dim, c->bits,
                           c->scale);
}

static int ref_sqmd_distance_fn(const SqmdWorkItem *work, float *coarse, void *ctx) {
    *coarse = ref_sqmd_coarse(ctx, work->neighbor_id);
    return 0;
}

/* The exact SQ + residual distance: the vector decoded by the reference's
 * retrieval_final_refine_decode_sq (l1 and residual levels, each
 * (code * scale) + zero, summed), then the sequential f32 sum of squares. */
static float ref_sqmd_exact(const RefSqmdCtx *c, int id) {
    float v[8192], dist = 0.0f;
    const float *qp = c->res_qp + 2 * (size_t)id;
    if (retrieval_final_refine_decode_sq(c->codes + (size_t)id * c->cb, c->bits, c->scale,
                                         c->zero, c->res + (size_t)id * c->res_cb,
                                         c->res_bits, qp[0], qp[1], c->dim, v))
        return FLT_MAX;
    for (int d = 0; d < c->dim; d++) {
        float diff = c->query[d] - v[d];
        dist += diff * diff;
    }
    return dist;
}

static float ref_sqmd_correction_fn(int id, int source, float coarse, void *ctx) {
    (void)source;
    (void)coarse;
    return ref_sqmd_exact(ctx, id);
}

/* retrieval_sqmd_make_seed_ids + seed distances + the fake-DSP traversal,
 * as sqmd_search_one_query_fake_dsp (retrieval_rpc_demo_pipeline.c:
 * 1060-1125) composes them.  correction 0: SQ code distances throughout;
 * 1: exact SQ + residual distances for seeds and results. */
int ref_sqmd_search(const int32_t *graph, int degree, int n, const uint8_t *codes,
                    int dim, int bits, float scale, float zero, const uint8_t *res,
                    const float *res_qp, int res_bits, int correction,
                    const float *query, int entry_points, int graph_entry,
                    unsigned int seed, int expand_count, int pool_capacity,
                    int lanes, int32_t *out_ids, float *out_dists, int out_cap,
                    int32_t *out_seed_ids) {
    RefSqmdCtx c;
    uint8_t qsq[8192];
    int seed_ids[2048];
    float seed_dists[2048];
    int count, rc;
    if (entry_points > 2048 || dim > 8192) return -100;
    memset(&c, 0, sizeof(c));
    c.codes = codes;
    c.dim = dim;
    c.bits = bits;
    c.cb = bits == 4 ? (dim + 1) / 2 : dim;
    c.scale = scale;
    c.zero = zero;
    c.res = res;
    c.res_qp = res_qp;
    c.res_bits = res_bits;
    c.res_cb = res_bits == 4 ? (dim + 1) / 2 : dim;
    c.query = query;
    c.qsq = qsq;
    if (ref_quantize_query(query, dim, bits, scale, zero, qsq)) return -101;
    count = retrieval_sqmd_make_seed_ids(n, entry_points, graph_entry, seed, seed_ids,
                                         entry_points);
    if (count <= 0) return -102;
    for (int i = 0; i < count; i++) {
        seed_dists[i] = correction ? ref_sqmd_exact(&c, seed_ids[i])
                                   : ref_sqmd_coarse(&c, seed_ids[i]);
        if (out_seed_ids) out_seed_ids[i] = seed_ids[i];
    }
    rc = sqmd_cpu_state_run_fake_dsp(graph, degree, n, pool_capacity, entry_points,
                                     expand_count, seed_ids, seed_dists, count, lanes,
                                     ref_sqmd_distance_fn, &c,
                                     correction ? ref_sqmd_correction_fn : NULL, &c,
                                     out_ids, out_dists, out_cap);
    return rc;
}

/* ---- the split distance (retrieval_rpc_demo_pipeline.c:906-1000) ----
 * The query on the 17x-refined SQ grid through the reference's
 * retrieval_quantize_query_l1_aligned_u8; coarse distances through
 * retrieval_compute_u8_u4_l1_distance_scalar (the DSP's split contract,
 * retrieval_async_dsp_search.c:1004-1009); the delta through
 * retrieval_compute_u8_u{4,8}_residual_dot_scalar and the bias, composed as
 * sqmd_split_residual_delta_fn / sqmd_split_distance_for_id compose them
 * (those two are static in the pipeline file); the traversal is the
 * fake-DSP turn loop with coarse + delta as the applied distance. */
typedef struct {
    RefSqmdCtx base;
    const float *bias;
    uint8_t qu8[8192];
    uint64_t sum_q;
} RefSplitCtx;

static float ref_split_delta(const RefSplitCtx *c, int id) {
    const float *qp = c->base.res_qp + 2 * (size_t)id;
    const uint8_t *res = c->base.res + (size_t)id * c->base.res_cb;
    double dot = c->base.res_bits == 8
        ? retrieval_compute_u8_u8_residual_dot_scalar(c->qu8, c->sum_q, res, c->base.dim,
                                                      c->base.scale, c->base.zero, qp[0], qp[1])
        : retrieval_compute_u8_u4_residual_dot_scalar(c->qu8, c->sum_q, res, c->base.dim,
                                                      c->base.scale, c->base.zero, qp[0], qp[1]);
    double delta = (double)c->bias[id] - 2.0 * dot;
    return isfinite(delta) ? (float)delta : NAN;
}

static float ref_split_coarse(const RefSplitCtx *c, int id) {
    return retrieval_compute_u8_u4_l1_distance_scalar(
        c->qu8, c->base.codes + (size_t)id * c->base.cb, c->base.dim, c->base.scale);
}

static int ref_split_distance_fn(const SqmdWorkItem *work, float *coarse, void *ctx) {
    *coarse = ref_split_coarse(ctx, work->neighbor_id);
    return 0;
}

static float ref_split_apply(float coarse, float delta) {
    float d = coarse + delta;
    if (d < 0.0f && d > -1.0e-4f) d = 0.0f;
    return d;
}

static float ref_split_correction_fn(int id, int source, float coarse, void *ctx) {
    (void)source;
    return ref_split_apply(coarse, ref_split_delta(ctx, id));
}

/* Returns the result count, or < 0 (-101: the query does not quantize,
 * -103: a seed's split distance fails, as sqmd_make_seed rejects it). */
int ref_sqmd_split_search(const int32_t *graph, int degree, int n, const uint8_t *codes,
                          int dim, float scale, float zero, const uint8_t *res,
                          const float *res_qp, int res_bits, const float *bias,
                          const float *query, int entry_points, int graph_entry,
                          unsigned int seed, int expand_count, int pool_capacity, int lanes,
                          int32_t *out_ids, float *out_dists, int out_cap) {
    RefSplitCtx c;
    RetrievalQueryU8Stats qst;
    int seed_ids[2048];
    float seed_dists[2048];
    int count;
    if (entry_points > 2048 || dim > 8192) return -100;
    memset(&c, 0, sizeof(c));
    c.base.codes = codes;
    c.base.dim = dim;
    c.base.bits = 4;
    c.base.cb = (dim + 1) / 2;
    c.base.scale = scale;
    c.base.zero = zero;
    c.base.res = res;
    c.base.res_qp = res_qp;
    c.base.res_bits = res_bits;
    c.base.res_cb = res_bits == 4 ? (dim + 1) / 2 : dim;
    c.base.query = query;
    c.bias = bias;
    if (retrieval_quantize_query_l1_aligned_u8(query, dim, scale, zero, c.qu8, &qst)) return -101;
    c.sum_q = qst.sum_q;
    count = retrieval_sqmd_make_seed_ids(n, entry_points, graph_entry, seed, seed_ids,
                                         entry_points);
    if (count <= 0) return -102;
    for (int i = 0; i < count; i++) {
        float coarse = ref_split_coarse(&c, seed_ids[i]);
        float delta = ref_split_delta(&c, seed_ids[i]);
        float d = (!isfinite(coarse) || !isfinite(delta)) ? FLT_MAX : ref_split_apply(coarse, delta);
        if (!isfinite(d) || d < 0.0f) return -103;
        seed_dists[i] = d;
    }
    return sqmd_cpu_state_run_fake_dsp(graph, degree, n, pool_capacity, entry_points,
                                       expand_count, seed_ids, seed_dists, count, lanes,
                                       ref_split_distance_fn, &c, ref_split_correction_fn, &c,
                                       out_ids, out_dists, out_cap);
}

/* ---- on-disk formats, written and read by the reference ---- */

static void ref_fill_meta(RetrievalPqArtifactMetadata *m, const uint8_t *uuid, int n,
                          int dim, int pq_m, int ks, int trainer, int sample,
                          int train_count, int iters, unsigned int seed,
                          const uint8_t *dataset_fp, const uint8_t *layout_fp) {
    memset(m, 0, sizeof(*m));
    memcpy(m->index_uuid, uuid, 16);
    m->num_vectors = (uint64_t)n;
    m->dim = (uint32_t)dim;
    m->m = (uint32_t)pq_m;
    m->ks = (uint32_t)ks;
    m->subdim = (uint32_t)(dim / pq_m);
    m->code_bytes = (uint32_t)retrieval_pq_code_bytes(pq_m, ks);
    m->trainer_kind = (uint32_t)trainer;
    m->sample_strategy = (uint32_t)sample;
    m->train_count = (uint64_t)train_count;
    m->kmeans_iters = (uint32_t)iters;
    m->seed = seed;
    memcpy(m->dataset_fingerprint, dataset_fp, 32);
    memcpy(m->layout_fingerprint, layout_fp, 32);
}

/* PUREPQ01 .pqbook/.pqcodes through retrieval_pq_artifacts_write. */
int ref_pq_artifacts_write(const char *prefix, const uint8_t *uuid, int n, int dim,
                           int pq_m, int ks, int trainer, int sample, int train_count,
                           int iters, unsigned int seed, const uint8_t *dataset_fp,
                           const uint8_t *layout_fp, const float *codebook,
                           const uint8_t *codes) {
    RetrievalPqArtifactMetadata m;
    ref_fill_meta(&m, uuid, n, dim, pq_m, ks, trainer, sample, train_count, iters, seed,
                  dataset_fp, layout_fp);
    return retrieval_pq_artifacts_write(prefix, &m, codebook, codes);
}

/* retrieval_pq_artifacts_load: returns 0 and copies the payloads out, or the
 * loader's error. */
int ref_pq_artifacts_load(const char *prefix, int *out_n, int *out_dim, int *out_m,
                          int *out_ks, float *codebook, size_t codebook_cap,
                          uint8_t *codes, size_t codes_cap) {
    RetrievalPqArtifacts a;
    int rc;
    memset(&a, 0, sizeof(a));
    rc = retrieval_pq_artifacts_load(prefix, &a);
    if (rc) return rc;
    *out_n = (int)a.metadata.num_vectors;
    *out_dim = (int)a.metadata.dim;
    *out_m = (int)a.metadata.m;
    *out_ks = (int)a.metadata.ks;
    if (a.codebook_values <= codebook_cap && a.codes_bytes <= codes_cap) {
        memcpy(codebook, a.codebook, a.codebook_values * sizeof(float));
        memcpy(codes, a.codes, a.codes_bytes);
    } else {
        rc = -50;
    }
    retrieval_pq_artifacts_release(&a);
    return rc;
}

/* The prefix under which ref_residual_open's PQRES01 files lie. */
const char *ref_residual_prefix(void *handle) {
    RefResidual *r = handle;
    return r ? r->prefix : NULL;
}

/* Opens PQRES01 files at `prefix` (written elsewhere) with the reference's
 * reader over the given PQ metadata / codebook / codes, and decodes `count`
 * vectors (ids) into out [count][dim].  Returns 0 or the open / decode
 * error. */
int ref_pq_residual_read(const char *prefix, const uint8_t *uuid, int n, int dim,
                         int pq_m, int ks, int train_count, int iters,
                         const uint8_t *dataset_fp, const uint8_t *layout_fp,
                         const float *codebook, const uint8_t *codes, int bits,
                         const int32_t *ids, int count, float *out) {
    RetrievalPqArtifactMetadata m;
    RetrievalPqResidualOpenRequest req;
    RetrievalPqResidualIndex *index = NULL;
    RetrievalPqResidualQuery *query = NULL;
    const uint8_t *chunks[1] = {codes};
    int counts[1] = {n};
    int rc;
    ref_fill_meta(&m, uuid, n, dim, pq_m, ks, 0, 0, train_count, iters, 0, dataset_fp,
                  layout_fp);
    memset(&req, 0, sizeof(req));
    req.artifact_prefix = prefix;
    req.residual_bits = bits;
    req.cache_slots = 8;
    req.pq_metadata = &m;
    req.codebook = codebook;
    req.code_chunks = chunks;
    req.chunk_vector_counts = counts;
    req.chunk_count = 1;
    req.chunk_vectors = n;
    rc = retrieval_pq_residual_index_open(&req, &index);
    if (rc) return -1000 + rc;
    rc = retrieval_pq_residual_query_create(index, &query);
    for (int i = 0; !rc && i < count; i++)
        rc = retrieval_pq_residual_query_decode(query, ids[i], out + (size_t)i * dim);
    retrieval_pq_residual_query_destroy(query);
    retrieval_pq_residual_index_close(index);
    return rc;
}

/* results.jsonl through rover_results_open / rover_write_result /
 * rover_results_close (config.c:129-155); order_map may be NULL. */
int ref_results_write(const char *path, int n, const char *order_map, int nq,
                      const int32_t *ids, const float *dists, const int32_t *counts,
                      int stride) {
    int rc = 0;
    rover_config.output = path;
    rover_config.n = n;
    rover_config.order_map = order_map ? order_map : "";
    if (rover_results_open()) return -1;
    for (int q = 0; q < nq && !rc; q++)
        rc = rover_write_result(q, ids + (size_t)q * stride, dists + (size_t)q * stride,
                                counts[q]);
    if (rover_results_close()) rc = -2;
    return rc;
}

/* retrieval_nsg_load_fixed: graph [n][width], stats {file width, entry,
 * truncated nodes, truncated edges}. */
int ref_nsg_load(const char *path, int n, int width, int32_t *graph, int64_t *stats) {
    RetrievalNsgLoadStats st;
    char err[256];
    int rc = retrieval_nsg_load_fixed(path, n, width, graph, &st, err, sizeof(err));
    stats[0] = st.file_width;
    stats[1] = st.entry_point;
    stats[2] = (int64_t)st.truncated_nodes;
    stats[3] = (int64_t)st.truncated_edges;
    return rc;
}

/* ---- graph construction: retrieval_search_build_graph_ex with the exact
 * provider (brute-force candidates), one iteration ---- */
int ref_build_exact(const float *base, int n, int dim, int r, int candidates,
                    unsigned int seed, int prune_mode, int threads, int32_t *out) {
    RetrievalSearchBuildConfig cfg;
    RetrievalSearchBuildRuntime rt;
    RetrievalRnnGraph g;
    int rc;
    memset(&cfg, 0, sizeof(cfg));
    memset(&rt, 0, sizeof(rt));
    memset(&g, 0, sizeof(g));
    cfg.r = r;
    cfg.iterations = 1;
    cfg.candidate_count = candidates;
    cfg.entry_points = 1;
    cfg.expand_count = 1;
    cfg.pool_size = candidates;
    cfg.seed = seed;
    cfg.prune_mode = prune_mode;
    cfg.batch_align = 1;
    rt.threads = threads;
    rt.search = retrieval_search_build_exact_provider;
    rc = retrieval_search_build_graph_ex(base, n, dim, &cfg, &rt, &g, NULL);
    if (rc) return rc;
    rc = retrieval_rnn_export_topk(&g, r, out);
    retrieval_rnn_graph_free(&g);
    return rc;
}

/* RNN-Descent: retrieval_rnn_build_graph_ex with `threads` workers and one
 * distance per batch entry (the CPU provider), then
 * retrieval_rnn_export_topk; out [n][top_k]. */
int ref_rnn_build(const float *base, int n, int dim, int s, int r, int t1, int t2,
                  unsigned int seed, int threads, int top_k, int32_t *out) {
    RetrievalRnnBuildConfig cfg;
    RetrievalRnnBuildRuntime rt;
    RetrievalRnnGraph g;
    int rc;
    memset(&cfg, 0, sizeof(cfg));
    memset(&rt, 0, sizeof(rt));
    memset(&g, 0, sizeof(g));
    cfg.s = s;
    cfg.r = r;
    cfg.t1 = t1;
    cfg.t2 = t2;
    cfg.seed = seed;
    rt.threads = threads;
    rt.batch_size = 1;
    rt.min_batch = 1;
    rc = retrieval_rnn_build_graph_ex(base, n, dim, &cfg, &rt, &g, NULL);
    if (rc) return rc;
    rc = retrieval_rnn_export_topk(&g, top_k, out);
    retrieval_rnn_graph_free(&g);
    return rc;
}
