/*
 * rover_host.c -- C host side of the MI355X quantized graph search.
 *
 * Implements the calculator.idl operator surface, the reference's library
 * search call and the batch extensions of include/rover_gpu.h on top of the
 * thin HIP ABI in rover_hip.h.  There is no CPU search path here: every
 * search runs the HIP kernels, and a missing or failing device is an error.
 */
#define _GNU_SOURCE
#include "rover_ctx.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

_Static_assert(ROVER_NOALLOC_MAX_SEEDS == RV_NOALLOC_MAX_SEEDS &&
                   ROVER_NOALLOC_SCRATCH_BYTES == RV_NOALLOC_BYTES,
               "no-allocation workspace limits");

static __thread char g_err[512];

void rover_set_err(const char *what) {
    const char *detail = rv_last_error();
    snprintf(g_err, sizeof(g_err), "%s%s%s", what,
             detail && detail[0] ? ": " : "", detail ? detail : "");
}

void rover_set_errf(const char *msg) { snprintf(g_err, sizeof(g_err), "%s", msg); }

const char *rover_last_error(void) { return g_err; }

#define set_err rover_set_err
#define ctx_of rover_ctx_of
#define dfree rover_dfree

RoverCtx *rover_ctx_of(remote_handle64 h) {
    RoverCtx *c = (RoverCtx *)(uintptr_t)h;
    if (!c || c->magic != ROVER_CTX_MAGIC) return NULL;
    return c;
}

void rover_dfree(void *p) {
    if (p) rv_free(p);
}

int rover_stream_turn(RoverCtx *c, void *stream) {
    if (c->launched && c->last_stream != stream &&
        rv_stream_wait_event(stream, c->ev_last)) {
        set_err("stream wait");
        return -1;
    }
    return 0;
}

int rover_stream_done(RoverCtx *c, void *stream) {
    if (rv_event_record(c->ev_last, stream)) {
        set_err("stream event");
        return -1;
    }
    c->last_stream = stream;
    c->launched = 1;
    return 0;
}

static int dalloc_copy(void **d, const void *src, size_t bytes) {
    *d = NULL;
    if (rv_malloc(d, bytes)) return -1;
    if (src && rv_memcpy_h2d(*d, src, bytes)) {
        rv_free(*d);
        *d = NULL;
        return -1;
    }
    return 0;
}

int rover_device_count(void) { return rv_device_count(); }

int calculator_open(const char *uri, remote_handle64 *handle) {
    RoverCtx *c;
    (void)uri;
    if (!handle) return ROVER_EBADPARM;
    *handle = 0;
    if (rv_device_count() <= 0) {
        snprintf(g_err, sizeof(g_err), "calculator_open: no HIP device");
        return ROVER_EFAILED;
    }
    c = calloc(1, sizeof(*c));
    if (!c) return ROVER_ENOMEMORY;
    c->magic = ROVER_CTX_MAGIC;
    c->device = 0;
    c->l1_bits = 8;
    rover_sha256(NULL, 0, c->orig_maps_sha);
    pthread_mutex_init(&c->lock, NULL);
    if (rv_set_device(0) || rv_malloc((void **)&c->d_queue, 64) ||
        rv_event_create_untimed(&c->ev_last) || rv_stream_create(&c->stream) ||
        rv_noalloc_scratch(0, &c->d_noalloc)) {
        set_err("calculator_open");
        dfree(c->d_queue);
        if (c->ev_last) rv_event_destroy(c->ev_last);
        if (c->stream) rv_stream_destroy(c->stream);
        free(c);
        return ROVER_EFAILED;
    }
    *handle = (remote_handle64)(uintptr_t)c;
    return ROVER_SUCCESS;
}

static void learned_free(RoverLearned *l) {
    dfree(l->d_w);
    dfree(l->d_f);
    free(l->h_params);
    memset(l, 0, sizeof(*l));
}

AEEResult calculator_retrieval_pure_pq_unload_dma(remote_handle64 h) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    dfree(c->d_codes);
    dfree(c->d_codebook);
    dfree(c->d_graph);
    dfree(c->d_res);
    dfree(c->d_qparams);
    dfree(c->d_orig);
    rover_pure_pq_error4_unload(h);
    rover_sha256(NULL, 0, c->orig_maps_sha);
    dfree(c->d_router_cent);
    dfree(c->d_router_ids);
    c->d_router_cent = NULL;
    c->d_router_ids = NULL;
    c->router_n = 0;
    c->d_codes = NULL;
    c->d_codebook = NULL;
    c->d_graph = NULL;
    c->d_res = NULL;
    c->d_qparams = NULL;
    c->d_orig = NULL;
    c->pq_loaded = 0;
    c->graph_n = 0;
    c->res_bits = 0;
    return ROVER_SUCCESS;
}

AEEResult calculator_retrieval_unload(remote_handle64 h) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    dfree(c->d_sq_graph);
    dfree(c->d_sq_codes);
    dfree(c->d_sq_res);
    dfree(c->d_sq_res_qp);
    dfree(c->d_sqmd_seeds);
    dfree(c->d_split_bias);
    c->d_split_bias = NULL;
    rover_sq_error4_unload(h);
    c->d_sq_graph = NULL;
    c->d_sq_codes = NULL;
    c->d_sq_res = NULL;
    c->d_sq_res_qp = NULL;
    c->d_sqmd_seeds = NULL;
    c->sqmd_seeds_count = 0;
    c->sq_res_bits = 0;
    c->sq_loaded = 0;
    c->sq_dma = 0;
    return ROVER_SUCCESS;
}

int calculator_close(remote_handle64 h) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    calculator_retrieval_pure_pq_unload_dma(h);
    calculator_retrieval_unload(h);
    learned_free(&c->learned);
    dfree(c->d_queue);
    dfree(c->d_entries);
    dfree(c->d_stage);
    dfree(c->d_route_dist);
    dfree(c->d_route_ent);
    dfree(c->d_sqmd_vis);
    dfree(c->d_visited);
    dfree(c->d_susp);
    dfree(c->d_sqmd_front);
    if (c->stream) {
        rv_stream_synchronize(c->stream);
        rv_stream_destroy(c->stream);
    }
    if (c->ev_last) rv_event_destroy(c->ev_last);
    c->magic = 0;
    pthread_mutex_destroy(&c->lock);
    free(c);
    return ROVER_SUCCESS;
}

AEEResult calculator_rover_configure(remote_handle64 h, int32_t abi_version,
                                     int32_t l1_bits) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    if (abi_version != ROVER_ABI_VERSION || (l1_bits != 4 && l1_bits != 8))
        return ROVER_EBADPARM;
    c->l1_bits = l1_bits;
    return ROVER_SUCCESS;
}

/* retrieval_async_dsp_search.c:698-799: validates the chunk table and the
 * PQ shape, copies the codebook; the code chunks are concatenated into one
 * dense [num_vectors][code_bytes] array in HBM. */
AEEResult calculator_retrieval_pure_pq_load_dma(
    remote_handle64 h, const uint8_t *const *pq_codes, int pq_codesLen,
    const int32_t *counts, int countsLen, int32_t chunk_vectors,
    const float *codebook, int codebookLen, int32_t num_vectors, int32_t dim,
    int32_t m, int32_t ks, int32_t subdim, int32_t code_bytes,
    const uint8_t *uuid, int uuidLen) {
    RoverCtx *c = ctx_of(h);
    long total = 0;
    int expect_bytes;
    if (!c) return ROVER_EBADPARM;
    (void)uuid;
    expect_bytes = (ks == 16) ? m / 2 + m % 2 : m;
    if (!pq_codes || !counts || !codebook || pq_codesLen <= 0 ||
        pq_codesLen != countsLen || chunk_vectors <= 0 || num_vectors <= 0 ||
        dim <= 0 || dim > 4096 || m <= 0 || (ks != 16 && ks != 256) ||
        subdim <= 0 || m * subdim != dim || code_bytes != expect_bytes ||
        codebookLen != m * ks * subdim || (uuid && uuidLen != 16))
        return ROVER_EBADPARM;
    for (int i = 0; i < pq_codesLen; i++) {
        if (!pq_codes[i] || counts[i] <= 0 || counts[i] > chunk_vectors ||
            (i + 1 < pq_codesLen && counts[i] != chunk_vectors))
            return ROVER_EBADPARM;
        total += counts[i];
    }
    if (total != num_vectors) return ROVER_EBADPARM;
    calculator_retrieval_pure_pq_unload_dma(h);
    if (rv_malloc((void **)&c->d_codes, (size_t)num_vectors * code_bytes) ||
        dalloc_copy((void **)&c->d_codebook, codebook,
                    (size_t)codebookLen * sizeof(float))) {
        set_err("pure_pq_load");
        calculator_retrieval_pure_pq_unload_dma(h);
        return ROVER_ENOMEMORY;
    }
    {
        size_t off = 0;
        for (int i = 0; i < pq_codesLen; i++) {
            size_t bytes = (size_t)counts[i] * code_bytes;
            if (rv_memcpy_h2d(c->d_codes + off, pq_codes[i], bytes)) {
                set_err("pure_pq_load");
                calculator_retrieval_pure_pq_unload_dma(h);
                return ROVER_EFAILED;
            }
            off += bytes;
        }
    }
    c->n = num_vectors;
    c->dim = dim;
    c->m = m;
    c->ks = ks;
    c->subdim = subdim;
    c->code_bytes = code_bytes;
    c->pq_loaded = 1;
    return ROVER_SUCCESS;
}

AEEResult rover_pure_pq_graph_load(remote_handle64 h, const int32_t *graph,
                                   int64_t graphLen, int32_t num_vectors,
                                   int32_t graph_k, int32_t entry) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    if (!c->pq_loaded || !graph || num_vectors != c->n || graph_k <= 0 ||
        graph_k > ROVER_MAX_DEGREE || graphLen != (int64_t)num_vectors * graph_k ||
        entry >= num_vectors)
        return ROVER_EBADPARM;
    dfree(c->d_graph);
    if (dalloc_copy((void **)&c->d_graph, graph,
                    (size_t)graphLen * sizeof(int32_t))) {
        set_err("graph_load");
        c->d_graph = NULL;
        return ROVER_ENOMEMORY;
    }
    c->graph_n = num_vectors;
    c->degree = graph_k;
    c->entry = entry;
    c->entries_count = 0; /* seed cache depends on the entry point */
    return ROVER_SUCCESS;
}

AEEResult rover_pure_pq_residual_load(remote_handle64 h, const uint8_t *codes,
                                      int64_t codesLen, const float *qparams,
                                      int64_t qparamsLen, int32_t bits) {
    RoverCtx *c = ctx_of(h);
    int64_t rb;
    if (!c) return ROVER_EBADPARM;
    if (!c->pq_loaded || !codes || !qparams || (bits != 4 && bits != 8))
        return ROVER_EBADPARM;
    rb = bits == 4 ? c->dim / 2 + c->dim % 2 : c->dim;
    if (codesLen != rb * c->n || qparamsLen != 2 * (int64_t)c->n)
        return ROVER_EBADPARM;
    dfree(c->d_res);
    dfree(c->d_qparams);
    c->d_res = NULL;
    c->d_qparams = NULL;
    if (dalloc_copy((void **)&c->d_res, codes, (size_t)codesLen) ||
        dalloc_copy((void **)&c->d_qparams, qparams,
                    (size_t)qparamsLen * sizeof(float))) {
        set_err("residual_load");
        return ROVER_ENOMEMORY;
    }
    c->res_bits = bits;
    return ROVER_SUCCESS;
}

/* The id-map digest an error4 artifact is bound to
 * (retrieval_final_refine.c:167-180): SHA-256 of graph_to_original then
 * graph_to_storage as little-endian int32 (the PQ path has no storage map,
 * so the identity), or of nothing when there is no map. */
static int maps_digest(const int32_t *map, int n, uint8_t out[32]) {
    int32_t *maps;
    if (!map) {
        rover_sha256(NULL, 0, out);
        return 0;
    }
    maps = malloc((size_t)n * 2 * sizeof(int32_t));
    if (!maps) return -1;
    for (int i = 0; i < n; i++) {
        maps[i] = map[i];
        maps[(size_t)n + i] = i;
    }
    rover_sha256(maps, (size_t)n * 2 * sizeof(int32_t), out);
    free(maps);
    return 0;
}

AEEResult rover_pure_pq_original_ids(remote_handle64 h, const int32_t *map,
                                     int mapLen) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    dfree(c->d_orig);
    c->d_orig = NULL;
    maps_digest(NULL, 0, c->orig_maps_sha);
    if (!map) return ROVER_SUCCESS;
    if (!c->pq_loaded || mapLen != c->n) return ROVER_EBADPARM;
    {
        /* must be a permutation (retrieval_final_refine.c:68-79) */
        uint8_t *seen = calloc((size_t)c->n, 1);
        if (!seen) return ROVER_ENOMEMORY;
        for (int i = 0; i < c->n; i++) {
            if (map[i] < 0 || map[i] >= c->n || seen[map[i]]) {
                free(seen);
                return ROVER_EBADPARM;
            }
            seen[map[i]] = 1;
        }
        free(seen);
    }
    if (maps_digest(map, c->n, c->orig_maps_sha)) return ROVER_ENOMEMORY;
    if (dalloc_copy((void **)&c->d_orig, map, (size_t)mapLen * sizeof(int32_t))) {
        maps_digest(NULL, 0, c->orig_maps_sha);
        return ROVER_ENOMEMORY;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_sq_residual_load(remote_handle64 h, const uint8_t *codes,
                                 int64_t codesLen, const float *qparams,
                                 int64_t qparamsLen, int32_t bits) {
    RoverCtx *c = ctx_of(h);
    int64_t rb;
    if (!c) return ROVER_EBADPARM;
    if (!c->sq_loaded || !codes || !qparams || (bits != 4 && bits != 8))
        return ROVER_EBADPARM;
    rb = bits == 4 ? (c->sq_dim + 1) / 2 : c->sq_dim;
    if (codesLen != rb * c->sq_n || qparamsLen != 2 * (int64_t)c->sq_n)
        return ROVER_EBADPARM;
    pthread_mutex_lock(&c->lock);
    dfree(c->d_sq_res);
    dfree(c->d_sq_res_qp);
    dfree(c->d_split_bias);  /* a bias is built from one residual */
    c->d_sq_res = NULL;
    c->d_sq_res_qp = NULL;
    c->d_split_bias = NULL;
    c->sq_res_bits = 0;
    if (dalloc_copy((void **)&c->d_sq_res, codes, (size_t)codesLen) ||
        dalloc_copy((void **)&c->d_sq_res_qp, qparams, (size_t)qparamsLen * sizeof(float))) {
        dfree(c->d_sq_res);
        c->d_sq_res = NULL;
        pthread_mutex_unlock(&c->lock);
        set_err("sq_residual_load");
        return ROVER_ENOMEMORY;
    }
    c->sq_res_bits = bits;
    pthread_mutex_unlock(&c->lock);
    return ROVER_SUCCESS;
}

static uint32_t le32(const uint8_t *p) {
    return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 |
           (uint32_t)p[3] << 24;
}

static uint64_t le64(const uint8_t *p) { return le32(p) | (uint64_t)le32(p + 4) << 32; }

/* retrieval_rpc_demo_support.c:227-235 (sqmd_fnv1a64) */
static uint64_t fnv1a64(const uint8_t *p, size_t n) {
    uint64_t v = 14695981039346656037ull;
    for (size_t i = 0; i < n; i++) {
        v ^= p[i];
        v *= 1099511628211ull;
    }
    return v;
}

/* retrieval_rpc_demo_support.c:237-299 (sqmd_load_correction_bias): the
 * 64-byte SQMDBIAS header (magic, version 1, header bytes 64, vectors,
 * dim, l1 bits, residual bits, payload bytes, payload FNV-1a64, source
 * fingerprint) checked against the loaded SQ index and residual, then the
 * f32 payload, every value finite. */
AEEResult rover_sq_split_bias_load(remote_handle64 h, const uint8_t *image, int64_t imageLen,
                                   const uint8_t *source_fingerprint) {
    RoverCtx *c = ctx_of(h);
    static const uint8_t magic[8] = {'S', 'Q', 'M', 'D', 'B', 'I', 'A', 'S'};
    const float *bias;
    size_t payload;
    if (!c || !image) return ROVER_EBADPARM;
    if (!c->sq_loaded || !c->d_sq_res) {
        snprintf(g_err, sizeof(g_err), "split bias: load the SQ index and its residual first");
        return ROVER_EBADPARM;
    }
    payload = (size_t)c->sq_n * sizeof(float);
    if (imageLen != (int64_t)(64 + payload) || memcmp(image, magic, 8) != 0 ||
        le32(image + 8) != 1u || le32(image + 12) != 64u || le32(image + 16) != (uint32_t)c->sq_n ||
        le32(image + 20) != (uint32_t)c->sq_dim || le32(image + 24) != (uint32_t)c->sq_bits ||
        c->sq_bits != 4 || le32(image + 28) != (uint32_t)c->sq_res_bits ||
        le64(image + 32) != (uint64_t)payload ||
        (source_fingerprint && memcmp(image + 48, source_fingerprint, 16) != 0) ||
        fnv1a64(image + 64, payload) != le64(image + 40)) {
        snprintf(g_err, sizeof(g_err), "split bias: header, fingerprint or payload check failed");
        return ROVER_EBADPARM;
    }
    bias = (const float *)(image + 64);
    for (int i = 0; i < c->sq_n; i++) {
        float v;
        memcpy(&v, bias + i, sizeof(v));
        if (!isfinite(v)) {
            snprintf(g_err, sizeof(g_err), "split bias: non-finite value");
            return ROVER_EBADPARM;
        }
    }
    pthread_mutex_lock(&c->lock);
    dfree(c->d_split_bias);
    c->d_split_bias = NULL;
    if (dalloc_copy((void **)&c->d_split_bias, image + 64, payload)) {
        pthread_mutex_unlock(&c->lock);
        set_err("split bias upload");
        return ROVER_ENOMEMORY;
    }
    pthread_mutex_unlock(&c->lock);
    return ROVER_SUCCESS;
}

AEEResult rover_pure_pq_error4_unload(remote_handle64 h) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    dfree(c->d_e4_blocks);
    dfree(c->d_e4_qparams);
    c->d_e4_blocks = NULL;
    c->d_e4_qparams = NULL;
    c->e4_loaded = 0;
    return ROVER_SUCCESS;
}

/* The checks of refine_open in ERROR4 mode (retrieval_final_refine.c:
 * 128-160 and 181-186) on an in-memory image with the given magic, for an
 * index of n vectors of dim dimensions: header fields, section sizes, the
 * whole-file digest when expected_sha256 is given, the payload digest, every
 * (scale, zero) finite with scale >= 0.  Fills the block geometry. */
static AEEResult check_error4_image(const uint8_t *a, int64_t len, const char *magic,
                                    const char *other_magic, int n_vec, int dim,
                                    const uint8_t *expected_sha256, int *out_vpb,
                                    int *out_cb, uint64_t *out_qb, uint64_t *out_cbytes) {
    uint8_t digest[32];
    const uint64_t n = (uint64_t)n_vec;
    const int cb = (dim + 1) / 2;
    const int vpb = RV_E4_BLOCK / cb;
    uint64_t qb, cbytes;
    if (vpb <= 0 || len < 256) goto bad;
    qb = n * 8;
    cbytes = (n + (uint64_t)vpb - 1) / (uint64_t)vpb * RV_E4_BLOCK;
    if (memcmp(a, magic, 8)) {
        snprintf(g_err, sizeof(g_err), "error4_load: %s",
                 memcmp(a, other_magic, 8) ? "not an error4 artifact"
                                           : "artifact of the other index kind");
        return ROVER_EBADPARM;
    }
    if (le32(a + 8) != 1 || le32(a + 12) != (uint32_t)n || le32(a + 16) != (uint32_t)dim ||
        (le32(a + 20) != 4 && le32(a + 20) != 8) ||
        (le32(a + 24) != 4 && le32(a + 24) != 8) || le32(a + 28) != RV_E4_BLOCK ||
        le32(a + 32) != 1 || le32(a + 36) != 1 || le64(a + 40) != qb ||
        le64(a + 48) != cbytes || (uint64_t)len != 256 + qb + cbytes)
        goto bad;
    if (expected_sha256) {
        rover_sha256(a, (size_t)len, digest);
        if (memcmp(digest, expected_sha256, 32)) {
            snprintf(g_err, sizeof(g_err), "error4_load: file digest mismatch");
            return ROVER_EBADPARM;
        }
    }
    rover_sha256(a + 256, (size_t)(qb + cbytes), digest);
    if (memcmp(digest, a + 56, 32)) {
        snprintf(g_err, sizeof(g_err), "error4_load: payload digest mismatch");
        return ROVER_EBADPARM;
    }
    for (uint64_t i = 0; i < n; i++) {
        float sz[2];
        memcpy(sz, a + 256 + 8 * i, 8);
        if (!isfinite(sz[0]) || sz[0] < 0.0f || !isfinite(sz[1])) {
            snprintf(g_err, sizeof(g_err), "error4_load: bad qparams at row %llu",
                     (unsigned long long)i);
            return ROVER_EBADPARM;
        }
    }
    *out_vpb = vpb;
    *out_cb = cb;
    *out_qb = qb;
    *out_cbytes = cbytes;
    return ROVER_SUCCESS;
bad:
    snprintf(g_err, sizeof(g_err), "error4_load: header or size mismatch");
    return ROVER_EBADPARM;
}

AEEResult rover_pure_pq_error4_load(remote_handle64 h, const uint8_t *a,
                                    int64_t len, const uint8_t *expected_sha256) {
    RoverCtx *c = ctx_of(h);
    uint64_t qb, cbytes;
    int cb, vpb;
    AEEResult rc;
    if (!c || !a) return ROVER_EBADPARM;
    if (!c->pq_loaded) {
        snprintf(g_err, sizeof(g_err), "error4_load: no PQ index loaded");
        return ROVER_EBADPARM;
    }
    rc = check_error4_image(a, len, "PQERR4V1", "SQERR4V1", c->n, c->dim, expected_sha256,
                            &vpb, &cb, &qb, &cbytes);
    if (rc != ROVER_SUCCESS) return rc;
    pthread_mutex_lock(&c->lock);
    rover_pure_pq_error4_unload(h);
    if (dalloc_copy((void **)&c->d_e4_qparams, a + 256, (size_t)qb) ||
        dalloc_copy((void **)&c->d_e4_blocks, a + 256 + qb, (size_t)cbytes)) {
        rover_pure_pq_error4_unload(h);
        pthread_mutex_unlock(&c->lock);
        set_err("error4_load");
        return ROVER_ENOMEMORY;
    }
    memcpy(c->e4_maps_sha, a + 216, 32);
    c->e4_vpb = vpb;
    c->e4_code_bytes = cb;
    c->e4_loaded = 1;
    pthread_mutex_unlock(&c->lock);
    return ROVER_SUCCESS;
}

AEEResult rover_sq_error4_unload(remote_handle64 h) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    dfree(c->d_sq_e4_blocks);
    dfree(c->d_sq_e4_qparams);
    c->d_sq_e4_blocks = NULL;
    c->d_sq_e4_qparams = NULL;
    c->sq_e4_loaded = 0;
    return ROVER_SUCCESS;
}

/* SQERR4V1 (retrieval_final_refine_open without a codebook, the SQ path of
 * retrieval_rpc_demo_pipeline.c:3026-3045): the same checks, against the
 * loaded SQ index; the SQ index has no id maps, so the artifact must be
 * bound to none (its +216 digest is that of no bytes). */
AEEResult rover_sq_error4_load(remote_handle64 h, const uint8_t *a, int64_t len,
                               const uint8_t *expected_sha256) {
    RoverCtx *c = ctx_of(h);
    uint8_t none[32];
    uint64_t qb, cbytes;
    int cb, vpb;
    AEEResult rc;
    if (!c || !a) return ROVER_EBADPARM;
    if (!c->sq_loaded) {
        snprintf(g_err, sizeof(g_err), "error4_load: no SQ index loaded");
        return ROVER_EBADPARM;
    }
    rc = check_error4_image(a, len, "SQERR4V1", "PQERR4V1", c->sq_n, c->sq_dim,
                            expected_sha256, &vpb, &cb, &qb, &cbytes);
    if (rc != ROVER_SUCCESS) return rc;
    rover_sha256(NULL, 0, none);
    if (memcmp(a + 216, none, 32)) {
        snprintf(g_err, sizeof(g_err), "error4_load: artifact is bound to id maps");
        return ROVER_EBADPARM;
    }
    pthread_mutex_lock(&c->lock);
    rover_sq_error4_unload(h);
    if (dalloc_copy((void **)&c->d_sq_e4_qparams, a + 256, (size_t)qb) ||
        dalloc_copy((void **)&c->d_sq_e4_blocks, a + 256 + qb, (size_t)cbytes)) {
        rover_sq_error4_unload(h);
        pthread_mutex_unlock(&c->lock);
        set_err("error4_load");
        return ROVER_ENOMEMORY;
    }
    c->sq_e4_vpb = vpb;
    c->sq_e4_code_bytes = cb;
    c->sq_e4_loaded = 1;
    pthread_mutex_unlock(&c->lock);
    return ROVER_SUCCESS;
}

AEEResult calculator_retrieval_load(remote_handle64 h, const int32_t *graph,
                                    int graphLen, const uint8_t *l1_codes,
                                    int l1_codesLen, int32_t num_vectors,
                                    int32_t dim, int32_t graph_k,
                                    int32_t entry, float l1_scale,
                                    float l1_zero) {
    RoverCtx *c = ctx_of(h);
    int cb;
    if (!c) return ROVER_EBADPARM;
    cb = c->l1_bits == 4 ? (dim + 1) / 2 : dim;
    if (!graph || !l1_codes || num_vectors <= 0 || dim <= 0 || dim > 4096 ||
        graph_k <= 0 || graph_k > ROVER_MAX_DEGREE ||
        (int64_t)graphLen != (int64_t)num_vectors * graph_k ||
        (int64_t)l1_codesLen != (int64_t)num_vectors * cb ||
        !isfinite(l1_scale) || !(l1_scale > 0.0f) || !isfinite(l1_zero) ||
        entry >= num_vectors)
        return ROVER_EBADPARM;
    calculator_retrieval_unload(h);
    if (dalloc_copy((void **)&c->d_sq_graph, graph, (size_t)graphLen * 4) ||
        dalloc_copy((void **)&c->d_sq_codes, l1_codes, (size_t)l1_codesLen)) {
        set_err("retrieval_load");
        calculator_retrieval_unload(h);
        return ROVER_ENOMEMORY;
    }
    c->sq_n = num_vectors;
    c->sq_dim = dim;
    c->sq_degree = graph_k;
    c->sq_entry = entry;
    c->sq_bits = c->l1_bits;
    c->sq_scale = l1_scale;
    c->sq_zero = l1_zero;
    c->sq_loaded = 1;
    c->entries_count = 0;
    return ROVER_SUCCESS;
}

/* calculator.idl:59-70 (retrieval_load_dma; retrieval_async_dsp_search.c:
 * 558-666): the SQ index from chunked l1 codes.  The IDL's file descriptors
 * become pointers: `graph` (NULL where the reference passes
 * RETRIEVAL_DMA_NO_GRAPH_FD: the graph stays with the caller, to be given by
 * rover_sq_graph_load before a search) and the chunk base pointers, chunk c
 * holding min(l1_chunk_vectors, num_vectors - c * l1_chunk_vectors) rows.
 * The data is copied into HBM.  As in the reference, both SQ loads are
 * dropped first and an entry point outside [0, num_vectors) means none. */
AEEResult calculator_retrieval_load_dma(remote_handle64 h, const int32_t *graph,
                                        const uint8_t *const *l1_codes, int l1_codesLen,
                                        int32_t l1_chunk_vectors, int32_t num_vectors,
                                        int32_t dim, int32_t graph_k,
                                        int32_t graph_entry_point, float l1_scale,
                                        float l1_zero_point) {
    RoverCtx *c = ctx_of(h);
    size_t cb;
    if (!c) return ROVER_EBADPARM;
    calculator_retrieval_unload(h);
    if (num_vectors <= 0 || dim <= 0 || dim > 4096 || graph_k <= 0 || graph_k > ROVER_MAX_DEGREE ||
        !l1_codes || l1_codesLen < 1 || l1_codesLen > ROVER_MAX_L1_CHUNKS ||
        l1_chunk_vectors <= 0 ||
        (int64_t)l1_codesLen * l1_chunk_vectors < (int64_t)num_vectors ||
        (int64_t)(l1_codesLen - 1) * l1_chunk_vectors >= (int64_t)num_vectors ||
        !isfinite(l1_scale) || !(l1_scale > 0.0f) || !isfinite(l1_zero_point))
        return ROVER_EBADPARM;
    for (int i = 0; i < l1_codesLen; i++)
        if (!l1_codes[i]) return ROVER_EBADPARM;
    cb = (size_t)(c->l1_bits == 4 ? (dim + 1) / 2 : dim);
    if (rv_malloc((void **)&c->d_sq_codes, (size_t)num_vectors * cb) ||
        (graph && dalloc_copy((void **)&c->d_sq_graph, graph,
                              (size_t)num_vectors * graph_k * 4))) {
        set_err("retrieval_load_dma");
        calculator_retrieval_unload(h);
        return ROVER_ENOMEMORY;
    }
    for (int i = 0; i < l1_codesLen; i++) {
        size_t first = (size_t)i * l1_chunk_vectors;
        size_t rows = (size_t)num_vectors - first < (size_t)l1_chunk_vectors
                          ? (size_t)num_vectors - first
                          : (size_t)l1_chunk_vectors;
        if (rv_memcpy_h2d(c->d_sq_codes + first * cb, l1_codes[i], rows * cb)) {
            set_err("retrieval_load_dma upload");
            calculator_retrieval_unload(h);
            return ROVER_EFAILED;
        }
    }
    c->sq_n = num_vectors;
    c->sq_dim = dim;
    c->sq_degree = graph_k;
    c->sq_entry = graph_entry_point >= 0 && graph_entry_point < num_vectors
                      ? graph_entry_point : -1;
    c->sq_bits = c->l1_bits;
    c->sq_scale = l1_scale;
    c->sq_zero = l1_zero_point;
    c->sq_loaded = 1;
    c->sq_dma = 1;
    c->entries_count = 0;
    return ROVER_SUCCESS;
}

/* calculator.idl:71 (retrieval_unload_dma): drops an index that
 * retrieval_load_dma loaded; an index from retrieval_load stays, as the
 * reference leaves a copied index to retrieval_unload. */
AEEResult calculator_retrieval_unload_dma(remote_handle64 h) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    if (c->sq_dma) calculator_retrieval_unload(h);
    return ROVER_SUCCESS;
}

/* The SQ graph in HBM for an index loaded without one (the reference keeps
 * that graph on the host: retrieval_rpc_demo_pipeline.c:3066-3077). */
AEEResult rover_sq_graph_load(remote_handle64 h, const int32_t *graph, int64_t graphLen,
                              int32_t num_vectors, int32_t graph_k,
                              int32_t graph_entry_point) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    if (!c->sq_loaded) return ROVER_EFAILED;
    if (!graph || num_vectors != c->sq_n || graph_k <= 0 || graph_k > ROVER_MAX_DEGREE ||
        graphLen != (int64_t)num_vectors * graph_k || graph_entry_point >= num_vectors)
        return ROVER_EBADPARM;
    dfree(c->d_sq_graph);
    c->d_sq_graph = NULL;
    if (dalloc_copy((void **)&c->d_sq_graph, graph, (size_t)graphLen * 4)) {
        set_err("rover_sq_graph_load");
        return ROVER_ENOMEMORY;
    }
    c->sq_degree = graph_k;
    c->sq_entry = graph_entry_point;
    c->entries_count = 0;
    return ROVER_SUCCESS;
}

/* The reference demo's known-answer hash of search results
 * (pure_pq_result_fingerprint, retrieval_rpc_demo_pure_pq.c:387-400):
 * FNV-1a 64 over each result as 8 little-endian bytes, the id's 4 then the
 * distance's bit pattern's 4, chained from `hash`. */
uint64_t rover_result_fingerprint(uint64_t hash, const int32_t *ids, const float *dists,
                                  int count) {
    const uint64_t prime = 0x100000001b3ull;
    for (int i = 0; i < count; i++) {
        uint32_t dbits;
        uint64_t rec;
        memcpy(&dbits, &dists[i], 4);
        rec = (uint64_t)(uint32_t)ids[i] | (uint64_t)dbits << 32;
        for (int b = 0; b < 64; b += 8) hash = (hash ^ ((rec >> b) & 0xffu)) * prime;
    }
    return hash;
}

/* rpc_demo_pure_pq.c:192-208 (pure_pq_make_seed_ids); the DSP search uses
 * the same recurrence (retrieval_async_dsp_search.c:219-230). */
int rover_make_seed_ids(int n, int entry_count, int graph_entry, uint32_t seed,
                        int32_t *out) {
    int count = entry_count < n ? entry_count : n;
    for (int i = 0; i < count; i++) {
        if (i == 0 && graph_entry >= 0 && graph_entry < n) {
            out[i] = graph_entry;
        } else {
            seed = seed * 1103515245u + 12345u;
            out[i] = (int32_t)(((uint64_t)seed * (uint64_t)n) >> 32);
        }
    }
    return count;
}

/* Uploads the seed list once per (n, entry_points, graph entry, seed). */
static int ensure_entries(RoverCtx *c, int n, int entry_points, int graph_entry,
                          uint32_t seed) {
    int32_t *ids;
    int count;
    if (c->entries_count > 0 && c->entries_n == n &&
        c->entries_graph_entry == graph_entry && c->entries_seed == seed &&
        c->entries_count == (entry_points < n ? entry_points : n))
        return 0;
    if (entry_points > c->entries_cap) {
        dfree(c->d_entries);
        c->d_entries = NULL;
        c->entries_cap = 0;
        if (rv_malloc((void **)&c->d_entries, (size_t)entry_points * 4)) {
            set_err("entries");
            return -1;
        }
        c->entries_cap = entry_points;
    }
    ids = malloc((size_t)entry_points * sizeof(int32_t));
    if (!ids) return -1;
    count = rover_make_seed_ids(n, entry_points, graph_entry, seed, ids);
    if (rv_memcpy_h2d(c->d_entries, ids, (size_t)count * 4)) {
        free(ids);
        set_err("entries upload");
        return -1;
    }
    free(ids);
    c->entries_n = n;
    c->entries_graph_entry = graph_entry;
    c->entries_seed = seed;
    c->entries_count = count;
    return 0;
}

static int ilog2_ceil(int v) {
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

/* Default visited-hash size: about 16 slots per pool entry, 256..4096. */
static int default_hash_log2(int max_pool) {
    int l = ilog2_ceil(max_pool * 16);
    if (l < 8) l = 8;
    if (l > 12) l = 12;
    return l;
}

static int fill_index(const RoverCtx *c, int method, int correction,
                      int refine, RvIndex *ix) {
    memset(ix, 0, sizeof(*ix));
    if (method == ROVER_METHOD_SQ) {
        if (!c->sq_loaded || !c->d_sq_graph) return -1;
        ix->graph = c->d_sq_graph;
        ix->degree = c->sq_degree;
        ix->n = c->sq_n;
        ix->graph_entry = c->sq_entry;
        ix->dim = c->sq_dim;
        ix->sq_codes = c->d_sq_codes;
        ix->sq_bits = c->sq_bits;
        ix->sq_code_bytes = c->sq_bits == 4 ? (c->sq_dim + 1) / 2 : c->sq_dim;
        ix->sq_scale = c->sq_scale;
        ix->sq_zero = c->sq_zero;
        return 0;
    }
    if (!c->pq_loaded || !c->d_graph) return -1;
    if ((correction || refine) && !c->d_res) return -1;
    ix->graph = c->d_graph;
    ix->degree = c->degree;
    ix->n = c->n;
    ix->graph_entry = c->entry;
    ix->codes = c->d_codes;
    ix->code_bytes = c->code_bytes;
    ix->codebook = c->d_codebook;
    ix->dim = c->dim;
    ix->m = c->m;
    ix->ks = c->ks;
    ix->subdim = c->subdim;
    ix->res_codes = c->d_res;
    ix->res_qparams = c->d_qparams;
    ix->res_bits = c->res_bits;
    ix->res_code_bytes = c->res_bits == 4 ? c->dim / 2 + c->dim % 2 : c->dim;
    ix->graph_to_original = c->d_orig;
    if (c->e4_loaded) {
        ix->e4_blocks = c->d_e4_blocks;
        ix->e4_qparams = c->d_e4_qparams;
        ix->e4_vpb = c->e4_vpb;
        ix->e4_code_bytes = c->e4_code_bytes;
    }
    return 0;
}

static int grow(void **d, size_t *cap, size_t bytes) {
    if (bytes <= *cap) return 0;
    dfree(*d);
    *d = NULL;
    *cap = 0;
    if (rv_malloc(d, bytes)) return -1;
    *cap = bytes;
    return 0;
}

/* Per persistent wave a counter word and one 64-bit word per 32 ids.  The
 * sets rely on the buffer's contents (stamp 0 = untouched, a stamp counter
 * per wave), so the buffer is zeroed whenever it is (re)allocated or was
 * last laid out for another index size. */
int rover_attach_visited(uint64_t **buf, size_t *cap, size_t *layout, const RvIndex *ix,
                         RvBatch *rb, int grid, void *stream) {
    const size_t words = ((size_t)ix->n + 31) / 32;
    const size_t bytes = (size_t)(grid > 0 ? grid : 1) * (words + 1) * sizeof(uint64_t);
    const size_t key = words + 1;
    if (bytes > *cap) {
        dfree(*buf);
        *buf = NULL;
        *cap = 0;
        *layout = 0;
        if (rv_malloc((void **)buf, bytes)) return -1;
        *cap = bytes;
    }
    if (*layout != key) {
        if (rv_fill_zero(*buf, *cap, stream)) return -1;
        *layout = key;
    }
    rb->visited = *buf;
    rb->visited_words = (int)words;
    return 0;
}

/* Suspend and resume (RoverSearchBatch.reorder_step): one exact visited set
 * per query, the state block carved into the RvSusp arrays, then phase 1
 * (searches still running at the step save their state), the resumption
 * order, and phase 2 (the saved searches, longest expected first). */
static int rover_launch_reordered(RoverCtx *c, const RvIndex *ix, RvBatch *rb, int grid,
                                  void *stream) {
    const RvSuspLayout sl = rv_susp_layout((size_t)rb->nq, (size_t)rb->max_pool);
    int32_t *count;
    if (rover_attach_visited(&c->d_visited, &c->visited_cap, &c->visited_layout, ix, rb,
                             rb->nq, stream))
        return -1;
    if (grow(&c->d_susp, &c->susp_cap, sl.bytes)) return -1;
    rb->susp = (unsigned char *)c->d_susp;
    count = (int32_t *)(rb->susp + sl.count);
    if (rv_fill_zero(count, sizeof(int32_t), stream)) return -1;
    rb->phase = 1;
    if (rv_search_launch(ix, rb, grid, stream)) return -1;
    if (rv_susp_order(rb->susp, rb->nq, rb->max_pool, stream)) return -1;
    {
        RvBatch r2 = *rb;
        r2.phase = 2;
        r2.nq_dev = count;
        if (rv_search_launch(ix, &r2, grid, stream)) return -1;
    }
    rb->phase = 0;
    return 0;
}

int rover_launch(RoverCtx *c, const RvIndex *ix, RvBatch *rb, int grid_waves, void *stream) {
    if (rb->visited_exact && rb->nq > 0) {
        int grid = grid_waves > 0 ? grid_waves : rv_search_prepare(ix, rb);
        if (grid <= 0) return -1;
        if (grid > rb->nq) grid = rb->nq;
        if (rb->suspend_step > 0) return rover_launch_reordered(c, ix, rb, grid, stream);
        if (rover_attach_visited(&c->d_visited, &c->visited_cap, &c->visited_layout, ix, rb,
                                 grid, stream))
            return -1;
        grid_waves = grid;
    }
    return rv_search_launch(ix, rb, grid_waves, stream);
}

AEEResult rover_prepare_launch(RoverCtx *c, const RoverSearchBatch *b,
                               int max_pool, int explicit_entries, RvIndex *ix,
                               RvBatch *rb) {
    uint32_t seed = b->seed ? b->seed : ROVER_ENTRY_POINT_SEED;
    if (b->entry_ids && b->entry_count <= 0 && b->entry_stride <= 0)
        return ROVER_EBADPARM;
    if (b->num_queries < 0 || (!b->entry_ids && b->entry_points <= 0) ||
        (b->hash_log2 < 0 && b->hash_log2 != ROVER_HASH_EXACT) ||
        b->expand_count <= 0 ||
        max_pool <= 0 || max_pool > ROVER_MAX_POOL || b->out_stride <= 0 ||
        !b->out_ids || !b->out_dists || !b->out_counts ||
        (b->num_queries > 0 && !b->queries) || b->refine_candidates < 0 ||
        (b->method != ROVER_METHOD_PQ && b->method != ROVER_METHOD_SQ) ||
        (b->method == ROVER_METHOD_SQ && (b->correction || b->refine_candidates)) ||
        (b->refine_mode != 0 && b->refine_mode != ROVER_REFINE_CURRENT &&
         b->refine_mode != ROVER_REFINE_ERROR4) ||
        (b->refine_mode == ROVER_REFINE_ERROR4 && b->refine_candidates <= 0) ||
        /* the SQ traversals run through launch_sqmd, never here */
        (b->method == ROVER_METHOD_SQ && b->traversal != ROVER_TRAVERSAL_BEST_FIRST) ||
        (b->method == ROVER_METHOD_PQ && b->traversal != ROVER_TRAVERSAL_DEFAULT &&
         b->traversal != ROVER_TRAVERSAL_BEST_FIRST) ||
        b->source_window < 0 || b->source_window > 2 ||
        (b->waves_per_query != 0 && b->waves_per_query != 1 && b->waves_per_query != 2 &&
         b->waves_per_query != 4) ||
        (b->candidate_tile != 0 && b->candidate_tile != 16 && b->candidate_tile != 32 &&
         b->candidate_tile != 64) ||
        b->reorder_step < 0 ||
        (b->reorder_step > 0 &&
         (b->hash_log2 != ROVER_HASH_EXACT || b->source_window > 1 || b->waves_per_query ||
          b->method != ROVER_METHOD_PQ)))
        return ROVER_EBADPARM;
    if (b->refine_mode == ROVER_REFINE_ERROR4) {
        if (!c->e4_loaded) {
            snprintf(g_err, sizeof(g_err), "search: no error4 artifact loaded");
            return ROVER_EFAILED;
        }
        /* the reference refuses to open an artifact bound to other id maps */
        if (memcmp(c->e4_maps_sha, c->orig_maps_sha, 32)) {
            snprintf(g_err, sizeof(g_err),
                     "search: error4 artifact is bound to another id map");
            return ROVER_EBADPARM;
        }
    }
    if (fill_index(c, b->method, b->correction, b->refine_candidates, ix)) {
        snprintf(g_err, sizeof(g_err), "search: index or its graph not loaded");
        return ROVER_EFAILED;
    }
    if ((b->expand_count < ix->degree ? b->expand_count : ix->degree) > 64) {
        /* one wavefront reads and scores one expansion's neighbors */
        snprintf(g_err, sizeof(g_err),
                 "search: at most 64 neighbors per expansion (expand_count %d, graph "
                 "degree %d)",
                 b->expand_count, ix->degree);
        return ROVER_EBADPARM;
    }
    if (!explicit_entries && !b->entry_ids &&
        ensure_entries(c, ix->n, b->entry_points, ix->graph_entry, seed))
        return ROVER_EFAILED;
    memset(rb, 0, sizeof(*rb));
    rb->queries = b->queries;
    rb->nq = b->num_queries;
    rb->pool_sizes = b->pool_sizes;
    rb->num_outs = b->num_outs;
    rb->pool_size = b->pool_size;
    rb->num_out = b->num_out;
    rb->max_pool = max_pool;
    rb->entries = b->entry_ids ? b->entry_ids : c->d_entries;
    rb->entry_stride = b->entry_ids && b->entry_stride > 0 ? b->entry_stride : 0;
    rb->n_entries = b->entry_ids ? (rb->entry_stride ? rb->entry_stride : b->entry_count)
                                 : c->entries_count;
    rb->expand_count = b->expand_count;
    rb->window = b->source_window == 2 ? 2 : 1;
    rb->waves_per_query = b->waves_per_query;
    rb->cand_tile = b->candidate_tile ? b->candidate_tile : 32;
    rb->scorer = b->method == ROVER_METHOD_SQ
                     ? RV_SCORE_SQ
                     : (b->correction ? RV_SCORE_ADC_ONLINE : RV_SCORE_ADC);
    rb->refine_candidates = b->refine_candidates;
    rb->refine_error4 = b->refine_mode == ROVER_REFINE_ERROR4;
    if (b->hash_log2 == ROVER_HASH_EXACT) {
        rb->hash_log2 = 4; /* the LDS hash is not used */
        rb->visited_exact = 1;
    } else {
        rb->hash_log2 = b->hash_log2 > 0 ? b->hash_log2 : default_hash_log2(max_pool);
    }
    rb->probe_limit = b->probe_limit > 0 ? b->probe_limit : 8;
    rb->out_ids = b->out_ids;
    rb->out_dists = b->out_dists;
    rb->out_counts = b->out_counts;
    rb->out_stride = b->out_stride;
    rb->out_stats = b->out_stats;
    rb->queue = c->d_queue;
    rb->stop_label_k = c->learned.label_k > 0 ? c->learned.label_k : 10;
    rb->stop_every = c->learned.stop_every > 0 ? c->learned.stop_every : 1;
    if (b->learned_stop || b->reorder_step > 0) {
        if (!c->learned.ready) {
            snprintf(g_err, sizeof(g_err), "search: learned stop model not set");
            return ROVER_EFAILED;
        }
        rb->stop_mode = 1;
        /* without the learned stop the predictor only orders the resumption */
        rb->stop_probe_only = !b->learned_stop;
        rb->suspend_step = b->reorder_step;
        rb->stop_threshold = b->stop_threshold > 0.0f ? b->stop_threshold : 0.9f;
        rb->stop_min_steps = c->learned.min_steps;
        rb->mlp.w1 = c->learned.d_w;
        rb->mlp.w2 = c->learned.d_w + RV_MLP_HIDDEN * RV_MLP_IN;
        rb->mlp.w3 = c->learned.d_w + RV_MLP_HIDDEN * RV_MLP_IN +
                     RV_MLP_HIDDEN * RV_MLP_HIDDEN;
        rb->mlp.b1 = c->learned.d_f;
        rb->mlp.b2 = c->learned.d_f + RV_MLP_HIDDEN;
        rb->mlp.b3 = c->learned.d_f + 2 * RV_MLP_HIDDEN;
        rb->mlp.feat_mean = c->learned.d_f + 2 * RV_MLP_HIDDEN + 1;
        rb->mlp.feat_inv_std = c->learned.d_f + 2 * RV_MLP_HIDDEN + 1 + RV_MLP_IN;
    }
    return ROVER_SUCCESS;
}


/* The device router runs at the sizes its GPU tests cover: at most
 * ROVER_ROUTER_MAX_CENTROIDS centroids, queries in launches of at most
 * ROUTE_CHUNK.  (One bench run that routed 10,000 queries over 4,096
 * centroids in one launch went silent; a CPU run of both kernels and this
 * chunking at that size stayed in bounds and selected what brute force
 * selects, so no cause is known, and larger sizes are refused.) */
#define ROUTE_CHUNK 2048
AEEResult rover_route_into(RoverCtx *c, const float *d_queries, int nq, int S,
                           int32_t *d_out, void *stream) {
    const int chunk = ROUTE_CHUNK;
    if (!c->d_router_cent || c->router_n <= 0) {
        snprintf(g_err, sizeof(g_err), "route: no entry router loaded");
        return ROVER_EFAILED;
    }
    if (S <= 0 || S > 16 || S > c->router_n || nq < 0) return ROVER_EBADPARM;
    if (c->router_n > ROVER_ROUTER_MAX_CENTROIDS) {
        snprintf(g_err, sizeof(g_err), "route: %d centroids exceed the router's limit %d",
                 c->router_n, ROVER_ROUTER_MAX_CENTROIDS);
        return ROVER_EBADPARM;
    }
    if (grow((void **)&c->d_route_dist, &c->route_dist_cap,
             (size_t)(nq < chunk ? nq : chunk) * c->router_n * sizeof(float))) {
        set_err("route scratch");
        return ROVER_ENOMEMORY;
    }
    for (int q0 = 0; q0 < nq; q0 += chunk) {
        int cnt = nq - q0 < chunk ? nq - q0 : chunk;
        if (rv_route_entries(d_queries + (size_t)q0 * c->dim, cnt, c->dim,
                             c->d_router_cent, c->router_n, c->d_router_ids, S,
                             c->d_route_dist, d_out + (size_t)q0 * S, stream)) {
            set_err("route");
            return ROVER_EFAILED;
        }
    }
    return ROVER_SUCCESS;
}

AEEResult rover_route_entries(remote_handle64 h, const float *d_queries,
                              int num_queries, int S, int32_t *d_out,
                              void *stream) {
    RoverCtx *c = ctx_of(h);
    AEEResult rc;
    if (!c || !d_queries || !d_out) return ROVER_EBADPARM;
    pthread_mutex_lock(&c->lock);
    rc = rover_route_into(c, d_queries, num_queries, S, d_out, stream);
    pthread_mutex_unlock(&c->lock);
    return rc;
}

AEEResult rover_pure_pq_router_load(remote_handle64 h, const float *centroids,
                                    int centroidsLen, const int32_t *rep_ids,
                                    int count) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    if (!c->pq_loaded || !centroids || !rep_ids || count <= 0 ||
        (int64_t)centroidsLen != (int64_t)count * c->dim)
        return ROVER_EBADPARM;
    for (int i = 0; i < count; i++)
        if (rep_ids[i] < 0 || rep_ids[i] >= c->n) return ROVER_EBADPARM;
    dfree(c->d_router_cent);
    dfree(c->d_router_ids);
    c->d_router_cent = NULL;
    c->d_router_ids = NULL;
    c->router_n = 0;
    if (dalloc_copy((void **)&c->d_router_cent, centroids,
                    (size_t)centroidsLen * sizeof(float)) ||
        dalloc_copy((void **)&c->d_router_ids, rep_ids, (size_t)count * sizeof(int32_t))) {
        set_err("router_load");
        return ROVER_ENOMEMORY;
    }
    c->router_n = count;
    return ROVER_SUCCESS;
}

int rover_sqmd_seed_ids(int n, int entry_points, int graph_entry, uint32_t seed,
                        int32_t *out, int capacity) {
    uint64_t limit, attempts = 0;
    int target, count = 0;
    if (n <= 0 || entry_points <= 0 || !out || capacity <= 0) return -1;
    target = entry_points < capacity ? entry_points : capacity;
    if (target > n) target = n;
    if (graph_entry >= 0 && graph_entry < n) out[count++] = graph_entry;
    limit = (uint64_t)target * 64u + 1024u;
    while (count < target && attempts < limit) {
        int id, dup = 0;
        seed = seed * 1103515245u + 12345u;
        id = (int)((seed / 65536u) % (uint32_t)n);
        attempts++;
        for (int i = 0; i < count && !dup; i++) dup = out[i] == id;
        if (!dup) out[count++] = id;
    }
    for (int id = 0; id < n && count < target; id++) {
        int dup = 0;
        for (int i = 0; i < count && !dup; i++) dup = out[i] == id;
        if (!dup) out[count++] = id;
    }
    return count;
}

/* Uploads the live SQ path's seed list once per (n, entry_points, entry,
 * seed). */
static int ensure_sqmd_seeds(RoverCtx *c, int n, int entry_points, int graph_entry,
                             uint32_t seed) {
    int32_t *ids;
    int count;
    if (c->d_sqmd_seeds && c->sqmd_seeds_n == n && c->sqmd_seeds_ep == entry_points &&
        c->sqmd_seeds_entry == graph_entry && c->sqmd_seeds_seed == seed)
        return 0;
    dfree(c->d_sqmd_seeds);
    c->d_sqmd_seeds = NULL;
    ids = malloc((size_t)entry_points * sizeof(int32_t));
    if (!ids) return -1;
    count = rover_sqmd_seed_ids(n, entry_points, graph_entry, seed, ids, entry_points);
    if (count <= 0 || dalloc_copy((void **)&c->d_sqmd_seeds, ids, (size_t)count * 4)) {
        free(ids);
        set_err("sqmd seeds");
        return -1;
    }
    free(ids);
    c->sqmd_seeds_count = count;
    c->sqmd_seeds_n = n;
    c->sqmd_seeds_ep = entry_points;
    c->sqmd_seeds_entry = graph_entry;
    c->sqmd_seeds_seed = seed;
    return 0;
}

/* The SQ traversals with distance-only heaps (ROVER_TRAVERSAL_SQMD and
 * ROVER_TRAVERSAL_SQ_DSP). */
static AEEResult launch_sqmd(RoverCtx *c, const RoverSearchBatch *b, int max_pool,
                             int traversal, void *stream) {
    RvIndex ix;
    RvSqmdBatch sb;
    uint32_t seed = b->seed ? b->seed : ROVER_ENTRY_POINT_SEED;
    const int dsp = traversal == ROVER_TRAVERSAL_SQ_DSP;
    int grid, front_lds = 0, words;
    if (b->num_queries < 0 || (b->num_queries > 0 && !b->queries) || b->expand_count <= 0 ||
        max_pool <= 0 || max_pool > ROVER_MAX_POOL || b->out_stride <= 0 || !b->out_ids ||
        !b->out_dists || !b->out_counts || b->refine_candidates < 0 || b->learned_stop ||
        (b->correction < 0 || b->correction > 2) || (dsp && b->correction) ||
        b->entry_points <= 0 ||
        (b->refine_mode != 0 && b->refine_mode != ROVER_REFINE_CURRENT &&
         b->refine_mode != ROVER_REFINE_ERROR4) ||
        (b->refine_mode == ROVER_REFINE_ERROR4 && b->refine_candidates <= 0))
        return ROVER_EBADPARM;
    /* Caller-given seed lists and expansions of more than 32 neighbors are
     * switched off for the SQ traversals: a GPU run that used them (with an
     * out-of-range seed and a degree-48 graph) stopped responding, and no
     * cause has been shown yet. */
    if (b->entry_ids || b->route_entries) {
        snprintf(g_err, sizeof(g_err),
                 "search: the SQ traversals take their own seeds (entry_points)");
        return ROVER_EBADPARM;
    }
    if (fill_index(c, ROVER_METHOD_SQ, 0, 0, &ix)) {
        snprintf(g_err, sizeof(g_err), "search: SQ index or its graph not loaded");
        return ROVER_EFAILED;
    }
    if ((b->expand_count < ix.degree ? b->expand_count : ix.degree) > 32) {
        snprintf(g_err, sizeof(g_err),
                 "search: the SQ traversals expand at most 32 neighbors per step");
        return ROVER_EBADPARM;
    }
    if (b->correction || b->refine_candidates > 0) {
        /* the corrected distances and the final refine decode SQ + residual */
        if (!c->d_sq_res) {
            snprintf(g_err, sizeof(g_err), "search: no SQ residual loaded");
            return ROVER_EFAILED;
        }
        ix.res_codes = c->d_sq_res;
        ix.res_qparams = c->d_sq_res_qp;
        ix.res_bits = c->sq_res_bits;
        ix.res_code_bytes = c->sq_res_bits == 4 ? (c->sq_dim + 1) / 2 : c->sq_dim;
    }
    if (b->correction == 2) {
        /* the split distance needs SQ4 codes and a loaded bias */
        if (!c->d_split_bias || c->sq_bits != 4) {
            snprintf(g_err, sizeof(g_err), "search: the split distance needs SQ4 codes and "
                                           "rover_sq_split_bias_load");
            return ROVER_EBADPARM;
        }
        ix.split_bias = c->d_split_bias;
    }
    if (b->refine_mode == ROVER_REFINE_ERROR4) {
        if (!c->sq_e4_loaded) {
            snprintf(g_err, sizeof(g_err), "search: no SQERR4V1 artifact loaded");
            return ROVER_EFAILED;
        }
        ix.e4_blocks = c->d_sq_e4_blocks;
        ix.e4_qparams = c->d_sq_e4_qparams;
        ix.e4_vpb = c->sq_e4_vpb;
        ix.e4_code_bytes = c->sq_e4_code_bytes;
    }
    memset(&sb, 0, sizeof(sb));
    sb.refine_candidates = b->refine_candidates;
    sb.refine_error4 = b->refine_mode == ROVER_REFINE_ERROR4;
    if (dsp) {
        if (ensure_entries(c, ix.n, b->entry_points, ix.graph_entry, seed))
            return ROVER_EFAILED;
        sb.seeds = c->d_entries;
        sb.n_seeds = c->entries_count;
    } else {
        if (ensure_sqmd_seeds(c, ix.n, b->entry_points, ix.graph_entry, seed))
            return ROVER_EFAILED;
        sb.seeds = c->d_sqmd_seeds;
        sb.n_seeds = c->sqmd_seeds_count;
    }
    sb.queries = b->queries;
    sb.nq = b->num_queries;
    sb.pool_sizes = b->pool_sizes;
    sb.num_outs = b->num_outs;
    sb.pool_size = b->pool_size;
    sb.num_out = b->num_out;
    sb.max_pool = max_pool;
    /* the frontier capacity is sized from the entry points and the expand
     * count, as the reference sizes it */
    sb.fcap_entries = b->entry_points;
    sb.fcap_expand = dsp && b->expand_count > ix.degree ? ix.degree : b->expand_count;
    sb.expand_count = b->expand_count;
    sb.max_work = 32;
    sb.correction = b->correction;
    sb.out_ids = b->out_ids;
    sb.out_dists = b->out_dists;
    sb.out_counts = b->out_counts;
    sb.out_stride = b->out_stride;
    sb.out_stats = b->out_stats;
    sb.queue = c->d_queue;
    if (sb.nq == 0) return ROVER_SUCCESS;
    grid = rv_sqmd_grid(&ix, &sb, &front_lds);
    if (grid < 0) {
        set_err("sqmd grid");
        return ROVER_EFAILED;
    }
    if (b->grid_waves > 0 && b->grid_waves < grid) grid = b->grid_waves;
    words = ((ix.n + 31) / 32 + 3) / 4 * 4;
    sb.visited_words = words;
    if (grow((void **)&c->d_sqmd_vis, &c->sqmd_vis_cap, (size_t)grid * words * 4)) {
        set_err("sqmd visited scratch");
        return ROVER_ENOMEMORY;
    }
    sb.visited = c->d_sqmd_vis;
    if (!front_lds) {
        sb.frontier_phys = sb.fcap_entries + max_pool * sb.fcap_expand;
        if (grow(&c->d_sqmd_front, &c->sqmd_front_cap,
                 (size_t)grid * sb.frontier_phys * sizeof(RvSqmdCand))) {
            set_err("sqmd frontier scratch");
            return ROVER_ENOMEMORY;
        }
        sb.frontier = (RvSqmdCand *)c->d_sqmd_front;
    }
    if (rv_sqmd_launch(&ix, &sb, grid, stream)) {
        set_err("sqmd launch");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

static AEEResult launch_batch_on(RoverCtx *c, const RoverSearchBatch *b_in,
                                 int max_pool, int explicit_entries, void *stream);

/* One eager batch on `stream`, ordered after the session's previous launch
 * (the work counter and scratch are the session's). */
static AEEResult launch_batch(RoverCtx *c, const RoverSearchBatch *b_in,
                              int max_pool, int explicit_entries,
                              void *stream) {
    RoverSearchBatch routed;
    const RoverSearchBatch *b = b_in;
    AEEResult rc;
    if (rover_stream_turn(c, stream)) return ROVER_EFAILED;
    rc = launch_batch_on(c, b_in, max_pool, explicit_entries, stream);
    if (rover_stream_done(c, stream) && rc == ROVER_SUCCESS) rc = ROVER_EFAILED;
    return rc;
}

static AEEResult launch_batch_on(RoverCtx *c, const RoverSearchBatch *b_in,
                                 int max_pool, int explicit_entries, void *stream) {
    RvIndex ix;
    RvBatch rb;
    RoverSearchBatch routed;
    const RoverSearchBatch *b = b_in;
    AEEResult rc;
    if (b_in->method == ROVER_METHOD_SQ && b_in->traversal != ROVER_TRAVERSAL_BEST_FIRST) {
        if (b_in->traversal != ROVER_TRAVERSAL_DEFAULT &&
            b_in->traversal != ROVER_TRAVERSAL_SQMD && b_in->traversal != ROVER_TRAVERSAL_SQ_DSP)
            return ROVER_EBADPARM;
        if (b_in->source_window > 1) {
            set_err("search: the SQ heap traversals have no source window");
            return ROVER_EBADPARM;
        }
        return launch_sqmd(c, b_in, max_pool,
                           b_in->traversal == ROVER_TRAVERSAL_SQ_DSP ? ROVER_TRAVERSAL_SQ_DSP
                                                                     : ROVER_TRAVERSAL_SQMD,
                           stream);
    }
    if (b_in->route_entries > 0) {
        /* per-query entries from the router, written to a session buffer */
        int S = b_in->route_entries;
        if (b_in->num_queries > 0 && !b_in->queries) return ROVER_EBADPARM;
        if (grow((void **)&c->d_route_ent, &c->route_ent_cap,
                 (size_t)(b_in->num_queries > 0 ? b_in->num_queries : 1) * S * 4)) {
            set_err("route entries");
            return ROVER_ENOMEMORY;
        }
        rc = rover_route_into(c, b_in->queries, b_in->num_queries, S, c->d_route_ent,
                              stream);
        if (rc != ROVER_SUCCESS) return rc;
        routed = *b_in;
        routed.entry_ids = c->d_route_ent;
        routed.entry_count = S;
        routed.entry_stride = S;
        b = &routed;
    }
    rc = rover_prepare_launch(c, b, max_pool, explicit_entries, &ix, &rb);
    if (rc != ROVER_SUCCESS) return rc;
    if (rover_launch(c, &ix, &rb, b->grid_waves, stream)) {
        set_err("search launch");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_search_batch_device(remote_handle64 h, const RoverSearchBatch *b,
                                    void *stream) {
    RoverCtx *c = ctx_of(h);
    AEEResult rc;
    if (!c || !b) return ROVER_EBADPARM;
    pthread_mutex_lock(&c->lock);
    rc = launch_batch(c, b, b->pool_size, 0, stream);
    pthread_mutex_unlock(&c->lock);
    return rc;
}

/* rover_search_batch_device with every workspace in the device's fixed
 * no-allocation array (rv_noalloc_scratch, resolved at calculator_open): the
 * work counter and the graph entry + LCG seed list, which a kernel writes
 * there.  Supports the best-first traversal (PQ, and SQ with
 * ROVER_TRAVERSAL_BEST_FIRST); routed entries and the SQ heap traversals need
 * scratch that grows with the batch or the index and are refused. */
AEEResult rover_search_batch_device_noalloc(remote_handle64 h, const RoverSearchBatch *b,
                                            void *stream) {
    RoverCtx *c = ctx_of(h);
    RvIndex ix;
    RvBatch rb;
    AEEResult rc;
    uint32_t seed;
    int32_t *work, *seeds;
    if (!c || !b) return ROVER_EBADPARM;
    if (!c->d_noalloc) {
        snprintf(g_err, sizeof(g_err), "search_noalloc: no workspace on this device");
        return ROVER_EFAILED;
    }
    if (b->route_entries > 0) {
        snprintf(g_err, sizeof(g_err),
                 "search_noalloc: routed entries need a routing scratch; pass entry_ids");
        return ROVER_EBADPARM;
    }
    if (b->method == ROVER_METHOD_SQ && b->traversal != ROVER_TRAVERSAL_BEST_FIRST) {
        snprintf(g_err, sizeof(g_err),
                 "search_noalloc: SQ batches need traversal ROVER_TRAVERSAL_BEST_FIRST");
        return ROVER_EBADPARM;
    }
    if (!b->entry_ids && b->entry_points > RV_NOALLOC_MAX_SEEDS) {
        snprintf(g_err, sizeof(g_err),
                 "search_noalloc: %d entry points exceed the workspace's %d seed ids",
                 b->entry_points, RV_NOALLOC_MAX_SEEDS);
        return ROVER_EBADPARM;
    }
    work = (int32_t *)c->d_noalloc;
    seeds = (int32_t *)((char *)c->d_noalloc + RV_NOALLOC_SEED_OFFSET);
    seed = b->seed ? b->seed : ROVER_ENTRY_POINT_SEED;
    pthread_mutex_lock(&c->lock);
    rc = rover_prepare_launch(c, b, b->pool_size, 1, &ix, &rb);
    if (rc != ROVER_SUCCESS) goto out;
    if (rb.visited_exact) {
        snprintf(g_err, sizeof(g_err),
                 "search_noalloc: exact visited sets need a scratch; use an LDS hash");
        rc = ROVER_EBADPARM;
        goto out;
    }
    rc = ROVER_EFAILED;
    rb.queue = work;
    if (!b->entry_ids) {
        int count = b->entry_points < ix.n ? b->entry_points : ix.n;
        rb.entries = seeds;
        rb.n_entries = count;
        rb.entry_stride = 0;
        if (b->num_queries > 0 &&
            rv_seed_list(seeds, ix.n, count, ix.graph_entry, seed, stream)) {
            set_err("search_noalloc seeds");
            goto out;
        }
    }
    if (rv_search_launch(&ix, &rb, b->grid_waves, stream)) {
        set_err("search_noalloc launch");
        goto out;
    }
    rc = ROVER_SUCCESS;
out:
    pthread_mutex_unlock(&c->lock);
    return rc;
}

/* ------------------------------------------------------------ graphs */

#define ROVER_GRAPH_MAGIC 0x52564750u /* "RVGP" */

typedef struct RoverGraph {
    uint32_t magic;
    void *exec;         /* instantiated HIP graph                    */
    int32_t *d_queue;   /* the graph's own work counter              */
    int32_t *d_entries; /* copy of the LCG seed list, when used      */
    uint64_t *d_visited; /* the graph's own exact visited sets       */
    size_t visited_cap, visited_layout;
} RoverGraph;

static void graph_free(RoverGraph *g) {
    if (!g) return;
    if (g->exec) rv_graph_destroy(g->exec);
    dfree(g->d_queue);
    dfree(g->d_entries);
    dfree(g->d_visited);
    g->magic = 0;
    free(g);
}

AEEResult rover_search_graph_create(remote_handle64 h, const RoverSearchBatch *b,
                                    const int32_t *d_num_queries, void *stream,
                                    void **graph) {
    RoverCtx *c = ctx_of(h);
    RvIndex ix;
    RvBatch rb;
    RoverGraph *g;
    AEEResult rc;
    int grid;
    if (!c || !b || !stream || !graph || b->route_entries > 0 || b->num_queries <= 0)
        return ROVER_EBADPARM;
    *graph = NULL;
    g = calloc(1, sizeof(*g));
    if (!g) return ROVER_ENOMEMORY;
    pthread_mutex_lock(&c->lock);
    rc = rover_prepare_launch(c, b, b->pool_size, 0, &ix, &rb);
    if (rc != ROVER_SUCCESS) goto fail;
    rc = ROVER_ENOMEMORY;
    if (rv_malloc((void **)&g->d_queue, 64)) {
        set_err("graph work counter");
        goto fail;
    }
    if (!b->entry_ids) {
        /* the session's seed list may be replaced later: keep a copy */
        if (rv_malloc((void **)&g->d_entries, (size_t)rb.n_entries * sizeof(int32_t)) ||
            rv_memcpy_d2d(g->d_entries, rb.entries, (size_t)rb.n_entries * sizeof(int32_t))) {
            set_err("graph seed list");
            goto fail;
        }
        rb.entries = g->d_entries;
    }
    rb.queue = g->d_queue;
    rb.nq_dev = d_num_queries;
    rc = ROVER_EFAILED;
    /* attributes and grid are settled before the capture begins */
    grid = rv_search_prepare(&ix, &rb);
    if (grid <= 0) {
        set_err("graph prepare");
        goto fail;
    }
    if (b->grid_waves > 0) grid = b->grid_waves;
    if (grid > rb.nq) grid = rb.nq;
    if (rb.visited_exact &&
        (rover_attach_visited(&g->d_visited, &g->visited_cap, &g->visited_layout, &ix, &rb,
                              grid, stream) ||
         rv_stream_synchronize(stream))) {
        set_err("graph visited sets");
        goto fail;
    }
    if (rv_graph_begin(stream)) {
        set_err("graph capture");
        goto fail;
    }
    {
        const int launch_rc = rv_search_launch(&ix, &rb, grid, stream);
        const int end_rc = rv_graph_end(stream, &g->exec);
        if (launch_rc || end_rc) {
            set_err("graph capture");
            goto fail;
        }
    }
    g->magic = ROVER_GRAPH_MAGIC;
    pthread_mutex_unlock(&c->lock);
    *graph = g;
    return ROVER_SUCCESS;
fail:
    pthread_mutex_unlock(&c->lock);
    graph_free(g);
    return rc;
}

AEEResult rover_search_graph_launch(void *graph, void *stream) {
    RoverGraph *g = (RoverGraph *)graph;
    if (!g || g->magic != ROVER_GRAPH_MAGIC || !stream) return ROVER_EBADPARM;
    if (rv_graph_launch(g->exec, stream)) {
        set_err("graph launch");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_search_graph_destroy(void *graph) {
    RoverGraph *g = (RoverGraph *)graph;
    if (!g || g->magic != ROVER_GRAPH_MAGIC) return ROVER_EBADPARM;
    graph_free(g);
    return ROVER_SUCCESS;
}

/* ------------------------------------------------------------ autotuner */

/* Mean launch time of `repeats` launches after one warm-up, or -1. */
static float time_launches(RoverCtx *c, const RvIndex *ix, RvBatch *rb, int grid,
                           void *stream, int repeats, void *e0, void *e1) {
    float ms = 0.0f;
    if (rover_launch(c, ix, rb, grid, stream) || rv_event_record(e0, stream)) return -1.0f;
    for (int r = 0; r < repeats; r++)
        if (rover_launch(c, ix, rb, grid, stream)) return -1.0f;
    if (rv_event_record(e1, stream) || rv_event_elapsed(e0, e1, &ms)) return -1.0f;
    return ms / (float)repeats;
}

AEEResult rover_autotune(remote_handle64 h, const RoverSearchBatch *b, void *stream,
                         int repeats, RoverTuning *out) {
    /* 0: the default LDS hash; ROVER_HASH_EXACT: exact sets in HBM */
    static const int hashes[6] = {0, 9, 10, 12, 14, ROVER_HASH_EXACT};
    static const int probes[3] = {4, 8, 16};
    static const int fill_num[3] = {4, 3, 2}; /* grid fill in quarters */
    RoverCtx *c = ctx_of(h);
    RvIndex ix;
    RvBatch rb, cand;
    AEEResult rc;
    void *e0 = NULL, *e1 = NULL;
    int32_t *ref_ids = NULL, *got_ids = NULL, *ref_cnt = NULL, *got_cnt = NULL;
    float *ref_d = NULL, *got_d = NULL;
    size_t nout, bytes_out;
    if (!c || !b || !out || b->num_queries <= 0 || b->route_entries > 0)
        return ROVER_EBADPARM;
    if (repeats <= 0) repeats = 3;
    memset(out, 0, sizeof(*out));
    nout = (size_t)b->num_queries * (size_t)b->out_stride;
    bytes_out = nout * sizeof(int32_t);
    pthread_mutex_lock(&c->lock);
    rc = rover_prepare_launch(c, b, b->pool_size, 0, &ix, &rb);
    if (rc != ROVER_SUCCESS) goto out;
    if (rover_stream_turn(c, stream)) {
        rc = ROVER_EFAILED;
        goto out;
    }
    rc = ROVER_ENOMEMORY;
    ref_ids = malloc(bytes_out);
    got_ids = malloc(bytes_out);
    ref_d = malloc(bytes_out);
    got_d = malloc(bytes_out);
    ref_cnt = malloc((size_t)b->num_queries * sizeof(int32_t));
    got_cnt = malloc((size_t)b->num_queries * sizeof(int32_t));
    if (!ref_ids || !got_ids || !ref_d || !got_d || !ref_cnt || !got_cnt) goto out;
    rc = ROVER_EFAILED;
    if (rv_event_create(&e0) || rv_event_create(&e1)) {
        set_err("autotune events");
        goto out;
    }
    /* the default configuration sets the reference results and time */
    out->hash_log2 = rb.visited_exact ? ROVER_HASH_EXACT : rb.hash_log2;
    out->probe_limit = rb.probe_limit;
    out->grid_waves = rv_search_prepare(&ix, &rb);
    if (out->grid_waves <= 0) {
        set_err("autotune prepare");
        goto out;
    }
    if (out->grid_waves > b->num_queries) out->grid_waves = b->num_queries;
    out->default_ms = time_launches(c, &ix, &rb, out->grid_waves, stream, repeats, e0, e1);
    if (out->default_ms < 0.0f || rv_stream_synchronize(stream) ||
        rv_memcpy_d2h(ref_ids, rb.out_ids, bytes_out) ||
        rv_memcpy_d2h(ref_d, rb.out_dists, bytes_out) ||
        rv_memcpy_d2h(ref_cnt, rb.out_counts, (size_t)b->num_queries * sizeof(int32_t))) {
        set_err("autotune reference run");
        goto out;
    }
    out->ms = out->default_ms;
    out->tried = 1;
    for (int hi = 0; hi < 6; hi++) {
        for (int pi = 0; pi < 3; pi++) {
            for (int fi = 0; fi < 3; fi++) {
                int full, grid;
                float ms;
                const int exact = hashes[hi] == ROVER_HASH_EXACT;
                if (exact && pi > 0) continue; /* no probes in an exact set */
                if (!exact && rb.suspend_step > 0) continue; /* reordering needs exact sets */
                cand = rb;
                cand.visited_exact = exact;
                if (exact) cand.hash_log2 = 4;
                else if (hashes[hi]) cand.hash_log2 = hashes[hi];
                else if (rb.visited_exact) continue; /* no LDS default to try */
                cand.probe_limit = exact ? rb.probe_limit : probes[pi];
                if (cand.visited_exact == rb.visited_exact &&
                    cand.hash_log2 == rb.hash_log2 && cand.probe_limit == rb.probe_limit &&
                    fi == 0)
                    continue; /* the default, already timed */
                full = rv_search_prepare(&ix, &cand);
                if (full <= 0) continue; /* does not fit in LDS */
                grid = full * fill_num[fi] / 4;
                if (grid < 1) grid = 1;
                if (grid > b->num_queries) grid = b->num_queries;
                ms = time_launches(c, &ix, &cand, grid, stream, repeats, e0, e1);
                if (ms < 0.0f || rv_stream_synchronize(stream) ||
                    rv_memcpy_d2h(got_ids, cand.out_ids, bytes_out) ||
                    rv_memcpy_d2h(got_d, cand.out_dists, bytes_out) ||
                    rv_memcpy_d2h(got_cnt, cand.out_counts,
                                  (size_t)b->num_queries * sizeof(int32_t))) {
                    set_err("autotune run");
                    goto out;
                }
                out->tried++;
                if (memcmp(got_ids, ref_ids, bytes_out) || memcmp(got_d, ref_d, bytes_out) ||
                    memcmp(got_cnt, ref_cnt, (size_t)b->num_queries * sizeof(int32_t))) {
                    out->rejected++;
                    continue;
                }
                if (ms < out->ms) {
                    out->ms = ms;
                    out->hash_log2 = exact ? ROVER_HASH_EXACT : cand.hash_log2;
                    out->probe_limit = cand.probe_limit;
                    out->grid_waves = grid;
                }
            }
        }
    }
    /* waves per query x candidate tile (the multi-wave kernel) with the
       visited set chosen above, every CU filled */
    out->waves_per_query = rb.waves_per_query;
    out->candidate_tile = rb.waves_per_query ? rb.cand_tile : 0;
    if (rb.waves_per_query == 0) {
        static const int waves[3] = {1, 2, 4};
        static const int tiles[3] = {16, 32, 64};
        for (int wi = 0; wi < 3; wi++) {
            for (int ti = 0; ti < 3; ti++) {
                int grid;
                float ms;
                cand = rb;
                cand.visited_exact = out->hash_log2 == ROVER_HASH_EXACT;
                cand.hash_log2 = cand.visited_exact ? 4 : out->hash_log2;
                cand.probe_limit = out->probe_limit;
                cand.waves_per_query = waves[wi];
                cand.cand_tile = tiles[ti];
                grid = rv_search_prepare(&ix, &cand);
                if (grid <= 0) {
                    /* not this batch's shape (or no fit): the axis does not apply */
                    rv_set_error("");
                    continue;
                }
                if (grid > b->num_queries) grid = b->num_queries;
                ms = time_launches(c, &ix, &cand, grid, stream, repeats, e0, e1);
                if (ms < 0.0f || rv_stream_synchronize(stream) ||
                    rv_memcpy_d2h(got_ids, cand.out_ids, bytes_out) ||
                    rv_memcpy_d2h(got_d, cand.out_dists, bytes_out) ||
                    rv_memcpy_d2h(got_cnt, cand.out_counts,
                                  (size_t)b->num_queries * sizeof(int32_t))) {
                    set_err("autotune run");
                    goto out;
                }
                out->tried++;
                if (memcmp(got_ids, ref_ids, bytes_out) || memcmp(got_d, ref_d, bytes_out) ||
                    memcmp(got_cnt, ref_cnt, (size_t)b->num_queries * sizeof(int32_t))) {
                    out->rejected++;
                    continue;
                }
                if (ms < out->ms) {
                    out->ms = ms;
                    out->grid_waves = grid;
                    out->waves_per_query = waves[wi];
                    out->candidate_tile = tiles[ti];
                }
            }
        }
    }
    /* leave the chosen configuration's results in the output buffers */
    cand = rb;
    cand.visited_exact = out->hash_log2 == ROVER_HASH_EXACT;
    cand.hash_log2 = cand.visited_exact ? 4 : out->hash_log2;
    cand.probe_limit = out->probe_limit;
    cand.waves_per_query = out->waves_per_query;
    cand.cand_tile = out->candidate_tile ? out->candidate_tile : 32;
    if (rover_launch(c, &ix, &cand, out->grid_waves, stream) || rv_stream_synchronize(stream)) {
        set_err("autotune final run");
        goto out;
    }
    rc = ROVER_SUCCESS;
out:
    if (c->launched || rc == ROVER_SUCCESS) rover_stream_done(c, stream);
    pthread_mutex_unlock(&c->lock);
    if (e0) rv_event_destroy(e0);
    if (e1) rv_event_destroy(e1);
    free(ref_ids);
    free(got_ids);
    free(ref_d);
    free(got_d);
    free(ref_cnt);
    free(got_cnt);
    return rc;
}

static int ensure_stage(RoverCtx *c, size_t bytes) {
    if (bytes <= c->stage_cap) return 0;
    dfree(c->d_stage);
    c->d_stage = NULL;
    c->stage_cap = 0;
    if (rv_malloc(&c->d_stage, bytes)) {
        set_err("stage alloc");
        return -1;
    }
    c->stage_cap = bytes;
    return 0;
}

static size_t up256(size_t v) { return (v + 255) & ~(size_t)255; }

static AEEResult search_host(RoverCtx *c, const RoverSearchBatch *b,
                             int explicit_entries) {
    RoverSearchBatch d;
    int dim, max_pool;
    size_t oq, ol, ok, oi, od, oc, os, oe, total, entries_bytes;
    AEEResult rc = ROVER_EFAILED;
    if (!c || !b) return ROVER_EBADPARM;
    if (b->num_queries < 0 || (b->num_queries > 0 && !b->queries) ||
        b->out_stride <= 0 || !b->out_ids || !b->out_dists || !b->out_counts)
        return ROVER_EBADPARM;
    if (b->num_queries == 0) return ROVER_SUCCESS;
    dim = b->method == ROVER_METHOD_SQ ? c->sq_dim : c->dim;
    max_pool = b->pool_size;
    if (b->pool_sizes)
        for (int i = 0; i < b->num_queries; i++)
            if (b->pool_sizes[i] > max_pool) max_pool = b->pool_sizes[i];
    pthread_mutex_lock(&c->lock);
    oq = 0;
    ol = up256(oq + (size_t)b->num_queries * dim * sizeof(float));
    ok = up256(ol + (b->pool_sizes ? (size_t)b->num_queries * 4 : 0));
    oi = up256(ok + (b->num_outs ? (size_t)b->num_queries * 4 : 0));
    od = up256(oi + (size_t)b->num_queries * b->out_stride * 4);
    oc = up256(od + (size_t)b->num_queries * b->out_stride * 4);
    os = up256(oc + (size_t)b->num_queries * 4);
    oe = up256(os + (b->out_stats ? (size_t)b->num_queries * 32 : 0));
    entries_bytes = b->entry_ids ? (size_t)(b->entry_stride > 0
                                                ? b->entry_stride * (size_t)b->num_queries
                                                : (size_t)b->entry_count) * 4
                                 : 0;
    total = up256(oe + entries_bytes);
    if (ensure_stage(c, total)) goto out;
    {
        char *base = (char *)c->d_stage;
        d = *b;
        d.queries = (const float *)(base + oq);
        d.pool_sizes = b->pool_sizes ? (const int32_t *)(base + ol) : NULL;
        d.num_outs = b->num_outs ? (const int32_t *)(base + ok) : NULL;
        d.out_ids = (int32_t *)(base + oi);
        d.out_dists = (float *)(base + od);
        d.out_counts = (int32_t *)(base + oc);
        d.out_stats = b->out_stats ? (uint64_t *)(base + os) : NULL;
        d.entry_ids = b->entry_ids ? (const int32_t *)(base + oe) : NULL;
        /* uploads, the search and downloads are ordered on the session's
           stream; from page-locked host memory (rover_host_alloc) the copies
           are DMA transfers, from pageable memory HIP stages them */
        void *s = c->stream;
        if (rover_stream_turn(c, s)) goto out;
        if ((entries_bytes && rv_memcpy_h2d_async(base + oe, b->entry_ids, entries_bytes, s)) ||
            rv_memcpy_h2d_async(base + oq, b->queries,
                                (size_t)b->num_queries * dim * sizeof(float), s) ||
            (b->pool_sizes &&
             rv_memcpy_h2d_async(base + ol, b->pool_sizes, (size_t)b->num_queries * 4, s)) ||
            (b->num_outs &&
             rv_memcpy_h2d_async(base + ok, b->num_outs, (size_t)b->num_queries * 4, s))) {
            set_err("batch upload");
            rv_stream_synchronize(s);
            goto out;
        }
        rc = launch_batch(c, &d, max_pool, explicit_entries, s);
        if (rc != ROVER_SUCCESS) {
            rv_stream_synchronize(s);
            goto out;
        }
        rc = ROVER_EFAILED;
        if (rv_memcpy_d2h_async(b->out_ids, base + oi,
                                (size_t)b->num_queries * b->out_stride * 4, s) ||
            rv_memcpy_d2h_async(b->out_dists, base + od,
                                (size_t)b->num_queries * b->out_stride * 4, s) ||
            rv_memcpy_d2h_async(b->out_counts, base + oc, (size_t)b->num_queries * 4, s) ||
            (b->out_stats &&
             rv_memcpy_d2h_async(b->out_stats, base + os, (size_t)b->num_queries * 32, s)) ||
            rv_stream_synchronize(s)) {
            set_err("batch download");
            rv_stream_synchronize(s);
            goto out;
        }
        rc = ROVER_SUCCESS;
    }
out:
    pthread_mutex_unlock(&c->lock);
    return rc;
}

AEEResult rover_search_batch(remote_handle64 h, const RoverSearchBatch *b) {
    RoverCtx *c = ctx_of(h);
    if (!c) return ROVER_EBADPARM;
    return search_host(c, b, 0);
}

/* calculator.idl:46-55 on the SQ index: one query of the DSP search op's
 * traversal (retrieval_async_dsp_search.c:368-509) through the batch path. */
AEEResult calculator_retrieval_search(remote_handle64 h, const float *query,
                                      int queryLen, int32_t entry_points,
                                      int32_t expand_count, int32_t pool_size,
                                      int32_t num_out, int32_t *ids, int idsLen,
                                      float *dists, int distsLen,
                                      int32_t *actual_count) {
    RoverCtx *c = ctx_of(h);
    RoverSearchBatch b;
    int32_t count = 0;
    AEEResult rc;
    if (!actual_count) return ROVER_EBADPARM;
    *actual_count = 0;
    if (!c) return ROVER_EBADPARM;
    if (!c->sq_loaded) return ROVER_EFAILED;
    if (!query || queryLen < c->sq_dim || pool_size <= 0 || num_out <= 0 ||
        !ids || !dists || idsLen < num_out || distsLen < num_out ||
        entry_points <= 0 || expand_count <= 0)
        return ROVER_EBADPARM;
    if (pool_size > ROVER_MAX_POOL) pool_size = ROVER_MAX_POOL;
    if (num_out > pool_size) num_out = pool_size;
    memset(&b, 0, sizeof(b));
    b.queries = query;
    b.num_queries = 1;
    b.pool_size = pool_size;
    b.num_out = num_out;
    b.method = ROVER_METHOD_SQ;
    b.traversal = ROVER_TRAVERSAL_SQ_DSP;
    b.entry_points = entry_points;
    b.seed = ROVER_ENTRY_POINT_SEED;
    b.expand_count = expand_count;
    b.out_ids = ids;
    b.out_dists = dists;
    b.out_counts = &count;
    b.out_stride = num_out;
    rc = rover_search_batch(h, &b);
    if (rc != ROVER_SUCCESS) return rc;
    if (count < 0) return ROVER_EFAILED;
    *actual_count = count;
    return ROVER_SUCCESS;
}

/* ------------------------------------------------ library search (face 2) */

int rover_pq_residual_evaluator(void *context, int vector_id,
                                float coarse_distance, float *out_distance) {
    (void)context;
    (void)vector_id;
    (void)coarse_distance;
    (void)out_distance;
    return -1;
}

static pthread_mutex_t g_view_lock = PTHREAD_MUTEX_INITIALIZER;
static remote_handle64 g_view_handle;
static struct {
    const int32_t *graph;
    const void *chunk0;
    const float *codebook;
    const uint8_t *res_codes;
    int n, degree, dim, m, ks, chunk_count, chunk_vectors, res_bits;
} g_view_key;

void rover_pure_pq_view_release(void) {
    pthread_mutex_lock(&g_view_lock);
    if (g_view_handle) calculator_close(g_view_handle);
    g_view_handle = 0;
    memset(&g_view_key, 0, sizeof(g_view_key));
    pthread_mutex_unlock(&g_view_lock);
}

int retrieval_pure_pq_search_scalar_with_policy(
    const PurePqSearchIndexView *ix, const PurePqSearchRequest *rq,
    const PurePqDistancePolicy *policy, int *out_ids, float *out_distances,
    int out_capacity, PurePqSearchStats *out_stats) {
    const RoverResidualView *rv = NULL;
    int online = 0, result = -1, ks_bytes;
    if (out_stats) memset(out_stats, 0, sizeof(*out_stats));
    /* retrieval_sqmd_pure_pq.c:251-278 (pure_pq_validate_search) */
    if (!ix || !rq || !out_ids || !out_distances || out_capacity <= 0 ||
        !ix->graph || ix->graph_degree <= 0 || ix->graph_degree > ROVER_MAX_DEGREE ||
        ix->num_vectors <= 0 || !ix->code_chunks || !ix->chunk_vector_counts ||
        ix->chunk_count <= 0 || ix->chunk_vectors <= 0 || !ix->codebook ||
        ix->dim <= 0 || ix->m <= 0 || (ix->ks != 16 && ix->ks != 256) ||
        ix->subdim <= 0 || ix->m * ix->subdim != ix->dim || !rq->query ||
        !rq->entry_ids || rq->entry_count <= 0 || rq->expand_count <= 0 ||
        rq->pool_capacity <= 0 || rq->num_out <= 0 ||
        rq->num_out > rq->pool_capacity || rq->pool_capacity > ROVER_MAX_POOL)
        return -1;
    for (int i = 0; i < ix->chunk_count; i++)
        if (!ix->code_chunks[i] || ix->chunk_vector_counts[i] <= 0 ||
            ix->chunk_vector_counts[i] > ix->chunk_vectors)
            return -1;
    /* retrieval_sqmd_pure_pq.c:31-38: OFF, or ONLINE with an evaluator;
     * the GPU runs only the residual evaluator it knows. */
    if (policy && policy->mode != PURE_PQ_CORRECTION_OFF) {
        if (policy->mode != PURE_PQ_CORRECTION_ONLINE ||
            policy->evaluator != rover_pq_residual_evaluator || !policy->context)
            return -1;
        rv = (const RoverResidualView *)policy->context;
        if (!rv->codes || !rv->qparams || (rv->bits != 4 && rv->bits != 8))
            return -1;
        online = 1;
    }
    ks_bytes = ix->ks == 16 ? ix->m / 2 + ix->m % 2 : ix->m;
    pthread_mutex_lock(&g_view_lock);
    if (!g_view_handle || g_view_key.graph != ix->graph ||
        g_view_key.chunk0 != ix->code_chunks[0] ||
        g_view_key.codebook != ix->codebook || g_view_key.n != ix->num_vectors ||
        g_view_key.degree != ix->graph_degree || g_view_key.dim != ix->dim ||
        g_view_key.m != ix->m || g_view_key.ks != ix->ks ||
        g_view_key.chunk_count != ix->chunk_count ||
        g_view_key.chunk_vectors != ix->chunk_vectors ||
        (online && (g_view_key.res_codes != rv->codes ||
                    g_view_key.res_bits != rv->bits))) {
        if (g_view_handle) calculator_close(g_view_handle);
        g_view_handle = 0;
        memset(&g_view_key, 0, sizeof(g_view_key));
        if (calculator_open(NULL, &g_view_handle) != ROVER_SUCCESS ||
            calculator_retrieval_pure_pq_load_dma(
                g_view_handle, ix->code_chunks, ix->chunk_count,
                ix->chunk_vector_counts, ix->chunk_count, ix->chunk_vectors,
                ix->codebook, ix->m * ix->ks * ix->subdim, ix->num_vectors,
                ix->dim, ix->m, ix->ks, ix->subdim, ks_bytes, NULL, 0) ||
            rover_pure_pq_graph_load(g_view_handle, ix->graph,
                                     (int64_t)ix->num_vectors * ix->graph_degree,
                                     ix->num_vectors, ix->graph_degree, -1) ||
            (online && rover_pure_pq_residual_load(
                           g_view_handle, rv->codes,
                           (int64_t)ix->num_vectors *
                               (rv->bits == 4 ? ix->dim / 2 + ix->dim % 2 : ix->dim),
                           rv->qparams, 2 * (int64_t)ix->num_vectors, rv->bits))) {
            if (g_view_handle) calculator_close(g_view_handle);
            g_view_handle = 0;
            goto out;
        }
        g_view_key.graph = ix->graph;
        g_view_key.chunk0 = ix->code_chunks[0];
        g_view_key.codebook = ix->codebook;
        g_view_key.n = ix->num_vectors;
        g_view_key.degree = ix->graph_degree;
        g_view_key.dim = ix->dim;
        g_view_key.m = ix->m;
        g_view_key.ks = ix->ks;
        g_view_key.chunk_count = ix->chunk_count;
        g_view_key.chunk_vectors = ix->chunk_vectors;
        g_view_key.res_codes = online ? rv->codes : NULL;
        g_view_key.res_bits = online ? rv->bits : 0;
    }
    {
        /* explicit entry ids (not the LCG): upload them as the seed list */
        RoverCtx *c = ctx_of(g_view_handle);
        RoverSearchBatch b;
        int32_t count = 0;
        uint64_t st[4] = {0, 0, 0, 0};
        int num_out = rq->num_out < out_capacity ? rq->num_out : out_capacity;
        if (rq->entry_count > c->entries_cap) {
            dfree(c->d_entries);
            c->d_entries = NULL;
            c->entries_cap = 0;
            if (rv_malloc((void **)&c->d_entries, (size_t)rq->entry_count * 4))
                goto out;
            c->entries_cap = rq->entry_count;
        }
        if (rv_memcpy_h2d(c->d_entries, rq->entry_ids, (size_t)rq->entry_count * 4))
            goto out;
        /* pin the cache key to something ensure_entries never matches */
        c->entries_n = -1;
        c->entries_count = rq->entry_count;
        memset(&b, 0, sizeof(b));
        b.queries = rq->query;
        b.num_queries = 1;
        b.pool_size = rq->pool_capacity;
        b.num_out = num_out;
        b.method = ROVER_METHOD_PQ;
        b.correction = online;
        b.entry_points = rq->entry_count;
        b.expand_count = rq->expand_count;
        b.out_ids = out_ids;
        b.out_dists = out_distances;
        b.out_counts = &count;
        b.out_stride = num_out;
        b.out_stats = st;
        if (search_host(c, &b, 1) != ROVER_SUCCESS || count < 0)
            goto out;
        if (out_stats) {
            out_stats->pq_scored_candidates = st[0];
            out_stats->corrected_candidates = st[1];
            out_stats->expanded_nodes = st[2];
            out_stats->visited_dedup_skips = st[3];
        }
        result = count;
    }
out:
    pthread_mutex_unlock(&g_view_lock);
    return result;
}

int retrieval_pure_pq_search_scalar(const PurePqSearchIndexView *index,
                                    const PurePqSearchRequest *request,
                                    int *out_ids, float *out_distances,
                                    int out_capacity,
                                    PurePqSearchStats *out_stats) {
    return retrieval_pure_pq_search_scalar_with_policy(
        index, request, NULL, out_ids, out_distances, out_capacity, out_stats);
}

/* ---------------------------------------------------------- device memory */

AEEResult rover_device_alloc(void **d_ptr, size_t bytes) {
    if (!d_ptr) return ROVER_EBADPARM;
    if (rv_malloc(d_ptr, bytes)) {
        set_err("rover_device_alloc");
        return ROVER_ENOMEMORY;
    }
    return ROVER_SUCCESS;
}
/* Decoded vectors of `count` ids into out [count][dim] (host buffers). */
static AEEResult decode_ids(remote_handle64 h, int sq, const int32_t *ids, int count,
                            float *out) {
    RoverCtx *c = ctx_of(h);
    RvIndex ix;
    int32_t *d_ids = NULL;
    float *d_out = NULL;
    AEEResult rc = ROVER_EFAILED;
    int n, dim;
    if (!c || count < 0 || (count > 0 && (!ids || !out))) return ROVER_EBADPARM;
    pthread_mutex_lock(&c->lock);
    if (sq) {
        if (fill_index(c, ROVER_METHOD_SQ, 0, 0, &ix) || !c->d_sq_res) {
            snprintf(g_err, sizeof(g_err), "decode: no SQ index with residuals loaded");
            goto out;
        }
        ix.res_codes = c->d_sq_res;
        ix.res_qparams = c->d_sq_res_qp;
        ix.res_bits = c->sq_res_bits;
        ix.res_code_bytes = c->sq_res_bits == 4 ? (c->sq_dim + 1) / 2 : c->sq_dim;
    } else if (fill_index(c, ROVER_METHOD_PQ, 1, 0, &ix)) {
        snprintf(g_err, sizeof(g_err), "decode: no PQ index with residuals loaded");
        goto out;
    }
    n = ix.n;
    dim = ix.dim;
    for (int i = 0; i < count; i++)
        if (ids[i] < 0 || ids[i] >= n) {
            rc = ROVER_EBADPARM;
            goto out;
        }
    if (count == 0) {
        rc = ROVER_SUCCESS;
        goto out;
    }
    if (rv_malloc((void **)&d_ids, (size_t)count * 4) ||
        rv_malloc((void **)&d_out, (size_t)count * dim * 4) ||
        rv_memcpy_h2d(d_ids, ids, (size_t)count * 4) ||
        rv_decode_vectors(&ix, sq, d_ids, count, d_out, NULL) || rv_synchronize() ||
        rv_memcpy_d2h(out, d_out, (size_t)count * dim * 4)) {
        set_err("decode");
        goto out;
    }
    rc = ROVER_SUCCESS;
out:
    pthread_mutex_unlock(&c->lock);
    dfree(d_ids);
    dfree(d_out);
    return rc;
}

AEEResult rover_pure_pq_decode(remote_handle64 h, const int32_t *ids, int count,
                               float *out) {
    return decode_ids(h, 0, ids, count, out);
}

AEEResult rover_sq_decode(remote_handle64 h, const int32_t *ids, int count, float *out) {
    return decode_ids(h, 1, ids, count, out);
}

/* ------------------------------------------------------------ traces */

#define ROVER_TRACE_MAGIC 0x52565452u /* "RVTR" */

typedef struct RoverTrace {
    uint32_t magic;
    FILE *f;
    int format;
} RoverTrace;

AEEResult rover_trace_open(const char *path, int format, void **trace) {
    RoverTrace *t;
    if (!path || !trace || (format != ROVER_TRACE_TEXT && format != ROVER_TRACE_P15))
        return ROVER_EBADPARM;
    *trace = NULL;
    t = calloc(1, sizeof(*t));
    if (!t) return ROVER_ENOMEMORY;
    /* the text trace is opened "w" (rpc_demo_pure_pq.c:535-538), the binary
       one "wbx" and starts with its magic (pipeline.c:3047-3050) */
    t->f = fopen(path, format == ROVER_TRACE_TEXT ? "w" : "wbx");
    if (!t->f || (format == ROVER_TRACE_P15 && fwrite("P15TRC01", 1, 8, t->f) != 8)) {
        if (t->f) fclose(t->f);
        free(t);
        snprintf(g_err, sizeof(g_err), "trace: cannot create %s", path);
        return ROVER_EFAILED;
    }
    t->magic = ROVER_TRACE_MAGIC;
    t->format = format;
    *trace = t;
    return ROVER_SUCCESS;
}

AEEResult rover_trace_query(void *trace, int entry_points, int pool_size, int query_index,
                            int refine_budget, const int32_t *ids, const float *dists,
                            int count, const float *decoded, int decoded_count, int dim,
                            const int32_t *out_ids, const float *out_dists, int out_count) {
    RoverTrace *t = trace;
    int ok = 1;
    if (!t || t->magic != ROVER_TRACE_MAGIC || count < 0 || out_count < 0 ||
        (count > 0 && (!ids || !dists)) || (out_count > 0 && (!out_ids || !out_dists)))
        return ROVER_EBADPARM;
    if (t->format == ROVER_TRACE_TEXT) {
        /* Q pool query count budget; C id dist per search result; V id and
           the decoded vector for the first min(count, budget or top-k)
           results; O id dist per output (rpc_demo_pure_pq.c:615-645) */
        if (decoded_count < 0 || (decoded_count > 0 && (!decoded || dim <= 0)) ||
            decoded_count > count)
            return ROVER_EBADPARM;
        ok = fprintf(t->f, "Q %d %d %d %d\n", pool_size, query_index, count, refine_budget) > 0;
        for (int i = 0; ok && i < count; i++)
            ok = fprintf(t->f, "C %d %a\n", ids[i], (double)dists[i]) > 0;
        for (int i = 0; ok && i < decoded_count; i++) {
            ok = fprintf(t->f, "V %d", ids[i]) > 0;
            for (int d = 0; ok && d < dim; d++)
                ok = fprintf(t->f, " %a", (double)decoded[(size_t)i * dim + d]) > 0;
            ok = ok && fputc('\n', t->f) != EOF;
        }
        for (int i = 0; ok && i < out_count; i++)
            ok = fprintf(t->f, "O %d %a\n", out_ids[i], (double)out_dists[i]) > 0;
    } else {
        /* u32 (entry points, pool, query, count), ids, distances; then
           u32 n and the refined ids and distances (pipeline.c:3276-3308) */
        const uint32_t hdr[4] = {(uint32_t)entry_points, (uint32_t)pool_size,
                                 (uint32_t)query_index, (uint32_t)count};
        const uint32_t n = (uint32_t)out_count;
        ok = fwrite(hdr, sizeof(hdr), 1, t->f) == 1 &&
             fwrite(ids, 4, (size_t)count, t->f) == (size_t)count &&
             fwrite(dists, 4, (size_t)count, t->f) == (size_t)count &&
             fwrite(&n, 4, 1, t->f) == 1 &&
             fwrite(out_ids, 4, (size_t)n, t->f) == (size_t)n &&
             fwrite(out_dists, 4, (size_t)n, t->f) == (size_t)n;
    }
    if (!ok || ferror(t->f)) {
        snprintf(g_err, sizeof(g_err), "trace: write failed");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_trace_close(void *trace) {
    RoverTrace *t = trace;
    int rc;
    if (!t || t->magic != ROVER_TRACE_MAGIC) return ROVER_EBADPARM;
    rc = fclose(t->f);
    t->magic = 0;
    free(t);
    return rc ? ROVER_EFAILED : ROVER_SUCCESS;
}

AEEResult rover_host_alloc(void **ptr, size_t bytes) {
    if (!ptr) return ROVER_EBADPARM;
    *ptr = NULL;
    if (rv_host_alloc(ptr, bytes)) {
        set_err("rover_host_alloc");
        return ROVER_ENOMEMORY;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_host_free(void *ptr) {
    if (ptr && rv_host_free(ptr)) {
        set_err("rover_host_free");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_device_free(void *d_ptr) {
    return rv_free(d_ptr) ? ROVER_EFAILED : ROVER_SUCCESS;
}
AEEResult rover_device_upload(void *d_dst, const void *src, size_t bytes) {
    if (rv_memcpy_h2d(d_dst, src, bytes)) {
        set_err("rover_device_upload");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}
AEEResult rover_device_download(void *dst, const void *d_src, size_t bytes) {
    if (rv_memcpy_d2h(dst, d_src, bytes)) {
        set_err("rover_device_download");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}
AEEResult rover_device_synchronize(void) {
    if (rv_synchronize()) {
        set_err("rover_device_synchronize");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_build_pq_encode(const float *d_x, int n, int dim, int m, int ks,
                                const float *d_codebook, uint8_t *d_codes) {
    if (!d_x || !d_codebook || !d_codes || n <= 0 || m <= 0 || dim % m)
        return ROVER_EBADPARM;
    if (rv_pq_encode(d_x, n, dim, m, ks, d_codebook, d_codes, NULL)) {
        set_err("rover_build_pq_encode");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_build_error4_encode(const float *d_x, const uint8_t *d_codes,
                                    const uint8_t *d_res,
                                    const float *d_res_qparams, int n, int dim,
                                    int m, int ks, const float *d_codebook,
                                    int res_bits, uint8_t *d_e4_codes,
                                    float *d_e4_qparams) {
    if (!d_x || !d_codes || !d_res || !d_res_qparams || !d_codebook ||
        !d_e4_codes || !d_e4_qparams || n <= 0 || dim <= 0 || dim > 8192 ||
        m <= 0 || dim % m || (ks != 16 && ks != 256) ||
        (res_bits != 4 && res_bits != 8))
        return ROVER_EBADPARM;
    if (rv_error4_encode(d_x, d_codes, d_res, d_res_qparams, n, dim, m, ks,
                         d_codebook, res_bits, d_e4_codes, d_e4_qparams, NULL) ||
        rv_synchronize()) {
        set_err("rover_build_error4_encode");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_build_graph_exact(const float *d_x, int n, int dim, int r,
                                  int candidates, uint32_t seed, int prune_mode,
                                  int32_t *d_graph) {
    if (!d_x || !d_graph || n <= 1 || dim <= 0 || dim > 4096 || r <= 0 || r >= n ||
        r > 64 || candidates < r || candidates > 128 || (prune_mode != 0 && prune_mode != 1))
        return ROVER_EBADPARM;
    if (rv_build_exact(d_x, n, dim, r, candidates, seed, prune_mode, d_graph, NULL)) {
        set_err("rover_build_graph_exact");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_build_graph_exact_clustered(const float *d_x, int n, int dim, int r,
                                            int candidates, uint32_t seed, int prune_mode,
                                            const int32_t *d_assign, const float *d_centroids,
                                            int num_clusters, int32_t *d_graph,
                                            int64_t *stats) {
    if (!d_x || !d_graph || !d_assign || !d_centroids || n <= 1 || dim <= 0 || dim > 4096 ||
        r <= 0 || r >= n || r > 64 || candidates < r || candidates > 128 || candidates >= n ||
        num_clusters <= 0 || num_clusters > ROVER_BUILD_MAX_CLUSTERS ||
        (prune_mode != 0 && prune_mode != 1))
        return ROVER_EBADPARM;
    if (rv_build_exact_clustered(d_x, n, dim, r, candidates, seed, prune_mode, d_assign,
                                 d_centroids, num_clusters, d_graph, stats, NULL)) {
        set_err("rover_build_graph_exact_clustered");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_build_graph_rnn(const float *d_x, int n, int dim, int s, int r, int t1,
                                int t2, uint32_t seed, int top_k, int32_t *d_graph) {
    if (!d_x || !d_graph || n <= 1 || dim <= 0 || s <= 0 || r <= 0 || r > 64 || s > r ||
        t1 <= 0 || t2 <= 0 || top_k <= 0 || top_k > r || (int64_t)n * r > 0x7fffffff)
        return ROVER_EBADPARM;
    if (rv_build_rnn(d_x, n, dim, s, r, t1, t2, seed, top_k, d_graph, NULL)) {
        set_err("rover_build_graph_rnn");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_build_residual_encode(const float *d_x, const uint8_t *d_codes,
                                      int n, int dim, int m, int ks,
                                      const float *d_codebook, int bits,
                                      uint8_t *d_res, float *d_qparams) {
    if (!d_x || !d_codes || !d_codebook || !d_res || !d_qparams || n <= 0 ||
        m <= 0 || dim % m || (bits != 4 && bits != 8))
        return ROVER_EBADPARM;
    if (rv_residual_encode(d_x, d_codes, n, dim, m, ks, d_codebook, bits, d_res,
                           d_qparams, NULL)) {
        set_err("rover_build_residual_encode");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_event_create(void **event) {
    if (!event || rv_event_create(event)) {
        set_err("rover_event_create");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}
AEEResult rover_event_record(void *event, void *stream) {
    if (rv_event_record(event, stream)) {
        set_err("rover_event_record");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}
AEEResult rover_event_elapsed_ms(void *start, void *stop, float *ms) {
    if (!ms || rv_event_elapsed(start, stop, ms)) {
        set_err("rover_event_elapsed_ms");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}
AEEResult rover_event_destroy(void *event) {
    return rv_event_destroy(event) ? ROVER_EFAILED : ROVER_SUCCESS;
}
AEEResult rover_stream_create(void **stream) {
    if (!stream || rv_stream_create(stream)) {
        set_err("rover_stream_create");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}
AEEResult rover_stream_destroy(void *stream) {
    return rv_stream_destroy(stream) ? ROVER_EFAILED : ROVER_SUCCESS;
}
AEEResult rover_stream_synchronize(void *stream) {
    if (rv_stream_synchronize(stream)) {
        set_err("rover_stream_synchronize");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}

AEEResult rover_device_copy(void *d_dst, const void *d_src, size_t bytes) {
    if (!d_dst || !d_src) return ROVER_EBADPARM;
    if (rv_copy(d_dst, d_src, bytes, NULL)) {
        set_err("rover_device_copy");
        return ROVER_EFAILED;
    }
    return ROVER_SUCCESS;
}
