/*
 * rover_oracle.c -- CPU restatement of Rover's scalar quantized graph search.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Compile with
 * -ffp-contract=off: every float operation below is rounded separately, in
 * the order the reference performs it.  Citations are to the Rover source
 * tree (src/dsp_demo/... unless stated).
 */
#define _GNU_SOURCE
#include "rover_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int id;
    float distance;
} OrcCand;

/* retrieval_pq_core.c:46-52 (retrieval_pq_code_bytes) */
int orc_code_bytes(int m, int ks) {
    if (m <= 0 || ks <= 0 || ks > 256) return -1;
    return ks <= 16 ? m / 2 + m % 2 : m;
}

/* retrieval_pq_residual.c:173-175 (pq_residual_code_bytes) */
int orc_residual_code_bytes(int dim, int bits) {
    return bits == 4 ? dim / 2 + dim % 2 : dim;
}

/* retrieval_pq_core.c:54-64 (retrieval_pq_code_get): low nibble = even m */
static int orc_code_get(const uint8_t *code, int subspace, int ks) {
    if (ks <= 16) {
        uint8_t packed = code[subspace >> 1];
        return (subspace & 1) ? (packed >> 4) : (packed & 0x0f);
    }
    return code[subspace];
}

/* retrieval_pq_core.c:16-25 (pq_sub_l2): sequential f32 sum of squares */
static float orc_sub_l2(const float *a, const float *b, int subdim) {
    float acc = 0.0f;
    for (int d = 0; d < subdim; d++) {
        float diff = a[d] - b[d];
        acc += diff * diff;
    }
    return acc;
}

/* retrieval_pq_core.c:362-390 (retrieval_pq_build_lut) */
int orc_build_lut(const float *query, const float *codebook, int m, int ks,
                  int subdim, float *lut) {
    if (!query || !codebook || !lut || m <= 0 || ks <= 0 || ks > 256 ||
        subdim <= 0)
        return -1;
    for (int mi = 0; mi < m; mi++) {
        const float *q = query + (size_t)mi * subdim;
        for (int k = 0; k < ks; k++) {
            const float *c = codebook + ((size_t)mi * ks + k) * subdim;
            float dist = orc_sub_l2(q, c, subdim);
            if (!isfinite(dist)) return -1;
            lut[(size_t)mi * ks + k] = dist;
        }
    }
    return 0;
}

/* retrieval_pq_core.c:392-416 (retrieval_pq_adc_scalar): sequential m order,
 * FLT_MAX on any non-finite value. */
float orc_adc(const uint8_t *code, const float *lut, int m, int ks) {
    float acc = 0.0f;
    for (int mi = 0; mi < m; mi++) {
        int c = orc_code_get(code, mi, ks);
        if (c >= ks) return FLT_MAX;
        float v = lut[(size_t)mi * ks + c];
        if (!isfinite(v)) return FLT_MAX;
        acc += v;
        if (!isfinite(acc)) return FLT_MAX;
    }
    return acc;
}

/* retrieval_pq_residual.c:314-385 (pq_residual_quantize_vector): per-vector
 * min/max of (x - centroid), scale=(max-min)/levels (floored at 1e-12),
 * code=clamp(roundf((r-zero)*(1/scale))). */
int orc_residual_encode(const float *vector, const uint8_t *pq_code,
                        const float *codebook, int m, int ks, int subdim,
                        int bits, uint8_t *out_code, float *out_qparams) {
    float minimum = FLT_MAX, maximum = -FLT_MAX, scale, inv;
    int levels = bits == 4 ? 15 : 255;
    int dim = m * subdim;
    int d = 0;
    if (bits != 4 && bits != 8) return -1;
    for (int s = 0; s < m; s++) {
        const float *c = codebook +
            ((size_t)s * ks + orc_code_get(pq_code, s, ks)) * subdim;
        for (int i = 0; i < subdim; i++, d++) {
            float r = vector[d] - c[i];
            if (!isfinite(r)) return -1;
            if (r < minimum) minimum = r;
            if (r > maximum) maximum = r;
        }
    }
    scale = (maximum - minimum) / (float)levels;
    if (scale < 1.0e-12f) scale = 1.0e-12f;
    inv = 1.0f / scale;
    memset(out_code, 0, (size_t)orc_residual_code_bytes(dim, bits));
    d = 0;
    for (int s = 0; s < m; s++) {
        const float *c = codebook +
            ((size_t)s * ks + orc_code_get(pq_code, s, ks)) * subdim;
        for (int i = 0; i < subdim; i++, d++) {
            float r = vector[d] - c[i];
            int q = (int)roundf((r - minimum) * inv);
            if (q < 0) q = 0;
            if (q > levels) q = levels;
            if (bits == 4) {
                if (d & 1) out_code[d >> 1] |= (uint8_t)(q << 4);
                else out_code[d >> 1] = (uint8_t)q;
            } else {
                out_code[d] = (uint8_t)q;
            }
        }
    }
    out_qparams[0] = scale;
    out_qparams[1] = minimum;
    return 0;
}

/* retrieval_pq_residual.c:946-952 (pq_residual_decode_value): three
 * separately rounded f32 operations, ((c + code*scale) + zero). */
static float orc_decode_value(float centroid, int code, float scale,
                              float zero) {
    volatile float product = (float)code * scale;
    volatile float sum = centroid + product;
    volatile float value = sum + zero;
    return value;
}

static int orc_res_code(const uint8_t *res, int bits, int d) {
    if (bits == 4) {
        uint8_t packed = res[d >> 1];
        return (d & 1) ? (packed >> 4) : (packed & 0x0f);
    }
    return res[d];
}

/* retrieval_pq_residual.c:983-1042 (retrieval_pq_residual_query_evaluate) */
int orc_residual_evaluate(const OrcIndex *ix, const float *query,
                          int vector_id, float *out_distance) {
    const uint8_t *pq, *res;
    const float *qp;
    float distance = 0.0f;
    int d = 0;
    if (!ix || !query || !out_distance || vector_id < 0 ||
        vector_id >= ix->num_vectors || !ix->res_codes || !ix->res_qparams)
        return -1;
    pq = ix->codes + (size_t)vector_id * orc_code_bytes(ix->m, ix->ks);
    res = ix->res_codes +
          (size_t)vector_id * orc_residual_code_bytes(ix->dim, ix->res_bits);
    qp = ix->res_qparams + 2 * (size_t)vector_id;
    if (!isfinite(qp[0]) || !isfinite(qp[1]) || qp[0] <= 0.0f) return -1;
    for (int s = 0; s < ix->m; s++) {
        const float *c = ix->codebook +
            ((size_t)s * ix->ks + orc_code_get(pq, s, ix->ks)) * ix->subdim;
        for (int i = 0; i < ix->subdim; i++, d++) {
            float rec = orc_decode_value(c[i], orc_res_code(res, ix->res_bits, d),
                                         qp[0], qp[1]);
            float diff = query[d] - rec;
            distance += diff * diff;
        }
    }
    if (!isfinite(distance)) return -1;
    *out_distance = distance;
    return 0;
}

/* retrieval_pq_residual.c:954-981 (retrieval_pq_residual_query_decode) */
int orc_residual_decode(const OrcIndex *ix, int vector_id, float *out) {
    const uint8_t *pq, *res;
    const float *qp;
    int d = 0;
    if (!ix || !out || vector_id < 0 || vector_id >= ix->num_vectors ||
        !ix->res_codes || !ix->res_qparams)
        return -1;
    qp = ix->res_qparams + 2 * (size_t)vector_id;
    if (!isfinite(qp[0]) || qp[0] <= 0 || !isfinite(qp[1])) return -1;
    pq = ix->codes + (size_t)vector_id * orc_code_bytes(ix->m, ix->ks);
    res = ix->res_codes +
          (size_t)vector_id * orc_residual_code_bytes(ix->dim, ix->res_bits);
    for (int s = 0; s < ix->m; s++) {
        const float *c = ix->codebook +
            ((size_t)s * ix->ks + orc_code_get(pq, s, ix->ks)) * ix->subdim;
        for (int i = 0; i < ix->subdim; i++, d++) {
            out[d] = orc_decode_value(c[i], orc_res_code(res, ix->res_bits, d),
                                      qp[0], qp[1]);
            if (!isfinite(out[d])) return -1;
        }
    }
    return 0;
}

/* retrieval_build_artifacts.c:25-66 (retrieval_build_error4), one row; the
 * published generator pq_refine_artifacts.py:47-65 with
 * generate_final_refine.py:36-44 (quantize_error) performs the same float32
 * steps: e = x - recon; lo = min e, hi = max e (first strict extreme wins);
 * scale = max(hi - lo, 1e-12) / 15; inv = 1 / scale;
 * code = clamp(floor((e - lo) * inv + 0.5), 0, 15), low nibble = even d. */
int orc_error4_encode(const float *x, const float *recon, int dim,
                      uint8_t *out_code, float *out_qparams) {
    float *e = malloc((size_t)dim * sizeof(float));
    float lo = INFINITY, hi = -INFINITY, scale, inv, span;
    if (!e) return -1;
    for (int d = 0; d < dim; d++) {
        volatile float diff = x[d] - recon[d];
        e[d] = diff;
        if (!isfinite(e[d])) {
            free(e);
            return -1;
        }
        if (e[d] < lo) lo = e[d];
        if (e[d] > hi) hi = e[d];
    }
    {
        volatile float sp = hi - lo;
        span = sp;
    }
    scale = fmaxf(span, 1e-12f) / 15.0f;
    inv = 1.0f / scale;
    memset(out_code, 0, (size_t)(dim + 1) / 2);
    for (int d = 0; d < dim; d++) {
        volatile float shifted = e[d] - lo;
        volatile float scaled = shifted * inv;
        volatile float v = scaled + 0.5f;
        int code = (int)floorf(v);
        if (code < 0) code = 0;
        if (code > 15) code = 15;
        out_code[d / 2] |= (uint8_t)(code << ((d % 2) * 4));
    }
    out_qparams[0] = scale;
    out_qparams[1] = lo;
    free(e);
    return 0;
}

/* retrieval_rpc_demo_support.c:573-627 (demo_set_qparams_from_range and
 * demo_compute_range): one global scale/zero over all values. */
int orc_sq_range_params(const float *data, long count, int bits,
                        float *out_scale, float *out_zero) {
    float lo = FLT_MAX, hi = -FLT_MAX, range;
    if (!data || count <= 0 || (bits != 4 && bits != 8)) return -1;
    for (long i = 0; i < count; i++) {
        if (data[i] < lo) lo = data[i];
        if (data[i] > hi) hi = data[i];
    }
    range = hi - lo;
    if (range < 1e-12f) range = 1e-12f;
    *out_scale = range / (float)((1 << bits) - 1);
    *out_zero = lo;
    return 0;
}

/* retrieval_rpc_demo_support.c:664-700 (demo_quantize_l1_vector): 4-bit
 * codes packed low nibble first. */
int orc_sq_encode(const float *v, int dim, int bits, float scale, float zero,
                  uint8_t *out) {
    const float inv = 1.0f / scale;
    const int levels = (1 << bits) - 1;
    if (bits != 4 && bits != 8) return -1;
    memset(out, 0, (size_t)(bits == 4 ? (dim + 1) / 2 : dim));
    for (int i = 0; i < dim; i++) {
        int q = (int)roundf((v[i] - zero) * inv);
        if (q < 0) q = 0;
        else if (q > levels) q = levels;
        if (bits == 4) {
            if (i & 1) out[i >> 1] |= (uint8_t)(q << 4);
            else out[i >> 1] = (uint8_t)q;
        } else {
            out[i] = (uint8_t)q;
        }
    }
    return 0;
}

/* retrieval_l1_int8_core.c:13-60 (retrieval_quantize_query_u8/_u4) */
int orc_sq_quantize_query(const float *query, int dim, int bits, float scale,
                          float zero, uint8_t *out) {
    int hi = bits == 4 ? 15 : 255;
    float inv;
    if (!query || !out || dim <= 0 || scale == 0.0f) return -1;
    inv = 1.0f / scale;
    for (int i = 0; i < dim; i++) {
        int v = (int)roundf((query[i] - zero) * inv);
        if (v < 0) v = 0;
        else if (v > hi) v = hi;
        out[i] = (uint8_t)v;
    }
    return 0;
}

/* retrieval_l1_int8_core.c:119-165: integer sum of squared code
 * differences, then (float)acc * (scale*scale) with scale*scale in f32. */
float orc_sq_distance(const uint8_t *q, const uint8_t *code, int dim, int bits,
                      float scale) {
    int64_t acc = 0;
    float scale_sq = scale * scale;
    for (int i = 0; i < dim; i++) {
        int c = bits == 4 ? ((i & 1) ? (code[i >> 1] >> 4) : (code[i >> 1] & 15))
                          : code[i];
        int qv = bits == 4 ? (q[i] & 15) : q[i];
        int diff = qv - c;
        acc += (int64_t)(diff * diff);
    }
    return (float)acc * scale_sq;
}

/* retrieval_rpc_demo_pure_pq.c:192-208 (pure_pq_make_seed_ids) */
int orc_seed_ids(int num_vectors, int entry_count, int graph_entry,
                 uint32_t seed, int *out) {
    int count = entry_count < num_vectors ? entry_count : num_vectors;
    for (int i = 0; i < count; i++) {
        if (i == 0 && graph_entry >= 0 && graph_entry < num_vectors) {
            out[i] = graph_entry;
        } else {
            seed = seed * 1103515245u + 12345u;
            out[i] = (int)(((uint64_t)seed * (uint64_t)num_vectors) >> 32);
        }
    }
    return count;
}

/* retrieval_sqmd_pure_pq.c:97-112 (pure_pq_order): distance, then id */
static int orc_order(float ld, int lid, float rd, int rid) {
    if (ld < rd) return -1;
    if (ld > rd) return 1;
    if (lid < rid) return -1;
    if (lid > rid) return 1;
    return 0;
}

static int orc_cand_cmp(const void *a, const void *b) {
    const OrcCand *x = a, *y = b;
    return orc_order(x->distance, x->id, y->distance, y->id);
}

/* retrieval_sqmd_pure_pq.c:121-176: bounded max-heap of the best L */
static void orc_pool_sift_up(OrcCand *pool, int pos) {
    OrcCand c = pool[pos];
    while (pos > 0) {
        int parent = (pos - 1) / 2;
        if (orc_order(c.distance, c.id, pool[parent].distance,
                      pool[parent].id) <= 0)
            break;
        pool[pos] = pool[parent];
        pos = parent;
    }
    pool[pos] = c;
}

static void orc_pool_sift_down(OrcCand *pool, int size, int pos) {
    OrcCand c = pool[pos];
    for (;;) {
        int child = 2 * pos + 1;
        if (child >= size) break;
        if (child + 1 < size &&
            orc_order(pool[child + 1].distance, pool[child + 1].id,
                      pool[child].distance, pool[child].id) > 0)
            child++;
        if (orc_order(c.distance, c.id, pool[child].distance,
                      pool[child].id) >= 0)
            break;
        pool[pos] = pool[child];
        pos = child;
    }
    pool[pos] = c;
}

static int orc_pool_insert(OrcCand *pool, int *size, int cap, int id,
                           float dist) {
    if (*size < cap) {
        pool[*size].id = id;
        pool[*size].distance = dist;
        orc_pool_sift_up(pool, *size);
        (*size)++;
        return 1;
    }
    if (orc_order(dist, id, pool[0].distance, pool[0].id) < 0) {
        pool[0].id = id;
        pool[0].distance = dist;
        orc_pool_sift_down(pool, *size, 0);
        return 1;
    }
    return 0;
}

/* retrieval_sqmd_pure_pq.c:178-238: unbounded min-heap frontier */
static void orc_front_push(OrcCand *f, int *size, int id, float dist) {
    int pos = (*size)++;
    OrcCand c = {id, dist};
    while (pos > 0) {
        int parent = (pos - 1) / 2;
        if (orc_order(c.distance, c.id, f[parent].distance, f[parent].id) >= 0)
            break;
        f[pos] = f[parent];
        pos = parent;
    }
    f[pos] = c;
}

static OrcCand orc_front_pop(OrcCand *f, int *size) {
    OrcCand top = f[0];
    (*size)--;
    if (*size > 0) {
        OrcCand c = f[*size];
        int pos = 0;
        for (;;) {
            int child = 2 * pos + 1;
            if (child >= *size) break;
            if (child + 1 < *size &&
                orc_order(f[child + 1].distance, f[child + 1].id,
                          f[child].distance, f[child].id) < 0)
                child++;
            if (orc_order(c.distance, c.id, f[child].distance,
                          f[child].id) <= 0)
                break;
            f[pos] = f[child];
            pos = child;
        }
        f[pos] = c;
    }
    return top;
}

/* retrieval_sqmd_pure_pq.c:240-249 (pure_pq_pool_contains) */
static int orc_pool_contains(const OrcCand *pool, int size, int id,
                             float dist) {
    for (int i = 0; i < size; i++)
        if (pool[i].id == id && pool[i].distance == dist) return 1;
    return 0;
}

/* retrieval_sqmd_pure_pq.c:59-87 (pure_pq_graph_row_count): -1 padding only
 * as a suffix; any other out-of-range id makes the row invalid. */
static int orc_row_count(const int32_t *row, int slots, int n) {
    for (int i = 0; i < slots; i++) {
        int id = row[i];
        if (id == -1) {
            for (int j = i + 1; j < slots; j++)
                if (row[j] != -1) return -1;
            return i;
        }
        if (id < 0 || id >= n) return -1;
    }
    return slots;
}

typedef struct {
    const OrcIndex *ix;
    int scorer;
    const float *query;
    const float *lut;
    uint8_t qsq[4096];
} OrcScorer;

/* One candidate's final distance: ADC (retrieval_pq_core.c:280-330, FLT_MAX
 * or non-finite is an error), then the distance policy
 * (retrieval_sqmd_pure_pq.c:40-57).  For SQ the policy evaluator returns the
 * SQ distance and the coarse ADC is discarded. */
static int orc_score(const OrcScorer *s, int id, float *out) {
    const OrcIndex *ix = s->ix;
    if (s->scorer == ORC_SCORE_SQ) {
        int cb = ix->sq_bits == 4 ? (ix->dim + 1) / 2 : ix->dim;
        float d = orc_sq_distance(s->qsq, ix->sq_codes + (size_t)id * cb,
                                  ix->dim, ix->sq_bits, ix->sq_scale);
        if (!isfinite(d)) return -1;
        *out = d;
        return 0;
    }
    float d = orc_adc(ix->codes + (size_t)id * orc_code_bytes(ix->m, ix->ks),
                      s->lut, ix->m, ix->ks);
    if (!isfinite(d) || d == FLT_MAX) return -1;
    if (s->scorer == ORC_SCORE_ADC_ONLINE) {
        if (orc_residual_evaluate(ix, s->query, id, &d) != 0 || !isfinite(d))
            return -1;
    }
    *out = d;
    return 0;
}

/* retrieval_sqmd_pure_pq.c:793-969
 * (retrieval_pure_pq_search_scalar_with_policy) */
int orc_search(const OrcIndex *ix, int scorer, const float *query,
               const int *entry_ids, int entry_count, int expand_count,
               int pool_capacity, int num_out, int *out_ids,
               float *out_distances, int out_capacity, OrcStats *out_stats) {
    OrcCand *pool = NULL, *front = NULL;
    uint32_t *visited = NULL;
    float *lut = NULL;
    int *batch = NULL;
    int pool_size = 0, front_size = 0, result = -1;
    OrcStats st;
    OrcScorer sc;
    memset(&st, 0, sizeof(st));
    if (out_stats) memset(out_stats, 0, sizeof(*out_stats));
    if (!ix || !query || !entry_ids || entry_count <= 0 || expand_count <= 0 ||
        pool_capacity <= 0 || num_out <= 0 || num_out > pool_capacity ||
        !out_ids || !out_distances || out_capacity <= 0 || !ix->graph ||
        ix->degree <= 0 || ix->num_vectors <= 0)
        return -1;
    if (scorer != ORC_SCORE_SQ &&
        (!ix->codes || !ix->codebook || ix->m <= 0 ||
         (ix->ks != 16 && ix->ks != 256) || ix->subdim <= 0 ||
         ix->m * ix->subdim != ix->dim))
        return -1;
    if (scorer == ORC_SCORE_ADC_ONLINE && (!ix->res_codes || !ix->res_qparams))
        return -1;
    if (scorer == ORC_SCORE_SQ &&
        (!ix->sq_codes || ix->dim > 4096 || (ix->sq_bits != 4 && ix->sq_bits != 8)))
        return -1;
    if (expand_count > ix->degree) expand_count = ix->degree;
    memset(&sc, 0, sizeof(sc));
    sc.ix = ix;
    sc.scorer = scorer;
    sc.query = query;
    pool = malloc((size_t)pool_capacity * sizeof(*pool));
    front = malloc((size_t)ix->num_vectors * sizeof(*front));
    visited = calloc(((size_t)ix->num_vectors + 31) / 32, sizeof(*visited));
    batch = malloc((size_t)expand_count * sizeof(*batch));
    if (!pool || !front || !visited || !batch) goto done;
    if (scorer == ORC_SCORE_SQ) {
        if (orc_sq_quantize_query(query, ix->dim, ix->sq_bits, ix->sq_scale,
                                  ix->sq_zero, sc.qsq) != 0)
            goto done;
    } else {
        lut = malloc((size_t)ix->m * ix->ks * sizeof(*lut));
        if (!lut || orc_build_lut(query, ix->codebook, ix->m, ix->ks,
                                  ix->subdim, lut) != 0)
            goto done;
        sc.lut = lut;
    }

    for (int e = 0; e < entry_count; e++) {
        int id = entry_ids[e];
        float dist;
        if (id < 0 || id >= ix->num_vectors) goto done;
        if (visited[id >> 5] & (1u << (id & 31))) {
            st.visited_dedup_skips++;
            continue;
        }
        visited[id >> 5] |= 1u << (id & 31);
        if (orc_score(&sc, id, &dist) != 0) goto done;
        st.pq_scored_candidates++;
        if (scorer != ORC_SCORE_ADC) st.corrected_candidates++;
        if (orc_pool_insert(pool, &pool_size, pool_capacity, id, dist))
            orc_front_push(front, &front_size, id, dist);
    }

    while (front_size > 0) {
        OrcCand src;
        const int32_t *row;
        int nb = 0, valid;
        if (pool_size == pool_capacity &&
            orc_order(front[0].distance, front[0].id, pool[0].distance,
                      pool[0].id) > 0)
            break;
        src = orc_front_pop(front, &front_size);
        if (!orc_pool_contains(pool, pool_size, src.id, src.distance)) continue;
        row = ix->graph + (size_t)src.id * ix->degree;
        valid = orc_row_count(row, expand_count, ix->num_vectors);
        if (valid < 0) goto done;
        for (int j = 0; j < valid; j++) {
            int id = row[j];
            if (visited[id >> 5] & (1u << (id & 31))) {
                st.visited_dedup_skips++;
                continue;
            }
            visited[id >> 5] |= 1u << (id & 31);
            batch[nb++] = id;
        }
        st.expanded_nodes++;
        if (nb == 0) continue;
        /* scores of the whole batch are computed before any insertion */
        float dists[256];
        float *bd = nb <= 256 ? dists : malloc((size_t)nb * sizeof(float));
        if (!bd) goto done;
        for (int j = 0; j < nb; j++) {
            if (orc_score(&sc, batch[j], &bd[j]) != 0) {
                if (bd != dists) free(bd);
                goto done;
            }
        }
        st.pq_scored_candidates += (uint64_t)nb;
        for (int j = 0; j < nb; j++) {
            if (scorer != ORC_SCORE_ADC) st.corrected_candidates++;
            if (orc_pool_insert(pool, &pool_size, pool_capacity, batch[j], bd[j]))
                orc_front_push(front, &front_size, batch[j], bd[j]);
        }
        if (bd != dists) free(bd);
    }

    qsort(pool, (size_t)pool_size, sizeof(*pool), orc_cand_cmp);
    result = num_out < pool_size ? num_out : pool_size;
    if (result > out_capacity) result = out_capacity;
    for (int i = 0; i < result; i++) {
        out_ids[i] = pool[i].id;
        out_distances[i] = pool[i].distance;
    }
    if (out_stats) *out_stats = st;
done:
    free(pool);
    free(front);
    free(visited);
    free(lut);
    free(batch);
    return result;
}

/* The live pipeline's source window (window 1 or 2 sources in flight):
 * pure_pq_source_pipeline_step (retrieval_sqmd_pure_pq.c:722-762) over
 * pure_pq_cpu_search_dispatch (:495-594) and pure_pq_cpu_search_collect
 * (:596-706), seeded by pure_pq_cpu_search_seed (:442-472), as the demo
 * drives one query (rpc_demo_pure_pq.c:283-321).  Each step first
 * dispatches until `window` turns are pending -- pop the frontier to the
 * best source still in the pool (stopping, without popping, when the pool
 * is full and the frontier top is worse than its worst), mark the source's
 * unvisited neighbours visited and queue them as the turn's work -- and then
 * collects the oldest turn: its work is scored, sorted by (distance, id)
 * and inserted.  The search ends when nothing is pending and nothing can be
 * dispatched.  With window 1 every turn is collected before the next
 * dispatch, which is orc_search's order.  The reference's work items are
 * scored by ADC on the DSP and then passed through the distance policy on
 * collect; the scorer does both here.  The DSP lane buffers bound a turn's
 * work at 32 items per lane; this restatement has no lanes and no bound.
 * Stats: expanded_nodes counts dispatched turns, pq_scored_candidates and
 * corrected_candidates the seeds and work items scored (corrected only with
 * a correcting scorer), visited_dedup_skips the seeds and neighbours found
 * visited.  *high_water receives the most turns pending at once. */
int orc_search_window(const OrcIndex *ix, int scorer, const float *query,
                      const int *entry_ids, int entry_count, int expand_count,
                      int pool_capacity, int num_out, int window, int *out_ids,
                      float *out_distances, int out_capacity, OrcStats *out_stats,
                      int *high_water) {
    OrcCand *pool = NULL, *front = NULL;
    uint32_t *visited = NULL;
    float *lut = NULL;
    int *turn_ids[2] = {NULL, NULL};
    int turn_n[2] = {0, 0};
    OrcCand *items = NULL;
    int pool_size = 0, front_size = 0, result = -1;
    int head = 0, pending = 0, hw = 0;
    OrcStats st;
    OrcScorer sc;
    memset(&st, 0, sizeof(st));
    if (out_stats) memset(out_stats, 0, sizeof(*out_stats));
    if (high_water) *high_water = 0;
    if (!ix || !query || !entry_ids || entry_count <= 0 || expand_count <= 0 ||
        pool_capacity <= 0 || num_out <= 0 || num_out > pool_capacity ||
        window < 1 || window > 2 || !out_ids || !out_distances ||
        out_capacity <= 0 || !ix->graph || ix->degree <= 0 || ix->num_vectors <= 0)
        return -1;
    if (scorer != ORC_SCORE_SQ &&
        (!ix->codes || !ix->codebook || ix->m <= 0 ||
         (ix->ks != 16 && ix->ks != 256) || ix->subdim <= 0 ||
         ix->m * ix->subdim != ix->dim))
        return -1;
    if (scorer == ORC_SCORE_ADC_ONLINE && (!ix->res_codes || !ix->res_qparams))
        return -1;
    if (scorer == ORC_SCORE_SQ &&
        (!ix->sq_codes || ix->dim > 4096 || (ix->sq_bits != 4 && ix->sq_bits != 8)))
        return -1;
    if (expand_count > ix->degree) expand_count = ix->degree;
    memset(&sc, 0, sizeof(sc));
    sc.ix = ix;
    sc.scorer = scorer;
    sc.query = query;
    pool = malloc((size_t)pool_capacity * sizeof(*pool));
    front = malloc((size_t)ix->num_vectors * sizeof(*front));
    visited = calloc(((size_t)ix->num_vectors + 31) / 32, sizeof(*visited));
    turn_ids[0] = malloc((size_t)expand_count * sizeof(int));
    turn_ids[1] = malloc((size_t)expand_count * sizeof(int));
    items = malloc((size_t)expand_count * sizeof(*items));
    if (!pool || !front || !visited || !turn_ids[0] || !turn_ids[1] || !items) goto done;
    if (scorer == ORC_SCORE_SQ) {
        if (orc_sq_quantize_query(query, ix->dim, ix->sq_bits, ix->sq_scale,
                                  ix->sq_zero, sc.qsq) != 0)
            goto done;
    } else {
        lut = malloc((size_t)ix->m * ix->ks * sizeof(*lut));
        if (!lut || orc_build_lut(query, ix->codebook, ix->m, ix->ks,
                                  ix->subdim, lut) != 0)
            goto done;
        sc.lut = lut;
    }

    /* seeds (pure_pq_cpu_search_seed): duplicates skipped, each scored */
    for (int e = 0; e < entry_count; e++) {
        int id = entry_ids[e];
        float dist;
        if (id < 0 || id >= ix->num_vectors) goto done;
        if (visited[id >> 5] & (1u << (id & 31))) {
            st.visited_dedup_skips++;
            continue;
        }
        visited[id >> 5] |= 1u << (id & 31);
        if (orc_score(&sc, id, &dist) != 0) goto done;
        st.pq_scored_candidates++;
        if (scorer != ORC_SCORE_ADC) st.corrected_candidates++;
        if (orc_pool_insert(pool, &pool_size, pool_capacity, id, dist))
            orc_front_push(front, &front_size, id, dist);
    }
    if (pool_size == 0) goto done;

    for (;;) {
        /* dispatch up to the window */
        while (pending < window) {
            OrcCand src;
            int found = 0, valid, slot, nb = 0;
            const int32_t *row;
            while (front_size > 0) {
                if (pool_size == pool_capacity &&
                    orc_order(front[0].distance, front[0].id, pool[0].distance,
                              pool[0].id) > 0)
                    break;
                src = orc_front_pop(front, &front_size);
                if (orc_pool_contains(pool, pool_size, src.id, src.distance)) {
                    found = 1;
                    break;
                }
            }
            if (!found) break;
            row = ix->graph + (size_t)src.id * ix->degree;
            valid = orc_row_count(row, expand_count, ix->num_vectors);
            if (valid < 0) goto done;
            slot = (head + pending) % 2;
            for (int j = 0; j < valid; j++) {
                int id = row[j];
                if (visited[id >> 5] & (1u << (id & 31))) {
                    st.visited_dedup_skips++;
                    continue;
                }
                visited[id >> 5] |= 1u << (id & 31);
                turn_ids[slot][nb++] = id;
            }
            turn_n[slot] = nb;
            st.expanded_nodes++;
            pending++;
            if (pending > hw) hw = pending;
        }
        if (pending == 0) break;
        /* collect the oldest turn: score, sort, insert */
        {
            int nb = turn_n[head];
            for (int j = 0; j < nb; j++) {
                items[j].id = turn_ids[head][j];
                if (orc_score(&sc, items[j].id, &items[j].distance) != 0) goto done;
            }
            st.pq_scored_candidates += (uint64_t)nb;
            if (scorer != ORC_SCORE_ADC) st.corrected_candidates += (uint64_t)nb;
            qsort(items, (size_t)nb, sizeof(*items), orc_cand_cmp);
            for (int j = 0; j < nb; j++)
                if (orc_pool_insert(pool, &pool_size, pool_capacity, items[j].id,
                                    items[j].distance))
                    orc_front_push(front, &front_size, items[j].id, items[j].distance);
            head = (head + 1) % 2;
            pending--;
        }
    }

    qsort(pool, (size_t)pool_size, sizeof(*pool), orc_cand_cmp);
    result = num_out < pool_size ? num_out : pool_size;
    if (result > out_capacity) result = out_capacity;
    for (int i = 0; i < result; i++) {
        out_ids[i] = pool[i].id;
        out_distances[i] = pool[i].distance;
    }
    if (out_stats) *out_stats = st;
    if (high_water) *high_water = hw;
done:
    free(pool);
    free(front);
    free(visited);
    free(lut);
    free(turn_ids[0]);
    free(turn_ids[1]);
    free(items);
    return result;
}

typedef struct {
    int graph, original;
    double distance;
} OrcRefineItem;

static int orc_refine_cmp(const void *a, const void *b) {
    const OrcRefineItem *x = a, *y = b;
    if (x->distance != y->distance) return x->distance < y->distance ? -1 : 1;
    return (x->original > y->original) - (x->original < y->original);
}

/* retrieval_final_refine.c:55-67 (retrieval_final_refine_decode_sq): per
 * dimension add_f32(decode_value(l1 code, scale, zero), decode_value(residual
 * code, rscale, rzero)), decode_value = (code * scale) + zero, every step a
 * rounded f32 (:46-54).  -1 for parameters the reference refuses, -4 for a
 * non-finite value. */
int orc_sq_decode(const OrcIndex *ix, int id, float *out) {
    const int lcb = ix->sq_bits == 4 ? (ix->dim + 1) / 2 : ix->dim;
    const int rcb = ix->res_bits == 4 ? (ix->dim + 1) / 2 : ix->dim;
    const uint8_t *l1, *res;
    float rs, rz;
    if (!ix->sq_codes || !ix->res_codes || id < 0 || id >= ix->num_vectors ||
        (ix->sq_bits != 4 && ix->sq_bits != 8) || (ix->res_bits != 4 && ix->res_bits != 8))
        return -1;
    l1 = ix->sq_codes + (size_t)id * lcb;
    res = ix->res_codes + (size_t)id * rcb;
    rs = ix->res_qparams[2 * (size_t)id];
    rz = ix->res_qparams[2 * (size_t)id + 1];
    if (!isfinite(ix->sq_scale) || !(ix->sq_scale > 0.0f) || !isfinite(ix->sq_zero) ||
        !isfinite(rs) || rs < 0.0f || !isfinite(rz))
        return -1;
    for (int d = 0; d < ix->dim; d++) {
        const unsigned lc = ix->sq_bits == 4 ? (l1[d >> 1] >> (4 * (d & 1))) & 15u : l1[d];
        const unsigned rc = ix->res_bits == 4 ? (res[d >> 1] >> (4 * (d & 1))) & 15u : res[d];
        volatile float lp = (float)lc * ix->sq_scale;
        volatile float lv = lp + ix->sq_zero;
        volatile float rp = (float)rc * rs;
        volatile float rv = rp + rz;
        volatile float v = lv + rv;
        out[d] = v;
        if (!isfinite(out[d])) return -4;
    }
    return 0;
}

/* retrieval_final_refine.c:256-292 (refine_impl) with the PQ residual
 * decoder (retrieval_pq_residual.c:954-981) as the vector source, plus the
 * error4 level when the index carries one (ERROR4 mode):
 * dist = sum_d ((double)q - (double)v)^2 sequentially in f64, sorted by
 * (dist, original id), top output_k returned as (float)dist. */
int orc_refine(const OrcIndex *ix, const float *query, const int32_t *ids,
               int count, int output_k, const int32_t *graph_to_original,
               int32_t *out_ids, float *out_distances) {
    OrcRefineItem *items;
    float *vec;
    int rc;
    if (!ix || !query || !ids || count < 0 || output_k <= 0) return -1;
    for (int d = 0; d < ix->dim; d++)
        if (!isfinite(query[d])) return -4;
    items = malloc((size_t)(count > 0 ? count : 1) * sizeof(*items));
    vec = malloc((size_t)ix->dim * sizeof(float));
    if (!items || !vec) {
        free(items);
        free(vec);
        return -1;
    }
    for (int i = 0; i < count; i++) {
        int id = ids[i];
        if (id < 0 || id >= ix->num_vectors) {
            rc = -1;
            goto out;
        }
        for (int j = 0; j < i; j++)
            if (ids[j] == id) {
                rc = -1; /* duplicate storage row */
                goto out;
            }
        items[i].graph = id;
        items[i].original = graph_to_original ? graph_to_original[id] : id;
        /* the SQ index (no PQ codebook) refines from the SQ decoder, the
         * PQ index from the residual decoder */
        rc = ix->codebook ? orc_residual_decode(ix, id, vec) : orc_sq_decode(ix, id, vec);
        if (rc != 0) {
            rc = rc == -4 ? -4 : -1;
            goto out;
        }
        if (ix->e4_codes) {
            /* retrieval_final_refine.c:221-246 (decode_vector, ERROR4):
             * v = v + ((code * scale) + zero), three rounded f32 steps; the
             * row is the graph id (PQ has no storage map) */
            const uint8_t *c4 = ix->e4_codes + (size_t)id * ((ix->dim + 1) / 2);
            const float s4 = ix->e4_qparams[2 * (size_t)id];
            const float z4 = ix->e4_qparams[2 * (size_t)id + 1];
            for (int d = 0; d < ix->dim; d++) {
                volatile float product = (float)((c4[d >> 1] >> (4 * (d & 1))) & 15) * s4;
                volatile float level = product + z4;
                volatile float value = vec[d] + level;
                vec[d] = value;
                if (!isfinite(vec[d])) {
                    rc = -4;
                    goto out;
                }
            }
        }
        double acc = 0.0;
        for (int d = 0; d < ix->dim; d++) {
            double delta = (double)query[d] - (double)vec[d];
            acc += delta * delta;
        }
        if (!isfinite(acc) || acc >= FLT_MAX) {
            rc = -4;
            goto out;
        }
        items[i].distance = acc;
    }
    qsort(items, (size_t)count, sizeof(*items), orc_refine_cmp);
    rc = count < output_k ? count : output_k;
    for (int i = 0; i < rc; i++) {
        out_ids[i] = items[i].graph;
        out_distances[i] = (float)items[i].distance;
    }
out:
    free(items);
    free(vec);
    return rc;
}

/* ---- the live SQ traversal (retrieval_sqmd_common.c) ---- */

typedef struct {
    int id;
    float distance;
} OrcDc;

/* :9-47 -- pool: max-heap on distance alone */
static void orc_sq_pool_up(OrcDc *p, int i) {
    OrcDc t = p[i];
    while (i > 0) {
        int parent = (i - 1) / 2;
        if (t.distance <= p[parent].distance) break;
        p[i] = p[parent];
        i = parent;
    }
    p[i] = t;
}

static void orc_sq_pool_down(OrcDc *p, int size, int i) {
    OrcDc t = p[i];
    for (;;) {
        int child = 2 * i + 1;
        if (child >= size) break;
        if (child + 1 < size && p[child + 1].distance > p[child].distance) child++;
        if (t.distance >= p[child].distance) break;
        p[i] = p[child];
        i = child;
    }
    p[i] = t;
}

/* :51-79 -- admit below capacity, else only strictly below the worst */
static int orc_sq_pool_insert(OrcDc *p, int *size, int cap, int id, float d) {
    if (*size < cap) {
        p[*size].id = id;
        p[*size].distance = d;
        orc_sq_pool_up(p, *size);
        (*size)++;
        return 1;
    }
    if (d < p[0].distance) {
        p[0].id = id;
        p[0].distance = d;
        orc_sq_pool_down(p, *size, 0);
        return 1;
    }
    return 0;
}

/* :81-147 -- frontier: min-heap on distance alone, bounded push */
static void orc_sq_front_up(OrcDc *f, int i) {
    OrcDc t = f[i];
    while (i > 0) {
        int parent = (i - 1) / 2;
        if (t.distance >= f[parent].distance) break;
        f[i] = f[parent];
        i = parent;
    }
    f[i] = t;
}

static void orc_sq_front_down(OrcDc *f, int size, int i) {
    OrcDc t = f[i];
    for (;;) {
        int child = 2 * i + 1;
        if (child >= size) break;
        if (child + 1 < size && f[child + 1].distance < f[child].distance) child++;
        if (t.distance <= f[child].distance) break;
        f[i] = f[child];
        i = child;
    }
    f[i] = t;
}

static void orc_sq_front_push(OrcDc *f, int *size, int cap, int id, float d) {
    if (*size >= cap) return;
    f[*size].id = id;
    f[*size].distance = d;
    orc_sq_front_up(f, *size);
    (*size)++;
}

static int orc_sq_front_pop(OrcDc *f, int *size) {
    int top = f[0].id;
    (*size)--;
    if (*size > 0) {
        f[0] = f[*size];
        orc_sq_front_down(f, *size, 0);
    }
    return top;
}

float orc_sq_residual_distance(const OrcIndex *ix, const float *query, int id) {
    const int cb = ix->sq_bits == 4 ? (ix->dim + 1) / 2 : ix->dim;
    const uint8_t *l1 = ix->sq_codes + (size_t)id * cb;
    const uint8_t *res = NULL;
    const float *qp = NULL;
    float dist = 0.0f;
    if (ix->res_codes && ix->res_qparams) {
        res = ix->res_codes + (size_t)id * orc_residual_code_bytes(ix->dim, ix->res_bits);
        qp = ix->res_qparams + 2 * (size_t)id;
    }
    for (int i = 0; i < ix->dim; i++) {
        int c = ix->sq_bits == 4 ? ((i & 1) ? (l1[i >> 1] >> 4) : (l1[i >> 1] & 15)) : l1[i];
        volatile float prod = (float)c * ix->sq_scale;
        volatile float value = prod + ix->sq_zero;
        if (res) {
            volatile float rp = (float)orc_res_code(res, ix->res_bits, i) * qp[0];
            volatile float rl = rp + qp[1];
            value = value + rl;
        }
        volatile float diff = query[i] - value;
        volatile float sq = diff * diff;
        dist = dist + sq;
    }
    return dist;
}

/* retrieval_l1_int8_core.c:65-110 (retrieval_quantize_query_l1_aligned_u8):
 * step scale / 17 in double, round half away from zero, clip to [0, 255]. */
int orc_sq_query_aligned_u8(const float *query, int dim, float scale, float zero,
                            uint8_t *out, uint64_t *sum_q) {
    const double qs = (double)scale / 17.0;
    uint64_t sum = 0;
    if (!query || !out || dim <= 0 || !isfinite(scale) || !isfinite(zero) || scale <= 0.0f ||
        !isfinite(qs) || qs <= 0.0)
        return -1;
    for (int i = 0; i < dim; i++) {
        const double raw = ((double)query[i] - (double)zero) / qs;
        long code;
        if (!isfinite(query[i]) || !isfinite(raw)) return -1;
        code = lround(raw);
        if (code < 0) code = 0;
        else if (code > 255) code = 255;
        out[i] = (uint8_t)code;
        sum += (uint64_t)code;
    }
    if (sum_q) *sum_q = sum;
    return 0;
}

/* The split distance of retrieval_rpc_demo_pipeline.c:936-1000
 * (sqmd_split_residual_delta_fn, sqmd_split_distance_for_id) with the
 * scalar kernels it calls: retrieval_l1_int8_core.c:171-191 (coarse: the
 * int64 sum of (q - 17 c)^2 times (scale / 17)^2 in double, to float) and
 * :193-260 (the residual dot from integer sums sum q v and sum v, combined
 * in double); delta = (double)bias - 2 dot, to float; distance = coarse +
 * delta in float, with (-1e-4, 0) clamped to 0.  Returns NAN where the
 * reference would fail the candidate (non-finite or negative). */
float orc_sq_split_distance(const OrcIndex *ix, const uint8_t *qu8, uint64_t sum_q, int id) {
    const int cb = (ix->dim + 1) / 2;
    const uint8_t *l1 = ix->sq_codes + (size_t)id * cb;
    const uint8_t *res = ix->res_codes + (size_t)id * orc_residual_code_bytes(ix->dim, ix->res_bits);
    const double qs = (double)ix->sq_scale / 17.0;
    const float rs = ix->res_qparams[2 * (size_t)id], rz = ix->res_qparams[2 * (size_t)id + 1];
    int64_t acc = 0;
    uint64_t sum_qv = 0, sum_v = 0;
    double dot, delta;
    float coarse, dl, dist;
    for (int i = 0; i < ix->dim; i++) {
        const int c = (i & 1) ? (l1[i >> 1] >> 4) : (l1[i >> 1] & 15);
        const int32_t diff = (int32_t)qu8[i] - 17 * c;
        const uint32_t v = (uint32_t)orc_res_code(res, ix->res_bits, i);
        acc += (int64_t)diff * (int64_t)diff;
        sum_qv += (uint64_t)qu8[i] * (uint64_t)v;
        sum_v += (uint64_t)v;
    }
    coarse = (float)((double)acc * qs * qs);
    if (!isfinite(rs) || !isfinite(rz) || rs < 0.0f) return NAN;
    dot = qs * (double)rs * (double)sum_qv + qs * (double)rz * (double)sum_q +
          (double)ix->sq_zero * (double)rs * (double)sum_v +
          (double)ix->dim * (double)ix->sq_zero * (double)rz;
    delta = (double)ix->split_bias[id] - 2.0 * dot;
    if (!isfinite(delta)) return NAN;
    dl = (float)delta;
    dist = coarse + dl;
    if (dist < 0.0f && dist > -1.0e-4f) dist = 0.0f;
    if (!isfinite(dist) || dist < 0.0f) return NAN;
    return dist;
}

int orc_sqmd_seed_ids(int n, int entry_points, int graph_entry, uint32_t seed,
                      int *out, int capacity) {
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
        id = (int)((seed / 65536u) % (unsigned)n);
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

/* One query of the fake-DSP turn loop (:2052-2152): seeds (:446-476), the
 * source pop that discards frontier tops worse than a full pool's worst
 * (:478-523), the turn's work list in neighbor order (invalid ids and the
 * source skipped, duplicates and visited nodes skipped, then all marked
 * visited; :283-383, :720-773), results applied in that order through the
 * correction (:638-708), and the final heap sort (:40-49). */
int orc_sqmd_search(const OrcIndex *ix, int correction, const float *query,
                    const int *seed_ids, int seed_count, int entry_points,
                    int expand_count, int pool_capacity, int *out_ids,
                    float *out_distances, int out_capacity, OrcSqmdStats *stats) {
    OrcSqmdStats st;
    OrcDc *pool = NULL, *front = NULL;
    uint32_t *visited = NULL;
    uint8_t *qsq = NULL;
    int psize = 0, fsize = 0, rc = -1, fcap, cb;
    uint64_t sum_q = 0;
    memset(&st, 0, sizeof(st));
    if (!ix || !query || !seed_ids || seed_count <= 0 || entry_points <= 0 ||
        expand_count <= 0 || pool_capacity <= 0 || !out_ids || !out_distances ||
        out_capacity <= 0 || !ix->sq_codes)
        return -1;
    if (correction == 2 && (ix->sq_bits != 4 || !ix->split_bias || !ix->res_codes ||
                            !ix->res_qparams))
        return -1;
    fcap = entry_points + pool_capacity * expand_count;
    cb = ix->sq_bits == 4 ? (ix->dim + 1) / 2 : ix->dim;
    pool = malloc((size_t)pool_capacity * sizeof(*pool));
    front = malloc((size_t)fcap * sizeof(*front));
    visited = calloc(((size_t)ix->num_vectors + 31) / 32, sizeof(uint32_t));
    qsq = malloc((size_t)ix->dim);
    if (!pool || !front || !visited || !qsq) goto out;
    if (correction == 2) {
        if (orc_sq_query_aligned_u8(query, ix->dim, ix->sq_scale, ix->sq_zero, qsq, &sum_q))
            goto out;
    } else {
        orc_sq_quantize_query(query, ix->dim, ix->sq_bits, ix->sq_scale, ix->sq_zero, qsq);
    }
#define ORC_DIST(id)                                                             \
    (correction == 2 ? orc_sq_split_distance(ix, qsq, sum_q, (id))               \
     : correction ? orc_sq_residual_distance(ix, query, (id))                    \
                : orc_sq_distance(qsq, ix->sq_codes + (size_t)(id) * cb, ix->dim, \
                                  ix->sq_bits, ix->sq_scale))
#define ORC_VIS(id) (visited[(id) >> 5] & (1u << ((id) & 31)))
    for (int i = 0; i < seed_count; i++) {
        int id = seed_ids[i];
        float d;
        if (id < 0 || id >= ix->num_vectors) goto out;
        if (ORC_VIS(id)) continue;
        visited[id >> 5] |= 1u << (id & 31);
        d = ORC_DIST(id);
        if (isnan(d)) goto out;
        if (orc_sq_pool_insert(pool, &psize, pool_capacity, id, d))
            orc_sq_front_push(front, &fsize, fcap, id, d);
    }
    for (;;) {
        int work[32], wc = 0, src, ec;
        const int32_t *row;
        while (fsize > 0 && psize >= pool_capacity && front[0].distance > pool[0].distance)
            orc_sq_front_pop(front, &fsize);
        if (fsize <= 0) break;
        src = orc_sq_front_pop(front, &fsize);
        row = ix->graph + (size_t)src * ix->degree;
        ec = expand_count < ix->degree ? expand_count : ix->degree;
        if (ec > 32) goto out; /* SQMD_STATUS_TOO_MUCH_WORK */
        for (int j = 0; j < ec; j++) {
            int nb = row[j], dup = 0;
            if (nb < 0 || nb >= ix->num_vectors || nb == src) {
                st.invalid_skips++;
                continue;
            }
            for (int i = 0; i < wc && !dup; i++) dup = work[i] == nb;
            if (dup || ORC_VIS(nb)) {
                st.dedup_skips++;
                continue;
            }
            work[wc++] = nb;
        }
        st.turns++;
        st.work_items += (uint64_t)wc;
        if (wc == 0) st.empty_turns++;
        for (int i = 0; i < wc; i++) visited[work[i] >> 5] |= 1u << (work[i] & 31);
        for (int i = 0; i < wc; i++) {
            float d = ORC_DIST(work[i]);
            if (isnan(d)) goto out;
            if (orc_sq_pool_insert(pool, &psize, pool_capacity, work[i], d))
                orc_sq_front_push(front, &fsize, fcap, work[i], d);
            st.applied++;
        }
    }
#undef ORC_DIST
#undef ORC_VIS
    for (int i = psize - 1; i > 0; i--) {
        OrcDc t = pool[0];
        pool[0] = pool[i];
        pool[i] = t;
        orc_sq_pool_down(pool, i, 0);
    }
    rc = psize < out_capacity ? psize : out_capacity;
    for (int i = 0; i < rc; i++) {
        out_ids[i] = pool[i].id;
        out_distances[i] = pool[i].distance;
    }
out:
    if (stats) *stats = st;
    free(pool);
    free(front);
    free(visited);
    free(qsq);
    return rc;
}

/* ---- graph construction (retrieval_search_build_core.c) ---- */

float orc_l2(const float *base, int dim, int a, int b) {
    const float *va = base + (size_t)a * dim, *vb = base + (size_t)b * dim;
    float acc = 0.0f;
    for (int d = 0; d < dim; d++) {
        float diff = va[d] - vb[d];
        acc += diff * diff;
    }
    return acc;
}

static uint32_t orc_build_lcg(uint32_t x) { return x * 1664525u + 1013904223u; }

/* :108-113 (sb_seed_for_node) */
static uint32_t orc_node_seed(uint32_t seed, int node) {
    uint32_t x = seed ^ (uint32_t)(node + 1) * 0x9e3779b9u;
    x = orc_build_lcg(x);
    x ^= x >> 16;
    return orc_build_lcg(x);
}

typedef struct {
    int id;
    float d;
} OrcEdge;

/* (distance, id) ascending (:115-131 and rnn_edge_is_better) */
static int orc_edge_cmp(const void *a, const void *b) {
    const OrcEdge *x = a, *y = b;
    if (x->d < y->d) return -1;
    if (x->d > y->d) return 1;
    return (x->id > y->id) - (x->id < y->id);
}

static int orc_edges_contain(const OrcEdge *e, int count, int id) {
    for (int i = 0; i < count; i++)
        if (e[i].id == id) return 1;
    return 0;
}

/* Draws from the node's LCG until `row` holds r distinct ids other than u
 * (at most n*n draws), then fills with the lowest unused ids (:682-776 and
 * :781-831 share this shape). */
static void orc_random_fill(const float *base, int n, int dim, int r, uint32_t state,
                            int u, OrcEdge *row, int *count) {
    uint64_t attempts = 0, limit = (uint64_t)(uint32_t)n * (uint64_t)(uint32_t)n;
    while (*count < r && attempts < limit) {
        int v;
        state = orc_build_lcg(state);
        v = (int)(state % (uint32_t)n);
        if (v != u && !orc_edges_contain(row, *count, v)) {
            row[*count].id = v;
            row[*count].d = orc_l2(base, dim, u, v);
            (*count)++;
        }
        attempts++;
    }
    for (int v = 0; v < n && *count < r; v++) {
        if (v == u || orc_edges_contain(row, *count, v)) continue;
        row[*count].id = v;
        row[*count].d = orc_l2(base, dim, u, v);
        (*count)++;
    }
}

int orc_build_exact(const float *base, int n, int dim, int r, int candidates,
                    uint32_t seed, int prune_mode, int32_t *out) {
    OrcEdge *all, *init, *items, *row;
    int rc = -1;
    if (!base || n <= 1 || dim <= 0 || r <= 0 || r >= n || candidates < r ||
        (prune_mode != 0 && prune_mode != 1) || !out)
        return -1;
    all = malloc((size_t)n * sizeof(*all));
    init = malloc((size_t)r * sizeof(*init));
    items = malloc((size_t)(candidates + r) * sizeof(*items));
    row = malloc((size_t)r * sizeof(*row));
    if (!all || !init || !items || !row) goto done;
    for (int u = 0; u < n; u++) {
        int ni = 0, na = 0, count = 0, deg = 0, cap = candidates + r;
        /* random initial row, sorted */
        orc_random_fill(base, n, dim, r, orc_node_seed(seed, u), u, init, &ni);
        qsort(init, (size_t)ni, sizeof(*init), orc_edge_cmp);
        /* exact candidates: every other node by (distance, id), first C */
        for (int v = 0; v < n; v++) {
            if (v == u) continue;
            all[na].id = v;
            all[na].d = orc_l2(base, dim, u, v);
            na++;
        }
        qsort(all, (size_t)na, sizeof(*all), orc_edge_cmp);
        if (na > candidates) na = candidates;
        /* candidates then the previous row, first occurrence kept */
        for (int i = 0; i < na && count < cap; i++)
            if (!orc_edges_contain(items, count, all[i].id)) items[count++] = all[i];
        for (int i = 0; i < ni && count < cap; i++)
            if (!orc_edges_contain(items, count, init[i].id)) items[count++] = init[i];
        qsort(items, (size_t)count, sizeof(*items), orc_edge_cmp);
        if (prune_mode == 0) {
            for (int i = 0; i < count && deg < r; i++) row[deg++] = items[i];
        } else {
            for (int i = 0; i < count && deg < r; i++) {
                int occluded = 0;
                for (int j = 0; j < deg && !occluded; j++)
                    occluded = orc_l2(base, dim, row[j].id, items[i].id) < items[i].d;
                if (!occluded) row[deg++] = items[i];
            }
            for (int i = 0; i < count && deg < r; i++)
                if (!orc_edges_contain(row, deg, items[i].id)) row[deg++] = items[i];
        }
        if (deg < r)
            orc_random_fill(base, n, dim, r, orc_node_seed(seed ^ 0xa5a55a5au, u), u, row,
                            &deg);
        qsort(row, (size_t)deg, sizeof(*row), orc_edge_cmp);
        for (int i = 0; i < r; i++) out[(size_t)u * r + i] = i < deg ? row[i].id : -1;
    }
    rc = 0;
done:
    free(all);
    free(init);
    free(items);
    free(row);
    return rc;
}

/* retrieval_rpc_demo_support.c:1709-1728 (demo_recall_hits_at_k) */
int orc_recall_hits(const int *res, int res_count, const int32_t *gt,
                    int gt_k, int k) {
    int hits = 0;
    for (int i = 0; i < k && i < res_count; i++)
        for (int j = 0; j < gt_k && j < k; j++)
            if (res[i] == gt[j]) {
                hits++;
                break;
            }
    return hits;
}

/* retrieval_rpc_demo_pure_pq.c:387-400 (pure_pq_result_fingerprint):
 * FNV-1a 64 over (id, distance bits) little-endian bytes. */
uint64_t orc_fingerprint(uint64_t hash, const int *ids, const float *dists,
                         int count) {
    for (int i = 0; i < count; i++) {
        uint32_t v[2];
        v[0] = (uint32_t)ids[i];
        memcpy(&v[1], &dists[i], 4);
        for (int b = 0; b < 8; b++) {
            hash ^= (v[b / 4] >> (8 * (b % 4))) & 255u;
            hash *= UINT64_C(1099511628211);
        }
    }
    return hash;
}

typedef struct {
    const OrcIndex *ix;
    int scorer;
    const float *queries;
    int nq;
    const int *entries;
    int n_entries, entry_stride, expand;
    const int *pool, *num_out;
    int stride, refine;
    int *out_ids;
    float *out_d;
    int *out_counts;
    OrcStats *stats;
    int next;
    int failed;
    int window;
} OrcBatch;

static void *orc_batch_worker(void *arg) {
    OrcBatch *b = arg;
    int *ids = NULL;
    float *dists = NULL;
    int cap = 0;
    for (;;) {
        int q = __atomic_fetch_add(&b->next, 1, __ATOMIC_RELAXED);
        if (q >= b->nq) break;
        int L = b->pool[q];
        if (L > cap) {
            free(ids);
            free(dists);
            cap = L;
            ids = malloc((size_t)cap * sizeof(int));
            dists = malloc((size_t)cap * sizeof(float));
            if (!ids || !dists) {
                __atomic_store_n(&b->failed, 1, __ATOMIC_RELAXED);
                b->out_counts[q] = -1;
                cap = 0;
                continue;
            }
        }
        const float *qv = b->queries + (size_t)q * b->ix->dim;
        int want = b->refine > 0 ? (b->refine < L ? b->refine : L) : b->num_out[q];
        OrcStats st;
        const int *ents = b->entries + (b->entry_stride > 0 ? (size_t)q * b->entry_stride : 0);
        int ne = b->entry_stride > 0 ? b->entry_stride : b->n_entries;
        int n = b->window > 1
                    ? orc_search_window(b->ix, b->scorer, qv, ents, ne, b->expand, L,
                                        want, b->window, ids, dists, cap, &st, NULL)
                    : orc_search(b->ix, b->scorer, qv, ents, ne, b->expand, L, want,
                                 ids, dists, cap, &st);
        if (n >= 0 && b->refine > 0) {
            n = orc_refine(b->ix, qv, ids, n, b->num_out[q], NULL, ids, dists);
        }
        if (n < 0) __atomic_store_n(&b->failed, 1, __ATOMIC_RELAXED);
        b->out_counts[q] = n;
        for (int i = 0; i < n && i < b->stride; i++) {
            b->out_ids[(size_t)q * b->stride + i] = ids[i];
            b->out_d[(size_t)q * b->stride + i] = dists[i];
        }
        if (b->stats) b->stats[q] = st;
    }
    free(ids);
    free(dists);
    return NULL;
}

int orc_search_batch(const OrcIndex *ix, int scorer, const float *queries,
                     int nq, const int *entries, int n_entries,
                     int entry_stride, int expand, const int *pool,
                     const int *num_out, int stride, int refine, int threads,
                     int *out_ids, float *out_d, int *out_counts,
                     OrcStats *stats) {
    return orc_search_batch_window(ix, scorer, queries, nq, entries, n_entries,
                                   entry_stride, expand, pool, num_out, stride,
                                   refine, 1, threads, out_ids, out_d, out_counts,
                                   stats);
}

int orc_search_batch_window(const OrcIndex *ix, int scorer, const float *queries,
                            int nq, const int *entries, int n_entries,
                            int entry_stride, int expand, const int *pool,
                            const int *num_out, int stride, int refine, int window,
                            int threads, int *out_ids, float *out_d, int *out_counts,
                            OrcStats *stats) {
    OrcBatch b = {ix, scorer, queries, nq, entries, n_entries, entry_stride,
                  expand, pool, num_out, stride, refine, out_ids, out_d,
                  out_counts, stats, 0, 0, window};
    pthread_t tid[256];
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    for (int t = 0; t < threads; t++)
        if (pthread_create(&tid[t], NULL, orc_batch_worker, &b) != 0) {
            threads = t;
            b.failed = 1;
            break;
        }
    for (int t = 0; t < threads; t++) pthread_join(tid[t], NULL);
    return b.failed ? -1 : 0;
}

/* ---- RNN-Descent (retrieval_rnn_descent_core.c:1432-1548) ----
 * Rows hold up to r (id, distance, new-flag) entries.  Initial rows: s
 * distinct random others per node from the node's LCG (:571-738); then t1
 * rounds of t2 update passes (:901-1161), with reverse edges added between
 * rounds (:1163-1305); rows sorted by (distance, id) and exported with -1
 * padding (:438-470, :475-570). */

typedef struct {
    int n, r;
    int *deg;
    int *ids;
    float *dist;
    uint8_t *fresh;
} OrcRnn;

typedef struct {
    int src, dst;
    float dist;
} OrcIns;

static int orc_rnn_better(float nd, int ni, float od, int oi) {
    return nd < od || (nd == od && ni < oi);
}

static int orc_rnn_has(const OrcRnn *g, int src, int dst) {
    for (int i = 0; i < g->deg[src]; i++)
        if (g->ids[(size_t)src * g->r + i] == dst) return 1;
    return 0;
}

/* :304-375 */
static int orc_rnn_insert(OrcRnn *g, int src, int dst, float d) {
    size_t row = (size_t)src * g->r;
    int far = -1;
    if (src == dst || orc_rnn_has(g, src, dst)) return 0;
    if (g->deg[src] < g->r) {
        g->ids[row + g->deg[src]] = dst;
        g->dist[row + g->deg[src]] = d;
        g->fresh[row + g->deg[src]] = 1;
        g->deg[src]++;
        return 1;
    }
    for (int i = 0; i < g->deg[src]; i++)
        if (far < 0 || orc_rnn_better(g->dist[row + far], g->ids[row + far], g->dist[row + i],
                                      g->ids[row + i]))
            far = i;
    if (far >= 0 && orc_rnn_better(d, dst, g->dist[row + far], g->ids[row + far])) {
        g->ids[row + far] = dst;
        g->dist[row + far] = d;
        g->fresh[row + far] = 1;
        return 1;
    }
    return 0;
}

/* :438-470, insertion sort by (distance, id) */
static void orc_rnn_sort_row(OrcRnn *g, int u) {
    size_t row = (size_t)u * g->r;
    for (int i = 1; i < g->deg[u]; i++) {
        int id = g->ids[row + i], j = i - 1;
        float d = g->dist[row + i];
        uint8_t f = g->fresh[row + i];
        while (j >= 0 && orc_rnn_better(d, id, g->dist[row + j], g->ids[row + j])) {
            g->ids[row + j + 1] = g->ids[row + j];
            g->dist[row + j + 1] = g->dist[row + j];
            g->fresh[row + j + 1] = g->fresh[row + j];
            j--;
        }
        g->ids[row + j + 1] = id;
        g->dist[row + j + 1] = d;
        g->fresh[row + j + 1] = f;
    }
}

static int orc_ins_cmp(const void *a, const void *b) {
    const OrcIns *x = a, *y = b;
    if (x->src != y->src) return x->src < y->src ? -1 : 1;
    if (x->dst != y->dst) return x->dst < y->dst ? -1 : 1;
    if (x->dist < y->dist) return -1;
    if (x->dist > y->dist) return 1;
    return 0;
}

static void orc_rnn_apply(OrcRnn *g, OrcIns *items, size_t count) {
    if (count > 1) qsort(items, count, sizeof(*items), orc_ins_cmp);
    for (size_t i = 0; i < count; i++) orc_rnn_insert(g, items[i].src, items[i].dst, items[i].dist);
}

int orc_rnn_build(const float *base, int n, int dim, int s, int r, int t1, int t2,
                  uint32_t seed, int top_k, int32_t *out) {
    OrcRnn g;
    OrcIns *pend = NULL;
    int *cand = NULL, rc = -1;
    size_t np = 0;
    if (!base || !out || n <= 1 || dim <= 0 || s <= 0 || r <= 0 || t1 <= 0 || t2 <= 0 ||
        s > r || top_k <= 0)
        return -1;
    g.n = n;
    g.r = r;
    g.deg = calloc((size_t)n, sizeof(int));
    g.ids = malloc((size_t)n * r * sizeof(int));
    g.dist = malloc((size_t)n * r * sizeof(float));
    g.fresh = calloc((size_t)n * r, 1);
    pend = malloc((size_t)n * r * sizeof(OrcIns));
    cand = malloc((size_t)r * sizeof(int));
    if (!g.deg || !g.ids || !g.dist || !g.fresh || !pend || !cand) goto out;
    if (s > n - 1) s = n - 1;
    for (int u = 0; u < n; u++) {
        uint32_t st = seed ^ (uint32_t)(u + 1) * 0x9e3779b9u;
        uint64_t attempts = 0, limit = (uint64_t)(uint32_t)n * (uint64_t)(uint32_t)n;
        int cnt = 0;
        st = st * 1664525u + 1013904223u;
        st ^= st >> 16;
        st = st * 1664525u + 1013904223u;
        while (cnt < s && attempts < limit) {
            int v, dup = 0;
            st = st * 1664525u + 1013904223u;
            v = (int)(st % (uint32_t)n);
            if (u != v) {
                for (int i = 0; i < cnt && !dup; i++) dup = cand[i] == v;
                if (!dup) cand[cnt++] = v;
            }
            attempts++;
        }
        for (int v = 0; v < n && cnt < s; v++) {
            int dup = v == u;
            for (int i = 0; i < cnt && !dup; i++) dup = cand[i] == v;
            if (!dup) cand[cnt++] = v;
        }
        for (int i = 0; i < cnt; i++) orc_rnn_insert(&g, u, cand[i], orc_l2(base, dim, u, cand[i]));
    }
    for (int a = 0; a < t1; a++) {
        for (int b = 0; b < t2; b++) {
            np = 0;
            for (int u = 0; u < n; u++) {
                size_t row = (size_t)u * r;
                int i = 0;
                orc_rnn_sort_row(&g, u);
                while (i < g.deg[u]) {
                    const int v = g.ids[row + i];
                    const float duv = g.dist[row + i];
                    const uint8_t fv = g.fresh[row + i];
                    int removed = 0;
                    for (int w = 0; w < i; w++) {
                        const int wid = g.ids[row + w];
                        float dvw;
                        if (!fv && !g.fresh[row + w]) continue;
                        dvw = orc_l2(base, dim, v, wid);
                        if (duv >= dvw) {
                            pend[np].src = wid;
                            pend[np].dst = v;
                            pend[np].dist = dvw;
                            np++;
                            removed = 1;
                            break;
                        }
                    }
                    if (removed) {
                        for (int k = i + 1; k < g.deg[u]; k++) {
                            g.ids[row + k - 1] = g.ids[row + k];
                            g.dist[row + k - 1] = g.dist[row + k];
                            g.fresh[row + k - 1] = g.fresh[row + k];
                        }
                        g.deg[u]--;
                    } else {
                        i++;
                    }
                }
                for (int k = 0; k < g.deg[u]; k++) g.fresh[row + k] = 0;
            }
            orc_rnn_apply(&g, pend, np);
        }
        if (a + 1 < t1) {
            np = 0;
            for (int u = 0; u < n; u++)
                for (int k = 0; k < g.deg[u]; k++) {
                    pend[np].src = g.ids[(size_t)u * r + k];
                    pend[np].dst = u;
                    pend[np].dist = g.dist[(size_t)u * r + k];
                    np++;
                }
            orc_rnn_apply(&g, pend, np);
        }
    }
    for (int u = 0; u < n; u++) {
        orc_rnn_sort_row(&g, u);
        for (int j = 0; j < top_k; j++)
            out[(size_t)u * top_k + j] = j < g.deg[u] ? g.ids[(size_t)u * r + j] : -1;
    }
    rc = 0;
out:
    free(g.deg);
    free(g.ids);
    free(g.dist);
    free(g.fresh);
    free(pend);
    free(cand);
    return rc;
}
