/*
 * rover_oracle.h -- CPU restatement of Rover's scalar quantized graph search.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the HIP search path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.  Nothing in rover_amd/ links or calls it.
 *
 * Every function restates one reference function (cited file:line below,
 * paths relative to the Rover source tree) with the same floating-point
 * operation order, so results are bit-identical when both are compiled
 * without FMA contraction (-ffp-contract=off).
 */
#ifndef ROVER_ORACLE_H
#define ROVER_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* How a scored candidate's distance is produced. */
enum {
    ORC_SCORE_ADC = 0,        /* PQ ADC, policy OFF                         */
    ORC_SCORE_ADC_ONLINE = 1, /* PQ ADC replaced by residual reconstruction */
    ORC_SCORE_SQ = 2          /* global-scale SQ8/SQ4 code distance         */
};

typedef struct OrcIndex {
    const int32_t *graph; /* [num_vectors][degree], -1 suffix padding */
    int degree;
    int num_vectors;
    /* PQ */
    const uint8_t *codes;    /* [num_vectors][code_bytes] */
    const float *codebook;   /* [m][ks][subdim] */
    int dim, m, ks, subdim;
    /* residual (ONLINE correction) */
    const uint8_t *res_codes;  /* [num_vectors][res_code_bytes] */
    const float *res_qparams;  /* [num_vectors][2] = (scale, zero_point) */
    int res_bits;              /* 4 or 8 */
    /* SQ (global scale) */
    const uint8_t *sq_codes;   /* [num_vectors][sq_code_bytes] */
    int sq_bits;               /* 4 or 8 */
    float sq_scale, sq_zero;
    /* third-level refine level (error4): 4-bit codes
     * [num_vectors][(dim+1)/2], low nibble = even d, and (scale, zero)
     * [num_vectors][2]; NULL refines without it */
    const uint8_t *e4_codes;
    const float *e4_qparams;
    /* SQ split distance: the per-vector bias sum_d (2 a_d r_d + r_d^2) of an
     * SQMDBIAS artifact, f32 [num_vectors]; NULL without one */
    const float *split_bias;
} OrcIndex;

typedef struct OrcStats {
    uint64_t pq_scored_candidates;
    uint64_t corrected_candidates;
    uint64_t expanded_nodes;
    uint64_t visited_dedup_skips;
} OrcStats;

int orc_code_bytes(int m, int ks);
int orc_residual_code_bytes(int dim, int bits);

int orc_build_lut(const float *query, const float *codebook, int m, int ks,
                  int subdim, float *lut);
float orc_adc(const uint8_t *code, const float *lut, int m, int ks);

int orc_residual_encode(const float *vector, const uint8_t *pq_code,
                        const float *codebook, int m, int ks, int subdim,
                        int bits, uint8_t *out_code, float *out_qparams);
int orc_residual_evaluate(const OrcIndex *index, const float *query,
                          int vector_id, float *out_distance);
int orc_residual_decode(const OrcIndex *index, int vector_id,
                        float *out_vector);

/* One vector's error4 code from the raw vector x and its reconstruction
 * recon (the residual decode): out_code [(dim+1)/2], out_qparams (scale,
 * zero).  Returns 0, or -1 on a non-finite error. */
int orc_error4_encode(const float *x, const float *recon, int dim,
                      uint8_t *out_code, float *out_qparams);

int orc_sq_range_params(const float *data, long count, int bits,
                        float *out_scale, float *out_zero);
int orc_sq_encode(const float *vector, int dim, int bits, float scale,
                  float zero, uint8_t *out_code);
int orc_sq_quantize_query(const float *query, int dim, int bits, float scale,
                          float zero, uint8_t *out_query);
float orc_sq_distance(const uint8_t *query_q, const uint8_t *code, int dim,
                      int bits, float scale);

int orc_seed_ids(int num_vectors, int entry_count, int graph_entry,
                 uint32_t seed, int *out_ids);

int orc_search(const OrcIndex *index, int scorer, const float *query,
               const int *entry_ids, int entry_count, int expand_count,
               int pool_capacity, int num_out, int *out_ids,
               float *out_distances, int out_capacity, OrcStats *out_stats);

/* The live pipeline with `window` (1 or 2) source turns in flight
 * (pure_pq_source_pipeline_step, retrieval_sqmd_pure_pq.c:722-762): a turn
 * is dispatched from the pool as it stands before the previous turn's work
 * is inserted.  window 1 gives orc_search's results.  expanded_nodes counts
 * dispatched turns; *high_water (may be NULL) the most pending at once. */
int orc_search_window(const OrcIndex *index, int scorer, const float *query,
                      const int *entry_ids, int entry_count, int expand_count,
                      int pool_capacity, int num_out, int window, int *out_ids,
                      float *out_distances, int out_capacity, OrcStats *out_stats,
                      int *high_water);

/* The SQ index's reconstruction (retrieval_final_refine_decode_sq):
 * sq_codes + residual; 0, -1 (bad parameters) or -4 (non-finite). */
int orc_sq_decode(const OrcIndex *index, int id, float *out);

/* With a PQ codebook the vectors come from the residual decoder, without
 * one (an SQ index) from orc_sq_decode. */
int orc_refine(const OrcIndex *index, const float *query, const int32_t *ids,
               int count, int output_k, const int32_t *graph_to_original,
               int32_t *out_ids, float *out_distances);

/* SQ traversal of the reference's live SQ path (retrieval_sqmd_common.c:
 * sqmd_cpu_state_* with the fake-DSP turn driver, :2052-2152): distance-only
 * heaps, results applied in work order.  correction = 0 ranks by the SQ code
 * distance, 1 by the exact SQ + residual distance (res_* of the index over
 * sq_codes).  Returns the output count, or -1 (bad seed id, expand > 32). */
typedef struct OrcSqmdStats {
    uint64_t work_items, dedup_skips, invalid_skips, applied, turns, empty_turns;
} OrcSqmdStats;

int orc_sqmd_search(const OrcIndex *index, int correction, const float *query,
                    const int *seed_ids, int seed_count, int entry_points,
                    int expand_count, int pool_capacity, int *out_ids,
                    float *out_distances, int out_capacity, OrcSqmdStats *stats);

/* retrieval_sqmd_seed.c:15-63 (retrieval_sqmd_make_seed_ids). */
int orc_sqmd_seed_ids(int num_vectors, int entry_points, int graph_entry,
                      uint32_t seed, int *out_ids, int capacity);

/* retrieval_rpc_demo_support.c:1397-1443
 * (retrieval_demo_exact_distance_residual_scalar). */
float orc_sq_residual_distance(const OrcIndex *index, const float *query, int id);

/* SQ split distance (SQ4 codes, 4- or 8-bit residuals, split_bias set):
 * the query quantized on the SQ grid refined 17x (u8), and for candidate id
 * the coarse u8 x u4 distance plus (bias - 2 <query_u8, residual>).
 * orc_sq_query_aligned_u8 returns 0 or -1 (bad query or scale). */
int orc_sq_query_aligned_u8(const float *query, int dim, float scale, float zero,
                            uint8_t *out, uint64_t *sum_q);
float orc_sq_split_distance(const OrcIndex *index, const uint8_t *query_u8,
                            uint64_t sum_q, int id);

/* Graph construction as retrieval_search_build_graph_ex builds it with the
 * exact provider and one iteration (retrieval_search_build_core.c:682-776
 * random initial rows, :314-361 exact candidates, :833-968 merge / prune /
 * fill, :781-831 random fill, retrieval_rnn_descent_core.c:438-470 row
 * order).  prune_mode 0 keeps the closest r candidates, 1 applies the RNG
 * occlusion rule.  out [n][r].  Returns 0 or -1. */
int orc_build_exact(const float *base, int n, int dim, int r, int candidates,
                    uint32_t seed, int prune_mode, int32_t *out);

/* RNN-Descent graph construction (retrieval_rnn_build_graph_ex,
 * retrieval_rnn_descent_core.c:1432-1548, then retrieval_rnn_export_topk):
 * s random initial neighbors per node, t1 rounds of t2 update passes with
 * reverse edges between rounds, rows of at most r; out [n][top_k] by
 * (distance, id), -1 padded.  Returns 0 or -1. */
int orc_rnn_build(const float *base, int n, int dim, int s, int r, int t1, int t2,
                  uint32_t seed, int top_k, int32_t *out);

/* retrieval_rnn_descent_core.c:192-212: sequential f32 sum of squared
 * differences. */
float orc_l2(const float *base, int dim, int a, int b);

int orc_recall_hits(const int *result_ids, int result_count,
                    const int32_t *gt_row, int gt_k, int k);
uint64_t orc_fingerprint(uint64_t hash, const int *ids,
                         const float *distances, int count);

/* Many queries on a pool of host threads (one query per thread at a time).
 * refine_k > 0 applies orc_refine to the first refine_candidates results.
 * entry_stride > 0: entry_ids holds one row of entry_stride ids per query
 * (entry_count is then ignored).
 * Returns 0, or -1 if any query failed (its count is then -1). */
int orc_search_batch(const OrcIndex *index, int scorer, const float *queries,
                     int num_queries, const int *entry_ids, int entry_count,
                     int entry_stride,
                     int expand_count, const int *pool_capacity,
                     const int *num_out, int out_stride, int refine_candidates,
                     int threads, int *out_ids, float *out_distances,
                     int *out_counts, OrcStats *out_stats);
/* The same with every query searched by orc_search_window at `window`. */
int orc_search_batch_window(const OrcIndex *index, int scorer, const float *queries,
                            int num_queries, const int *entry_ids, int entry_count,
                            int entry_stride, int expand_count,
                            const int *pool_capacity, const int *num_out,
                            int out_stride, int refine_candidates, int window,
                            int threads, int *out_ids, float *out_distances,
                            int *out_counts, OrcStats *out_stats);

#ifdef __cplusplus
}
#endif
#endif
