/*
 * rover_hip.h -- the thin C-ABI between the C host library (rover_host.c)
 * and the HIP device code (rover_search.hip).  Plain pointers and sizes
 * only.  Pointers named d_* are device (HBM) addresses.
 */
#ifndef ROVER_HIP_H
#define ROVER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    RV_SCORE_ADC = 0,        /* PQ ADC ranking (correction off)           */
    RV_SCORE_ADC_ONLINE = 1, /* residual-reconstructed L2 (online)        */
    RV_SCORE_SQ = 2,         /* global-scale SQ8/SQ4 integer code distance */
    RV_SCORE_ONLINE_PAIR32 = 3,   /* online, two lanes per candidate; chosen
                                     by rv_search_launch for M=32 x 4-dim,
                                     ks=256, 8-bit residuals, <= 32 neighbors */
    RV_SCORE_ONLINE_PAIR32_R4 = 4 /* the same with 4-bit residuals */
};

/* Status written per query to out_counts when negative. */
enum {
    RV_QUERY_BAD_ROW = -2,     /* graph row with -1 not as suffix / bad id  */
    RV_QUERY_BAD_ENTRY = -3,   /* entry id out of range                     */
    RV_QUERY_NONFINITE = -4,   /* non-finite or FLT_MAX distance            */
    RV_QUERY_BAD_PARAM = -5    /* L/k outside the launch's limits           */
};

/* error4 artifact code blocks are this many bytes */
#define RV_E4_BLOCK 4096

/* Device-resident index.  All arrays are row-major and dense. */
typedef struct RvIndex {
    const int32_t *graph;       /* [n][degree], -1 suffix padding         */
    int degree, n, graph_entry;
    const uint8_t *codes;       /* [n][code_bytes] PQ codes               */
    int code_bytes;
    const float *codebook;      /* [m][ks][subdim]                        */
    int dim, m, ks, subdim;
    const uint8_t *res_codes;   /* [n][res_code_bytes]                    */
    const float *res_qparams;   /* [n][2] (scale, zero_point)             */
    int res_bits, res_code_bytes;
    const uint8_t *sq_codes;    /* [n][sq_code_bytes]                     */
    int sq_bits, sq_code_bytes;
    float sq_scale, sq_zero;
    const int32_t *graph_to_original; /* optional permutation for refine  */
    /* third-level refine level (error4): the artifact's code blocks as
       stored (4096-byte blocks of e4_vpb rows of e4_code_bytes 4-bit codes,
       low nibble = even dim) and (scale, zero) per row                    */
    const uint8_t *e4_blocks;
    const float *e4_qparams;
    int e4_vpb, e4_code_bytes;
    /* SQ split distance: per-vector bias of an SQMDBIAS artifact [n]     */
    const float *split_bias;
} RvIndex;

/* Learned early-termination predictor: 16 -> 64 -> 64 -> 1 MLP,
 * BF16 weights (stored as raw uint16), FP32 biases, FP32 accumulation. */
#define RV_MLP_IN 16
/* Trace records: per query 2 counters + the scored steps of the final top-K
 * (K <= RV_TRACE_TOPK). */
#define RV_TRACE_TOPK 16
#define RV_TRACE_META (2 + RV_TRACE_TOPK)
#define RV_MLP_HIDDEN 64
typedef struct RvMlp {
    const uint16_t *w1;    /* bf16, grouped [4][64][4]: (input i, unit j)
                              at ((i / 4) * 64 + j) * 4 + i % 4          */
    const float *b1;       /* [64]           */
    const uint16_t *w2;    /* bf16, grouped [16][64][4], same rule       */
    const float *b2;       /* [64]           */
    const uint16_t *w3;    /* [64]      bf16 */
    const float *b3;       /* [1]            */
    const float *feat_mean;/* [16]           */
    const float *feat_inv_std; /* [16]       */
} RvMlp;

/* Suspended searches (RvBatch.phase): the state a query saves when it is
 * suspended and resumes from, per query of the batch, in one block of
 * rv_susp_layout(nq, max_pool).bytes at RvBatch.susp:
 *   keys  [nq][max_pool] u64  pool keys
 *   flags [nq][max_pool] u8   expanded flags
 *   meta  [nq][RV_SUSP_META] i32: pool size, scan position, step, scored /
 *         expanded / duplicate counts, learned-stop state
 *   prob  [nq] f32  predictor's stop probability when suspended
 *   list  [nq] i32  suspended ids in suspension order
 *   order [nq] i32  resumption order
 *   count i32       queries suspended */
#define RV_SUSP_META 16
typedef struct RvSuspLayout {
    size_t flags, meta, prob, list, order, count, bytes;
} RvSuspLayout;
static inline
#ifdef __HIPCC__
__host__ __device__
#endif
RvSuspLayout rv_susp_layout(size_t nq, size_t max_pool) {
    RvSuspLayout l;
    l.flags = nq * max_pool * 8;
    l.meta = (l.flags + nq * max_pool + 15) / 16 * 16;
    l.prob = l.meta + nq * RV_SUSP_META * 4;
    l.list = l.prob + nq * 4;
    l.order = l.list + nq * 4;
    l.count = l.order + nq * 4;
    l.bytes = l.count + 16;
    return l;
}

typedef struct RvBatch {
    const float *queries;        /* [nq][dim]                               */
    int nq;
    const int32_t *pool_sizes;   /* [nq] per-query L, or NULL -> pool_size  */
    const int32_t *num_outs;     /* [nq] per-query k, or NULL -> num_out    */
    int pool_size, num_out;
    int max_pool;                /* LDS is sized for this L                 */
    const int32_t *entries;      /* [n_entries] seed ids, shared by all, or
                                    [nq][entry_stride] per query            */
    int n_entries;
    int entry_stride;            /* 0: shared list; else per-query rows     */
    int expand_count;
    int scorer;                  /* RV_SCORE_*                              */
    int refine_candidates;       /* 0: off; F: refine the top F, output k   */
    int refine_error4;           /* refine adds the error4 level            */
    int hash_log2;               /* visited hash has 1<<hash_log2 slots     */
    int probe_limit;
    uint64_t *visited;           /* exact visited sets in HBM in place of the
                                    LDS hash, or NULL: per persistent wave
                                    (blockIdx.x) a stamp counter word, then
                                    visited_words words of (query stamp << 32
                                    | 32 visited bits)                      */
    int visited_words;           /* ceil(n / 32)                            */
    int visited_exact;           /* host: attach exact sets before launch   */
    int32_t *out_ids;            /* [nq][out_stride]                        */
    float *out_dists;            /* [nq][out_stride]                        */
    int32_t *out_counts;         /* [nq] count or RV_QUERY_* status         */
    int out_stride;
    uint64_t *out_stats;         /* [nq][4] scored, corrected, expanded,
                                    dedup; may be NULL                      */
    int32_t *queue;              /* device work counter, zeroed per launch  */
    const int32_t *nq_dev;       /* optional: live query count in device
                                    memory (<= nq), read at kernel start, so a
                                    captured graph serves any batch size    */
    /* learned stop (stop_mode 0 = off) */
    int stop_mode;               /* 1: infer and stop; 2: record traces     */
    RvMlp mlp;
    float stop_threshold;        /* stop when sigmoid(out) >= threshold     */
    int stop_every;              /* evaluate every stop_every expansions    */
    int stop_min_steps;
    int stop_label_k;            /* top-K whose stability is predicted      */
    float *trace_feats;          /* [nq][trace_cap][16]                     */
    int32_t *trace_steps;        /* [nq][trace_cap]                         */
    int32_t *trace_meta;         /* [nq][RV_TRACE_META]: records, last top-K
                                    change step, then the steps at which the
                                    final top-K members were scored, ascending
                                    (min(K, pool) values, padded with the last) */
    int trace_cap;
    int window;                  /* source turns in flight: 0/1, or 2 (the
                                    live pipeline's lookahead); not with
                                    trace recording                         */
    int waves_per_query;         /* 0: one wave per query, adaptive scoring;
                                    1, 2, 4: the multi-wave kernel           */
    int cand_tile;               /* multi-wave kernel: unvisited neighbors
                                    scored per pass (16, 32, 64)             */
    int phase;                   /* 0: whole searches.  1: a query still
                                    running after suspend_step expansions
                                    saves its state in susp and leaves; 2:
                                    the saved queries resume in susp.order
                                    (queue length *susp.count).  Phases 1
                                    and 2 keep one exact visited set per
                                    query: visited holds nq sets            */
    int suspend_step;
    int stop_probe_only;         /* learned stop: the predictor runs only at
                                    the suspension (its probability orders
                                    the resumption) and stops no search     */
    unsigned char *susp;         /* suspended searches' state block         */
} RvBatch;

/* The reference's SQ traversal with distance-only heaps: the live SQ path
 * (retrieval_sqmd_common.c, fake-DSP turn loop) and the DSP search op
 * (retrieval_async_dsp_search.c:368-509).  One wavefront per query. */
typedef struct RvSqmdCand {
    int32_t id;
    float d;
} RvSqmdCand;

typedef struct RvSqmdBatch {
    const float *queries;        /* [nq][dim]                               */
    int nq;
    const int32_t *pool_sizes;   /* [nq] per-query L, or NULL -> pool_size  */
    const int32_t *num_outs;     /* [nq] per-query k, or NULL -> num_out    */
    int pool_size, num_out, max_pool;
    const int32_t *seeds;        /* shared list, or [nq][seed_stride] rows  */
    int n_seeds, seed_stride;
    int fcap_entries;            /* frontier capacity = fcap_entries +
                                    L * fcap_expand (the reference sizes it
                                    from its entry_points / expand)         */
    int fcap_expand;
    int expand_count;            /* neighbors read per expansion            */
    int max_work;                /* > this many neighbors per turn: the
                                    query fails (the live path's limit 32)  */
    int correction;              /* 0: SQ code distance; 1: SQ + residual   */
    int refine_candidates;       /* > 0: final refine of the first F pool
                                    members from the SQ + residual decode
                                    (retrieval_final_refine_decode_sq)      */
    int refine_error4;           /* the refine adds the error4 level        */
    int32_t *out_ids;            /* [nq][out_stride]                        */
    float *out_dists;
    int32_t *out_counts;         /* count or RV_QUERY_* status              */
    int out_stride;
    uint64_t *out_stats;         /* [nq][4]: work items scored, invalid
                                    neighbor skips, expansions, visited /
                                    duplicate skips; may be NULL            */
    int32_t *queue;
    const int32_t *nq_dev;
    uint32_t *visited;           /* [grid][visited_words] exact bitmaps     */
    int visited_words;           /* multiple of 4                           */
    RvSqmdCand *frontier;        /* [grid][frontier_phys] in HBM, or NULL:
                                    the frontier lives in LDS              */
    int frontier_phys;           /* entries one wave can hold               */
} RvSqmdBatch;

/* LDS bytes per wave for the SQ traversal with the frontier in LDS (0 when
 * it does not fit), the waves per CU, and the launch.  grid_waves <= 0
 * picks the occupancy-derived count; the batch's visited / frontier
 * scratch must cover that many waves (rv_sqmd_grid gives it). */
int rv_sqmd_grid(const RvIndex *index, const RvSqmdBatch *batch, int *lds_frontier);
int rv_sqmd_launch(const RvIndex *index, const RvSqmdBatch *batch, int grid_waves,
                   void *stream);

/* Graph construction on the device, bit-identical to
 * retrieval_search_build_graph_ex with the exact provider and one iteration:
 * exact candidates (the `candidates` closest by (distance, id)), the
 * node's random initial row, RNG occlusion (prune_mode 1) or the closest r
 * (0), fill, random fill, rows by (distance, id).  r <= 64, candidates
 * <= 128.  d_graph [n][r].  Allocates a scratch of n * 512 (candidates <= 64)
 * or n * 1024 bytes, freed after a synchronize of `stream`. */
int rv_build_exact(const float *d_x, int n, int dim, int r, int candidates,
                   uint32_t seed, int prune_mode, int32_t *d_graph, void *stream);
/* The rows of the build above from precomputed candidate keys d_keys
 * [n][keep] ((distance bits << 32) | id, ascending; the first `candidates`
 * used), enqueued on `stream`. */
int rv_rows_from_keys(const float *d_x, int n, int dim, int r, int candidates,
                      uint32_t seed, int prune_mode, const uint64_t *d_keys, int keep,
                      int32_t *d_graph, void *stream);
/* rv_build_exact's graph computed over a caller-supplied clustering (d_assign
 * [n] cluster ids in [0, nc), d_centroids [nc][dim]): own-cluster candidate
 * lists, then every other cluster that a triangle-inequality bound does not
 * rule out, so the graph is the brute-force one for any clustering.
 * out_stats (may be NULL) [4]: node blocks, cluster pairs screened in,
 * (block, cluster) pairs searched, blocks with such pairs.  Synchronous. */
int rv_build_exact_clustered(const float *d_x, int n, int dim, int r, int candidates,
                             uint32_t seed, int prune_mode, const int32_t *d_assign,
                             const float *d_centroids, int nc, int32_t *d_graph,
                             int64_t *out_stats, void *stream);
/* RNN-Descent on the device (synchronous): out d_graph [n][top_k]. */
int rv_build_rnn(const float *d_x, int n, int dim, int s, int r, int t1, int t2,
                 uint32_t seed, int top_k, int32_t *d_graph, void *stream);
void rv_set_error(const char *msg);

int rv_device_count(void);
int rv_set_device(int device);
int rv_malloc(void **d_ptr, size_t bytes);
int rv_free(void *d_ptr);
int rv_memcpy_h2d(void *d_dst, const void *src, size_t bytes);
int rv_memcpy_d2h(void *dst, const void *d_src, size_t bytes);
int rv_memcpy_d2d(void *d_dst, const void *d_src, size_t bytes);
int rv_synchronize(void);
const char *rv_last_error(void);
int rv_event_create(void **ev);
int rv_event_record(void *ev, void *stream);
int rv_event_elapsed(void *start, void *stop, float *ms);
int rv_event_destroy(void *ev);
int rv_stream_create(void **stream);
int rv_stream_destroy(void *stream);
int rv_stream_synchronize(void *stream);
/* Copies ordered on `stream` (asynchronous for page-locked host memory). */
int rv_memcpy_h2d_async(void *d_dst, const void *src, size_t bytes, void *stream);
int rv_memcpy_d2h_async(void *dst, const void *d_src, size_t bytes, void *stream);
/* Page-locked host memory. */
int rv_host_alloc(void **p, size_t bytes);
int rv_host_free(void *p);
/* An event without timing data, and a device-side wait of `stream` for it. */
int rv_event_create_untimed(void **ev);
int rv_stream_wait_event(void *stream, void *ev);

/* Decoded vectors d_out [count][dim] of the ids d_ids: sq = 0 the PQ +
 * residual reconstruction, sq = 1 the SQ + residual one.  Asynchronous. */
int rv_decode_vectors(const RvIndex *index, int sq, const int32_t *d_ids, int count,
                      float *d_out, void *stream);

/* Fixed per-device workspace for the no-allocation batch entry point: the
 * work counter at offset 0, then at most RV_NOALLOC_MAX_SEEDS seed ids at
 * RV_NOALLOC_SEED_OFFSET.  rv_noalloc_scratch makes `device` current once,
 * resolves the workspace's device address with hipGetSymbolAddress, keeps
 * it, and returns it (0) or -1. */
#define RV_NOALLOC_SEED_OFFSET 256
#define RV_NOALLOC_MAX_SEEDS 16384
#define RV_NOALLOC_BYTES (RV_NOALLOC_SEED_OFFSET + 4 * RV_NOALLOC_MAX_SEEDS)
int rv_noalloc_scratch(int device, void **d_scratch);
/* The graph entry + LCG seed list (rpc_demo_pure_pq.c:192-208) of `count`
 * (<= n) ids, written to d_out by a kernel on `stream`. */
/* Resumption order of the suspended queries: ascending stop probability
 * (the longest expected first), 4096 probability buckets. */
int rv_susp_order(unsigned char *susp, int nq, int max_pool, void *stream);

/* Stable LSD radix sort of n 64-bit keys on bits [begin_bit, end_bit)
 * (bits outside are ignored for the order), with optional 32-bit values;
 * the inputs are left as they are.  scratch: rv_radix_scratch_bytes(n,
 * values given) bytes of device memory. */
size_t rv_radix_scratch_bytes(size_t n, int with_values);
int rv_radix_sort(const uint64_t *keys_in, uint64_t *keys_out, const uint32_t *vals_in,
                  uint32_t *vals_out, size_t n, int begin_bit, int end_bit, void *scratch,
                  size_t scratch_bytes, void *stream);

int rv_seed_list(int32_t *d_out, int n, int count, int graph_entry, uint32_t seed,
                 void *stream);

/* Bytes of LDS one query slot needs, and how many slots fit per CU. */
size_t rv_search_lds_bytes(const RvIndex *index, const RvBatch *batch);
int rv_search_slots_per_cu(const RvIndex *index, const RvBatch *batch);

/* Launches the persistent search over the batch on `stream` (NULL = the
 * default stream).  grid_waves <= 0 picks the occupancy-derived count.  The
 * call is asynchronous; it zeroes batch->queue with a kernel first. */
int rv_search_launch(const RvIndex *index, const RvBatch *batch,
                     int grid_waves, void *stream);

/* Encoders for index building on the device (bit-identical to the
 * reference encoders). */
int rv_pq_encode(const float *d_x, int n, int dim, int m, int ks,
                 const float *d_codebook, uint8_t *d_codes, void *stream);
int rv_residual_encode(const float *d_x, const uint8_t *d_codes, int n,
                       int dim, int m, int ks, const float *d_codebook,
                       int bits, uint8_t *d_res, float *d_qparams,
                       void *stream);
int rv_error4_encode(const float *d_x, const uint8_t *d_codes,
                     const uint8_t *d_res, const float *d_res_qparams, int n,
                     int dim, int m, int ks, const float *d_codebook,
                     int res_bits, uint8_t *d_e4, float *d_e4_qparams,
                     void *stream);

/* MLP training kernels (FP32 master weights, BF16-rounded forward).
 * Flat parameter layout: w1[64][16] | b1[64] | w2[64][64] | b2[64] | w3[64] |
 * b3[1].  Gradients and losses are accumulated (callers zero them). */
int rv_mlp_param_count(void);
/* d_slab: the caller's partial-sum workspace of rv_mlp_slab_floats() floats;
 * calls that may overlap need separate slabs.  d_idx (optional) gathers the
 * samples. */
#define RV_MLP_SLAB_ROWS 1024
size_t rv_mlp_slab_floats(void);
int rv_mlp_fwd_bwd_sampled(const float *d_params, const float *d_x,
                           const float *d_y, const float *d_weight,
                           const int32_t *d_idx, int n, float *d_grad,
                           float *d_loss, float *d_out, float *d_slab, void *stream);
int rv_mlp_adam(float *d_params, const float *d_grad, float *d_m1,
                float *d_m2, int step, float lr, float beta1, float beta2,
                float eps, float grad_scale, void *stream);
int rv_mlp_to_bf16(const float *d_params, uint16_t *d_bf16, void *stream);
/* Labels/weights from search traces: y = 1 once all but `slack` of the
 * final top-K members have been scored (slack 0: the top-K is final),
 * weight = pos_w or neg_w for valid records, 0 for padding. */
int rv_trace_labels(const int32_t *d_steps, const int32_t *d_meta, int nq,
                    int cap, int label_k, int slack, float pos_w, float neg_w,
                    float *d_y, float *d_weight, void *stream);
int rv_normalize(float *d_x, size_t n, const float *d_mean,
                 const float *d_inv_std, void *stream);
/* n pseudo-random sample ids in [0, population) from (seed, step). */
int rv_sample_indices(int32_t *d_idx, int n, size_t population, uint32_t seed,
                      uint32_t step, void *stream);
const char *rv_learned_error(void);
uint16_t rv_bf16_round_host(float f);
int rv_fill_zero(void *d_ptr, size_t bytes, void *stream);
int rv_copy(void *d_dst, const void *d_src, size_t bytes, void *stream);

/* Entry router: for each query the S (<= 16) nearest of nc centroids by
 * exact squared L2, written as their representative node ids
 * (out[nq][S], nearest first).  scratch holds nq*nc floats. */
/* HIP graph capture of a launch sequence on `stream` (not the default
 * stream): begin, launch as usual, end -> instantiated executable. */
int rv_graph_begin(void *stream);
int rv_graph_end(void *stream, void **exec);
int rv_graph_launch(void *exec, void *stream);
int rv_graph_destroy(void *exec);
/* Sets the search kernel's attributes for this batch ahead of a capture and
 * returns the occupancy-derived grid (waves) it would use. */
int rv_search_prepare(const RvIndex *index, const RvBatch *batch);

int rv_route_entries(const float *d_queries, int nq, int dim,
                     const float *d_centroids, int nc, const int32_t *d_rep_ids,
                     int S, float *d_scratch, int32_t *d_out, void *stream);

#ifdef __cplusplus
}
#endif
#endif
