/*
 * rover_gpu.h -- C-ABI of the MI355X quantized graph search.
 *
 * Three faces, all plain C (pointers and sizes, no torch or HIP types):
 *
 *  1. The reference's FastRPC operator surface (calculator.idl), with the
 *     qaic C mapping: `in sequence<T> x` -> `const T *x, int xLen`, `rout`
 *     sequences -> `T *x, int xLen`, `rout int32` -> `int32_t *`.  Ops that
 *     take file descriptors of shared buffers in the reference take host
 *     pointers here (the data is copied into HBM once, at load).
 *  2. The reference's library call for one query,
 *     retrieval_pure_pq_search_scalar[_with_policy]
 *     (retrieval_sqmd_pure_pq.h:264-278), same structs and return contract.
 *  3. Batch extensions with no reference counterpart: many queries with
 *     per-query beam width L and k in one persistent kernel, from host or
 *     device buffers, plus the learned early-termination model.
 *
 * Return values of calculator_* and rover_* ops are AEEResult-style codes
 * (ROVER_SUCCESS == AEE_SUCCESS == 0); the library search returns a count or
 * -1, as the reference does.
 */
#ifndef ROVER_GPU_H
#define ROVER_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int AEEResult;
typedef uint64_t remote_handle64;

#define ROVER_SUCCESS 0    /* AEE_SUCCESS   */
#define ROVER_EFAILED 1    /* AEE_EFAILED   */
#define ROVER_ENOMEMORY 2  /* AEE_ENOMEMORY */
#define ROVER_EBADPARM 14  /* AEE_EBADPARM  */

#define ROVER_ABI_VERSION 1        /* search-only, as calculator_rover_configure */
#define ROVER_MAX_POOL 2000        /* rover_limits.h:40 DSP_SEARCH_MAX_POOL */
/* Graph rows of up to this many slots load (the reference checks none);
   a search reads min(expand_count, degree) of them per expansion, at most
   64 (one wavefront's lanes): a wider expansion is refused with
   ROVER_EBADPARM, the SQ heap traversals take at most 32. */
#define ROVER_MAX_DEGREE 1024
#define ROVER_ENTRY_POINT_SEED 42u /* rover_limits.h:38, config.c:18 */

/* ---------------------------------------------------------------- 1. IDL */

/* Session open/close (src/app/dsp.c:18-31).  `uri` is ignored beyond
 * selecting device 0; each handle owns its own loaded indexes. */
int calculator_open(const char *uri, remote_handle64 *handle);
int calculator_close(remote_handle64 handle);

/* calculator.idl:15; main.c:28 -> dsp.c:8-17.  abi_version must be
 * ROVER_ABI_VERSION; l1_bits (4 or 8) selects the SQ code width. */
AEEResult calculator_rover_configure(remote_handle64 h, int32_t abi_version,
                                     int32_t l1_bits);

/* calculator.idl:36-45 (retrieval_load): SQ index = dense graph
 * int32[num_vectors*graph_k] and global-scale l1 codes. */
AEEResult calculator_retrieval_load(remote_handle64 h, const int32_t *graph,
                                    int graphLen, const uint8_t *l1_codes,
                                    int l1_codesLen, int32_t num_vectors,
                                    int32_t dim, int32_t graph_k,
                                    int32_t graph_entry_point, float l1_scale,
                                    float l1_zero_point);

/* calculator.idl:46-55 (retrieval_search) on the SQ index.  Seeds are the
 * graph entry plus the LCG of retrieval_async_dsp_search.c:219-230 with
 * seed 42.  Candidates are ordered by (distance, id). */
AEEResult calculator_retrieval_search(remote_handle64 h, const float *query,
                                      int queryLen, int32_t entry_points,
                                      int32_t expand_count, int32_t pool_size,
                                      int32_t num_out, int32_t *candidate_ids,
                                      int candidate_idsLen,
                                      float *candidate_dists,
                                      int candidate_distsLen,
                                      int32_t *actual_count);

/* calculator.idl:56 */
AEEResult calculator_retrieval_unload(remote_handle64 h);

/* calculator.idl:59-70 (retrieval_load_dma; retrieval_async_dsp_search.c:
 * 558-666).  The IDL's file descriptors become host pointers, copied into
 * HBM: `graph` int32[num_vectors*graph_k], or NULL where the reference
 * passes RETRIEVAL_DMA_NO_GRAPH_FD (-1; the graph stays with the caller:
 * give it with rover_sq_graph_load before searching), and l1_codesLen
 * (1..ROVER_MAX_L1_CHUNKS) chunk base pointers, chunk c holding
 * min(l1_chunk_vectors, num_vectors - c*l1_chunk_vectors) code rows.  Both
 * SQ loads are dropped first; an entry point outside [0, num_vectors)
 * means none (seeds from the LCG only). */
#define ROVER_MAX_L1_CHUNKS 16 /* retrieval_rpc_demo.h:17-18 MAX_L1_CHUNKS */
AEEResult calculator_retrieval_load_dma(remote_handle64 h, const int32_t *graph,
                                        const uint8_t *const *l1_codes,
                                        int l1_codesLen, int32_t l1_chunk_vectors,
                                        int32_t num_vectors, int32_t dim,
                                        int32_t graph_k, int32_t graph_entry_point,
                                        float l1_scale, float l1_zero_point);

/* calculator.idl:71 (retrieval_unload_dma): drops an index loaded by
 * retrieval_load_dma (one from retrieval_load stays loaded). */
AEEResult calculator_retrieval_unload_dma(remote_handle64 h);

/* The SQ graph for an index loaded without one (no reference op: the
 * reference keeps that graph on the host, retrieval_rpc_demo_pipeline.c:
 * 3066-3077, and the GPU needs it in HBM). */
AEEResult rover_sq_graph_load(remote_handle64 h, const int32_t *graph,
                              int64_t graphLen, int32_t num_vectors,
                              int32_t graph_k, int32_t graph_entry_point);

/* calculator.idl:119-131 (retrieval_pure_pq_load_dma): the IDL's
 * `fd_pq_codes` sequence becomes the chunk base pointers. */
AEEResult calculator_retrieval_pure_pq_load_dma(
    remote_handle64 h, const uint8_t *const *pq_codes, int pq_codesLen,
    const int32_t *pq_chunk_vector_counts, int pq_chunk_vector_countsLen,
    int32_t pq_chunk_vectors, const float *pq_codebook, int pq_codebookLen,
    int32_t num_vectors, int32_t dim, int32_t pq_m, int32_t pq_ks,
    int32_t pq_subdim, int32_t pq_code_bytes, const uint8_t *index_uuid,
    int index_uuidLen);

/* calculator.idl:132 */
AEEResult calculator_retrieval_pure_pq_unload_dma(remote_handle64 h);

/* The pure-PQ graph stays host-owned in the reference
 * (rpc_demo_pure_pq.c:491-497 demo_resolve_graph_index); on the GPU it is
 * resident in HBM next to the codes. */
AEEResult rover_pure_pq_graph_load(remote_handle64 h, const int32_t *graph,
                                   int64_t graphLen, int32_t num_vectors,
                                   int32_t graph_k, int32_t graph_entry_point);

/* Residual index for online correction and final refine
 * (retrieval_pq_residual_index_open, retrieval_pq_residual.c:716-835),
 * given as flat arrays: codes [num_vectors][res_code_bytes] and
 * qparams [num_vectors][2] = (scale, zero_point). */
AEEResult rover_pure_pq_residual_load(remote_handle64 h,
                                      const uint8_t *residual_codes,
                                      int64_t residual_codesLen,
                                      const float *qparams, int64_t qparamsLen,
                                      int32_t residual_bits);

/* Per-vector residuals of the SQ index for the corrected SQ traversal:
 * codes [num_vectors][(dim+1)/2 or dim] (bits 4 / 8, low nibble = even dim)
 * and (scale, zero) [num_vectors][2] -- the residual and qparams payloads of
 * encode_l1_residual.py (the reference's residual block file flattened). */
AEEResult rover_sq_residual_load(remote_handle64 h, const uint8_t *codes,
                                 int64_t codesLen, const float *qparams,
                                 int64_t qparamsLen, int32_t bits);

/* Split-distance bias for the SQ traversal (replaces the reference's
 * sqmd_load_correction_bias, retrieval_rpc_demo_support.c:237-299, which
 * reads SQMD_CORRECTION_BIAS_PATH): the image of an SQMDBIAS file written by
 * generate_sqmd_correction_bias.py -- 64-byte header, then f32 bias
 * [num_vectors].  Needs the SQ index with 4-bit codes and its residual
 * loaded; the header must match them (vectors, dim, bits), the payload's
 * FNV-1a64 and, when source_fingerprint is not NULL, its 16 bytes (the
 * first 16 of SHA-256 over the SHA-256 digests of the l1, residual and
 * qparams files).  Searches with correction = 2 then rank by the split
 * distance.  Reloading the residual drops the bias. */
AEEResult rover_sq_split_bias_load(remote_handle64 h, const uint8_t *image,
                                   int64_t imageLen,
                                   const uint8_t *source_fingerprint);

/* Entry router (no reference counterpart): `count` centroids
 * [count][dim] and one representative node id per centroid.  A batch with
 * route_entries = S seeds each query with the representatives of its S
 * nearest centroids (exact squared L2, ties by centroid index), passed to
 * the search as that query's entry ids.  Routing refuses (ROVER_EBADPARM)
 * a router of more than ROVER_ROUTER_MAX_CENTROIDS centroids: larger ones
 * are not tested on the device; route larger sets on the host and pass
 * entry_ids rows. */
#define ROVER_ROUTER_MAX_CENTROIDS 256
AEEResult rover_pure_pq_router_load(remote_handle64 h, const float *centroids,
                                    int centroidsLen, const int32_t *rep_ids,
                                    int count);

/* The routing step alone, device pointers: out[nq][S] node ids. */
AEEResult rover_route_entries(remote_handle64 h, const float *d_queries,
                              int num_queries, int S, int32_t *d_out,
                              void *stream);

/* Optional graph->original id map used by the refine tie-break
 * (retrieval_final_refine.c:212-219); NULL restores the identity. */
AEEResult rover_pure_pq_original_ids(remote_handle64 h, const int32_t *map,
                                     int mapLen);

/* Third-level refine artifact (ERROR4 mode of retrieval_final_refine_open_pq,
 * retrieval_final_refine.c:98-190; written by retrieval_build_error4,
 * retrieval_build_artifacts.c:25-66, and pq_refine_artifacts.py:47-65),
 * given as the whole PQERR4V1 file image: 256-byte header, (scale, zero)
 * [N][2] at 256, then 4096-byte blocks of 4096 / ((dim+1)/2) rows of
 * (dim+1)/2 bytes of 4-bit codes (low nibble = even dim).  Checked as the
 * reference's open checks it: magic, version, N and dim against the loaded
 * PQ index, bit widths, block size, section sizes, the payload digest at
 * +56, every (scale, zero) finite with scale >= 0, and, when expected_sha256
 * is not NULL, the whole-file SHA-256 (RetrievalFinalRefineConfig.
 * artifact_sha256).  The digests of the caller's base files (+88) and
 * codebook (+184) bind the artifact to files and are the caller's to check;
 * the id-map digest (+216) is checked against rover_pure_pq_original_ids at
 * each ERROR4 launch.  Returns ROVER_EBADPARM for an artifact the reference
 * would reject. */
AEEResult rover_pure_pq_error4_load(remote_handle64 h, const uint8_t *artifact,
                                    int64_t artifactLen,
                                    const uint8_t *expected_sha256);
AEEResult rover_pure_pq_error4_unload(remote_handle64 h);

/* Third-level refine artifact of the SQ index (SQERR4V1; the SQ path opens
 * it with retrieval_final_refine_open and no codebook,
 * retrieval_rpc_demo_pipeline.c:3026-3045, retrieval_final_refine.c:98-190;
 * written by retrieval_build_error4 with pq = 0,
 * retrieval_build_artifacts.c:25-66): the whole file image, checked as for
 * PQERR4V1 against the loaded SQ index; the SQ index has no id maps, so the
 * artifact must be bound to none.  SQ traversal batches with
 * refine_candidates = F > 0 re-rank their first F results in FP64 from the
 * SQ + residual decode (retrieval_final_refine_decode_sq, needs
 * rover_sq_residual_load), refine_mode = ROVER_REFINE_ERROR4 adding this
 * artifact's level. */
AEEResult rover_sq_error4_load(remote_handle64 h, const uint8_t *artifact,
                               int64_t artifactLen, const uint8_t *expected_sha256);
AEEResult rover_sq_error4_unload(remote_handle64 h);

/* Reconstructed vectors out [count][dim] (host buffers) of the loaded
 * index: the PQ + residual decode (retrieval_pq_residual_query_decode,
 * retrieval_pq_residual.c:954-981) or the SQ + residual decode
 * (retrieval_final_refine_decode_sq, retrieval_final_refine.c:55-67), each
 * step a rounded f32 as in the reference.  Ids out of range give
 * ROVER_EBADPARM. */
AEEResult rover_pure_pq_decode(remote_handle64 h, const int32_t *ids, int count,
                               float *out);
AEEResult rover_sq_decode(remote_handle64 h, const int32_t *ids, int count,
                          float *out);

/* Refine traces of the reference demos.  ROVER_TRACE_TEXT is the pure-PQ
 * demo's text trace (retrieval_rpc_demo_pure_pq.c:535-538, 615-645): per
 * query "Q pool query count budget", a "C id dist" line per search result,
 * a "V id v0 v1 ..." line with the decoded vector of each of the first
 * decoded_count results (the demo decodes min(count, budget or top-k)), and
 * an "O id dist" line per output, numbers as C's %a of a double.
 * ROVER_TRACE_P15 is the SQ pipeline's binary trace (retrieval_rpc_demo_
 * pipeline.c:3047-3050, 3276-3308): "P15TRC01", then per query u32 (entry
 * points, pool, query, count), the ids and f32 distances of the search
 * results, u32 n and the n refined ids and distances.  No reference
 * counterpart as functions: the demos write them inline. */
enum { ROVER_TRACE_TEXT = 1, ROVER_TRACE_P15 = 2 };
AEEResult rover_trace_open(const char *path, int format, void **trace);
AEEResult rover_trace_query(void *trace, int entry_points, int pool_size,
                            int query_index, int refine_budget, const int32_t *ids,
                            const float *dists, int count, const float *decoded,
                            int decoded_count, int dim, const int32_t *out_ids,
                            const float *out_dists, int out_count);
AEEResult rover_trace_close(void *trace);

/* SHA-256 (FIPS 180-4) of a buffer, as the artifact digests use. */
void rover_sha256(const void *data, size_t bytes, uint8_t out[32]);

/* ---------------------------------------------------- 2. library search */

typedef struct PurePqSearchIndexView {
    const int32_t *graph;
    int graph_degree;
    int num_vectors;
    const uint8_t *const *code_chunks;
    const int *chunk_vector_counts;
    int chunk_count;
    int chunk_vectors;
    const float *codebook;
    int dim;
    int m;
    int ks;
    int subdim;
} PurePqSearchIndexView;

typedef struct PurePqSearchRequest {
    const float *query;
    const int *entry_ids;
    int entry_count;
    int expand_count;
    int pool_capacity;
    int num_out;
} PurePqSearchRequest;

typedef struct PurePqSearchStats {
    uint64_t pq_scored_candidates;
    uint64_t corrected_candidates;
    uint64_t expanded_nodes;
    uint64_t visited_dedup_skips;
} PurePqSearchStats;

typedef enum PurePqCorrectionMode {
    PURE_PQ_CORRECTION_OFF = 0,
    PURE_PQ_CORRECTION_ONLINE = 1
} PurePqCorrectionMode;

typedef int (*PurePqDistanceEvaluatorFn)(void *context, int vector_id,
                                         float coarse_distance,
                                         float *out_distance);

typedef struct PurePqDistancePolicy {
    PurePqCorrectionMode mode;
    PurePqDistanceEvaluatorFn evaluator;
    void *context;
} PurePqDistancePolicy;

/* The one ONLINE evaluator the GPU recognises: set policy.evaluator to
 * rover_pq_residual_evaluator and policy.context to a RoverResidualView.
 * The function itself only marks the policy; called on the host it
 * returns -1.  Any other evaluator makes the search return -1. */
typedef struct RoverResidualView {
    const uint8_t *codes;   /* [num_vectors][res_code_bytes] */
    const float *qparams;   /* [num_vectors][2] */
    int bits;               /* 4 or 8 */
} RoverResidualView;
int rover_pq_residual_evaluator(void *context, int vector_id,
                                float coarse_distance, float *out_distance);

/* retrieval_sqmd_pure_pq.h:264-278.  Runs on the GPU of a process-wide
 * session; the index view is copied to HBM on first use and kept while the
 * same view (pointers and shape) is passed again.  Call
 * rover_pure_pq_view_release() before changing an index in place. */
int retrieval_pure_pq_search_scalar_with_policy(
    const PurePqSearchIndexView *index, const PurePqSearchRequest *request,
    const PurePqDistancePolicy *policy, int *out_ids, float *out_distances,
    int out_capacity, PurePqSearchStats *out_stats);
int retrieval_pure_pq_search_scalar(const PurePqSearchIndexView *index,
                                    const PurePqSearchRequest *request,
                                    int *out_ids, float *out_distances,
                                    int out_capacity,
                                    PurePqSearchStats *out_stats);
void rover_pure_pq_view_release(void);

/* ------------------------------------------------------------ 3. batch */

enum { ROVER_METHOD_PQ = 0, ROVER_METHOD_SQ = 1 };

/* Per-query status in out_counts when negative. */
enum {
    ROVER_QUERY_BAD_ROW = -2,
    ROVER_QUERY_BAD_ENTRY = -3,
    ROVER_QUERY_NONFINITE = -4,
    ROVER_QUERY_BAD_PARAM = -5
};

/* RoverSearchBatch.hash_log2: exact visited sets in HBM -- per persistent
 * wave n/4 bytes of stamped words (query stamp, 32 visited bits), tested
 * and set by compare-and-swap; nothing is cleared between queries */
#define ROVER_HASH_EXACT (-1)

typedef struct RoverSearchBatch {
    const float *queries;        /* [num_queries][dim]                      */
    int num_queries;
    const int32_t *pool_sizes;   /* per-query L, or NULL -> pool_size       */
    const int32_t *num_outs;     /* per-query k, or NULL -> num_out         */
    int pool_size;
    int num_out;
    int method;                  /* ROVER_METHOD_PQ / ROVER_METHOD_SQ       */
    int correction;              /* PQ: 0 off (ADC ranking), 1 online       */
    int refine_candidates;       /* PQ: 0 off, else refine the top F        */
    int entry_points;            /* seeds: graph entry + LCG                */
    uint32_t seed;               /* 0 -> ROVER_ENTRY_POINT_SEED             */
    const int32_t *entry_ids;    /* explicit seeds (PurePqSearchRequest.
                                    entry_ids): one list shared by all
                                    queries, or [num_queries][entry_stride]
                                    rows; NULL -> the LCG seeds above       */
    int entry_count;
    int entry_stride;            /* 0: shared list; > 0: per-query rows     */
    int route_entries;           /* > 0: each query's entries are the S
                                    nearest router centroids' nodes
                                    (rover_pure_pq_router_load); overrides
                                    entry_ids                               */
    int expand_count;
    int32_t *out_ids;            /* [num_queries][out_stride]               */
    float *out_dists;            /* [num_queries][out_stride]               */
    int32_t *out_counts;         /* [num_queries]                           */
    int out_stride;
    uint64_t *out_stats;         /* [num_queries][4] or NULL                */
    /* launch shape; 0 -> tuned default.  hash_log2: the LDS visited hash
       has 2^hash_log2 slots (4..16); ROVER_HASH_EXACT: exact visited sets
       in HBM instead (kept by the session), so no candidate is scored
       twice                                                                 */
    int hash_log2;
    int probe_limit;
    int grid_waves;
    /* learned early termination (needs rover_learned_stop_set) */
    int learned_stop;
    float stop_threshold;        /* 0 -> 0.9 */
    /* refine level (RETRIEVAL_REFINE_CURRENT / _ERROR4,
       retrieval_final_refine.h:7-8): 0 or ROVER_REFINE_CURRENT decodes PQ +
       residual; ROVER_REFINE_ERROR4 adds the error4 level
       (needs rover_pure_pq_error4_load)                                    */
    int refine_mode;
    /* traversal (ROVER_TRAVERSAL_*): 0 = the method's reference traversal
       (PQ: best-first in (distance, id) order; SQ: the live SQ traversal)  */
    int traversal;
    /* source window of the best-first traversal: 0 or 1 = the scalar
       search (retrieval_sqmd_pure_pq.c:793-969); 2 = the live pipeline with
       two source turns in flight (pure_pq_source_pipeline_step,
       retrieval_sqmd_pure_pq.c:722-762, SQMD_MAX_INFLIGHT_SOURCES=2 at
       rpc_demo_pure_pq.c:422-431): each source is chosen before the previous
       turn's candidates are inserted, which gives that pipeline's results
       (not the scalar search's).  Not with the SQ heap traversals.          */
    int source_window;
    /* waves per query (no reference counterpart): 0 = one wavefront per
       query with adaptive scoring (the default kernel); 1, 2 or 4 = the
       multi-wave kernel (online correction on 32 x 4-dim PQ with at most 32
       neighbors per expansion, source window 1): a workgroup of that many
       wavefronts shares one query's pool, wave 0 expands and merges, and all
       waves score the unvisited neighbors, candidate_tile (16, 32 or 64; 0 =
       32) of them per pass.  Results are identical to the default kernel's.
       grid_waves then counts workgroups.                                    */
    int waves_per_query;
    int candidate_tile;
    /* resumption order (no reference counterpart): 0 = off; N > 0 = every
       search still running after N expansions is suspended (its pool, flags,
       counters and learned-stop state saved), and the suspended searches
       resume in order of the learned-stop predictor's stop probability at
       that point, lowest (the longest expected) first, so that long searches
       do not start last.  Results and counters are those of the unsuspended
       search.  Needs exact visited sets (hash_log2 = ROVER_HASH_EXACT; one
       set per query of the batch, ceil(n / 32) + 1 words of 8 bytes each),
       source window 1, one wave per query and a learned-stop model
       (rover_learned_stop_set or _train); with learned_stop = 0 the
       predictor only orders the resumption and stops no search.            */
    int reorder_step;
} RoverSearchBatch;

enum { ROVER_REFINE_CURRENT = 1, ROVER_REFINE_ERROR4 = 2 };

enum {
    ROVER_TRAVERSAL_DEFAULT = 0,
    /* best-first over a pool ordered by (distance, id)
       (retrieval_sqmd_pure_pq.c:793-969); any method                      */
    ROVER_TRAVERSAL_BEST_FIRST = 1,
    /* SQ only: the live SQ traversal (retrieval_sqmd_common.c, fake-DSP turn
       loop :2052-2152): distance-only heaps, seeds from
       retrieval_sqmd_make_seed_ids (entry_points, seed), at most 32
       neighbors per expansion; correction = 1 ranks by the exact SQ +
       residual distance (needs rover_sq_residual_load); correction = 2 by
       the split distance (retrieval_rpc_demo_pipeline.c:936-1000: the u8
       query on the SQ grid refined 17x, coarse u8 x u4 distance plus bias -
       2 <query, residual>; needs rover_sq_split_bias_load).  entry_ids and
       route_entries are not accepted (ROVER_EBADPARM).                    */
    ROVER_TRAVERSAL_SQMD = 2,
    /* SQ only: the DSP search op (retrieval_async_dsp_search.c:368-509):
       the same heaps, seeds graph entry + LCG (rover_make_seed_ids); here
       also at most 32 neighbors per expansion (ROVER_EBADPARM beyond)     */
    ROVER_TRAVERSAL_SQ_DSP = 3
};

/* Host buffers in and out; returns when results are in place. */
AEEResult rover_search_batch(remote_handle64 h, const RoverSearchBatch *b);

/* Every pointer of `b` is a device (HBM) address; asynchronous on `stream`
 * (a hipStream_t, NULL = default stream). */
AEEResult rover_search_batch_device(remote_handle64 h,
                                    const RoverSearchBatch *b, void *stream);

/* rover_search_batch_device that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every pointer of `b` is a device address, as above; the library's own
 * workspace (the work counter and the graph entry + LCG seed list, at most
 * ROVER_NOALLOC_MAX_SEEDS entry points) is a fixed array on each device,
 * ROVER_NOALLOC_SCRATCH_BYTES large, held until the process ends.  Calls on
 * one device share that array, and it keeps the last call's contents:
 * calls that overlap on one GPU (two streams, two threads, two handles)
 * race and must be ordered by the caller.  Best-first traversal with the
 * LDS visited hash only: routed entries (route_entries > 0), exact visited
 * sets (ROVER_HASH_EXACT) and the SQ heap traversals return ROVER_EBADPARM,
 * as do more entry points than the array holds. */
#define ROVER_NOALLOC_MAX_SEEDS 16384
#define ROVER_NOALLOC_SCRATCH_BYTES (256 + 4 * ROVER_NOALLOC_MAX_SEEDS)
AEEResult rover_search_batch_device_noalloc(remote_handle64 h,
                                            const RoverSearchBatch *b,
                                            void *stream);

/* Batch search captured in a HIP graph (no reference counterpart): the
 * work-queue reset and the search kernel of one device batch, captured on
 * `stream` (a stream from rover_stream_create, not the default stream) and
 * instantiated.  The graph keeps every pointer of `b` and reads the buffers
 * at each replay, so callers refill queries / per-query L and k / entry rows
 * in place and replay.  b->num_queries is the capacity; with
 * d_num_queries != NULL the kernel reads the live count (<= capacity) from
 * that device int32 at each replay, so one graph serves any batch size.
 * The graph has its own work counter; it stays valid while the session's
 * index is loaded.  route_entries is not supported (pass entry rows). */
AEEResult rover_search_graph_create(remote_handle64 h, const RoverSearchBatch *b,
                                    const int32_t *d_num_queries, void *stream,
                                    void **graph);
AEEResult rover_search_graph_launch(void *graph, void *stream);
AEEResult rover_search_graph_destroy(void *graph);

/* Autotuner (no reference counterpart): times the device batch `b` under
 * launch configurations -- visited set (LDS hash of 2^9, 2^10, 2^12, 2^14 slots or
 * the default, or ROVER_HASH_EXACT), probe limit (4, 8, 16) and grid fill (all
 * query slots of every CU, 3/4, 1/2); then, with the best of those and where
 * the batch allows it (b->waves_per_query == 0, online correction on
 * 32 x 4-dim PQ, source window 1), waves per query (1, 2, 4) x candidate
 * tile (16, 32, 64) of the multi-wave kernel -- `repeats` launches each
 * after one warm-up, keeps only configurations whose ids, distances and
 * counts equal the default configuration's, and returns the fastest.
 * Results are left in b's output buffers. */
typedef struct RoverTuning {
    int hash_log2;
    int probe_limit;
    int grid_waves;
    float ms;          /* mean launch time of the chosen configuration */
    float default_ms;  /* the same for the default configuration       */
    int tried;         /* configurations timed                         */
    int rejected;      /* configurations whose results differed        */
    int waves_per_query;  /* 0: the one-wave kernel; else 1, 2 or 4     */
    int candidate_tile;   /* with waves_per_query > 0: 16, 32 or 64     */
} RoverTuning;
AEEResult rover_autotune(remote_handle64 h, const RoverSearchBatch *b, void *stream,
                         int repeats, RoverTuning *out);

/* ------------------------------------------- learned early termination */

/* Trains the stop predictor (16 -> 64 -> 64 -> 1 MLP, BF16 weights, FP32
 * accumulation) on traces of plain searches of `queries` at beam width
 * pool_size, and installs it on the session.  A record taken every
 * stop_every expansions is labelled 1 when the search's top label_k is
 * already final.  No reference counterpart. */
typedef struct RoverLearnedStopTrain {
    const float *queries;     /* host [num_queries][dim] training queries */
    int num_queries;
    int pool_size;            /* L of the trace searches                  */
    int correction;           /* 0: ADC ranking, 1: online correction     */
    int label_k;              /* top-K whose stability is predicted (<= 16) */
    int label_slack;          /* label 1 once all but label_slack of the
                                 final top-K have been scored (0: final)   */
    int entry_points;
    const int32_t *entry_ids; /* host seed list, or NULL -> LCG seeds      */
    int entry_count;
    int entry_stride;         /* 0: entry_ids is one shared list; > 0: rows
                                 [num_queries][entry_stride]               */
    int route_entries;        /* > 0: route each training query (as in a
                                 search batch); overrides entry_ids        */
    int expand_count;
    int stop_every;           /* feature/decision cadence (expansions)    */
    int min_steps;            /* no decision before this many expansions  */
    int trace_cap;            /* records kept per query (<= 4096)         */
    int steps;                /* Adam steps                               */
    int batch;                /* samples per step                         */
    float lr;
    float pos_weight;         /* weight of label-1 samples (0 -> 1)       */
    uint32_t seed;
    /* outputs */
    float final_loss;         /* mean weighted BCE over the last steps    */
    long samples;             /* trace records                            */
    float positive_frac;
} RoverLearnedStopTrain;
AEEResult rover_learned_stop_train(remote_handle64 h, RoverLearnedStopTrain *t);

/* The plain search of a device batch (pointers in HBM, asynchronous on
 * `stream`) with the stop features recorded, for training predictors
 * outside the library and for checking one: every stop_every expansions the
 * 16 raw features the learned stop evaluates go to
 * d_feats [num_queries][trace_cap][16] and the expansion count to
 * d_steps [num_queries][trace_cap]; per query d_meta [num_queries][18] holds
 * the records kept, the last expansion that changed the top label_k
 * (<= 16), then the expansions at which the final top label_k members were
 * scored, ascending.  Results go to b's outputs as usual; learned_stop and
 * route_entries must be 0.  No reference counterpart. */
AEEResult rover_search_trace_device(remote_handle64 h, const RoverSearchBatch *b,
                                    int stop_every, int label_k, int trace_cap,
                                    float *d_feats, int32_t *d_steps,
                                    int32_t *d_meta, void *stream);

/* Installs / reads back the predictor: FP32 parameters in the flat order
 * w1[64][16] | b1[64] | w2[64][64] | b2[64] | w3[64] | b3[1] (rounded to
 * BF16 for inference) and the feature normalization (x - mean) * inv_std. */
AEEResult rover_learned_stop_set(remote_handle64 h, const float *params,
                                 int paramsLen, const float *mean,
                                 const float *inv_std, int stop_every,
                                 int label_k, int min_steps);
AEEResult rover_learned_stop_get(remote_handle64 h, float *params,
                                 int paramsLen, float *mean, float *inv_std);

/* The training kernels on their own (device pointers): adds the gradient of
 * sum_i w_i * BCE(sigmoid(f(x_i)), y_i) to d_grad and the loss to *d_loss,
 * logits to d_out (may be NULL); one Adam step scaled by grad_scale. */
int rover_mlp_param_count(void);
AEEResult rover_mlp_forward_backward(const float *d_params, const float *d_x,
                                     const float *d_y, const float *d_weight,
                                     int n, float *d_grad, float *d_loss,
                                     float *d_out);
AEEResult rover_mlp_adam_step(float *d_params, const float *d_grad, float *d_m1,
                              float *d_m2, int step, float lr,
                              float grad_scale);

/* Seeds of one query: index 0 is graph_entry (when in range), the rest the
 * LCG s = s*1103515245 + 12345, id = (s * n) >> 32
 * (rpc_demo_pure_pq.c:192-208).  Returns the count. */
int rover_make_seed_ids(int num_vectors, int entry_count, int graph_entry,
                        uint32_t seed, int32_t *out_ids);

/* The demo's result fingerprint (pure_pq_result_fingerprint,
 * retrieval_rpc_demo_pure_pq.c:387-400): FNV-1a 64 over each result's id
 * then distance bit pattern, 4 little-endian bytes each, chained from `hash`
 * (start from ROVER_FINGERPRINT_BASIS; the demo chains every query's whole
 * sorted pool in query order). */
#define ROVER_FINGERPRINT_BASIS 14695981039346656037ull
uint64_t rover_result_fingerprint(uint64_t hash, const int32_t *ids,
                                  const float *dists, int count);

/* Seeds of the live SQ path (retrieval_sqmd_seed.c:15-63): graph_entry
 * (when in range), then s = s*1103515245 + 12345, id = (s / 65536) % n,
 * skipping duplicates (at most 64 * target + 1024 draws), then the lowest
 * unused ids.  Returns the count (<= capacity), or -1. */
int rover_sqmd_seed_ids(int num_vectors, int entry_points, int graph_entry,
                        uint32_t seed, int32_t *out_ids, int capacity);

/* Index building on the device, bit-identical to the reference encoders:
 * PQ codes (retrieval_pq_core.c:226-268) and per-vector residual codes and
 * qparams (retrieval_pq_residual.c:314-385).  Device pointers; synchronous
 * with respect to the default stream. */
AEEResult rover_build_pq_encode(const float *d_x, int n, int dim, int m, int ks,
                                const float *d_codebook, uint8_t *d_codes);
AEEResult rover_build_residual_encode(const float *d_x, const uint8_t *d_codes,
                                      int n, int dim, int m, int ks,
                                      const float *d_codebook, int bits,
                                      uint8_t *d_res, float *d_qparams);

/* error4 codes on the device, bit-identical to retrieval_build_error4
 * (retrieval_build_artifacts.c:25-66): the reconstruction is the residual
 * decode (retrieval_pq_residual.c:954-981), e = x - recon,
 * scale = max(max e - min e, 1e-12) / 15, zero = min e,
 * code = clamp(floor((e - zero) * (1 / scale) + 0.5), 0, 15).  Flat outputs
 * [n][(dim+1)/2] and [n][2]; a row whose error is not finite gets a NaN
 * scale (the loader rejects it).  Device pointers; synchronous. */
AEEResult rover_build_error4_encode(const float *d_x, const uint8_t *d_codes,
                                    const uint8_t *d_res,
                                    const float *d_res_qparams, int n, int dim,
                                    int m, int ks, const float *d_codebook,
                                    int res_bits, uint8_t *d_e4_codes,
                                    float *d_e4_qparams);

/* Graph construction on the device, bit-identical to the reference's
 * search-build with the exact candidate provider and one iteration
 * (retrieval_search_build_graph_ex, retrieval_search_build_core.c:1286-1455
 * with retrieval_search_build_exact_provider, :314-361): per node the
 * `candidates` closest others by exact squared L2 (sequential f32 sum,
 * ties by id) merged with the node's random initial row (LCG from `seed`),
 * then the RNG occlusion rule (prune_mode 1) or the closest r (0), fill
 * from the remaining candidates, random fill, the row by (distance, id).
 * d_x [n][dim] and d_graph [n][r] are device pointers; r <= 64, r < n,
 * r <= candidates <= 128; synchronous. */
AEEResult rover_build_graph_exact(const float *d_x, int n, int dim, int r,
                                  int candidates, uint32_t seed, int prune_mode,
                                  int32_t *d_graph);

/* rover_build_graph_exact's graph (the same bits) for large n, computed over
 * a clustering the caller supplies: d_assign [n] cluster ids in
 * [0, num_clusters), d_centroids [num_clusters][dim] (any points; k-means
 * centroids make it fast).  Each node's candidates come from its own
 * cluster and from every other cluster that a triangle-inequality bound
 * (with margins covering f32 rounding) cannot rule out, so the candidate
 * lists, and the graph, equal the brute-force ones for any clustering.
 * The reference has one provider (retrieval_search_build_core.c:314-361,
 * an O(n^2) scan); this one keeps its results where n^2 is out of reach.
 * stats (may be NULL) [4]: node blocks (16 members of one cluster),
 * cluster pairs screened in, (block, cluster) pairs searched, blocks with
 * such pairs.  candidates < n; num_clusters <= ROVER_BUILD_MAX_CLUSTERS;
 * ROVER_EFAILED when the clustering leaves more than 2^24 cluster pairs or
 * 2^27 (block, cluster) pairs to search.  Synchronous. */
#define ROVER_BUILD_MAX_CLUSTERS 32768
AEEResult rover_build_graph_exact_clustered(const float *d_x, int n, int dim, int r,
                                            int candidates, uint32_t seed, int prune_mode,
                                            const int32_t *d_assign, const float *d_centroids,
                                            int num_clusters, int32_t *d_graph,
                                            int64_t *stats);

/* RNN-Descent graph construction on the device, bit-identical to the
 * reference's retrieval_rnn_build_graph_ex
 * (retrieval_rnn_descent_core.c:1432-1548) followed by
 * retrieval_rnn_export_topk (:475-570): s random initial neighbors per node
 * (LCG from `seed`), t1 rounds of t2 update passes (occlusion by earlier
 * row members, pending insertions applied per node in (src, dst) order),
 * reverse edges between rounds, rows of at most r; d_graph [n][top_k]
 * by (distance, id), -1 padded.  d_x [n][dim] and d_graph are device
 * pointers; s <= r <= 64, top_k <= r; synchronous. */
AEEResult rover_build_graph_rnn(const float *d_x, int n, int dim, int s, int r,
                                int t1, int t2, uint32_t seed, int top_k,
                                int32_t *d_graph);

/* Device memory helpers, so callers need no HIP headers. */
AEEResult rover_device_alloc(void **d_ptr, size_t bytes);
AEEResult rover_device_free(void *d_ptr);
/* Page-locked host memory: query and result buffers of rover_search_batch
 * in it are copied by DMA, overlapping nothing but skipping HIP's staging. */
AEEResult rover_host_alloc(void **ptr, size_t bytes);
AEEResult rover_host_free(void *ptr);
AEEResult rover_device_upload(void *d_dst, const void *src, size_t bytes);
AEEResult rover_device_download(void *dst, const void *d_src, size_t bytes);
AEEResult rover_device_synchronize(void);
/* Device-to-device copy by a plain vector-copy kernel (16-byte aligned
 * sizes and pointers); its traffic is exactly 2 * bytes. */
AEEResult rover_device_copy(void *d_dst, const void *d_src, size_t bytes);
int rover_device_count(void);
const char *rover_last_error(void);

/* Timing and streams (hipEvent_t / hipStream_t as opaque pointers).
 * Streams are created at the default priority. */
AEEResult rover_event_create(void **event);
AEEResult rover_event_record(void *event, void *stream);
AEEResult rover_event_elapsed_ms(void *start, void *stop, float *ms);
AEEResult rover_event_destroy(void *event);
AEEResult rover_stream_create(void **stream);
AEEResult rover_stream_destroy(void *stream);
AEEResult rover_stream_synchronize(void *stream);

#ifdef __cplusplus
}
#endif
#endif
