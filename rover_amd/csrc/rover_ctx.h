/*
 * rover_ctx.h -- session state shared by the C host sources.
 */
#ifndef ROVER_CTX_H
#define ROVER_CTX_H

#include <pthread.h>
#include <stddef.h>
#include <stdint.h>

#include "rover_gpu.h"
#include "rover_hip.h"

#define ROVER_CTX_MAGIC 0x52564731u /* "RVG1" */

typedef struct RoverLearned {
    int ready;
    uint16_t *d_w;     /* bf16 w1 | w2 | w3 */
    float *d_f;        /* b1 | b2 | b3 | mean | inv_std */
    float *h_params;   /* fp32 master copy [rv_mlp_param_count()] */
    float h_mean[RV_MLP_IN], h_inv_std[RV_MLP_IN];
    int stop_every, label_k, min_steps;
} RoverLearned;

typedef struct RoverCtx {
    uint32_t magic;
    int device;
    int l1_bits;
    /* pure-PQ index */
    int pq_loaded;
    int n, dim, m, ks, subdim, code_bytes;
    uint8_t *d_codes;
    float *d_codebook;
    int32_t *d_graph;
    int graph_n, degree, entry;
    uint8_t *d_res;
    float *d_qparams;
    int res_bits;
    int32_t *d_orig;
    uint8_t orig_maps_sha[32]; /* digest of the id maps an error4 artifact
                                  must be bound to (empty when no map)    */
    /* third-level refine artifact (error4): code blocks and qparams */
    uint8_t *d_e4_blocks;
    float *d_e4_qparams;
    int e4_loaded, e4_vpb, e4_code_bytes;
    uint8_t e4_maps_sha[32];   /* the artifact's header digest at +216   */
    /* SQ index (retrieval_load) */
    int sq_loaded;
    int32_t *d_sq_graph;
    uint8_t *d_sq_codes;
    int sq_n, sq_dim, sq_degree, sq_entry, sq_bits;
    int sq_dma;                /* loaded by retrieval_load_dma            */
    float sq_scale, sq_zero;
    uint8_t *d_sq_res;         /* SQ residual codes and qparams          */
    float *d_sq_res_qp;
    int sq_res_bits;
    float *d_split_bias;       /* SQMDBIAS payload for the split distance */
    uint8_t *d_sq_e4_blocks;   /* SQERR4V1 code blocks and qparams        */
    float *d_sq_e4_qparams;
    int sq_e4_loaded, sq_e4_vpb, sq_e4_code_bytes;
    /* SQ traversal scratch: exact visited bitmaps and HBM frontiers per
       wave, and the cached live-path seed list */
    uint32_t *d_sqmd_vis;
    uint64_t *d_visited;         /* exact visited sets of the search kernel  */
    size_t visited_cap;
    size_t visited_layout;       /* the (mode, words) the sets were set up for */
    void *d_susp;                /* suspended searches' state (one block)    */
    size_t susp_cap;
    size_t sqmd_vis_cap;
    void *d_sqmd_front;
    size_t sqmd_front_cap;
    int32_t *d_sqmd_seeds;
    int sqmd_seeds_count, sqmd_seeds_n, sqmd_seeds_ep, sqmd_seeds_entry;
    uint32_t sqmd_seeds_seed;
    /* launch workspace */
    int32_t *d_queue;
    void *last_stream;        /* stream of the last eager launch: a launch on
                                 another stream waits for it first, since the
                                 work counter and scratch are shared        */
    int launched;
    void *ev_last;            /* completion event of that launch           */
    void *stream;             /* the session's stream for host-buffer calls */
    void *d_noalloc;          /* this device's fixed no-allocation workspace */
    int32_t *d_entries;
    int entries_n, entries_count, entries_graph_entry;
    uint32_t entries_seed;
    int entries_cap;
    void *d_stage;
    size_t stage_cap;
    /* entry router: centroids [router_n][dim] and representative node ids */
    float *d_router_cent;
    int32_t *d_router_ids;
    int router_n;
    float *d_route_dist;      /* routing scratch */
    size_t route_dist_cap;
    int32_t *d_route_ent;     /* per-query entry rows of the last launch */
    size_t route_ent_cap;
    RoverLearned learned;
    pthread_mutex_t lock;
} RoverCtx;

RoverCtx *rover_ctx_of(remote_handle64 h);
void rover_sha256(const void *data, size_t bytes, uint8_t out[32]);
void rover_set_err(const char *what);
void rover_set_errf(const char *msg);
void rover_dfree(void *p);
/* Fills the kernel's index and batch descriptors for a batch request;
 * explicit_entries keeps the seed list already on the device. */
/* The exact visited sets (RvBatch.visited) of a launch of `grid` waves in
 * *buf, (re)allocated and zeroed on `stream` when too small. */
int rover_attach_visited(uint64_t **buf, size_t *cap, size_t *layout, const RvIndex *ix,
                         RvBatch *rb, int grid, void *stream);
/* rv_search_launch, with the session's exact visited sets attached first
 * when the batch asks for them. */
int rover_launch(RoverCtx *c, const RvIndex *ix, RvBatch *rb, int grid_waves, void *stream);
AEEResult rover_prepare_launch(RoverCtx *c, const RoverSearchBatch *b,
                               int max_pool, int explicit_entries, RvIndex *ix,
                               RvBatch *rb);
/* Eager launches of a session share its work counter and scratch.
 * rover_stream_turn, before enqueueing on `stream`, makes `stream` wait (on
 * the device) for the session's previous launch when that ran on another
 * stream; rover_stream_done, after, records the new launch's completion. */
int rover_stream_turn(RoverCtx *c, void *stream);
int rover_stream_done(RoverCtx *c, void *stream);
/* Routes device queries to S entries each (nearest router centroids) into
 * d_out[nq][S]; chunks the scratch.  Asynchronous on `stream`. */
AEEResult rover_route_into(RoverCtx *c, const float *d_queries, int nq, int S,
                           int32_t *d_out, void *stream);

#endif
