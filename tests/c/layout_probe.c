/* Sizes, field offsets and enum values of the library-search structs, as
 * one translation unit sees them.  Built twice: with LAYOUT_REFERENCE it
 * includes the reference's own retrieval_sqmd_pure_pq.h (from a Rover
 * source tree), otherwise rover_gpu.h; layout_main.c compares the tables. */
#include <stddef.h>
#include <stdint.h>

#ifdef LAYOUT_REFERENCE
#include "dsp_demo/retrieval_sqmd_pure_pq.h"
#define LAYOUT_TABLE layout_reference
#define LAYOUT_COUNT layout_reference_count
#else
#include "rover_gpu.h"
#define LAYOUT_TABLE layout_rover
#define LAYOUT_COUNT layout_rover_count
#endif

#define F(T, m) offsetof(T, m)
const size_t LAYOUT_TABLE[] = {
    sizeof(PurePqSearchIndexView), F(PurePqSearchIndexView, graph),
    F(PurePqSearchIndexView, graph_degree), F(PurePqSearchIndexView, num_vectors),
    F(PurePqSearchIndexView, code_chunks), F(PurePqSearchIndexView, chunk_vector_counts),
    F(PurePqSearchIndexView, chunk_count), F(PurePqSearchIndexView, chunk_vectors),
    F(PurePqSearchIndexView, codebook), F(PurePqSearchIndexView, dim),
    F(PurePqSearchIndexView, m), F(PurePqSearchIndexView, ks), F(PurePqSearchIndexView, subdim),
    sizeof(PurePqSearchRequest), F(PurePqSearchRequest, query), F(PurePqSearchRequest, entry_ids),
    F(PurePqSearchRequest, entry_count), F(PurePqSearchRequest, expand_count),
    F(PurePqSearchRequest, pool_capacity), F(PurePqSearchRequest, num_out),
    sizeof(PurePqSearchStats), F(PurePqSearchStats, pq_scored_candidates),
    F(PurePqSearchStats, corrected_candidates), F(PurePqSearchStats, expanded_nodes),
    F(PurePqSearchStats, visited_dedup_skips),
    sizeof(PurePqDistancePolicy), F(PurePqDistancePolicy, mode),
    F(PurePqDistancePolicy, evaluator), F(PurePqDistancePolicy, context),
    sizeof(PurePqCorrectionMode), (size_t)PURE_PQ_CORRECTION_OFF,
    (size_t)PURE_PQ_CORRECTION_ONLINE, sizeof(PurePqDistanceEvaluatorFn),
};
const size_t LAYOUT_COUNT = sizeof(LAYOUT_TABLE) / sizeof(LAYOUT_TABLE[0]);
