/* A C caller of the library search, built against rover_gpu.h alone and
 * linked with librover.so: reads a small index and one query from a binary
 * file, calls retrieval_pure_pq_search_scalar[_with_policy] exactly as the
 * reference's callers call it, and prints the count, the stats and each
 * result's id and distance bits.
 *
 * Input file (little-endian int32 header, then the arrays):
 *   n, degree, dim, m, ks, code_bytes, entry_count, expand, pool, num_out,
 *   online, res_bits, chunk_vectors;
 *   graph int32[n*degree], codes u8[n*code_bytes], codebook f32[m*ks*dim/m],
 *   query f32[dim], entries int32[entry_count],
 *   online: residual codes u8[n*res_bytes], qparams f32[2n]. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rover_gpu.h"

static void *take(FILE *f, size_t bytes) {
    void *p = malloc(bytes ? bytes : 1);
    if (!p || fread(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "short input\n");
        exit(2);
    }
    return p;
}

int main(int argc, char **argv) {
    int32_t h[13];
    FILE *f;
    if (argc != 2 || !(f = fopen(argv[1], "rb")) || fread(h, 4, 13, f) != 13) return 2;
    const int n = h[0], degree = h[1], dim = h[2], m = h[3], ks = h[4], code_bytes = h[5];
    const int entry_count = h[6], expand = h[7], pool = h[8], num_out = h[9];
    const int online = h[10], res_bits = h[11], cv = h[12];
    int32_t *graph = take(f, (size_t)n * degree * 4);
    uint8_t *codes = take(f, (size_t)n * code_bytes);
    float *codebook = take(f, (size_t)ks * dim * 4);
    float *query = take(f, (size_t)dim * 4);
    int *entries = take(f, (size_t)entry_count * 4);
    const int res_bytes = res_bits == 4 ? (dim + 1) / 2 : dim;
    uint8_t *res = online ? take(f, (size_t)n * res_bytes) : NULL;
    float *qp = online ? take(f, (size_t)n * 8) : NULL;
    fclose(f);

    const int chunks = (n + cv - 1) / cv;
    const uint8_t **chunk = malloc(sizeof(*chunk) * chunks);
    int *counts = malloc(sizeof(int) * chunks);
    for (int c = 0; c < chunks; c++) {
        chunk[c] = codes + (size_t)c * cv * code_bytes;
        counts[c] = n - c * cv < cv ? n - c * cv : cv;
    }
    PurePqSearchIndexView view = {graph, degree, n, chunk, counts, chunks, cv,
                                  codebook, dim, m, ks, dim / m};
    PurePqSearchRequest req = {query, entries, entry_count, expand, pool, num_out};
    PurePqSearchStats st;
    int *ids = malloc(sizeof(int) * num_out);
    float *dists = malloc(sizeof(float) * num_out);
    int count;
    if (online) {
        RoverResidualView rv = {res, qp, res_bits};
        PurePqDistancePolicy pol = {PURE_PQ_CORRECTION_ONLINE, rover_pq_residual_evaluator, &rv};
        count = retrieval_pure_pq_search_scalar_with_policy(&view, &req, &pol, ids, dists,
                                                            num_out, &st);
    } else {
        count = retrieval_pure_pq_search_scalar(&view, &req, ids, dists, num_out, &st);
    }
    printf("count %d stats %llu %llu %llu %llu\n", count,
           (unsigned long long)st.pq_scored_candidates,
           (unsigned long long)st.corrected_candidates,
           (unsigned long long)st.expanded_nodes, (unsigned long long)st.visited_dedup_skips);
    for (int i = 0; i < count; i++) {
        uint32_t b;
        memcpy(&b, &dists[i], 4);
        printf("%d %08x\n", ids[i], b);
    }
    rover_pure_pq_view_release();
    return count < 0;
}
