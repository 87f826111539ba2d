// rover_build.hip -- CDNA4 (gfx950) graph construction: the reference's
// search-build with the exact candidate provider and one iteration
// (retrieval_search_build_core.c:1286-1455), bit for bit.
//
//   rv_knn_exact_kernel   exact squared L2 from every node to every other
//                         (retrieval_rnn_descent_core.c:192-212: sequential
//                         f32 sum of (a - b)^2 over the dimensions, no
//                         contraction) and the 64 closest per node by
//                         (distance, id) (:314-361 sorts all of them).  A
//                         workgroup holds 16 nodes (4 per wavefront) and
//                         streams 64-node column tiles through LDS, 128
//                         dimensions at a time; each lane owns one column
//                         and computes its distance to the wave's 4 nodes
//                         with 16-byte LDS reads; the wave keeps each node's
//                         best 64 as one 64-bit key per lane, merging a tile
//                         by a bitonic sort and a 6-step bitonic merge
//                         (skipped when no key of the tile beats the current
//                         64th).
//   rv_build_rows_kernel  per node, one wavefront: the random initial row
//                         (:682-776), the candidate list (exact candidates
//                         then the initial row, first occurrence kept), the
//                         RNG occlusion rule or the closest r (:904-949),
//                         fill from the remaining candidates, random fill
//                         (:781-831), the row in (distance, id) order.
//   rv_rnn_*_kernel       RNN-Descent (retrieval_rnn_descent_core.c:1432-1548)
//                         on rows of at most r (id, distance, new-flag)
//                         entries in HBM, one wavefront per node: random
//                         initial rows (:571-738), update passes (:901-1014:
//                         the row sorted, the pair distances it needs
//                         computed across the lanes into LDS, then the
//                         sequential occlusion walk), reverse edges
//                         (:1163-1305), and the insertions each pass
//                         produces applied per source node in (src, dst)
//                         order after a device radix sort (:854-899).
//
// Bound: VALU (the exact distances are sequential per pair and cannot use
// the matrix cores without changing their rounding).
#include <hip/hip_runtime.h>

#include <float.h>
#include <stdint.h>
#include <stdio.h>

#include "rover_hip.h"
#include "rv_wave.h"

#define RB_WAVE 64
#define RB_WAVES 4         // wavefronts per workgroup in the kNN kernel
#define RB_ROWS 4          // nodes per wavefront in the kNN kernel
#define RB_DC 128          // dimensions per LDS chunk in the kNN kernel
#define RB_TILE 64         // column nodes per tile
#define RB_KEY_MAX 0xffffffffffffffffull

static int rb_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        char msg[256];
        snprintf(msg, sizeof(msg), "%s: %s", what, hipGetErrorString(e));
        rv_set_error(msg);
        return -1;
    }
    return 0;
}

__device__ __forceinline__ uint64_t rb_key(float d, int id) {
    return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)id;
}

// Merges 64 new keys into the ascending 64-key list `cur` (one key per
// lane), keeping the 64 smallest, ascending.
__device__ __forceinline__ uint64_t rb_merge64(uint64_t cur, uint64_t fresh) {
    fresh = rv_sort_wave(fresh);
    const uint64_t rev = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(fresh >> 32), 63 - __lane_id())
                          << 32) |
                         (uint32_t)__shfl((int)(uint32_t)fresh, 63 - __lane_id());
    uint64_t m = cur < rev ? cur : rev;  // bitonic: the 64 smallest
    m = rv_bitonic_step<64, 32>(m);
    m = rv_bitonic_step<64, 16>(m);
    m = rv_bitonic_step<64, 8>(m);
    m = rv_bitonic_step<64, 4>(m);
    m = rv_bitonic_step<64, 2>(m);
    m = rv_bitonic_step<64, 1>(m);
    return m;
}

// The bitonic sequence m (one key per lane) sorted ascending.
__device__ __forceinline__ uint64_t rb_bitonic_sort64(uint64_t m) {
    m = rv_bitonic_step<64, 32>(m);
    m = rv_bitonic_step<64, 16>(m);
    m = rv_bitonic_step<64, 8>(m);
    m = rv_bitonic_step<64, 4>(m);
    m = rv_bitonic_step<64, 2>(m);
    m = rv_bitonic_step<64, 1>(m);
    return m;
}

// Merges 64 new keys into the ascending 128-key list (lo: ranks 0..63, hi:
// ranks 64..127, one key of each per lane), keeping the 128 smallest.  The
// 64 smallest of lo and the new keys are the new lo (every hi key is at
// least every lo key); the 64 largest of them, merged with hi, give the new
// hi.
__device__ __forceinline__ void rb_merge128(uint64_t &lo, uint64_t &hi, uint64_t fresh) {
    fresh = rv_sort_wave(fresh);
    const int src = 63 - __lane_id();
    const uint64_t rev = ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(fresh >> 32), src) << 32) |
                         (uint32_t)__shfl((int)(uint32_t)fresh, src);
    const uint64_t a = lo < rev ? lo : rev;
    const uint64_t spill = rb_bitonic_sort64(lo < rev ? rev : lo);
    lo = rb_bitonic_sort64(a);
    const uint64_t srev =
        ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(spill >> 32), src) << 32) |
        (uint32_t)__shfl((int)(uint32_t)spill, src);
    hi = rb_bitonic_sort64(hi < srev ? hi : srev);
}

// Exact squared L2 of two rows in global memory, the reference's order.
__device__ __forceinline__ float rb_l2(const float *x, int dim, int a, int b) {
    const float *va = x + (size_t)a * dim, *vb = x + (size_t)b * dim;
    float acc = 0.0f;
    for (int d = 0; d < dim; d++) {
        const float diff = va[d] - vb[d];
        acc = acc + diff * diff;
    }
    return acc;
}

// ---------------------------------------------------------------- kNN

// Copies rows [r0, r0 + count) x dimensions [d0, d0 + dc) of x into LDS
// rows of `stride` floats (zeros past n).  Thread t handles element
// t + k * blockDim.x of the (row, dimension-group) grid; the row and group
// indexes advance by fixed steps, so the loop divides nothing.  VEC moves
// 16 bytes per element (dc % 4 == 0).
template <bool VEC>
__device__ __forceinline__ void rb_stage(const float *x, int n, int dim, int r0, int count,
                                         int d0, int dc, float *dst, int stride) {
    const int nv = VEC ? dc >> 2 : dc;
    const int total = count * nv;
    const int step_r = (int)blockDim.x / nv, step_e = (int)blockDim.x - step_r * nv;
    int r = (int)threadIdx.x / nv, e = (int)threadIdx.x - r * nv;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const int v = r0 + r;
        if (VEC) {
            float4 val = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (v < n) val = *reinterpret_cast<const float4 *>(x + (size_t)v * dim + d0 + 4 * e);
            *reinterpret_cast<float4 *>(dst + r * stride + 4 * e) = val;
        } else {
            dst[r * stride + e] = v < n ? x[(size_t)v * dim + d0 + e] : 0.0f;
        }
        r += step_r;
        e += step_e;
        if (e >= nv) {
            e -= nv;
            r++;
        }
    }
}

// VEC: dimensions read 4 at a time (dim % 4 == 0); the tile row stride is
// RB_DC + 4 floats so the 16-byte reads of 16 consecutive lanes fall in
// distinct banks (RB_DC + 1 for single floats).  The dimensions pass through
// LDS in chunks of RB_DC in ascending order, and each lane carries its
// RB_ROWS partial sums from chunk to chunk, so every distance is still one
// sequential sum over the dimensions.
// KEEP = 64 or 128 closest per node (keys [n][KEEP]).
template <bool VEC, int KEEP>
__global__ __launch_bounds__(RB_WAVE *RB_WAVES) void rv_knn_exact_kernel(const float *x, int n,
                                                                         int dim,
                                                                         uint64_t *keys) {
    constexpr int NODES = RB_WAVES * RB_ROWS;
    constexpr int STRIDE = VEC ? RB_DC + 4 : RB_DC + 1;
    __shared__ __align__(16) float tile[RB_TILE * STRIDE];  // [64][STRIDE] columns
    __shared__ __align__(16) float rows[NODES * RB_DC];     // [NODES][RB_DC] the nodes
    const int lane = (int)__lane_id();
    const int wave = (int)threadIdx.x / RB_WAVE;
    const int u0 = blockIdx.x * NODES;
    const int nchunk = (dim + RB_DC - 1) / RB_DC;
    const float *col = tile + lane * STRIDE;
    const float *mine = rows + wave * RB_ROWS * RB_DC;
    uint64_t best[RB_ROWS], best_hi[RB_ROWS];
#pragma unroll
    for (int r = 0; r < RB_ROWS; r++) best[r] = best_hi[r] = RB_KEY_MAX;
    for (int c0 = 0; c0 < n; c0 += RB_TILE) {
        float acc[RB_ROWS];
#pragma unroll
        for (int r = 0; r < RB_ROWS; r++) acc[r] = 0.0f;
        for (int k = 0; k < nchunk; k++) {
            const int d0 = k * RB_DC, dc = min(RB_DC, dim - d0);
            __syncthreads();
            rb_stage<VEC>(x, n, dim, c0, RB_TILE, d0, dc, tile, STRIDE);
            if (nchunk > 1 || c0 == 0) rb_stage<VEC>(x, n, dim, u0, NODES, d0, dc, rows, RB_DC);
            __syncthreads();
            if (VEC) {
                for (int d = 0; d < dc; d += 4) {
                    const float4 c = *reinterpret_cast<const float4 *>(col + d);
#pragma unroll
                    for (int r = 0; r < RB_ROWS; r++) {
                        const float4 a = *reinterpret_cast<const float4 *>(mine + r * RB_DC + d);
                        float t;
                        t = a.x - c.x; acc[r] = acc[r] + t * t;
                        t = a.y - c.y; acc[r] = acc[r] + t * t;
                        t = a.z - c.z; acc[r] = acc[r] + t * t;
                        t = a.w - c.w; acc[r] = acc[r] + t * t;
                    }
                }
            } else {
                for (int d = 0; d < dc; d++) {
                    const float c = col[d];
#pragma unroll
                    for (int r = 0; r < RB_ROWS; r++) {
                        const float t = mine[r * RB_DC + d] - c;
                        acc[r] = acc[r] + t * t;
                    }
                }
            }
        }
        const int v = c0 + lane;
#pragma unroll
        for (int r = 0; r < RB_ROWS; r++) {
            const int u = u0 + wave * RB_ROWS + r;
            const uint64_t key = v < n && v != u ? rb_key(acc[r], v) : RB_KEY_MAX;
            // the current last key is lane 63's; merge only when the tile
            // can improve the list
            const uint64_t last = KEEP == 128 ? best_hi[r] : best[r];
            const uint64_t lim =
                ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(last >> 32), 63) << 32) |
                (uint32_t)__shfl((int)(uint32_t)last, 63);
            if (__ballot(key < lim)) {
                if (KEEP == 128) rb_merge128(best[r], best_hi[r], key);
                else best[r] = rb_merge64(best[r], key);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RB_ROWS; r++) {
        const int u = u0 + wave * RB_ROWS + r;
        if (u < n) {
            keys[(size_t)u * KEEP + lane] = best[r];
            if (KEEP == 128) keys[(size_t)u * KEEP + RB_WAVE + lane] = best_hi[r];
        }
    }
}

// -------------------------------------------------------------- rows

__device__ __forceinline__ uint32_t rb_lcg(uint32_t x) { return x * 1664525u + 1013904223u; }

// retrieval_search_build_core.c:108-113 (sb_seed_for_node)
__device__ __forceinline__ uint32_t rb_node_seed(uint32_t seed, int node) {
    uint32_t x = seed ^ (uint32_t)(node + 1) * 0x9e3779b9u;
    x = rb_lcg(x);
    x ^= x >> 16;
    return rb_lcg(x);
}

// Lane 0 only: draws from the node's LCG until ids[0..*count) holds r
// distinct ids other than u (at most n*n draws), then the lowest unused
// ids (:682-776 and :781-831 share this shape).  Distances are filled in
// by the caller.
__device__ static void rb_draw(int n, int r, uint32_t state, int u, int32_t *ids, int *count) {
    const uint64_t limit = (uint64_t)(uint32_t)n * (uint64_t)(uint32_t)n;
    uint64_t attempts = 0;
    while (*count < r && attempts < limit) {
        state = rb_lcg(state);
        const int v = (int)(state % (uint32_t)n);
        bool dup = v == u;
        for (int i = 0; i < *count && !dup; i++) dup = ids[i] == v;
        if (!dup) ids[(*count)++] = v;
        attempts++;
    }
    for (int v = 0; v < n && *count < r; v++) {
        bool dup = v == u;
        for (int i = 0; i < *count && !dup; i++) dup = ids[i] == v;
        if (!dup) ids[(*count)++] = v;
    }
}

__global__ __launch_bounds__(RB_WAVE) void rv_build_rows_kernel(const float *x, int n, int dim,
                                                                int r, int cand, int keep,
                                                                uint32_t seed, int prune_mode,
                                                                const uint64_t *knn,
                                                                int32_t *graph) {
    __shared__ int32_t s_ids[RB_WAVE];        // drawn ids (initial row / random fill)
    __shared__ uint64_t s_lists[3 * RB_WAVE];  // exact candidates [128] | initial row [64]
    __shared__ uint64_t s_items[3 * RB_WAVE];  // candidate keys, ascending
    __shared__ uint64_t s_row[RB_WAVE];        // accepted keys
    __shared__ int s_count;
    const int lane = (int)__lane_id();
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        // ---- the random initial row, sorted
        if (lane == 0) {
            int cnt = 0;
            rb_draw(n, r, rb_node_seed(seed, u), u, s_ids, &cnt);
            s_count = cnt;
        }
        __syncthreads();
        const int ni = s_count;
        uint64_t init = lane < ni ? rb_key(rb_l2(x, dim, u, s_ids[lane]), s_ids[lane])
                                  : RB_KEY_MAX;
        init = rv_sort_wave(init);
        const uint64_t prov = lane < cand ? knn[(size_t)u * keep + lane] : RB_KEY_MAX;
        const uint64_t prov2 =
            RB_WAVE + lane < cand ? knn[(size_t)u * keep + RB_WAVE + lane] : RB_KEY_MAX;
        __syncthreads();
        s_lists[lane] = prov;
        s_lists[RB_WAVE + lane] = prov2;
        s_lists[2 * RB_WAVE + lane] = init;
        __syncthreads();
        // ---- candidates: merge the two ascending lists, dropping the
        // initial row's ids that are exact candidates (same key)
        if (lane == 0) {
            int i = 0, j = 0, cnt = 0;
            while (i < cand || j < ni) {
                const uint64_t a = i < cand ? s_lists[i] : RB_KEY_MAX;
                const uint64_t b = j < ni ? s_lists[2 * RB_WAVE + j] : RB_KEY_MAX;
                if (a == RB_KEY_MAX && b == RB_KEY_MAX) break;
                if (a < b) {
                    s_items[cnt++] = a, i++;
                } else if (b < a) {
                    s_items[cnt++] = b, j++;
                } else {
                    s_items[cnt++] = a, i++, j++;
                }
            }
            s_count = cnt;
        }
        __syncthreads();
        const int count = s_count;
        int deg = 0;
        if (prune_mode == 0) {
            deg = count < r ? count : r;
            if (lane < deg) s_row[lane] = s_items[lane];
        } else {
            // RNG occlusion: a candidate is kept unless an accepted neighbor
            // is strictly closer to it than u is
            for (int i = 0; i < count && deg < r; i++) {
                const uint64_t key = s_items[i];
                const int c = (int)(uint32_t)key;
                const float dc = __uint_as_float((uint32_t)(key >> 32));
                bool occ = false;
                if (lane < deg) occ = rb_l2(x, dim, (int)(uint32_t)s_row[lane], c) < dc;
                if (__ballot(occ) == 0ull) {
                    if (lane == 0) s_row[deg] = key;
                    deg++;
                }
                __syncthreads();
            }
            // fill from the remaining candidates in order
            for (int i = 0; i < count && deg < r; i++) {
                const uint64_t key = s_items[i];
                const bool has = lane < deg && s_row[lane] == key;
                if (__ballot(has) == 0ull) {
                    if (lane == 0) s_row[deg] = key;
                    deg++;
                }
                __syncthreads();
            }
        }
        __syncthreads();
        // ---- random fill when fewer than r
        if (deg < r) {
            if (lane == 0) {
                int cnt = deg;
                for (int k = 0; k < deg; k++) s_ids[k] = (int)(uint32_t)s_row[k];
                rb_draw(n, r, rb_node_seed(seed ^ 0xa5a55a5au, u), u, s_ids, &cnt);
                s_count = cnt;
            }
            __syncthreads();
            const int total = s_count;
            if (lane >= deg && lane < total) s_row[lane] = rb_key(rb_l2(x, dim, u, s_ids[lane]),
                                                                  s_ids[lane]);
            deg = total;
            __syncthreads();
        }
        // ---- the row in (distance, id) order
        uint64_t rowk = lane < deg ? s_row[lane] : RB_KEY_MAX;
        rowk = rv_sort_wave(rowk);
        if (lane < r) graph[(size_t)u * r + lane] = lane < deg ? (int32_t)(uint32_t)rowk : -1;
        __syncthreads();
    }
}

// The rows of the search-build from precomputed exact candidate keys
// (keys [n][keep], the first `cand` of each row used), on `stream`.
extern "C" int rv_rows_from_keys(const float *x, int n, int dim, int r, int cand,
                                 uint32_t seed, int prune_mode, const uint64_t *keys,
                                 int keep, int32_t *graph, void *stream) {
    int dev = 0, cus = 0;
    if (cand > keep || cand > 2 * RB_WAVE || r > RB_WAVE || r > cand) {
        rv_set_error("rv_rows_from_keys: bad arguments");
        return -1;
    }
    if (rb_check(hipGetDevice(&dev), "hipGetDevice") ||
        rb_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                 "hipDeviceGetAttribute"))
        return -1;
    int grid = cus * 16;
    if (grid > n) grid = n;
    rv_build_rows_kernel<<<grid, RB_WAVE, 0, (hipStream_t)stream>>>(x, n, dim, r, cand, keep,
                                                                    seed, prune_mode, keys,
                                                                    graph);
    return rb_check(hipGetLastError(), "rv_build_rows_kernel");
}

extern "C" int rv_build_exact(const float *x, int n, int dim, int r, int cand, uint32_t seed,
                              int prune_mode, int32_t *graph, void *stream) {
    if (!x || !graph || n <= 1 || dim <= 0 || dim > 4096 || r <= 0 || r >= n ||
        r > RB_WAVE || cand < r || cand > 2 * RB_WAVE || (prune_mode != 0 && prune_mode != 1)) {
        rv_set_error("rv_build_exact: bad arguments");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    uint64_t *keys = nullptr;
    const int keep = cand > RB_WAVE ? 2 * RB_WAVE : RB_WAVE;
    if (rb_check(hipMalloc((void **)&keys, (size_t)n * keep * sizeof(uint64_t)),
                 "hipMalloc knn"))
        return -1;
    const int blocks = (n + RB_WAVES * RB_ROWS - 1) / (RB_WAVES * RB_ROWS);
    int rc = 0;
    if (dim % 4 == 0 && keep == RB_WAVE)
        rv_knn_exact_kernel<true, 64><<<blocks, RB_WAVE * RB_WAVES, 0, s>>>(x, n, dim, keys);
    else if (dim % 4 == 0)
        rv_knn_exact_kernel<true, 128><<<blocks, RB_WAVE * RB_WAVES, 0, s>>>(x, n, dim, keys);
    else if (keep == RB_WAVE)
        rv_knn_exact_kernel<false, 64><<<blocks, RB_WAVE * RB_WAVES, 0, s>>>(x, n, dim, keys);
    else
        rv_knn_exact_kernel<false, 128><<<blocks, RB_WAVE * RB_WAVES, 0, s>>>(x, n, dim, keys);
    rc = rb_check(hipGetLastError(), "rv_knn_exact_kernel");
    if (!rc) {
        int dev = 0, cus = 0;
        rc = rb_check(hipGetDevice(&dev), "hipGetDevice") ||
             rb_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                      "hipDeviceGetAttribute");
        if (!rc) {
            int grid = cus * 16;
            if (grid > n) grid = n;
            rv_build_rows_kernel<<<grid, RB_WAVE, 0, s>>>(x, n, dim, r, cand, keep, seed,
                                                          prune_mode, keys, graph);
            rc = rb_check(hipGetLastError(), "rv_build_rows_kernel");
        }
    }
    // the scratch is released once the kernels are done with it
    const int sync_rc = rb_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    (void)hipFree(keys);
    return rc ? rc : sync_rc;
}

// ------------------------------------------------------- RNN-Descent

__device__ __forceinline__ uint64_t rb_shfl_xor64(uint64_t v, int m) {
    const int lo = __shfl_xor((int)(uint32_t)v, m);
    const int hi = __shfl_xor((int)(uint32_t)(v >> 32), m);
    return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// :571-738: s distinct random others per node from the node's LCG (then
// the lowest unused ids), each inserted with its distance and marked new.
__global__ __launch_bounds__(RB_WAVE) void rv_rnn_init_kernel(const float *x, int n, int dim,
                                                              int s, int r, uint32_t seed,
                                                              int32_t *ids, float *dists,
                                                              uint8_t *fresh, int32_t *deg) {
    __shared__ int32_t s_ids[RB_WAVE];
    __shared__ int s_count;
    const int lane = (int)__lane_id();
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        if (lane == 0) {
            int cnt = 0;
            rb_draw(n, s, rb_node_seed(seed, u), u, s_ids, &cnt);
            s_count = cnt;
        }
        __syncthreads();
        const int cnt = s_count;
        if (lane < r) {
            const bool have = lane < cnt;
            const size_t slot = (size_t)u * r + lane;
            ids[slot] = have ? s_ids[lane] : -1;
            dists[slot] = have ? rb_l2(x, dim, u, s_ids[lane]) : FLT_MAX;
            fresh[slot] = have ? 1 : 0;
        }
        if (lane == 0) deg[u] = cnt;
        __syncthreads();
    }
}

// :901-1014, one pass for one node u: the row sorted by (distance, id);
// entry i (in the current row) is dropped when some earlier entry w, with
// v or w new, has dist(u, v) >= dist(v, w) -- the first such w in row
// order -- and the pending insertion (w, v, dist(v, w)) recorded; all flags
// then cleared.  Pending slots [u][r]: key src << 32 | dst, or all ones.
__global__ __launch_bounds__(RB_WAVE) void rv_rnn_update_kernel(const float *x, int n, int dim,
                                                                int r, int32_t *ids,
                                                                float *dists, uint8_t *fresh,
                                                                int32_t *deg, uint64_t *pkeys,
                                                                float *pvals) {
    __shared__ uint64_t s_k0[RB_WAVE];
    __shared__ uint8_t s_f0[RB_WAVE];
    __shared__ int32_t s_id[RB_WAVE];
    __shared__ float s_d[RB_WAVE];
    __shared__ uint8_t s_f[RB_WAVE];
    __shared__ int32_t s_ord[RB_WAVE];
    __shared__ float s_pair[RB_WAVE * RB_WAVE];  // [i][w], w < i, sorted positions
    const int lane = (int)__lane_id();
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const size_t row = (size_t)u * r;
        const int dg = deg[u];
        uint64_t key = RB_KEY_MAX;
        uint8_t fl = 0;
        if (lane < dg) {
            key = rb_key(dists[row + lane], ids[row + lane]);
            fl = fresh[row + lane];
        }
        s_k0[lane] = key;
        s_f0[lane] = fl;
        key = rv_sort_wave(key);
        __syncthreads();
        // ids are distinct, so each sorted key names one original entry
        uint8_t f = 0;
        if (lane < dg)
            for (int t = 0; t < dg; t++)
                if (s_k0[t] == key) f = s_f0[t];
        s_id[lane] = (int32_t)(uint32_t)key;
        s_d[lane] = __uint_as_float((uint32_t)(key >> 32));
        s_f[lane] = f;
        s_ord[lane] = lane;
        __syncthreads();
        const int pairs = dg * (dg - 1) / 2;
        for (int p = lane; p < pairs; p += RB_WAVE) {
            int i = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)p)) * 0.5f);
            while (i > 1 && i * (i - 1) / 2 > p) i--;
            while ((i + 1) * i / 2 <= p) i++;
            const int w = p - i * (i - 1) / 2;
            if (s_f[i] || s_f[w]) s_pair[i * RB_WAVE + w] = rb_l2(x, dim, s_id[i], s_id[w]);
        }
        __syncthreads();
        int cur = dg, i = 0, np = 0;
        while (i < cur) {
            const int oi = s_ord[i];
            const float duv = s_d[oi];
            const bool fv = s_f[oi] != 0;
            bool q = false;
            if (lane < i) {
                const int ow = s_ord[lane];
                q = (fv || s_f[ow]) && duv >= s_pair[oi * RB_WAVE + ow];
            }
            const uint64_t m = __ballot(q);
            if (m) {
                const int ow = s_ord[__ffsll((unsigned long long)m) - 1];
                if (lane == 0) {
                    pkeys[row + np] = ((uint64_t)(uint32_t)s_id[ow] << 32) | (uint32_t)s_id[oi];
                    pvals[row + np] = s_pair[oi * RB_WAVE + ow];
                }
                np++;
                const int nxt = lane >= i && lane + 1 < cur ? s_ord[lane + 1] : 0;
                __syncthreads();
                if (lane >= i && lane + 1 < cur) s_ord[lane] = nxt;
                __syncthreads();
                cur--;
            } else {
                i++;
            }
        }
        if (lane < r) {
            if (lane >= np) {
                pkeys[row + lane] = RB_KEY_MAX;
                pvals[row + lane] = 0.0f;
            }
            const bool have = lane < cur;
            const int o = have ? s_ord[lane] : 0;
            ids[row + lane] = have ? s_id[o] : -1;
            dists[row + lane] = have ? s_d[o] : FLT_MAX;
            fresh[row + lane] = 0;
        }
        if (lane == 0) deg[u] = cur;
        __syncthreads();
    }
}

// :1163-1205: every edge u -> v (distance d) as the pending insertion
// (v, u, d).
__global__ void rv_rnn_reverse_kernel(int n, int r, const int32_t *ids, const float *dists,
                                      const int32_t *deg, uint64_t *pkeys, float *pvals) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)n * r) return;
    const int u = (int)(t / r), j = (int)(t % r);
    if (j < deg[u]) {
        pkeys[t] = ((uint64_t)(uint32_t)ids[t] << 32) | (uint32_t)u;
        pvals[t] = dists[t];
    } else {
        pkeys[t] = RB_KEY_MAX;
        pvals[t] = 0.0f;
    }
}

__device__ __forceinline__ size_t rb_lower_bound(const uint64_t *a, size_t count, uint64_t k) {
    size_t lo = 0, hi = count;
    while (lo < hi) {
        const size_t mid = (lo + hi) >> 1;
        if (a[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// :854-899 with :304-375: the sorted insertions of node u applied in
// order: self edges and present edges skipped, appended while the row has
// room, else replacing the row's largest (distance, id) when the new edge
// is smaller; inserted entries are marked new.
__global__ __launch_bounds__(RB_WAVE) void rv_rnn_apply_kernel(int n, int r, int32_t *ids,
                                                               float *dists, uint8_t *fresh,
                                                               int32_t *deg,
                                                               const uint64_t *skeys,
                                                               const float *svals,
                                                               size_t count) {
    const int lane = (int)__lane_id();
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const size_t lo = rb_lower_bound(skeys, count, (uint64_t)(uint32_t)u << 32);
        const size_t hi = rb_lower_bound(skeys, count, (uint64_t)(uint32_t)(u + 1) << 32);
        if (lo == hi) continue;
        const size_t row = (size_t)u * r;
        int dg = deg[u];
        int id = -1;
        float d = FLT_MAX;
        uint8_t f = 0;
        if (lane < r) {
            id = ids[row + lane];
            d = dists[row + lane];
            f = fresh[row + lane];
        }
        for (size_t t = lo; t < hi; t++) {
            const int dst = (int)(uint32_t)skeys[t];
            const float nd = svals[t];
            if (dst == u) continue;
            if (__ballot(lane < dg && id == dst)) continue;
            if (dg < r) {
                if (lane == dg) {
                    id = dst;
                    d = nd;
                    f = 1;
                }
                dg++;
            } else {
                uint64_t mx = lane < dg ? rb_key(d, id) : 0ull;
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    const uint64_t o = rb_shfl_xor64(mx, off);
                    mx = o > mx ? o : mx;
                }
                const float fd = __uint_as_float((uint32_t)(mx >> 32));
                const int fid = (int)(uint32_t)mx;
                if ((nd < fd || (nd == fd && dst < fid)) && lane < dg && id == fid) {
                    id = dst;
                    d = nd;
                    f = 1;
                }
            }
        }
        if (lane < r) {
            ids[row + lane] = id;
            dists[row + lane] = d;
            fresh[row + lane] = f;
        }
        if (lane == 0) deg[u] = dg;
    }
}

// :438-470 and :475-494: rows by (distance, id), the first top_k ids, -1
// padded.
__global__ __launch_bounds__(RB_WAVE) void rv_rnn_export_kernel(int n, int r, int top_k,
                                                                const int32_t *ids,
                                                                const float *dists,
                                                                const int32_t *deg,
                                                                int32_t *graph) {
    const int lane = (int)__lane_id();
    for (int u = blockIdx.x; u < n; u += gridDim.x) {
        const size_t row = (size_t)u * r;
        const int dg = deg[u];
        uint64_t key = lane < dg ? rb_key(dists[row + lane], ids[row + lane]) : RB_KEY_MAX;
        key = rv_sort_wave(key);
        if (lane < top_k) graph[(size_t)u * top_k + lane] = lane < dg ? (int32_t)(uint32_t)key : -1;
    }
}

extern "C" int rv_build_rnn(const float *x, int n, int dim, int s, int r, int t1, int t2,
                            uint32_t seed, int top_k, int32_t *graph, void *stream) {
    if (!x || !graph || n <= 1 || dim <= 0 || s <= 0 || r <= 0 || r > RB_WAVE || s > r ||
        t1 <= 0 || t2 <= 0 || top_k <= 0 || top_k > r || (size_t)n * r > 0x7fffffffu) {
        rv_set_error("rv_build_rnn: bad arguments");
        return -1;
    }
    hipStream_t st = (hipStream_t)stream;
    const size_t slots = (size_t)n * r;
    int32_t *ids = nullptr, *deg = nullptr;
    float *dists = nullptr, *pv = nullptr, *pv2 = nullptr;
    uint8_t *fresh = nullptr;
    uint64_t *pk = nullptr, *pk2 = nullptr;
    void *tmp = nullptr;
    size_t tmp_bytes = 0;
    int rc = rb_check(hipMalloc((void **)&ids, slots * 4), "hipMalloc rnn ids") ||
             rb_check(hipMalloc((void **)&dists, slots * 4), "hipMalloc rnn dists") ||
             rb_check(hipMalloc((void **)&fresh, slots), "hipMalloc rnn flags") ||
             rb_check(hipMalloc((void **)&deg, (size_t)n * 4), "hipMalloc rnn degrees") ||
             rb_check(hipMalloc((void **)&pk, slots * 8), "hipMalloc rnn keys") ||
             rb_check(hipMalloc((void **)&pk2, slots * 8), "hipMalloc rnn keys") ||
             rb_check(hipMalloc((void **)&pv, slots * 4), "hipMalloc rnn values") ||
             rb_check(hipMalloc((void **)&pv2, slots * 4), "hipMalloc rnn values");
    tmp_bytes = rv_radix_scratch_bytes(slots, 1);
    if (!rc) rc = rb_check(hipMalloc(&tmp, tmp_bytes), "hipMalloc sort");
    int grid = 0;
    if (!rc) {
        int dev = 0, cus = 0;
        rc = rb_check(hipGetDevice(&dev), "hipGetDevice") ||
             rb_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                      "hipDeviceGetAttribute");
        grid = cus * 8 < n ? cus * 8 : n;
    }
    // sorts the pending insertions and applies them per source node
    auto apply = [&]() -> int {
        if (rv_radix_sort(pk, pk2, reinterpret_cast<const uint32_t *>(pv),
                          reinterpret_cast<uint32_t *>(pv2), slots, 0, 64, tmp, tmp_bytes, st))
            return -1;
        rv_rnn_apply_kernel<<<grid, RB_WAVE, 0, st>>>(n, r, ids, dists, fresh, deg, pk2, pv2,
                                                      slots);
        return rb_check(hipGetLastError(), "rv_rnn_apply_kernel");
    };
    if (!rc) {
        rv_rnn_init_kernel<<<grid, RB_WAVE, 0, st>>>(x, n, dim, s, r, seed, ids, dists, fresh,
                                                     deg);
        rc = rb_check(hipGetLastError(), "rv_rnn_init_kernel");
    }
    for (int a = 0; a < t1 && !rc; a++) {
        for (int b = 0; b < t2 && !rc; b++) {
            rv_rnn_update_kernel<<<grid, RB_WAVE, 0, st>>>(x, n, dim, r, ids, dists, fresh, deg,
                                                           pk, pv);
            rc = rb_check(hipGetLastError(), "rv_rnn_update_kernel") || apply();
        }
        if (!rc && a + 1 < t1) {
            const int threads = 256;
            rv_rnn_reverse_kernel<<<(int)((slots + threads - 1) / threads), threads, 0, st>>>(
                n, r, ids, dists, deg, pk, pv);
            rc = rb_check(hipGetLastError(), "rv_rnn_reverse_kernel") || apply();
        }
    }
    if (!rc) {
        rv_rnn_export_kernel<<<grid, RB_WAVE, 0, st>>>(n, r, top_k, ids, dists, deg, graph);
        rc = rb_check(hipGetLastError(), "rv_rnn_export_kernel");
    }
    // the scratch is released once the kernels are done with it
    const int sync_rc = rb_check(hipStreamSynchronize(st), "hipStreamSynchronize");
    (void)hipFree(ids);
    (void)hipFree(dists);
    (void)hipFree(fresh);
    (void)hipFree(deg);
    (void)hipFree(pk);
    (void)hipFree(pk2);
    (void)hipFree(pv);
    (void)hipFree(pv2);
    (void)hipFree(tmp);
    return rc ? rc : sync_rc;
}
