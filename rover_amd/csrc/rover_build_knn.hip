// rover_build_knn.hip -- the exact candidate provider of the reference's
// search-build (retrieval_search_build_core.c:314-361: every other node by
// exact squared L2, (distance, id) order, the first `candidates`) for large
// n, computed over clusters and certified, so the result is the brute-force
// one for any clustering the caller supplies:
//
//   1. nodes are grouped by cluster (counting sort on the device) and each
//      cluster's radius r_c = max over its members of |x - mu_c| taken;
//   2. every node's `keep` closest among its own cluster's members
//      (rv_knn_range_kernel: 16 nodes per workgroup against the cluster's
//      members in 64-node tiles, the distances sequential f32 sums in the
//      reference's order, keys (distance bits, original id));
//   3. certification: a node y of another cluster c is at least
//      (|x - mu_c| - r_c)^2 from x, so cluster c can hold one of x's
//      candidates only if that bound, shrunk by a margin that covers the f32
//      rounding of every quantity involved, is at most x's current
//      candidates-th distance (an upper bound of the final one).  Cluster
//      pairs are screened first with |mu_c' - mu_c| - r_c' - r_c
//      (rv_kn_cpair_kernel), then the members of the pairs left, node by
//      node (rv_kn_npair_kernel); every (node, cluster) pair not ruled out
//      is listed;
//   4. the listed clusters' members are merged into their node's list
//      (rv_knn_extra_kernel, one wavefront per node).
//
// Every node outside the searched clusters is then provably farther than
// the final candidates-th key, so the lists equal the brute-force lists key
// for key (ties by id included).  A clustering that matches the data
// (k-means) leaves few pairs for step 4.
#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "rover_hip.h"
#include "rv_wave.h"

#define KN_WAVE 64
#define KN_WAVES 4
#define KN_ROWS 4
#define KN_NODES (KN_WAVES * KN_ROWS)
#define KN_TILE 64
#define KN_DC 128
#define KN_KEY_MAX 0xffffffffffffffffull
#define KN_MAX_CLUSTERS 32768  // the cluster-pair screen runs nc^2 threads

static int kn_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        char msg[256];
        snprintf(msg, sizeof(msg), "%s: %s", what, hipGetErrorString(e));
        rv_set_error(msg);
        return -1;
    }
    return 0;
}

__device__ __forceinline__ uint64_t kn_key(float d, int id) {
    return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)id;
}

__device__ __forceinline__ uint64_t kn_lane_key(uint64_t v, int src) {
    return ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(v >> 32), src) << 32) |
           (uint32_t)__shfl((int)(uint32_t)v, src);
}

__device__ __forceinline__ uint64_t kn_bitonic_sort64(uint64_t m) {
    m = rv_bitonic_step<64, 32>(m);
    m = rv_bitonic_step<64, 16>(m);
    m = rv_bitonic_step<64, 8>(m);
    m = rv_bitonic_step<64, 4>(m);
    m = rv_bitonic_step<64, 2>(m);
    m = rv_bitonic_step<64, 1>(m);
    return m;
}

// The ascending 128-key list (lo: ranks 0..63, hi: 64..127) with 64 new
// keys merged in, keeping the 128 smallest.
__device__ __forceinline__ void kn_merge128(uint64_t &lo, uint64_t &hi, uint64_t fresh) {
    fresh = rv_sort_wave(fresh);
    const int src = 63 - __lane_id();
    const uint64_t rev = kn_lane_key(fresh, src);
    const uint64_t a = lo < rev ? lo : rev;
    const uint64_t spill = kn_bitonic_sort64(lo < rev ? rev : lo);
    lo = kn_bitonic_sort64(a);
    const uint64_t srev = kn_lane_key(spill, src);
    hi = kn_bitonic_sort64(hi < srev ? hi : srev);
}

// ------------------------------------------------------- grouping

__global__ void rv_kn_count_kernel(const int32_t *assign, int n, int nc, int32_t *counts,
                                   int32_t *bad) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)n) return;
    const int c = assign[i];
    if (c < 0 || c >= nc) {
        atomicOr(bad, 1);
        return;
    }
    atomicAdd(&counts[c], 1);
}

// perm[pos] = original id; xp[pos] = x[perm[pos]]
__global__ void rv_kn_scatter_kernel(const int32_t *assign, int n, int32_t *cursor,
                                     int32_t *perm) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)n) return;
    const int pos = atomicAdd(&cursor[assign[i]], 1);
    perm[pos] = (int32_t)i;
}

__global__ void rv_kn_gather_kernel(const float *x, const int32_t *perm, int n, int dim,
                                    float *xp) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)n * dim) return;
    const size_t row = t / dim, d = t % dim;
    xp[t] = x[(size_t)perm[row] * dim + d];
}

// One wavefront per cluster: r_c = max over members of the f32 distance to
// the centroid (a sequential sum, then sqrtf).
__global__ void rv_kn_radius_kernel(const float *xp, const int32_t *offsets, int nc, int dim,
                                    const float *cent, float *radius) {
    const int c = blockIdx.x;
    const int lane = (int)__lane_id();
    if (c >= nc) return;
    const float *mu = cent + (size_t)c * dim;
    float r = 0.0f;
    for (int i = offsets[c] + lane; i < offsets[c + 1]; i += KN_WAVE) {
        const float *v = xp + (size_t)i * dim;
        float acc = 0.0f;
        for (int d = 0; d < dim; d++) {
            const float t = v[d] - mu[d];
            acc = acc + t * t;
        }
        r = fmaxf(r, sqrtf(acc));
    }
    for (int off = 32; off > 0; off >>= 1) r = fmaxf(r, __shfl_xor(r, off));
    if (lane == 0) radius[c] = r;
}

// ------------------------------------------------- own-cluster lists

// A block: nodes [node0, node0 + nodes) of the grouped order against the
// columns [col0, col1) of the grouped order (the nodes' cluster).
struct KnBlock {
    int node0, nodes, col0, col1;
};

// The kNN tile loop of the brute-force builder over one column range:
// 16 nodes per workgroup (4 per wavefront), 64-column tiles staged through
// LDS 128 dimensions at a time, sequential f32 sums carried across the
// chunks; the best 128 keys of each node as two keys per lane.
__global__ __launch_bounds__(KN_WAVE *KN_WAVES) void rv_knn_range_kernel(
    const float *xp, const int32_t *perm, int dim, const KnBlock *blocks, int nblocks,
    uint64_t *keys) {
    __shared__ __align__(16) float tile[KN_TILE * (KN_DC + 1)];
    __shared__ __align__(16) float rows[KN_NODES * KN_DC];
    constexpr int STRIDE = KN_DC + 1;
    const int lane = (int)__lane_id();
    const int wave = (int)threadIdx.x / KN_WAVE;
    const int nchunk = (dim + KN_DC - 1) / KN_DC;
    for (int bi = blockIdx.x; bi < nblocks; bi += gridDim.x) {
        const KnBlock blk = blocks[bi];
        uint64_t lo[KN_ROWS], hi[KN_ROWS];
#pragma unroll
        for (int r = 0; r < KN_ROWS; r++) lo[r] = hi[r] = KN_KEY_MAX;
        for (int c0 = blk.col0; c0 < blk.col1; c0 += KN_TILE) {
            float acc[KN_ROWS];
#pragma unroll
            for (int r = 0; r < KN_ROWS; r++) acc[r] = 0.0f;
            for (int k = 0; k < nchunk; k++) {
                const int d0 = k * KN_DC, dc = min(KN_DC, dim - d0);
                __syncthreads();
                for (int i = threadIdx.x; i < KN_TILE * dc; i += blockDim.x) {
                    const int rr = i / dc, e = i - rr * dc;
                    const int v = c0 + rr;
                    tile[rr * STRIDE + e] = v < blk.col1 ? xp[(size_t)v * dim + d0 + e] : 0.0f;
                }
                for (int i = threadIdx.x; i < KN_NODES * dc; i += blockDim.x) {
                    const int rr = i / dc, e = i - rr * dc;
                    rows[rr * KN_DC + e] =
                        rr < blk.nodes ? xp[(size_t)(blk.node0 + rr) * dim + d0 + e] : 0.0f;
                }
                __syncthreads();
                const float *col = tile + lane * STRIDE;
                const float *mine = rows + wave * KN_ROWS * KN_DC;
                for (int d = 0; d < dc; d++) {
                    const float cv = col[d];
#pragma unroll
                    for (int r = 0; r < KN_ROWS; r++) {
                        const float t = mine[r * KN_DC + d] - cv;
                        acc[r] = acc[r] + t * t;
                    }
                }
            }
            const int v = c0 + lane;
            const int vid = v < blk.col1 ? perm[v] : -1;
#pragma unroll
            for (int r = 0; r < KN_ROWS; r++) {
                const int u = blk.node0 + wave * KN_ROWS + r;
                const bool in = wave * KN_ROWS + r < blk.nodes;
                const uint64_t key = in && v < blk.col1 && v != u ? kn_key(acc[r], vid)
                                                                 : KN_KEY_MAX;
                const uint64_t lim = kn_lane_key(hi[r], 63);
                if (__ballot(key < lim)) kn_merge128(lo[r], hi[r], key);
            }
        }
#pragma unroll
        for (int r = 0; r < KN_ROWS; r++) {
            if (wave * KN_ROWS + r < blk.nodes) {
                const size_t u = (size_t)perm[blk.node0 + wave * KN_ROWS + r];
                keys[u * 128 + lane] = lo[r];
                keys[u * 128 + KN_WAVE + lane] = hi[r];
            }
        }
    }
}

// ---------------------------------------------------- certification
//
// For y in cluster c', |x - y| >= |x - mu_c'| - r_c' and |x - mu_c'| >=
// |mu_c - mu_c'| - |x - mu_c| >= |mu_c - mu_c'| - r_c for x in cluster c.
// A node's list can gain a member of c' only if the bound, squared, is at
// most the list's cand-th distance (an upper bound of the final one).  The
// computed quantities are sequential f32 sums of squares and their square
// roots, each within a relative 1e-3 of the exact value for any dim here;
// the margins below are wider, so every skipped cluster is provably (under
// f32 rounding) farther than the list's last key.

__device__ __forceinline__ bool kn_may_hold(float dist_lb, float dk) {
    const double lb = (double)dist_lb * (1.0 - 1e-3) - 1e-3;
    return lb <= 0.0 || lb * lb <= (double)dk * (1.0 + 1e-3) + 1e-3;
}

// Per cluster: the largest cand-th distance of its members' lists, +inf
// when a list is not full (then nothing can be skipped).
__global__ void rv_kn_maxd_kernel(const int32_t *perm, const int32_t *offsets, int nc,
                                  const uint64_t *keys, int cand, float *maxd) {
    const int c = blockIdx.x;
    const int lane = (int)__lane_id();
    if (c >= nc) return;
    float m = 0.0f;
    for (int i = offsets[c] + lane; i < offsets[c + 1]; i += KN_WAVE) {
        const uint64_t k = keys[(size_t)perm[i] * 128 + cand - 1];
        m = fmaxf(m, k == KN_KEY_MAX ? INFINITY : __uint_as_float((uint32_t)(k >> 32)));
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if (lane == 0) maxd[c] = m;
}

// One thread per cluster pair (c, c'), c != c', both non-empty: listed when
// the cluster-level bound |mu_c - mu_c'| - r_c - r_c' does not rule c' out
// for every member of c.
__global__ void rv_kn_cpair_kernel(const float *cent, int nc, int dim, const int32_t *offsets,
                                   const float *radius, const float *maxd, int32_t *cpairs,
                                   unsigned long long cap, unsigned long long *count) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)nc * nc) return;
    const int c = (int)(t / nc), c2 = (int)(t % nc);
    if (c == c2 || offsets[c] == offsets[c + 1] || offsets[c2] == offsets[c2 + 1]) return;
    const float *a = cent + (size_t)c * dim, *b = cent + (size_t)c2 * dim;
    float acc = 0.0f;
    for (int d = 0; d < dim; d++) {
        const float s = a[d] - b[d];
        acc = acc + s * s;
    }
    const float lb = sqrtf(acc) * (1.0f - 1e-3f) - radius[c] * (1.0f + 1e-3f) -
                     radius[c2] * (1.0f + 1e-3f);
    if (isinf(maxd[c]) || kn_may_hold(lb, maxd[c])) {
        const unsigned long long k = atomicAdd(count, 1ull);
        if (k < cap) {
            cpairs[2 * k] = c;
            cpairs[2 * k + 1] = c2;
        }
    }
}

// One workgroup per listed cluster pair (c, c'): each member x of c whose
// own bound |x - mu_c'| - r_c' does not rule c' out gets the pair (x, c').
__global__ void rv_kn_npair_kernel(const float *xp, const int32_t *perm, const int32_t *offsets,
                                   int dim, const float *cent, const float *radius,
                                   const uint64_t *keys, int cand, const int32_t *cpairs,
                                   int ncp, int32_t *pairs, unsigned long long cap,
                                   unsigned long long *count) {
    for (int pi = blockIdx.x; pi < ncp; pi += gridDim.x) {
        const int c = cpairs[2 * pi], c2 = cpairs[2 * pi + 1];
        const float *mu = cent + (size_t)c2 * dim;
        for (int pos = offsets[c] + threadIdx.x; pos < offsets[c + 1]; pos += blockDim.x) {
            const uint64_t k = keys[(size_t)perm[pos] * 128 + cand - 1];
            bool need = k == KN_KEY_MAX;
            if (!need) {
                const float *xv = xp + (size_t)pos * dim;
                float acc = 0.0f;
                for (int d = 0; d < dim; d++) {
                    const float s = xv[d] - mu[d];
                    acc = acc + s * s;
                }
                need = kn_may_hold(sqrtf(acc) * (1.0f - 1e-3f) - radius[c2] * (1.0f + 1e-3f),
                                   __uint_as_float((uint32_t)(k >> 32)));
            }
            if (need) {
                const unsigned long long j = atomicAdd(count, 1ull);
                if (j < cap) {
                    pairs[2 * j] = pos;
                    pairs[2 * j + 1] = c2;
                }
            }
        }
    }
}

// ----------------------------------------------- certified additions

// One wavefront per listed node: its list [128 keys] merged with every
// member of each of its listed clusters.  xs: the node's position in the
// grouped order; ranges [pair0, pair1) of the pair list (sorted by node).
__global__ __launch_bounds__(KN_WAVE) void rv_knn_extra_kernel(
    const float *xp, const int32_t *perm, const int32_t *offsets, int dim,
    const int32_t *pairs, const int32_t *node_start, int nodes, uint64_t *keys) {
    __shared__ float q[4096];
    const int lane = (int)__lane_id();
    for (int w = blockIdx.x; w < nodes; w += gridDim.x) {
        const int p0 = node_start[w], p1 = node_start[w + 1];
        const int pos = pairs[2 * p0];
        const int u = perm[pos];
        for (int d = lane; d < dim; d += KN_WAVE) q[d] = xp[(size_t)pos * dim + d];
        __syncthreads();
        uint64_t lo = keys[(size_t)u * 128 + lane], hi = keys[(size_t)u * 128 + KN_WAVE + lane];
        for (int p = p0; p < p1; p++) {
            const int c = pairs[2 * p + 1];
            for (int v0 = offsets[c]; v0 < offsets[c + 1]; v0 += KN_WAVE) {
                const int v = v0 + lane;
                uint64_t key = KN_KEY_MAX;
                if (v < offsets[c + 1]) {
                    const float *yv = xp + (size_t)v * dim;
                    float acc = 0.0f;
                    for (int d = 0; d < dim; d++) {
                        const float s = q[d] - yv[d];
                        acc = acc + s * s;
                    }
                    key = kn_key(acc, perm[v]);
                }
                const uint64_t lim = kn_lane_key(hi, 63);
                if (__ballot(key < lim)) kn_merge128(lo, hi, key);
            }
        }
        keys[(size_t)u * 128 + lane] = lo;
        keys[(size_t)u * 128 + KN_WAVE + lane] = hi;
        __syncthreads();
    }
}

// ------------------------------------------------------------ driver

static int kn_cmp64(const void *a, const void *b) {
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return (x > y) - (x < y);
}

extern "C" int rv_build_exact_clustered(const float *x, int n, int dim, int r, int cand,
                                        uint32_t seed, int prune_mode, const int32_t *assign,
                                        const float *centroids, int nc, int32_t *graph,
                                        int64_t *out_stats, void *stream) {
    if (!x || !graph || !assign || !centroids || n <= 1 || dim <= 0 || dim > 4096 ||
        r <= 0 || r >= n || r > KN_WAVE || cand < r || cand > 128 || cand >= n || nc <= 0 ||
        nc > KN_MAX_CLUSTERS ||
        (prune_mode != 0 && prune_mode != 1)) {
        rv_set_error("rv_build_exact_clustered: bad arguments");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    int rc = -1;
    int32_t *d_counts = nullptr, *d_cursor = nullptr, *d_perm = nullptr;
    int32_t *d_off = nullptr, *d_bad = nullptr, *d_pairs = nullptr;
    unsigned long long *d_npairs = nullptr, npairs64 = 0, ncpairs64 = 0;
    int32_t *d_nstart = nullptr, *d_cpairs = nullptr;
    float *d_xp = nullptr, *d_rad = nullptr, *d_maxd = nullptr;
    uint64_t *d_keys = nullptr;
    KnBlock *d_blocks = nullptr;
    int32_t *h_off = nullptr, *h_pairs = nullptr, *h_nstart = nullptr;
    KnBlock *h_blocks = nullptr;
    int nblocks = 0, bad = 0, npairs = 0, ncpairs = 0, listed = 0;
    // at most every other cluster for every node, and 2^27 pairs (1 GiB)
    const size_t all_pairs = (size_t)n * (size_t)(nc - 1);
    const size_t cap = all_pairs < ((size_t)1 << 27) ? (all_pairs > 0 ? all_pairs : 1)
                                                     : ((size_t)1 << 27);
    const size_t ccap = (size_t)1 << 24;
    const unsigned g1 = (unsigned)((n + 255) / 256);
    if (kn_check(hipMalloc((void **)&d_counts, (size_t)nc * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_cursor, (size_t)nc * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_off, (size_t)(nc + 1) * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_perm, (size_t)n * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_bad, 8), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_xp, (size_t)n * dim * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_rad, (size_t)nc * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_keys, (size_t)n * 128 * 8), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_pairs, cap * 8), "hipMalloc"))
        goto done;
    h_off = (int32_t *)malloc((size_t)(nc + 1) * 4);
    if (!h_off) goto done;
    // 1. grouping by cluster
    if (rv_fill_zero(d_counts, (size_t)nc * 4, s) || rv_fill_zero(d_bad, 8, s)) goto done;
    rv_kn_count_kernel<<<g1, 256, 0, s>>>(assign, n, nc, d_counts, d_bad);
    if (kn_check(hipGetLastError(), "rv_kn_count_kernel") ||
        kn_check(hipMemcpyAsync(h_off + 1, d_counts, (size_t)nc * 4, hipMemcpyDeviceToHost, s),
                 "counts") ||
        kn_check(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, s), "bad") ||
        kn_check(hipStreamSynchronize(s), "sync"))
        goto done;
    if (bad) {
        rv_set_error("rv_build_exact_clustered: cluster id out of range");
        goto done;
    }
    h_off[0] = 0;
    for (int c = 0; c < nc; c++) h_off[c + 1] += h_off[c];
    for (int c = 0; c < nc; c++) {
        const int cnt = h_off[c + 1] - h_off[c];
        nblocks += (cnt + KN_NODES - 1) / KN_NODES;
    }
    h_blocks = (KnBlock *)malloc((size_t)(nblocks > 0 ? nblocks : 1) * sizeof(KnBlock));
    if (!h_blocks) goto done;
    {
        int b = 0;
        for (int c = 0; c < nc; c++)
            for (int v = h_off[c]; v < h_off[c + 1]; v += KN_NODES) {
                h_blocks[b].node0 = v;
                h_blocks[b].nodes = h_off[c + 1] - v < KN_NODES ? h_off[c + 1] - v : KN_NODES;
                h_blocks[b].col0 = h_off[c];
                h_blocks[b].col1 = h_off[c + 1];
                b++;
            }
    }
    if (kn_check(hipMalloc((void **)&d_blocks, (size_t)(nblocks > 0 ? nblocks : 1) *
                                                   sizeof(KnBlock)),
                 "hipMalloc") ||
        kn_check(hipMemcpyAsync(d_off, h_off, (size_t)(nc + 1) * 4, hipMemcpyHostToDevice, s),
                 "offsets") ||
        kn_check(hipMemcpyAsync(d_cursor, h_off, (size_t)nc * 4, hipMemcpyHostToDevice, s),
                 "cursor") ||
        kn_check(hipMemcpyAsync(d_blocks, h_blocks, (size_t)nblocks * sizeof(KnBlock),
                                hipMemcpyHostToDevice, s),
                 "blocks"))
        goto done;
    rv_kn_scatter_kernel<<<g1, 256, 0, s>>>(assign, n, d_cursor, d_perm);
    if (kn_check(hipGetLastError(), "rv_kn_scatter_kernel")) goto done;
    rv_kn_gather_kernel<<<(unsigned)(((size_t)n * dim + 255) / 256), 256, 0, s>>>(
        x, d_perm, n, dim, d_xp);
    if (kn_check(hipGetLastError(), "rv_kn_gather_kernel")) goto done;
    rv_kn_radius_kernel<<<nc, KN_WAVE, 0, s>>>(d_xp, d_off, nc, dim, centroids, d_rad);
    if (kn_check(hipGetLastError(), "rv_kn_radius_kernel")) goto done;
    // 2. own-cluster lists
    if (rv_fill_zero(d_keys, (size_t)n * 128 * 8, s)) goto done;
    {
        int grid = nblocks < 65536 ? nblocks : 65536;
        if (grid > 0) {
            rv_knn_range_kernel<<<grid, KN_WAVE * KN_WAVES, 0, s>>>(d_xp, d_perm, dim, d_blocks,
                                                                     nblocks, d_keys);
            if (kn_check(hipGetLastError(), "rv_knn_range_kernel")) goto done;
        }
    }
    // 3. certification: cluster pairs, then nodes of the listed pairs
    if (kn_check(hipMalloc((void **)&d_npairs, 8), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_maxd, (size_t)nc * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_cpairs, ccap * 8), "hipMalloc") ||
        rv_fill_zero(d_npairs, 8, s))
        goto done;
    rv_kn_maxd_kernel<<<nc, KN_WAVE, 0, s>>>(d_perm, d_off, nc, d_keys, cand, d_maxd);
    rv_kn_cpair_kernel<<<(unsigned)(((size_t)nc * nc + 255) / 256), 256, 0, s>>>(
        centroids, nc, dim, d_off, d_rad, d_maxd, d_cpairs, ccap, d_npairs);
    if (kn_check(hipGetLastError(), "rv_kn_cpair_kernel") ||
        kn_check(hipMemcpyAsync(&ncpairs64, d_npairs, 8, hipMemcpyDeviceToHost, s),
                 "ncpairs") ||
        kn_check(hipStreamSynchronize(s), "sync"))
        goto done;
    if (ncpairs64 > ccap) {
        rv_set_error("rv_build_exact_clustered: too many cluster pairs to check; the "
                     "clustering does not fit the data");
        goto done;
    }
    ncpairs = (int)ncpairs64;
    if (ncpairs > 0) {
        if (rv_fill_zero(d_npairs, 8, s)) goto done;
        rv_kn_npair_kernel<<<ncpairs < 65536 ? ncpairs : 65536, 256, 0, s>>>(
            d_xp, d_perm, d_off, dim, centroids, d_rad, d_keys, cand, d_cpairs, ncpairs,
            d_pairs, cap, d_npairs);
        if (kn_check(hipGetLastError(), "rv_kn_npair_kernel") ||
            kn_check(hipMemcpyAsync(&npairs64, d_npairs, 8, hipMemcpyDeviceToHost, s),
                     "npairs") ||
            kn_check(hipStreamSynchronize(s), "sync"))
            goto done;
    }
    if (npairs64 > cap) {
        rv_set_error("rv_build_exact_clustered: too many uncertified (node, cluster) pairs; "
                     "the clustering does not fit the data");
        goto done;
    }
    npairs = (int)npairs64;
    // 4. certified additions, pairs grouped by node
    if (npairs > 0) {
        h_pairs = (int32_t *)malloc((size_t)npairs * 8);
        h_nstart = (int32_t *)malloc((size_t)(npairs + 1) * 4);
        if (!h_pairs || !h_nstart ||
            kn_check(hipMemcpy(h_pairs, d_pairs, (size_t)npairs * 8, hipMemcpyDeviceToHost),
                     "pairs"))
            goto done;
        {
            // the pairs by (node position, cluster)
            int64_t *kv = (int64_t *)malloc((size_t)npairs * 8);
            if (!kv) goto done;
            for (int i = 0; i < npairs; i++)
                kv[i] = ((int64_t)h_pairs[2 * i] << 32) | (uint32_t)h_pairs[2 * i + 1];
            qsort(kv, (size_t)npairs, sizeof(int64_t), kn_cmp64);
            for (int i = 0; i < npairs; i++) {
                h_pairs[2 * i] = (int32_t)(kv[i] >> 32);
                h_pairs[2 * i + 1] = (int32_t)(uint32_t)kv[i];
                if (i == 0 || h_pairs[2 * i] != h_pairs[2 * i - 2]) h_nstart[listed++] = i;
            }
            h_nstart[listed] = npairs;
            free(kv);
        }
        if (kn_check(hipMalloc((void **)&d_nstart, (size_t)(listed + 1) * 4), "hipMalloc") ||
            kn_check(hipMemcpy(d_nstart, h_nstart, (size_t)(listed + 1) * 4,
                               hipMemcpyHostToDevice),
                     "nstart") ||
            kn_check(hipMemcpy(d_pairs, h_pairs, (size_t)npairs * 8, hipMemcpyHostToDevice),
                     "pairs"))
            goto done;
        rv_knn_extra_kernel<<<listed < 65536 ? listed : 65536, KN_WAVE, 0, s>>>(
            d_xp, d_perm, d_off, dim, d_pairs, d_nstart, listed, d_keys);
        if (kn_check(hipGetLastError(), "rv_knn_extra_kernel")) goto done;
    }
    // 5. the rows, as the brute-force builder forms them
    if (rv_rows_from_keys(x, n, dim, r, cand, seed, prune_mode, d_keys, 128, graph, s))
        goto done;
    if (out_stats) {
        out_stats[0] = nblocks;
        out_stats[1] = ncpairs;
        out_stats[2] = npairs;
        out_stats[3] = listed;
    }
    rc = kn_check(hipStreamSynchronize(s), "hipStreamSynchronize");
done:
    hipStreamSynchronize(s);
    hipFree(d_counts);
    hipFree(d_cursor);
    hipFree(d_off);
    hipFree(d_perm);
    hipFree(d_bad);
    hipFree(d_xp);
    hipFree(d_rad);
    hipFree(d_keys);
    hipFree(d_pairs);
    hipFree(d_npairs);
    hipFree(d_nstart);
    hipFree(d_blocks);
    hipFree(d_cpairs);
    hipFree(d_maxd);
    free(h_off);
    free(h_pairs);
    free(h_nstart);
    free(h_blocks);
    return rc;
}
