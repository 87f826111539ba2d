// rover_build_knn.hip -- the exact candidate provider of the reference's
// search-build (retrieval_search_build_core.c:314-361: every other node by
// exact squared L2, (distance, id) order, the first `candidates`) for large
// n, computed over clusters and certified, so the result is the brute-force
// one for any clustering the caller supplies:
//
//   1. nodes are grouped by cluster (counting sort on the device), cut into
//      blocks of 16 consecutive members of one cluster, and each cluster's
//      radius r_c = max over its members of |x - mu_c| taken;
//   2. every node's 128 closest among its own cluster's members
//      (rv_knn_range_kernel: a block per workgroup against the cluster's
//      members in 64-node tiles, the distances sequential f32 sums in the
//      reference's order, keys (distance bits, original id));
//   3. certification: a node y of another cluster c is at least
//      |x - mu_c| - r_c from x, so cluster c can hold one of x's candidates
//      only if that bound squared, shrunk by a margin that covers the f32
//      rounding of every quantity involved, is at most x's current
//      candidates-th distance (an upper bound of the final one).  Cluster
//      pairs are screened first with |mu_c' - mu_c| - r_c' - r_c
//      (rv_kn_cpair_kernel), then the members of the pairs left, node by
//      node (rv_kn_npair_kernel); a block is listed with cluster c when any
//      of its nodes is not ruled out;
//   4. the listed (block, cluster) pairs are sorted by block (device radix
//      sort) and every block's lists are merged with all members of its
//      listed clusters (rv_knn_range_kernel again, lists loaded from HBM).
//
// Every node outside the searched clusters is then provably farther than
// the final candidates-th key, and searching more clusters than a node
// needs only adds true candidates, so the lists equal the brute-force lists
// key for key (ties by id included).  A clustering that matches the data
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

// ------------------------------------------------- candidate lists

// A block: nodes [node0, node0 + nodes) of the grouped order, members of
// one cluster whose columns are [col0, col1) of the grouped order.
struct KnBlock {
    int node0, nodes, col0, col1;
};

// The kNN tile loop of the brute-force builder over column ranges: 16 nodes
// per workgroup (4 per wavefront), 64-column tiles staged through LDS 128
// dimensions at a time, sequential f32 sums carried across the chunks; the
// best 128 keys of each node as two keys per lane.  EXTRA = false: the
// block's own cluster, lists start empty; EXTRA = true: the clusters listed
// for the block (sorted pairs [pfirst, plast) of (block << 32 | cluster)),
// lists loaded from `keys`.
template <bool EXTRA>
__global__ __launch_bounds__(KN_WAVE *KN_WAVES) void rv_knn_range_kernel(
    const float *xp, const int32_t *perm, int dim, const KnBlock *blocks, int nblocks,
    const int32_t *offsets, const uint64_t *pairs, const int32_t *pfirst,
    const int32_t *plast, uint64_t *keys) {
    __shared__ __align__(16) float tile[KN_TILE * (KN_DC + 1)];
    __shared__ __align__(16) float rows[KN_NODES * KN_DC];
    constexpr int STRIDE = KN_DC + 1;
    const int lane = (int)__lane_id();
    const int wave = (int)threadIdx.x / KN_WAVE;
    const int nchunk = (dim + KN_DC - 1) / KN_DC;
    for (int bi = blockIdx.x; bi < nblocks; bi += gridDim.x) {
        const KnBlock blk = blocks[bi];
        const int p0 = EXTRA ? pfirst[bi] : 0, p1 = EXTRA ? plast[bi] : 1;
        if (p0 >= p1) continue;
        uint64_t lo[KN_ROWS], hi[KN_ROWS];
#pragma unroll
        for (int r = 0; r < KN_ROWS; r++) {
            lo[r] = hi[r] = KN_KEY_MAX;
            if (EXTRA && wave * KN_ROWS + r < blk.nodes) {
                const size_t u = (size_t)perm[blk.node0 + wave * KN_ROWS + r];
                lo[r] = keys[u * 128 + lane];
                hi[r] = keys[u * 128 + KN_WAVE + lane];
            }
        }
        for (int p = p0; p < p1; p++) {
            int col0 = blk.col0, col1 = blk.col1;
            if (EXTRA) {
                const int c = (int)(uint32_t)pairs[p];
                col0 = offsets[c];
                col1 = offsets[c + 1];
            }
            for (int c0 = col0; c0 < col1; c0 += KN_TILE) {
                float acc[KN_ROWS];
#pragma unroll
                for (int r = 0; r < KN_ROWS; r++) acc[r] = 0.0f;
                for (int k = 0; k < nchunk; k++) {
                    const int d0 = k * KN_DC, dc = min(KN_DC, dim - d0);
                    __syncthreads();
                    for (int i = threadIdx.x; i < KN_TILE * dc; i += blockDim.x) {
                        const int rr = i / dc, e = i - rr * dc;
                        const int v = c0 + rr;
                        tile[rr * STRIDE + e] = v < col1 ? xp[(size_t)v * dim + d0 + e] : 0.0f;
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
                const int vid = v < col1 ? perm[v] : -1;
#pragma unroll
                for (int r = 0; r < KN_ROWS; r++) {
                    const int u = blk.node0 + wave * KN_ROWS + r;
                    const bool in = wave * KN_ROWS + r < blk.nodes;
                    const uint64_t key = in && v < col1 && v != u ? kn_key(acc[r], vid)
                                                                  : KN_KEY_MAX;
                    const uint64_t lim = kn_lane_key(hi[r], 63);
                    if (__ballot(key < lim)) kn_merge128(lo[r], hi[r], key);
                }
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

// One workgroup (256 threads) per listed cluster pair (c, c'): member x of
// c needs c' when its own bound |x - mu_c'| - r_c' does not rule c' out.
// The 16 lanes of a lane group cover one block of c (blocks start at the
// cluster's first member and 256 is a multiple of 16); the group's first
// lane lists (block, c') when any of them needs it.
__global__ __launch_bounds__(256) void rv_kn_npair_kernel(
    const float *xp, const int32_t *perm, const int32_t *offsets, const int32_t *bstart,
    int dim, const float *cent, const float *radius, const uint64_t *keys, int cand,
    const int32_t *cpairs, int ncp, uint64_t *pairs, unsigned long long cap,
    unsigned long long *count) {
    const int lane = (int)__lane_id();
    for (int pi = blockIdx.x; pi < ncp; pi += gridDim.x) {
        const int c = cpairs[2 * pi], c2 = cpairs[2 * pi + 1];
        const float *mu = cent + (size_t)c2 * dim;
        const int first = offsets[c], end = offsets[c + 1];
        for (int base = first; base < end; base += 256) {
            const int pos = base + (int)threadIdx.x;
            bool need = false;
            if (pos < end) {
                const uint64_t k = keys[(size_t)perm[pos] * 128 + cand - 1];
                need = k == KN_KEY_MAX;
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
            }
            const uint64_t m = __ballot(need);
            const bool any = ((m >> (16 * (lane >> 4))) & 0xffffull) != 0;
            if (any && (lane & 15) == 0) {
                const unsigned long long j = atomicAdd(count, 1ull);
                if (j < cap)
                    pairs[j] = ((uint64_t)(uint32_t)(bstart[c] + (pos - first) / KN_NODES) << 32) |
                               (uint32_t)c2;
            }
        }
    }
}

// The sorted (block << 32 | cluster) pairs: each block's range [pfirst,
// plast) (zero-filled beforehand, so blocks without pairs get an empty
// range); `listed` counts the blocks with pairs.
__global__ void rv_kn_mark_kernel(const uint64_t *sorted, int n, int32_t *pfirst,
                                  int32_t *plast, int32_t *listed) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const uint32_t b = (uint32_t)(sorted[i] >> 32);
    if (i == 0 || (uint32_t)(sorted[i - 1] >> 32) != b) {
        pfirst[b] = i;
        atomicAdd(listed, 1);
    }
    if (i == n - 1 || (uint32_t)(sorted[i + 1] >> 32) != b) plast[b] = i + 1;
}

// ------------------------------------------------------------ driver

extern "C" int rv_build_exact_clustered(const float *x, int n, int dim, int r, int cand,
                                        uint32_t seed, int prune_mode, const int32_t *assign,
                                        const float *centroids, int nc, int32_t *graph,
                                        int64_t *out_stats, void *stream) {
    if (!x || !graph || !assign || !centroids || n <= 1 || dim <= 0 || dim > 4096 ||
        r <= 0 || r >= n || r > KN_WAVE || cand < r || cand > 128 || cand >= n || nc <= 0 ||
        nc > KN_MAX_CLUSTERS || (prune_mode != 0 && prune_mode != 1)) {
        rv_set_error("rv_build_exact_clustered: bad arguments");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    int rc = -1;
    int32_t *d_counts = nullptr, *d_cursor = nullptr, *d_perm = nullptr, *d_off = nullptr;
    int32_t *d_bad = nullptr, *d_bstart = nullptr, *d_cpairs = nullptr;
    int32_t *d_pfirst = nullptr, *d_plast = nullptr, *d_listed = nullptr;
    float *d_xp = nullptr, *d_rad = nullptr, *d_maxd = nullptr;
    uint64_t *d_keys = nullptr, *d_pairs = nullptr, *d_sorted = nullptr;
    unsigned long long *d_count = nullptr, npairs64 = 0, ncpairs64 = 0;
    void *d_tmp = nullptr;
    size_t tmp_bytes = 0;
    KnBlock *d_blocks = nullptr;
    int32_t *h_off = nullptr, *h_bstart = nullptr;
    KnBlock *h_blocks = nullptr;
    int nblocks = 0, bad = 0, npairs = 0, ncpairs = 0, listed = 0;
    const size_t ccap = (size_t)1 << 24;
    const unsigned g1 = (unsigned)((n + 255) / 256);
    if (kn_check(hipMalloc((void **)&d_counts, (size_t)nc * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_cursor, (size_t)nc * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_off, (size_t)(nc + 1) * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_bstart, (size_t)nc * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_perm, (size_t)n * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_bad, 8), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_xp, (size_t)n * dim * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_rad, (size_t)nc * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_keys, (size_t)n * 128 * 8), "hipMalloc"))
        goto done;
    h_off = (int32_t *)malloc((size_t)(nc + 1) * 4);
    h_bstart = (int32_t *)malloc((size_t)nc * 4);
    if (!h_off || !h_bstart) goto done;
    // 1. grouping by cluster, blocks of 16 members
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
        h_bstart[c] = nblocks;
        nblocks += (h_off[c + 1] - h_off[c] + KN_NODES - 1) / KN_NODES;
    }
    h_blocks = (KnBlock *)malloc((size_t)nblocks * sizeof(KnBlock));
    if (!h_blocks) goto done;
    for (int c = 0, b = 0; c < nc; c++)
        for (int v = h_off[c]; v < h_off[c + 1]; v += KN_NODES, b++) {
            h_blocks[b].node0 = v;
            h_blocks[b].nodes = h_off[c + 1] - v < KN_NODES ? h_off[c + 1] - v : KN_NODES;
            h_blocks[b].col0 = h_off[c];
            h_blocks[b].col1 = h_off[c + 1];
        }
    if (kn_check(hipMalloc((void **)&d_blocks, (size_t)nblocks * sizeof(KnBlock)),
                 "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_pfirst, (size_t)nblocks * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_plast, (size_t)nblocks * 4), "hipMalloc") ||
        kn_check(hipMemcpyAsync(d_off, h_off, (size_t)(nc + 1) * 4, hipMemcpyHostToDevice, s),
                 "offsets") ||
        kn_check(hipMemcpyAsync(d_cursor, h_off, (size_t)nc * 4, hipMemcpyHostToDevice, s),
                 "cursor") ||
        kn_check(hipMemcpyAsync(d_bstart, h_bstart, (size_t)nc * 4, hipMemcpyHostToDevice, s),
                 "block starts") ||
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
    rv_knn_range_kernel<false><<<nblocks < 65536 ? nblocks : 65536, KN_WAVE * KN_WAVES, 0, s>>>(
        d_xp, d_perm, dim, d_blocks, nblocks, d_off, nullptr, nullptr, nullptr, d_keys);
    if (kn_check(hipGetLastError(), "rv_knn_range_kernel")) goto done;
    // 3. certification: cluster pairs, then the blocks of the listed pairs
    if (kn_check(hipMalloc((void **)&d_count, 8), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_listed, 8), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_maxd, (size_t)nc * 4), "hipMalloc") ||
        kn_check(hipMalloc((void **)&d_cpairs, ccap * 8), "hipMalloc") ||
        rv_fill_zero(d_count, 8, s))
        goto done;
    rv_kn_maxd_kernel<<<nc, KN_WAVE, 0, s>>>(d_perm, d_off, nc, d_keys, cand, d_maxd);
    rv_kn_cpair_kernel<<<(unsigned)(((size_t)nc * nc + 255) / 256), 256, 0, s>>>(
        centroids, nc, dim, d_off, d_rad, d_maxd, d_cpairs, ccap, d_count);
    if (kn_check(hipGetLastError(), "rv_kn_cpair_kernel") ||
        kn_check(hipMemcpyAsync(&ncpairs64, d_count, 8, hipMemcpyDeviceToHost, s),
                 "cluster pairs") ||
        kn_check(hipStreamSynchronize(s), "sync"))
        goto done;
    if (ncpairs64 > ccap) {
        rv_set_error("rv_build_exact_clustered: too many cluster pairs to check; the "
                     "clustering does not fit the data");
        goto done;
    }
    ncpairs = (int)ncpairs64;
    if (ncpairs > 0) {
        // at most one pair per block of its first cluster for every listed
        // cluster pair, and 2^27 pairs (1 GiB)
        size_t cap = (size_t)ncpairs * (size_t)nblocks;
        if (cap > ((size_t)1 << 27)) cap = (size_t)1 << 27;
        if (kn_check(hipMalloc((void **)&d_pairs, cap * 8), "hipMalloc") ||
            rv_fill_zero(d_count, 8, s))
            goto done;
        rv_kn_npair_kernel<<<ncpairs < 65536 ? ncpairs : 65536, 256, 0, s>>>(
            d_xp, d_perm, d_off, d_bstart, dim, centroids, d_rad, d_keys, cand, d_cpairs,
            ncpairs, d_pairs, cap, d_count);
        if (kn_check(hipGetLastError(), "rv_kn_npair_kernel") ||
            kn_check(hipMemcpyAsync(&npairs64, d_count, 8, hipMemcpyDeviceToHost, s),
                     "pairs") ||
            kn_check(hipStreamSynchronize(s), "sync"))
            goto done;
        if (npairs64 > cap) {
            char msg[200];
            snprintf(msg, sizeof(msg),
                     "rv_build_exact_clustered: %llu (block, cluster) pairs left to search "
                     "(of %d cluster pairs), more than %llu; the clustering does not fit "
                     "the data",
                     npairs64, ncpairs, (unsigned long long)cap);
            rv_set_error(msg);
            goto done;
        }
        npairs = (int)npairs64;
    }
    // 4. certified additions: the pairs sorted by block, each block's lists
    // merged with its listed clusters
    if (npairs > 0) {
        // the key's high word is a block index below nblocks
        int end_bit = 33;
        while (end_bit < 64 && ((uint64_t)nblocks >> (end_bit - 32)) != 0) end_bit++;
        tmp_bytes = rv_radix_scratch_bytes((size_t)npairs, 0);
        if (kn_check(hipMalloc((void **)&d_sorted, (size_t)npairs * 8), "hipMalloc") ||
            kn_check(hipMalloc(&d_tmp, tmp_bytes), "hipMalloc"))
            goto done;
        if (rv_radix_sort(d_pairs, d_sorted, nullptr, nullptr, (size_t)npairs, 0, end_bit, d_tmp,
                          tmp_bytes, s) ||
            rv_fill_zero(d_pfirst, (size_t)nblocks * 4, s) ||
            rv_fill_zero(d_plast, (size_t)nblocks * 4, s) || rv_fill_zero(d_listed, 8, s))
            goto done;
        rv_kn_mark_kernel<<<(unsigned)((npairs + 255) / 256), 256, 0, s>>>(d_sorted, npairs,
                                                                           d_pfirst, d_plast,
                                                                           d_listed);
        if (kn_check(hipGetLastError(), "rv_kn_mark_kernel")) goto done;
        rv_knn_range_kernel<true>
            <<<nblocks < 65536 ? nblocks : 65536, KN_WAVE * KN_WAVES, 0, s>>>(
                d_xp, d_perm, dim, d_blocks, nblocks, d_off, d_sorted, d_pfirst, d_plast,
                d_keys);
        if (kn_check(hipGetLastError(), "rv_knn_range_kernel") ||
            kn_check(hipMemcpyAsync(&listed, d_listed, 4, hipMemcpyDeviceToHost, s),
                     "listed"))
            goto done;
    }
    // 5. the rows, as the brute-force builder forms them
    if (rv_rows_from_keys(x, n, dim, r, cand, seed, prune_mode, d_keys, 128, graph, s))
        goto done;
    rc = kn_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    if (!rc && out_stats) {
        out_stats[0] = nblocks;
        out_stats[1] = ncpairs;
        out_stats[2] = npairs;
        out_stats[3] = listed;
    }
done:
    hipStreamSynchronize(s);
    hipFree(d_counts);
    hipFree(d_cursor);
    hipFree(d_off);
    hipFree(d_bstart);
    hipFree(d_perm);
    hipFree(d_bad);
    hipFree(d_xp);
    hipFree(d_rad);
    hipFree(d_keys);
    hipFree(d_pairs);
    hipFree(d_sorted);
    hipFree(d_tmp);
    hipFree(d_count);
    hipFree(d_listed);
    hipFree(d_cpairs);
    hipFree(d_maxd);
    hipFree(d_blocks);
    hipFree(d_pfirst);
    hipFree(d_plast);
    free(h_off);
    free(h_bstart);
    free(h_blocks);
    return rc;
}
