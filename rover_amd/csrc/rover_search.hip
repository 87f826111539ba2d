// rover_search.hip -- CDNA4 (gfx950) kernels for Rover's quantized graph
// search: batched best-first beam search, one wavefront per query, run as a
// persistent kernel over a device work counter.
//
// Semantics follow the reference scalar search
// (retrieval_sqmd_pure_pq.c:793-969, retrieval_pure_pq_search_scalar_with_policy):
//   * candidates are ordered by (distance, id) (pure_pq_order, :97-112);
//   * the result pool keeps the best L candidates seen;
//   * the search repeatedly expands the best pool member that has not been
//     expanded yet and stops when none is left.  This expansion sequence is
//     the one the reference's frontier heap produces: a frontier entry that
//     was evicted from the pool is worse than the pool's worst member from
//     then on, so when it reaches the heap top the reference terminates
//     (:875-888); every other heap top is the best unexpanded pool member.
//   * a node is scored once; re-scoring a node that is not in the pool can
//     never change the pool (its key is worse than the pool's worst for the
//     rest of the search), so the LDS visited hash may drop entries when it
//     is full without changing results.  Duplicates of pool members are
//     removed exactly in the merge (identical key).
//
// All float arithmetic keeps the reference's operation order with no
// contraction (the file is compiled with -ffp-contract=off):
//   LUT    retrieval_pq_core.c:362-390 + 16-25 (sequential over subdim)
//   ADC    retrieval_pq_core.c:392-416 (sequential over m)
//   online retrieval_pq_residual.c:983-1042 (((c + code*s) + z), sequential)
//   SQ     retrieval_l1_int8_core.c:119-165 (integer sum, one f32 multiply)
//   refine retrieval_final_refine.c:256-292 (f64 sequential, (dist, orig))
//
// LDS holds, per query: the query vector, the ADC table (m*ks f32) or the
// quantized query, the sorted pool as one array of 64-bit keys (float
// bits << 32 | id; distances are >= +0 so unsigned order equals (distance,
// id) order) merged in place, expanded flags, and the visited hash.
#include <hip/hip_runtime.h>

#include <float.h>
#include <pthread.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "rover_hip.h"
#include "rv_wave.h"

#define RV_WAVE 64
#define RV_EMPTY_SLOT 0xffffffffu
#define RV_KEY_MAX 0xffffffffffffffffull

static __thread char g_rv_error[256];

static int rv_check(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        snprintf(g_rv_error, sizeof(g_rv_error), "%s: %s", what,
                 hipGetErrorString(e));
        return -1;
    }
    return 0;
}

extern "C" const char *rv_last_error(void) { return g_rv_error; }

extern "C" void rv_set_error(const char *msg) {
    snprintf(g_rv_error, sizeof(g_rv_error), "%s", msg);
}

// ---------------------------------------------------------------- wave utils

__device__ __forceinline__ int rv_lane() { return (int)__lane_id(); }

__device__ __forceinline__ uint64_t rv_key(float d, int id) {
    return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)id;
}
__device__ __forceinline__ float rv_key_dist(uint64_t k) {
    return __uint_as_float((uint32_t)(k >> 32));
}
__device__ __forceinline__ int rv_key_id(uint64_t k) {
    return (int)(uint32_t)(k & 0xffffffffu);
}

__device__ __forceinline__ uint64_t rv_uniform64(uint64_t v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ int rv_uniform(int v) {
    return __builtin_amdgcn_readfirstlane(v);
}

// Number of keys in sorted arr[0..n) strictly smaller than key.
__device__ __forceinline__ int rv_lower_bound(const uint64_t *arr, int n,
                                              uint64_t key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (arr[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Open-addressing visited set in LDS.  Returns true when id was not
// recorded before (newly inserted, or the probe budget ran out: then the
// node is simply scored again, which cannot change results).
__device__ __forceinline__ bool rv_visit(uint32_t *table, int log2n,
                                         int probe_limit, uint32_t id) {
    const uint32_t mask = (1u << log2n) - 1u;
    uint32_t h = (id * 0x9E3779B1u) >> (32 - log2n);
    for (int p = 0; p < probe_limit; p++) {
        uint32_t old = atomicCAS(&table[h], RV_EMPTY_SLOT, id);
        if (old == RV_EMPTY_SLOT) return true;
        if (old == id) return false;
        h = (h + 1u) & mask;
    }
    return true;
}

// Exact visited set in HBM: word i holds (stamp << 32) | the visited bits of
// ids 32i..32i+31 for the query stamped `stamp`; a word with another stamp
// counts as empty.  Test-and-set by compare-and-swap (lanes of one wave may
// hit one word); the first read bypasses the non-coherent L1.  Returns true
// when id was not visited by this query before.
__device__ __forceinline__ bool rv_visit_exact(uint64_t *words, uint32_t stamp, uint32_t id) {
    unsigned long long *w = reinterpret_cast<unsigned long long *>(words + (id >> 5));
    const uint64_t bit = 1ull << (id & 31u);
    uint64_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        const uint64_t bits = (uint32_t)(old >> 32) == stamp ? (old & 0xffffffffull) : 0ull;
        if (bits & bit) return false;
        const uint64_t want = ((uint64_t)stamp << 32) | bits | bit;
        const uint64_t prev = atomicCAS(w, (unsigned long long)old, (unsigned long long)want);
        if (prev == old) return true;
        old = prev;
    }
}

// ----------------------------------------------------------- distance ops

__device__ __forceinline__ int rv_code_get(const uint8_t *code, int s, int ks) {
    if (ks <= 16) {
        uint8_t packed = code[s >> 1];
        return (s & 1) ? (packed >> 4) : (packed & 0x0f);
    }
    return code[s];
}

// ADC: sequential f32 sum over subspaces of LUT[m][code_m].
__device__ __forceinline__ float rv_adc(const float *lut, const uint8_t *row,
                                        int m, int ks, int code_bytes) {
    float acc = 0.0f;
    if (ks == 256 && (code_bytes & 15) == 0) {
        for (int c = 0; c < code_bytes; c += 16) {
            const uint4 w = *reinterpret_cast<const uint4 *>(row + c);
            const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
            for (int b = 0; b < 16; b++) {
                const int code = (ws[b >> 2] >> (8 * (b & 3))) & 255;
                acc += lut[(c + b) * 256 + code];
            }
        }
    } else {
        for (int s = 0; s < m; s++) acc += lut[s * ks + rv_code_get(row, s, ks)];
    }
    return acc;
}

__device__ __forceinline__ int rv_res_code(const uint8_t *res, int bits, int d) {
    if (bits == 4) {
        uint8_t packed = res[d >> 1];
        return (d & 1) ? (packed >> 4) : (packed & 0x0f);
    }
    return res[d];
}

// Residual-reconstructed value ((centroid + code*scale) + zero).
__device__ __forceinline__ float rv_decode_value(float c, int code, float s,
                                                 float z) {
    float product = (float)code * s;
    float sum = c + product;
    return sum + z;
}

// Online correction for the common shapes: M subspaces of SUB dims with
// 8-bit PQ codes (ks = 256) and RB-bit residuals.  The candidate's code row,
// residual row and qparams are read with 16-byte (8-byte) vector loads and
// all centroid rows are loaded before the sequential accumulation, so a
// candidate costs two rounds of independent loads (row data, then codebook
// rows) instead of M dependent ones.  Arithmetic order is the reference's.
template <int M, int SUB, int RB>
__device__ __forceinline__ bool rv_online_fixed(const RvIndex &ix, const float *q,
                                                int id, float *out) {
    constexpr int D = M * SUB;
    constexpr int RBYTES = RB == 4 ? (D + 1) / 2 : D;
    static_assert(M % 16 == 0 && RBYTES % 16 == 0, "16-byte rows");
    const float2 qp = *reinterpret_cast<const float2 *>(ix.res_qparams + 2 * (size_t)id);
    uint32_t pw[M / 4];
    const uint4 *prow = reinterpret_cast<const uint4 *>(ix.codes + (size_t)id * M);
#pragma unroll
    for (int w = 0; w < M / 16; w++) {
        const uint4 v = prow[w];
        pw[4 * w + 0] = v.x;
        pw[4 * w + 1] = v.y;
        pw[4 * w + 2] = v.z;
        pw[4 * w + 3] = v.w;
    }
    uint32_t rw[RBYTES / 4];
    const uint4 *rrow = reinterpret_cast<const uint4 *>(ix.res_codes + (size_t)id * RBYTES);
#pragma unroll
    for (int w = 0; w < RBYTES / 16; w++) {
        const uint4 v = rrow[w];
        rw[4 * w + 0] = v.x;
        rw[4 * w + 1] = v.y;
        rw[4 * w + 2] = v.z;
        rw[4 * w + 3] = v.w;
    }
    float c[D];
#pragma unroll
    for (int m = 0; m < M; m++) {
        const int code = (pw[m >> 2] >> (8 * (m & 3))) & 255;
        const float *cent = ix.codebook + ((size_t)m * 256 + code) * SUB;
        if (SUB == 4) {
            const float4 v = *reinterpret_cast<const float4 *>(cent);
            c[m * 4 + 0] = v.x;
            c[m * 4 + 1] = v.y;
            c[m * 4 + 2] = v.z;
            c[m * 4 + 3] = v.w;
        } else if (SUB == 2) {
            const float2 v = *reinterpret_cast<const float2 *>(cent);
            c[m * 2 + 0] = v.x;
            c[m * 2 + 1] = v.y;
        } else {
#pragma unroll
            for (int i = 0; i < SUB; i++) c[m * SUB + i] = cent[i];
        }
    }
    const float s = qp.x, z = qp.y;
    if (!isfinite(s) || !isfinite(z) || !(s > 0.0f)) return false;
    float dist = 0.0f;
#pragma unroll
    for (int d = 0; d < D; d++) {
        int code;
        if (RB == 4) code = (rw[d >> 3] >> (4 * (d & 7))) & 15;
        else code = (rw[d >> 2] >> (8 * (d & 3))) & 255;
        const float rec = rv_decode_value(c[d], code, s, z);
        const float diff = q[d] - rec;
        dist += diff * diff;
    }
    *out = dist;
    return true;
}

// Online correction distance; returns false for invalid qparams.
__device__ bool rv_online(const RvIndex &ix, const float *q, int id,
                          float *out) {
    if (ix.ks == 256 && ix.m == 32 && ix.subdim == 4) {
        if (ix.res_bits == 8) return rv_online_fixed<32, 4, 8>(ix, q, id, out);
        if (ix.res_bits == 4) return rv_online_fixed<32, 4, 4>(ix, q, id, out);
    }
    if (ix.ks == 256 && ix.m == 48 && ix.subdim == 2) {
        if (ix.res_bits == 8) return rv_online_fixed<48, 2, 8>(ix, q, id, out);
        if (ix.res_bits == 4) return rv_online_fixed<48, 2, 4>(ix, q, id, out);
    }
    const uint8_t *pq = ix.codes + (size_t)id * ix.code_bytes;
    const uint8_t *res = ix.res_codes + (size_t)id * ix.res_code_bytes;
    const float s = ix.res_qparams[2 * (size_t)id];
    const float z = ix.res_qparams[2 * (size_t)id + 1];
    if (!isfinite(s) || !isfinite(z) || !(s > 0.0f)) return false;
    float dist = 0.0f;
    int d = 0;
    for (int sub = 0; sub < ix.m; sub++) {
        const float *cent = ix.codebook +
            ((size_t)sub * ix.ks + rv_code_get(pq, sub, ix.ks)) * ix.subdim;
        for (int i = 0; i < ix.subdim; i++, d++) {
            float rec = rv_decode_value(cent[i], rv_res_code(res, ix.res_bits, d), s, z);
            float diff = q[d] - rec;
            dist += diff * diff;
        }
    }
    *out = dist;
    return true;
}

// Online correction with two lanes per candidate (M = 32 subspaces of 4
// dims): lane c (< 32) and lane c + 32 score the same candidate.  Each lane
// decodes its half of the dimensions into squared terms; the low lane sums
// terms 0..63 sequentially, hands the partial sum to the high lane, which
// continues with terms 64..127.  The additions happen in exactly the
// reference's order, so the distance is bit-identical to the one-lane sum.
// Must be called by the whole wavefront; `active` marks lanes whose
// candidate is scored (the others score row 0, so that every lane runs the
// same instructions, and ignore the result).  The result is valid in the
// high lane.
template <int RB>
__device__ __forceinline__ bool rv_online_pair32(const RvIndex &ix, const float *q,
                                                 int id_in, bool active, bool high,
                                                 float *out) {
    constexpr int MH = 16, DH = 64;
    constexpr int RH = RB == 4 ? 32 : 64;  // residual bytes per half
    const int id = active ? id_in : 0;
    float t[DH];
    float s, z;
    {
        const float2 qp = *reinterpret_cast<const float2 *>(ix.res_qparams + 2 * (size_t)id);
        s = qp.x;
        z = qp.y;
        const uint4 pv = *reinterpret_cast<const uint4 *>(ix.codes + (size_t)id * 32 +
                                                          (high ? 16 : 0));
        const uint32_t pw[4] = {pv.x, pv.y, pv.z, pv.w};
        uint32_t rw[RH / 4];
        const uint4 *rrow = reinterpret_cast<const uint4 *>(
            ix.res_codes + (size_t)id * (2 * RH) + (high ? RH : 0));
#pragma unroll
        for (int w = 0; w < RH / 16; w++) {
            const uint4 v = rrow[w];
            rw[4 * w + 0] = v.x;
            rw[4 * w + 1] = v.y;
            rw[4 * w + 2] = v.z;
            rw[4 * w + 3] = v.w;
        }
        const int m0 = high ? MH : 0;
        float4 cv[MH];
#pragma unroll
        for (int m = 0; m < MH; m++) {
            const int code = (pw[m >> 2] >> (8 * (m & 3))) & 255;
            cv[m] = *reinterpret_cast<const float4 *>(ix.codebook +
                                                      ((size_t)(m0 + m) * 256 + code) * 4);
        }
        const float *qh = q + (high ? DH : 0);
#pragma unroll
        for (int d = 0; d < DH; d++) {
            int code;
            if (RB == 4) code = (rw[d >> 3] >> (4 * (d & 7))) & 15;
            else code = (rw[d >> 2] >> (8 * (d & 3))) & 255;
            const float4 c4 = cv[d >> 2];
            const float c = (d & 3) == 0 ? c4.x : (d & 3) == 1 ? c4.y : (d & 3) == 2 ? c4.z : c4.w;
            const float rec = rv_decode_value(c, code, s, z);
            const float diff = qh[d] - rec;
            t[d] = diff * diff;
        }
    }
    float acc = 0.0f;
#pragma unroll
    for (int d = 0; d < DH; d++) acc += t[d];
    // the high lane receives terms 0..63
    float dist = __uint_as_float(rv_low_half(__float_as_uint(acc)));
#pragma unroll
    for (int d = 0; d < DH; d++) dist += t[d];
    *out = dist;
    return isfinite(s) && isfinite(z) && s > 0.0f;
}

// G-lane online scoring of one candidate (G = 4, 8, 16; same 32-subspace
// x 4-dim shapes as above): the lane with sub-index j of the group decodes
// dimensions [j*128/G, (j+1)*128/G) and forms their squared differences,
// then the sequential sum runs through the group in dimension order, each
// lane adding its terms to the running sum handed up from the lane below.
// The distance is left in the group's last lane (sub == G - 1).  Used when
// an expansion has few unvisited neighbors, so that the per-dimension work
// is spread over the lanes the idle candidates would have held.
template <int G, int RB>
__device__ __forceinline__ bool rv_online_group(const RvIndex &ix, const float *q, int id_in,
                                                bool active, int sub, float *out) {
    constexpr int DG = 128 / G;       // dimensions per lane
    constexpr int MG = DG / 4;        // subspaces per lane
    constexpr int RBY = DG * RB / 8;  // residual bytes per lane
    const int id = active ? id_in : 0;
    const float2 qp = *reinterpret_cast<const float2 *>(ix.res_qparams + 2 * (size_t)id);
    const float s = qp.x, z = qp.y;
    const uint8_t *crow = ix.codes + (size_t)id * 32 + sub * MG;
    uint32_t cw[MG >= 4 ? MG / 4 : 1];
    if constexpr (MG == 2) {
        cw[0] = *reinterpret_cast<const uint16_t *>(crow);
    } else if constexpr (MG == 4) {
        cw[0] = *reinterpret_cast<const uint32_t *>(crow);
    } else if constexpr (MG == 8) {
        const uint2 v = *reinterpret_cast<const uint2 *>(crow);
        cw[0] = v.x;
        cw[1] = v.y;
    } else {  // MG == 16 (two lanes per candidate)
        static_assert(MG == 16, "group size");
        const uint4 v = *reinterpret_cast<const uint4 *>(crow);
        cw[0] = v.x;
        cw[1] = v.y;
        cw[2] = v.z;
        cw[3] = v.w;
    }
    const uint8_t *rrow = ix.res_codes + (size_t)id * (16 * RB) + sub * RBY;
    uint32_t rw[RBY / 4];
    if (RBY == 4) {
        rw[0] = *reinterpret_cast<const uint32_t *>(rrow);
    } else if (RBY == 8) {
        const uint2 v = *reinterpret_cast<const uint2 *>(rrow);
        rw[0] = v.x;
        rw[RBY / 4 - 1] = v.y;
    } else {
#pragma unroll
        for (int w = 0; w < RBY / 16; w++) {
            const uint4 v = reinterpret_cast<const uint4 *>(rrow)[w];
            rw[4 * w + 0] = v.x;
            rw[4 * w + 1] = v.y;
            rw[4 * w + 2] = v.z;
            rw[4 * w + 3] = v.w;
        }
    }
    const float *qs = q + sub * DG;
    float t[DG];
#pragma unroll
    for (int m = 0; m < MG; m++) {
        const int code = (cw[m >> 2] >> (8 * (m & 3))) & 255;
        const float4 c4 = *reinterpret_cast<const float4 *>(
            ix.codebook + ((size_t)(sub * MG + m) * 256 + code) * 4);
        const float cs[4] = {c4.x, c4.y, c4.z, c4.w};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int d = 4 * m + k;
            int rc;
            if (RB == 4) rc = (rw[d >> 3] >> (4 * (d & 7))) & 15;
            else rc = (rw[d >> 2] >> (8 * (d & 3))) & 255;
            const float rec = rv_decode_value(cs[k], rc, s, z);
            const float diff = qs[d] - rec;
            t[d] = diff * diff;
        }
    }
    // a group lies inside one 16-lane row (G divides 16), so the running
    // sum moves up by a row shift; the group's first lane ignores it
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < G; j++) {
        const float below = rv_row_prev(acc);
        if (sub == j) {
            float a = j == 0 ? 0.0f : below;
#pragma unroll
            for (int d = 0; d < DG; d++) a += t[d];
            acc = a;
        }
    }
    *out = acc;
    return isfinite(s) && isfinite(z) && s > 0.0f;
}

// One lane per candidate for the 32-subspace x 4-dim shape, loading as it
// goes (few registers; the multi-wave kernel's 64-candidate tile).  Same
// sequential order as rv_online_fixed.
template <int RB>
__device__ __forceinline__ bool rv_online_lane(const RvIndex &ix, const float *q, int id,
                                               float *out) {
    const float2 qp = *reinterpret_cast<const float2 *>(ix.res_qparams + 2 * (size_t)id);
    const float s = qp.x, z = qp.y;
    const uint4 *crow = reinterpret_cast<const uint4 *>(ix.codes + (size_t)id * 32);
    const uint8_t *rrow = ix.res_codes + (size_t)id * (16 * RB);
    float acc = 0.0f;
    for (int w = 0; w < 2; w++) {
        const uint4 cv = crow[w];
        const uint32_t cws[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll 2
        for (int m = 0; m < 16; m++) {
            const int sub = 16 * w + m;
            const int code = (cws[m >> 2] >> (8 * (m & 3))) & 255;
            const float4 c4 = *reinterpret_cast<const float4 *>(
                ix.codebook + ((size_t)sub * 256 + code) * 4);
            const float cs[4] = {c4.x, c4.y, c4.z, c4.w};
            uint32_t rw;
            if (RB == 8) rw = *reinterpret_cast<const uint32_t *>(rrow + 4 * sub);
            else rw = *reinterpret_cast<const uint16_t *>(rrow + 2 * sub);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const int rc = RB == 8 ? (int)((rw >> (8 * k)) & 255) : (int)((rw >> (4 * k)) & 15);
                const float rec = rv_decode_value(cs[k], rc, s, z);
                const float diff = q[4 * sub + k] - rec;
                acc += diff * diff;
            }
        }
    }
    *out = acc;
    return isfinite(s) && isfinite(z) && s > 0.0f;
}

// Two-lane online scoring is used for 32-subspace x 4-dim PQ with at most
// 32 neighbors per expansion.
__host__ __device__ inline bool rv_online_pairable(const RvIndex &ix, int expand) {
    return expand <= 32 && ix.ks == 256 && ix.m == 32 && ix.subdim == 4 &&
           (ix.res_bits == 8 || ix.res_bits == 4);
}

__device__ __forceinline__ float rv_sq(const uint8_t *qsq, const uint8_t *code,
                                       int dim, int bits, float scale_sq) {
    int acc = 0;
    if (bits == 8) {
        int i = 0;
        if ((dim & 15) == 0) {
            for (; i < dim; i += 16) {
                const uint4 w = *reinterpret_cast<const uint4 *>(code + i);
                const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
                for (int b = 0; b < 16; b++) {
                    int c = (ws[b >> 2] >> (8 * (b & 3))) & 255;
                    int diff = (int)qsq[i + b] - c;
                    acc += diff * diff;
                }
            }
        }
        for (; i < dim; i++) {
            int diff = (int)qsq[i] - (int)code[i];
            acc += diff * diff;
        }
    } else {
        for (int i = 0; i < dim; i++) {
            int c = (i & 1) ? (code[i >> 1] >> 4) : (code[i >> 1] & 15);
            int diff = (int)(qsq[i] & 15) - c;
            acc += diff * diff;
        }
    }
    return (float)acc * scale_sq;
}

// ----------------------------------------------------------- LDS layout

struct RvLayout {
    int off_q, off_tab, off_keys, off_flags;
    int off_born;    // trace mode: step each pool entry was scored
    int off_refine;  // final refine: f64 distance bits of the refined members
    int off_stop;              // learned-stop state of the current query
    int off_cand, off_hash;
    int off_mw;      // several waves per query: neighbor list, keys, control
    int total;
};

__host__ __device__ static inline int rv_align(int v, int a) {
    return (v + a - 1) / a * a;
}

__host__ __device__ static inline RvLayout rv_layout(int dim, int scorer, int m,
                                                    int ks, int max_pool,
                                                    int hash_log2, bool trace,
                                                    int max_refine, bool multi_wave) {
    RvLayout l;
    int off = 0;
    l.off_q = off;
    off = rv_align(off + 4 * dim, 16);
    l.off_tab = off;
    if (scorer == RV_SCORE_ADC) off = rv_align(off + 4 * m * ks, 16);
    else if (scorer == RV_SCORE_SQ) off = rv_align(off + dim, 16);
    l.off_keys = off;
    off = rv_align(off + 8 * max_pool, 16);
    l.off_flags = off;
    off = rv_align(off + max_pool, 16);
    l.off_born = -1;
    if (trace) {
        l.off_born = off;
        off = rv_align(off + 2 * max_pool, 16);
    }
    l.off_refine = -1;
    if (max_refine > 0) {
        l.off_refine = off;
        off = rv_align(off + 8 * (max_refine < max_pool ? max_refine : max_pool), 16);
    }
    l.off_stop = off;
    off = rv_align(off + 64, 16);
    l.off_cand = off;
    off = rv_align(off + 8 * RV_WAVE, 16);
    l.off_hash = off;
    off = rv_align(off + 4 * (1 << hash_log2), 16);
    l.off_mw = -1;
    if (multi_wave) {
        l.off_mw = off;
        off = rv_align(off + 12 * RV_WAVE + 16, 16);
    }
    l.total = off;
    return l;
}

// ------------------------------------------------------------ learned stop

__device__ __forceinline__ float rv_bf16(uint16_t v) {
    return __uint_as_float((uint32_t)v << 16);
}

// Features of the search state after `step` expansions (identical code
// path for trace recording and inference).  All inputs are wave-uniform.
// None depends on the beam width L, so a predictor trained on long (large-L)
// traces applies to searches of any L >= 2K.
struct RvStopState {
    float d_entry, d_entry2, top1_0, topk_0;
    float improve_ema;  // decaying rate of steps that changed the top-K
    int last_change, last_top1;
    float suspend_p;    // predictor probability when the search was suspended
};

// Feature i (0..15) of the search state; lane i computes feature i, so the
// vector takes one register per lane (lanes 16..63 return 0).
__device__ __forceinline__ float rv_feature(const RvStopState &st, const uint64_t *keys,
                                            int n, int label_k, int step, int next_pos,
                                            int i) {
    const float eps = 1e-30f;
    const int K = n < label_k ? n : label_k;
    const int k2 = n < 2 * label_k ? n : 2 * label_k;
    const float top1 = rv_key_dist(keys[0]);
    const float topk = rv_key_dist(keys[K - 1]);
    const float d2k = rv_key_dist(keys[k2 - 1]);
    const float fmin = next_pos < n ? rv_key_dist(keys[next_pos]) : d2k;
    const int np = next_pos < 4 * label_k ? next_pos : 4 * label_k;
    float v = 0.0f;
    switch (i) {
    case 0: v = st.d_entry2 > 0.0f ? st.d_entry / st.d_entry2 : 1.0f; break;
    case 1: v = __logf(top1 + 1.0f); break;
    case 2: v = __logf(topk + 1.0f); break;
    case 3: v = top1 / (st.d_entry + eps); break;
    case 4: v = topk / (st.d_entry + eps); break;
    case 5: v = top1 / (st.top1_0 + eps); break;
    case 6: v = topk / (st.topk_0 + eps); break;
    case 7: v = __log2f((float)step + 1.0f); break;
    case 8: v = __logf(fmin + 1.0f); break;
    case 9: v = fmin / (topk + eps); break;
    case 10: v = __log2f((float)(step - st.last_change) + 1.0f); break;
    case 11: v = st.improve_ema; break;
    case 12: v = __log2f((float)(step - st.last_top1) + 1.0f); break;
    case 13: v = (float)np / (float)label_k; break;
    case 14: v = d2k / (topk + eps); break;
    case 15: v = fmin / (d2k + eps); break;
    default: break;
    }
    return v;
}

// 16 -> 64 -> 64 -> 1 MLP, lane j owns hidden unit j.  FP32 accumulation
// over BF16 weights (inputs in ascending order); the weights of four
// consecutive inputs for unit j come in one 8-byte load (grouped layout,
// see rv_mlp_bf16_kernel); the first hidden layer is broadcast through hbuf
// (64 floats of LDS, 16-byte aligned).  Returns the logit (wave-uniform).
__device__ __forceinline__ float rv_bf16_lo(uint32_t v) { return __uint_as_float(v << 16); }
__device__ __forceinline__ float rv_bf16_hi(uint32_t v) { return __uint_as_float(v & 0xffff0000u); }

__device__ __forceinline__ float rv_lane_value(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}

// Orders the LDS accesses of one wavefront's lanes (a wavefront-scope fence
// and a wave barrier: no s_barrier, no wait on other waves).  The search
// kernels whose workgroup is one wavefront use __syncthreads for this; code
// that one wave of a larger workgroup runs alone uses this instead.
__device__ __forceinline__ void rv_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <bool BLOCK_SYNC = true>
__device__ float rv_mlp_logit(const RvMlp &mlp_in, float f_lane, float *hbuf) {
    const int j = rv_lane();
    // The weights are re-read at every evaluation: the empty asm hides the
    // pointers' loop invariance, so the loads are not hoisted out of the
    // search loop into registers the scoring needs.
    RvMlp mlp = mlp_in;
    asm volatile("" : "+s"(mlp.w1), "+s"(mlp.w2), "+s"(mlp.w3), "+s"(mlp.b1));
    asm volatile("" : "+s"(mlp.b2), "+s"(mlp.b3), "+s"(mlp.feat_mean), "+s"(mlp.feat_inv_std));
    // lane i < 16 holds input i, normalized; each is read as a wave-uniform
    // scalar when the layer needs it
    const int fi = j < RV_MLP_IN ? j : 0;
    const float x = (f_lane - mlp.feat_mean[fi]) * mlp.feat_inv_std[fi];
    const uint2 *w1 = reinterpret_cast<const uint2 *>(mlp.w1);
    float h1 = mlp.b1[j];
#pragma unroll
    for (int g = 0; g < RV_MLP_IN / 4; g++) {
        const uint2 v = w1[g * RV_MLP_HIDDEN + j];
        h1 += rv_bf16_lo(v.x) * rv_lane_value(x, 4 * g + 0);
        h1 += rv_bf16_hi(v.x) * rv_lane_value(x, 4 * g + 1);
        h1 += rv_bf16_lo(v.y) * rv_lane_value(x, 4 * g + 2);
        h1 += rv_bf16_hi(v.y) * rv_lane_value(x, 4 * g + 3);
    }
    h1 = h1 > 0.0f ? h1 : 0.0f;
    hbuf[j] = h1;
    if (BLOCK_SYNC) __syncthreads();
    else rv_wave_sync();
    float h2 = mlp.b2[j];
    const float4 *h4 = reinterpret_cast<const float4 *>(hbuf);
    const uint2 *w2 = reinterpret_cast<const uint2 *>(mlp.w2);
#pragma unroll 2
    for (int g = 0; g < RV_MLP_HIDDEN / 4; g++) {
        const float4 hv = h4[g];
        const uint2 v = w2[g * RV_MLP_HIDDEN + j];
        h2 += rv_bf16_lo(v.x) * hv.x;
        h2 += rv_bf16_hi(v.x) * hv.y;
        h2 += rv_bf16_lo(v.y) * hv.z;
        h2 += rv_bf16_hi(v.y) * hv.w;
    }
    if (BLOCK_SYNC) __syncthreads();  // hbuf is reused by the next merge
    else rv_wave_sync();
    h2 = h2 > 0.0f ? h2 : 0.0f;
    float o = rv_bf16(mlp.w3[j]) * h2;
    // butterfly sum over the wave (partners lane ^ 32, 16, ..., 1, in that
    // order) by permlane, swizzle and DPP exchanges
    o += __uint_as_float(rv_xor_lane<32>(__float_as_uint(o)));
    o += __uint_as_float(rv_xor_lane<16>(__float_as_uint(o)));
    o += __uint_as_float(rv_xor_lane<8>(__float_as_uint(o)));
    o += __uint_as_float(rv_xor_lane<4>(__float_as_uint(o)));
    o += __uint_as_float(rv_xor_lane<2>(__float_as_uint(o)));
    o += __uint_as_float(rv_xor_lane<1>(__float_as_uint(o)));
    return o + mlp.b3[0];
}

// ------------------------------------------------------------- the kernel

// STOP is the learned-stop mode (b.stop_mode): 0 plain, 1 predict and stop,
// 2 record traces, 3 as 1 with suspension and resumption (b.phase, for
// RoverSearchBatch.reorder_step; also the plain search, the predictor then
// only ordering the resumption).  Separate instantiations keep the plain search free of
// the predictor's code and registers.  WIN is the source window (b.window):
// 1 expands the best unexpanded pool member and merges its candidates before
// choosing the next (the scalar search); 2 is the live pipeline with two
// source turns in flight (pure_pq_source_pipeline_step,
// retrieval_sqmd_pure_pq.c:722-762): the next source is chosen, and its
// neighbor row requested, before the previous turn's candidates are merged,
// so the merge runs while the row load is in flight.
template <int SCORER, int STOP, int WIN>
// The plain two-lane online instantiation is held to 128 VGPRs (4 waves per
// SIMD); the few values that spill live outside the scoring loop.  Held to
// the same budget, the learned-stop instantiation spills into its feature
// and predictor code and runs slower than at 3 waves per SIMD.
__global__ __launch_bounds__(RV_WAVE)
__attribute__((amdgpu_waves_per_eu((SCORER == RV_SCORE_ONLINE_PAIR32 && STOP == 0) ? 4 : ((SCORER == RV_SCORE_ONLINE_PAIR32 || SCORER == RV_SCORE_ONLINE_PAIR32_R4) && (STOP == 1 || STOP == 3)) ? 3 : 1)))
void rv_search_kernel(RvIndex ix, RvBatch b, RvLayout lay) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *qv = reinterpret_cast<float *>(smem + lay.off_q);
    float *lut = reinterpret_cast<float *>(smem + lay.off_tab);
    uint8_t *qsq = reinterpret_cast<uint8_t *>(smem + lay.off_tab);
    // the pool: keys ascending, expanded flags, and (trace mode) the step
    // at which each entry was scored
    uint64_t *pool_keys = reinterpret_cast<uint64_t *>(smem + lay.off_keys);
    uint8_t *pool_flags = smem + lay.off_flags;
    constexpr bool tracing = STOP == 2;
    // STOP 3: the learned-stop search that can be suspended and resumed
    // (RvBatch.phase); the other instantiations carry none of that code
    constexpr bool susp = STOP == 3;
    uint16_t *pool_born = tracing ? reinterpret_cast<uint16_t *>(smem + lay.off_born) : nullptr;
    uint64_t *cand = reinterpret_cast<uint64_t *>(smem + lay.off_cand);
    uint32_t *hash = reinterpret_cast<uint32_t *>(smem + lay.off_hash);
    const int lane = rv_lane();
    const int hash_n = 1 << b.hash_log2;
    const int expand = b.expand_count < ix.degree ? b.expand_count : ix.degree;
    constexpr bool pairable =
        SCORER == RV_SCORE_ONLINE_PAIR32 || SCORER == RV_SCORE_ONLINE_PAIR32_R4;
    constexpr int pair_rb = SCORER == RV_SCORE_ONLINE_PAIR32_R4 ? 4 : 8;
    const float sq_scale_sq = ix.sq_scale * ix.sq_scale;

    int nq = b.nq;
    if (b.nq_dev) {
        const int live = *b.nq_dev;
        nq = live < nq ? (live > 0 ? live : 0) : nq;
    }
    nq = rv_uniform(nq);
    // exact visited sets: this wave's region and stamp counter (kept in a
    // register while the wave runs, stored back when it leaves)
    const bool gvis = b.visited != nullptr;
    // suspended and resumed searches (always launched in phase 1 or 2)
    // keep one set per query (the region of query q, its stamp in the first
    // word); the others one per wave
    constexpr bool vq = susp;
    uint64_t *vwords = gvis && !vq ? b.visited + (size_t)blockIdx.x * (b.visited_words + 1)
                                   : nullptr;
    uint32_t vstamp = gvis && !vq ? rv_uniform((int)(uint32_t)vwords[0]) : 0u;
    if (gvis && !vq) vwords += 1;
    auto visit = [&](uint32_t id) -> bool {
        return gvis ? rv_visit_exact(vwords, vstamp, id)
                    : rv_visit(hash, b.hash_log2, b.probe_limit, id);
    };
    for (;;) {
        int q = 0;
        if (lane == 0) q = atomicAdd(b.queue, 1);
        q = rv_uniform(__shfl(q, 0));
        if (q >= nq) break;
        const bool resume = susp && b.phase == 2;
        if (resume) {
            const RvSuspLayout sl = rv_susp_layout((size_t)b.nq, (size_t)b.max_pool);
            q = rv_uniform(reinterpret_cast<const int32_t *>(b.susp + sl.order)[q]);
        }

        const int L = b.pool_sizes ? b.pool_sizes[q] : b.pool_size;
        const int K = b.num_outs ? b.num_outs[q] : b.num_out;
        int32_t *oid = b.out_ids + (size_t)q * b.out_stride;
        float *odist = b.out_dists + (size_t)q * b.out_stride;
        const int refine = b.refine_candidates;
        if (L <= 0 || L > b.max_pool || K <= 0 || K > L || K > b.out_stride) {
            if (lane == 0) b.out_counts[q] = RV_QUERY_BAD_PARAM;
            if (tracing && lane < RV_TRACE_META)
                b.trace_meta[(size_t)q * RV_TRACE_META + lane] = 0;
            continue;
        }

        // ---- per-query setup: query vector, table, visited set
        const float *qg = b.queries + (size_t)q * ix.dim;
        for (int d = lane; d < ix.dim; d += RV_WAVE) qv[d] = qg[d];
        if (gvis && vq) {
            // the query's own set: a new stamp when it starts, the saved
            // one when it resumes
            vwords = b.visited + (size_t)q * (b.visited_words + 1);
            vstamp = rv_uniform((int)(uint32_t)vwords[0]);
            vwords += 1;
        }
        if (gvis && !resume) {
            // a fresh stamp empties the set; after 2^32 - 1 queries on this
            // wave (or this query's region) the stamps would repeat, so the
            // words are cleared first
            vstamp++;
            if (vstamp == 0u) {
                for (int i = lane; i < b.visited_words; i += RV_WAVE) vwords[i] = 0ull;
                __threadfence();
                vstamp = 1u;
            }
            if (vq && lane == 0) vwords[-1] = (uint64_t)vstamp;
        } else if (!gvis) {
            for (int i = lane; i < hash_n; i += RV_WAVE) hash[i] = RV_EMPTY_SLOT;
        }
        __syncthreads();
        if (SCORER == RV_SCORE_ADC) {
            const int entries = ix.m * ix.ks;
            for (int e = lane; e < entries; e += RV_WAVE) {
                const int sub = e / ix.ks;
                const float *qs = qv + sub * ix.subdim;
                const float *c = ix.codebook + (size_t)e * ix.subdim;
                float acc = 0.0f;
                for (int d = 0; d < ix.subdim; d++) {
                    float diff = qs[d] - c[d];
                    acc += diff * diff;
                }
                lut[e] = acc;
            }
        } else if (SCORER == RV_SCORE_SQ) {
            const float inv = 1.0f / ix.sq_scale;
            const int hi = ix.sq_bits == 4 ? 15 : 255;
            for (int d = lane; d < ix.dim; d += RV_WAVE) {
                int v = (int)roundf((qv[d] - ix.sq_zero) * inv);
                v = v < 0 ? 0 : (v > hi ? hi : v);
                qsq[d] = (uint8_t)v;
            }
        }
        __syncthreads();

        int n = 0;          // pool size
        int status = 0;
        uint64_t st_scored = 0, st_expanded = 0, st_dedup = 0;
        // learned-stop state lives in LDS: it is touched once per expansion
        // and kept out of the registers the scoring needs
        RvStopState &ss = *reinterpret_cast<RvStopState *>(smem + lay.off_stop);
        if (STOP != 0) {
            ss.d_entry = 0.0f;
            ss.d_entry2 = 0.0f;
            ss.top1_0 = 0.0f;
            ss.topk_0 = 0.0f;
            ss.improve_ema = 1.0f;
            ss.last_change = 0;
            ss.last_top1 = 0;
        }
        int born_step = 0;  // scored-at step given to new pool entries
        int trace_n = 0;

        // Merge one candidate per lane (key, or RV_KEY_MAX for none) into
        // the pool.  Returns the lowest position a new candidate landed at
        // (or INT_MAX) via *first_new, the count admitted, and the new
        // position of the pool entry at old position `track` via *tracked.
        // span: where the candidate keys are (0: lanes 0..31, 1: lanes
        // 32..63, 2: anywhere); the other lanes hold RV_KEY_MAX
        auto merge = [&](uint64_t key, int span, int track, int *first_new, int *admitted,
                         int *tracked) {
            if (n == L && key != RV_KEY_MAX && key >= pool_keys[L - 1]) key = RV_KEY_MAX;
            *first_new = 0x7fffffff;
            *admitted = 0;
            *tracked = track;
            if (__ballot(key != RV_KEY_MAX) == 0ull) return;
            if (span == 2) {
                key = rv_sort_wave(key);
            } else {
                key = rv_sort_halves(key);
                if (span == 1) {
                    // the sorted upper half moves to lanes 0..31
                    const uint64_t up = rv_xor_lane64<32>(key);
                    key = lane < 32 ? up : RV_KEY_MAX;
                }
            }
            const uint64_t prev = rv_prev_lane64(key);
            bool valid = key != RV_KEY_MAX && !(lane > 0 && prev == key);
            int lb = 0;
            if (valid) {
                lb = rv_lower_bound(pool_keys, n, key);
                if (lb < n && pool_keys[lb] == key) valid = false;  // already pooled
            }
            const uint64_t vmask = __ballot(valid);
            const int cv = __popcll(vmask);
            if (cv == 0) return;
            const int rank = __popcll(vmask & ((1ull << lane) - 1ull));
            if (track >= 0 && track < n) {
                const uint64_t tk = pool_keys[track];
                *tracked = track + rv_uniform(__popcll(__ballot(valid && key < tk)));
            }
            // Pool entry j moves right by #{valid c : lb_c <= j}: no
            // candidate equals a pool key, so candidate c sorts below entry
            // j exactly when lb_c <= j.  New positions increase with j, so
            // moving the pool in place, last 128-entry chunk first and each
            // chunk read before it is written, never overwrites an entry
            // that has not moved yet; entries left of the smallest lb stay.
            // Lane l moves entries j0 + 2l and j0 + 2l + 1, read with one
            // 16-byte load (the pair's flags with one 2-byte load).  The
            // valid lanes hold ascending keys and lbs; walking them from the
            // top, `above` counts those right of the current chunk.
            const int lb_min =
                __builtin_amdgcn_readlane(lb, __ffsll((unsigned long long)vmask) - 1);
            uint64_t rem = vmask;
            int above = 0;
            for (int j0 = ((n - 1) >> 7) << 7; j0 >= 0 && j0 + 2 * RV_WAVE > lb_min;
                 j0 -= 2 * RV_WAVE) {
                const int ja = j0 + 2 * lane;
                int right_a = above, right_b = above;
                while (rem) {
                    const int top = 63 - __clzll((long long)rem);
                    const int lbc = __builtin_amdgcn_readlane(lb, top);
                    if (lbc < j0) break;
                    right_a += lbc > ja ? 1 : 0;
                    right_b += lbc > ja + 1 ? 1 : 0;
                    above++;
                    rem &= ~(1ull << top);
                }
                // shift_b >= shift_a (entry ja + 1 has every candidate of
                // entry ja below it, and those landing between them)
                const int shift_a = cv - right_a, shift_b = cv - right_b;
                if (ja < n && shift_b > 0) {
                    // the pair is read whole; indices up to max_pool stay
                    // inside the LDS layout (each array is padded to 16
                    // bytes and followed by another)
                    const ulonglong2 pk = *reinterpret_cast<const ulonglong2 *>(pool_keys + ja);
                    const uint16_t pf = *reinterpret_cast<const uint16_t *>(pool_flags + ja);
                    const uint32_t pb =
                        tracing ? *reinterpret_cast<const uint32_t *>(pool_born + ja) : 0u;
                    const int npa = ja + shift_a, npb = ja + 1 + shift_b;
                    if (shift_a > 0 && npa < L) {
                        pool_keys[npa] = pk.x;
                        pool_flags[npa] = (uint8_t)pf;
                        if (tracing) pool_born[npa] = (uint16_t)pb;
                    }
                    if (ja + 1 < n && npb < L) {
                        pool_keys[npb] = pk.y;
                        pool_flags[npb] = (uint8_t)(pf >> 8);
                        if (tracing) pool_born[npb] = (uint16_t)(pb >> 16);
                    }
                }
            }
            int pos = 0x7fffffff;
            if (valid) {
                pos = lb + rank;
                if (pos < L) {
                    pool_keys[pos] = key;
                    pool_flags[pos] = 0;
                    if (tracing) pool_born[pos] = (uint16_t)born_step;
                } else {
                    pos = 0x7fffffff;
                }
            }
            // keys are sorted and the valid ones ranked in lane order, so the
            // lowest new position is the first valid lane's
            *first_new = __builtin_amdgcn_readlane(pos, __ffsll((unsigned long long)vmask) - 1);
            *admitted = rv_uniform(__popcll(__ballot(pos != 0x7fffffff)));
            n = (n + cv) < L ? (n + cv) : L;
            __syncthreads();
        };

        // ---- a suspended search resumes: pool, flags, counters, scan
        // position, step and learned-stop state as it left them
        int r_scan = 0, r_step = 0;
        if (resume) {
            const RvSuspLayout sl = rv_susp_layout((size_t)b.nq, (size_t)b.max_pool);
            const int32_t *mt =
                reinterpret_cast<const int32_t *>(b.susp + sl.meta) + (size_t)q * RV_SUSP_META;
            n = rv_uniform(mt[0]);
            r_scan = rv_uniform(mt[1]);
            r_step = rv_uniform(mt[2]);
            st_scored = (uint32_t)rv_uniform(mt[3]);
            st_expanded = (uint32_t)rv_uniform(mt[4]);
            st_dedup = (uint32_t)rv_uniform(mt[5]);
            const uint64_t *sk = reinterpret_cast<const uint64_t *>(b.susp) + (size_t)q * b.max_pool;
            const uint8_t *sf = b.susp + sl.flags + (size_t)q * b.max_pool;
            for (int i = lane; i < n; i += RV_WAVE) {
                pool_keys[i] = sk[i];
                pool_flags[i] = sf[i];
            }
            if (STOP != 0) {
                ss.d_entry = __int_as_float(mt[6]);
                ss.d_entry2 = __int_as_float(mt[7]);
                ss.top1_0 = __int_as_float(mt[8]);
                ss.topk_0 = __int_as_float(mt[9]);
                ss.improve_ema = __int_as_float(mt[10]);
                ss.last_change = mt[11];
                ss.last_top1 = mt[12];
            }
            __syncthreads();
        }

        // ---- seeds (rpc_demo_pure_pq.c:192-208 seeds; :843-873 scoring);
        // one shared list, or a row of entry_stride ids per query
        const int32_t *ents =
            b.entries + (b.entry_stride > 0 ? (size_t)q * b.entry_stride : 0);
        // with two-lane online scoring, 32 entries per round: lane slot owns
        // the entry (range and visited tests), lane slot + 32 its key
        const int per_round = pairable ? 32 : RV_WAVE;
        const int slot = pairable ? (lane & 31) : lane;
        const bool owner = !pairable || lane < 32;
        const int key_lane0 = pairable ? 32 : 0;
        for (int e0 = 0; !resume && e0 < b.n_entries && status == 0; e0 += per_round) {
            const int e = e0 + slot;
            uint64_t key = RV_KEY_MAX;
            bool bad = false, nonfinite = false, dup = false;
            if (pairable) {
                int id = -1, fresh = 0;
                if (owner && e < b.n_entries) {
                    id = ents[e];
                    if (id < 0 || id >= ix.n) bad = true;
                    else if (!visit((uint32_t)id)) dup = true;
                    else fresh = 1;
                }
                id = (int)rv_low_half((uint32_t)id);
                fresh = (int)rv_low_half((uint32_t)fresh);
                const bool active = fresh != 0;
                float d = 0.0f;
                const bool ok = rv_online_pair32<pair_rb>(ix, qv, id, active, lane >= 32, &d);
                if (lane >= 32 && active) {
                    if (!ok || !isfinite(d) || d == FLT_MAX) nonfinite = true;
                    else key = rv_key(d + 0.0f, id);
                }
            } else if (e < b.n_entries) {
                const int id = ents[e];
                if (id < 0 || id >= ix.n) {
                    bad = true;
                } else if (!visit((uint32_t)id)) {
                    dup = true;
                } else {
                    float d = 0.0f;
                    bool ok = true;
                    if (SCORER == RV_SCORE_ADC) {
                        d = rv_adc(lut, ix.codes + (size_t)id * ix.code_bytes, ix.m,
                                   ix.ks, ix.code_bytes);
                    } else if (SCORER == RV_SCORE_ADC_ONLINE) {
                        ok = rv_online(ix, qv, id, &d);
                    } else {
                        d = rv_sq(qsq, ix.sq_codes + (size_t)id * ix.sq_code_bytes,
                                  ix.dim, ix.sq_bits, sq_scale_sq);
                    }
                    if (!ok || !isfinite(d) || d == FLT_MAX) nonfinite = true;
                    else key = rv_key(d + 0.0f, id);
                }
            }
            // the reference validates entries one at a time: an invalid entry
            // fails the query whatever follows it
            if (__ballot(bad) != 0ull) status = RV_QUERY_BAD_ENTRY;
            else if (__ballot(nonfinite) != 0ull) status = RV_QUERY_NONFINITE;
            if (status) break;
            st_scored += __popcll(__ballot(key != RV_KEY_MAX));
            st_dedup += __popcll(__ballot(dup));
            if (STOP != 0 && e0 == 0 && lane == key_lane0 && key != RV_KEY_MAX)
                ss.d_entry = rv_key_dist(key);
            if (STOP != 0 && e0 == 0 && lane == key_lane0 + 1 && key != RV_KEY_MAX)
                ss.d_entry2 = rv_key_dist(key);
            int fn, adm, tr;
            merge(key, pairable ? 1 : 2, -1, &fn, &adm, &tr);
        }
        if (!resume && status == 0 && n == 0) status = RV_QUERY_BAD_ENTRY;
        // A long seed list would fill the visited hash with far-away seeds.
        // Keep only the pool members then: a seed outside the pool that the
        // search meets again is re-scored and rejected (its key is worse
        // than the pool's worst), so results do not change; only the
        // scored/dedup counters then differ from the reference's.
        if (status == 0 && !gvis && b.n_entries * 4 > hash_n) {
            for (int i = lane; i < hash_n; i += RV_WAVE) hash[i] = RV_EMPTY_SLOT;
            __syncthreads();
            const uint64_t *k0 = pool_keys;
            for (int i = lane; i < n; i += RV_WAVE)
                rv_visit(hash, b.hash_log2, b.probe_limit, (uint32_t)rv_key_id(k0[i]));
            __syncthreads();
        }
        if (STOP != 0 && status == 0 && !resume) {
            // the first two entries' distances were taken in the lanes that
            // held their keys; every lane computes features from them
            ss.d_entry = rv_lane_value(ss.d_entry, key_lane0);
            ss.d_entry2 = rv_lane_value(ss.d_entry2, key_lane0 + 1);
            const uint64_t *k0 = pool_keys;
            ss.top1_0 = rv_key_dist(k0[0]);
            ss.topk_0 = rv_key_dist(k0[(n < b.stop_label_k ? n : b.stop_label_k) - 1]);
        }

        // ---- main loop: expand the best unexpanded pool member
        int scan = r_scan;
        int step = r_step;
        bool stopped = false;
        // status while the search is suspended (internal; never output)
        constexpr int suspended_status = 1;
        // window 2: the last dispatched turn's candidate keys wait here and
        // are merged once the next source has been chosen
        uint64_t pend_key = RV_KEY_MAX;
        int pend_span = 0;
        bool pending = false;
        auto after_merge = [&](int first_new) {
            if (STOP != 0) {
                ss.improve_ema =
                    0.7f * ss.improve_ema + (first_new < b.stop_label_k ? 0.3f : 0.0f);
                if (first_new < b.stop_label_k) ss.last_change = step;
                if (first_new == 0) ss.last_top1 = step;
            }
        };
        while (status == 0) {
            const uint64_t *keys = pool_keys;
            uint8_t *flags = pool_flags;
            int next = -1;
            while (scan < n) {
                const int p = scan + lane;
                const bool open = p < n && flags[p] == 0;
                const uint64_t m = __ballot(open);
                if (m) {
                    next = scan + __ffsll((unsigned long long)m) - 1;
                    break;
                }
                scan += RV_WAVE;
            }
            next = rv_uniform(next);
            if (next < 0) {
                if (WIN == 1 || !pending) break;
                // nothing to dispatch: collect the pending turn and look again
                int first_new, admitted, unused;
                merge(pend_key, pend_span, -1, &first_new, &admitted, &unused);
                pending = false;
                after_merge(first_new);
                scan = first_new < scan ? first_new : scan;
                continue;
            }

            // learned stop: features of the current state
            const bool suspend_now = susp && b.phase == 1 && step >= b.suspend_step;
            float p_now = -1.0f;
            if (STOP == 2 && step > 0 && step % b.stop_every == 0) {
                const float f = rv_feature(ss, keys, n, b.stop_label_k, step, next, lane);
                if (trace_n < b.trace_cap) {
                    if (lane < RV_MLP_IN)
                        b.trace_feats[((size_t)q * b.trace_cap + trace_n) * RV_MLP_IN + lane] = f;
                    if (lane == 0) b.trace_steps[(size_t)q * b.trace_cap + trace_n] = step;
                    trace_n++;
                }
            } else if (STOP == 1 || STOP == 3) {
                const bool regular = !(susp && b.stop_probe_only) && step > 0 &&
                                     step % b.stop_every == 0 && step >= b.stop_min_steps;
                if (regular || suspend_now) {
                    // the reduction leaves the same logit in every lane
                    const float f = rv_feature(ss, keys, n, b.stop_label_k, step, next, lane);
                    const float logit =
                        rv_mlp_logit(b.mlp, f, reinterpret_cast<float *>(cand));
                    p_now = 1.0f / (1.0f + __expf(-logit));
                    if (regular && rv_uniform(p_now >= b.stop_threshold ? 1 : 0)) {
                        stopped = true;
                        break;
                    }
                }
            }
            if (suspend_now) {
                // leave the search as it stands (saved below); it resumes
                // in phase 2
                ss.suspend_p = p_now;
                status = suspended_status;
                break;
            }

            const uint64_t src_key = keys[next];
            const int src = rv_key_id(src_key);
            if (lane == 0) flags[next] = 1;

            // neighbor row and its validity (retrieval_sqmd_pure_pq.c:59-87)
            int nb = -1;
            if (lane < expand) nb = ix.graph[(size_t)src * ix.degree + lane];
            int src_pos = next;
            if (WIN == 2 && pending) {
                // collect the previous turn while the row is in flight; the
                // source (flagged) moves with the pool
                int first_new, admitted;
                merge(pend_key, pend_span, next, &first_new, &admitted, &src_pos);
                pending = false;
                after_merge(first_new);
                scan = first_new < src_pos + 1 ? first_new : src_pos + 1;
            } else if (WIN == 2) {
                scan = next + 1;
            }
            const uint64_t pad = __ballot(lane < expand && nb == -1);
            const int valid_n = pad ? (__ffsll((unsigned long long)pad) - 1) : expand;
            const uint64_t after = __ballot(lane >= valid_n && lane < expand && nb != -1);
            const uint64_t oob = __ballot(lane < valid_n && (nb < 0 || nb >= ix.n));
            if (after | oob) {
                status = RV_QUERY_BAD_ROW;
                break;
            }
            uint64_t key = RV_KEY_MAX;
            bool dup = false, nonfinite = false;
            int span = expand <= 32 ? 0 : 2;  // where the keys are, for the merge
            if (pairable) {
                // lane i < valid_n owns neighbor i's visited test
                int fresh = 0;
                if (lane < valid_n) {
                    fresh = visit((uint32_t)nb) ? 1 : 0;
                    dup = !fresh;
                }
                const uint64_t fmask = __ballot(fresh != 0);
                const int nf = __popcll(fmask);
                if (nf > 8) {
                    // neighbor i is scored by lanes i and i + 32; the high
                    // lane holds the key
                    const int nbs = (int)rv_low_half((uint32_t)nb);
                    const bool active = rv_low_half((uint32_t)fresh) != 0u;
                    float d = 0.0f;
                    const bool ok =
                        rv_online_pair32<pair_rb>(ix, qv, nbs, active, lane >= 32, &d);
                    if (lane >= 32 && active) {
                        if (!ok || !isfinite(d) || d == FLT_MAX) nonfinite = true;
                        else key = rv_key(d + 0.0f, nbs);
                    }
                    span = 1;
                } else if (nf > 0) {
                    // the unvisited neighbors, compacted: candidate r is
                    // scored by the G lanes r*G .. r*G + G - 1 and its key
                    // lands in lane r
                    int *cids = reinterpret_cast<int *>(cand);
                    if (fresh) cids[__popcll(fmask & ((1ull << lane) - 1ull))] = nb;
                    __syncthreads();
                    const int G = nf <= 4 ? 16 : 8;
                    const int r = lane / G;
                    const bool active = r < nf;
                    const int id = active ? cids[r] : 0;
                    float d = 0.0f;
                    bool ok;
                    if (G == 16) ok = rv_online_group<16, pair_rb>(ix, qv, id, active, lane & 15, &d);
                    else ok = rv_online_group<8, pair_rb>(ix, qv, id, active, lane & 7, &d);
                    const int from = lane < nf ? lane * G + G - 1 : 63;
                    const float dr = __shfl(d, from);
                    const int okr = __shfl(ok ? 1 : 0, from);
                    const int idr = __shfl(id, from);
                    if (lane < nf) {
                        if (!okr || !isfinite(dr) || dr == FLT_MAX) nonfinite = true;
                        else key = rv_key(dr + 0.0f, idr);
                    }
                    span = 0;
                    __syncthreads();  // cids is reused by the next expansion
                }
            } else if (lane < valid_n) {
                if (!visit((uint32_t)nb)) {
                    dup = true;
                } else {
                    float d = 0.0f;
                    bool ok = true;
                    if (SCORER == RV_SCORE_ADC) {
                        d = rv_adc(lut, ix.codes + (size_t)nb * ix.code_bytes, ix.m,
                                   ix.ks, ix.code_bytes);
                    } else if (SCORER == RV_SCORE_ADC_ONLINE) {
                        ok = rv_online(ix, qv, nb, &d);
                    } else {
                        d = rv_sq(qsq, ix.sq_codes + (size_t)nb * ix.sq_code_bytes,
                                  ix.dim, ix.sq_bits, sq_scale_sq);
                    }
                    if (!ok || !isfinite(d) || d == FLT_MAX) nonfinite = true;
                    else key = rv_key(d + 0.0f, nb);
                }
            }
            if (__ballot(nonfinite) != 0ull) {
                status = RV_QUERY_NONFINITE;
                break;
            }
            st_expanded++;
            st_scored += __popcll(__ballot(key != RV_KEY_MAX));
            st_dedup += __popcll(__ballot(dup));
            step++;

            if (WIN == 2) {
                // this turn's candidates wait for the next dispatch
                pend_key = key;
                pend_span = span;
                pending = true;
                continue;
            }
            int first_new, admitted;
            born_step = step < 65535 ? step : 65535;
            merge(key, span, next, &first_new, &admitted, &src_pos);
            (void)admitted;
            after_merge(first_new);
            // everything left of the expanded source is expanded, except the
            // new candidates; resume the scan at the first of either
            scan = first_new < src_pos + 1 ? first_new : src_pos + 1;
        }
        if (WIN == 2 && pending && status == 0) {
            // a stopped search still takes the scores of its last turn
            int first_new, admitted, unused;
            merge(pend_key, pend_span, -1, &first_new, &admitted, &unused);
        }
        (void)stopped;

        if (tracing) {
            // records, last top-K change, and the scored steps of the final
            // top-K ascending (padded with the largest); a failed query
            // keeps no records
            int32_t *mq = b.trace_meta + (size_t)q * RV_TRACE_META;
            const int kk = status == 0 ? (n < b.stop_label_k ? n : b.stop_label_k) : 0;
            const uint16_t *born = pool_born;
            const int v = lane < kk ? (int)born[lane] : -1;
            int rank = 0;
            for (int j = 0; j < kk; j++) {
                const int u = born[j];
                rank += (u < v) || (u == v && j < lane);
            }
            int mx = v;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const int o = __shfl_xor(mx, off);
                mx = o > mx ? o : mx;
            }
            if (lane < kk) mq[2 + rank] = v;
            else if (lane < RV_TRACE_TOPK) mq[2 + lane] = mx < 0 ? 0 : mx;
            if (lane == 0) {
                mq[0] = status == 0 ? trace_n : 0;
                mq[1] = ss.last_change;
            }
        }

        if (susp && status == suspended_status) {
            const RvSuspLayout sl = rv_susp_layout((size_t)b.nq, (size_t)b.max_pool);
            uint64_t *sk = reinterpret_cast<uint64_t *>(b.susp) + (size_t)q * b.max_pool;
            uint8_t *sf = b.susp + sl.flags + (size_t)q * b.max_pool;
            for (int i = lane; i < n; i += RV_WAVE) {
                sk[i] = pool_keys[i];
                sf[i] = pool_flags[i];
            }
            int32_t *mt = reinterpret_cast<int32_t *>(b.susp + sl.meta) + (size_t)q * RV_SUSP_META;
            if (lane == 0) {
                mt[0] = n;
                mt[1] = scan;
                mt[2] = step;
                mt[3] = (int32_t)st_scored;
                mt[4] = (int32_t)st_expanded;
                mt[5] = (int32_t)st_dedup;
                mt[6] = __float_as_int(ss.d_entry);
                mt[7] = __float_as_int(ss.d_entry2);
                mt[8] = __float_as_int(ss.top1_0);
                mt[9] = __float_as_int(ss.topk_0);
                mt[10] = __float_as_int(ss.improve_ema);
                mt[11] = ss.last_change;
                mt[12] = ss.last_top1;
                reinterpret_cast<float *>(b.susp + sl.prob)[q] = ss.suspend_p;
                const int at = atomicAdd(reinterpret_cast<int32_t *>(b.susp + sl.count), 1);
                reinterpret_cast<int32_t *>(b.susp + sl.list)[at] = q;
            }
            __syncthreads();
            continue;
        }

        // ---- output
        const uint64_t *keys = pool_keys;
        if (status != 0) {
            if (lane == 0) b.out_counts[q] = status;
        } else if (refine == 0) {
            const int cnt = K < n ? K : n;
            for (int i = lane; i < cnt; i += RV_WAVE) {
                oid[i] = rv_key_id(keys[i]);
                odist[i] = rv_key_dist(keys[i]);
            }
            if (lane == 0) b.out_counts[q] = cnt;
        } else {
            // final refine of the first F pool members
            // (retrieval_final_refine.c:256-292 with the residual decoder,
            // plus the error4 level of decode_vector, :221-246, when asked)
            const int F = refine < n ? refine : n;
            const bool e4 = b.refine_error4 != 0;
            uint64_t *dbits = reinterpret_cast<uint64_t *>(smem + lay.off_refine);
            bool bad = false;
            for (int i = lane; i < F; i += RV_WAVE) {
                const int id = rv_key_id(keys[i]);
                const uint8_t *pq = ix.codes + (size_t)id * ix.code_bytes;
                const uint8_t *res = ix.res_codes + (size_t)id * ix.res_code_bytes;
                const float s = ix.res_qparams[2 * (size_t)id];
                const float z = ix.res_qparams[2 * (size_t)id + 1];
                const uint8_t *c4 = nullptr;
                float s4 = 0.0f, z4 = 0.0f;
                if (e4) {
                    c4 = ix.e4_blocks + (size_t)(id / ix.e4_vpb) * RV_E4_BLOCK +
                         (size_t)(id % ix.e4_vpb) * ix.e4_code_bytes;
                    s4 = ix.e4_qparams[2 * (size_t)id];
                    z4 = ix.e4_qparams[2 * (size_t)id + 1];
                }
                double acc = 0.0;
                if (!isfinite(s) || !isfinite(z) || !(s > 0.0f)) bad = true;
                int d = 0;
                for (int sub = 0; sub < ix.m; sub++) {
                    const float *cent = ix.codebook +
                        ((size_t)sub * ix.ks + rv_code_get(pq, sub, ix.ks)) * ix.subdim;
                    for (int t = 0; t < ix.subdim; t++, d++) {
                        float v = rv_decode_value(cent[t], rv_res_code(res, ix.res_bits, d), s, z);
                        if (!isfinite(v)) bad = true;
                        if (e4) {
                            // add_f32(v, (code * scale) + zero), each rounded
                            const float level =
                                (float)((c4[d >> 1] >> (4 * (d & 1))) & 15) * s4 + z4;
                            v = v + level;
                            if (!isfinite(v)) bad = true;
                        }
                        double delta = (double)qv[d] - (double)v;
                        acc += delta * delta;
                    }
                }
                if (!isfinite(acc) || acc >= (double)FLT_MAX) bad = true;
                dbits[i] = (uint64_t)__double_as_longlong(acc);
            }
            if (__ballot(bad) != 0ull) {
                if (lane == 0) b.out_counts[q] = RV_QUERY_NONFINITE;
            } else {
                __syncthreads();
                for (int i = lane; i < F; i += RV_WAVE) {
                    const uint64_t di = dbits[i];
                    const int gi = rv_key_id(keys[i]);
                    const int oi = ix.graph_to_original ? ix.graph_to_original[gi] : gi;
                    int rank = 0;
                    for (int j = 0; j < F; j++) {
                        const uint64_t dj = dbits[j];
                        const int gj = rv_key_id(keys[j]);
                        const int oj = ix.graph_to_original ? ix.graph_to_original[gj] : gj;
                        rank += (dj < di) || (dj == di && oj < oi);
                    }
                    if (rank < K) {
                        oid[rank] = gi;
                        odist[rank] = (float)__longlong_as_double((long long)di);
                    }
                }
                if (lane == 0) b.out_counts[q] = K < F ? K : F;
            }
        }
        if (b.out_stats && lane == 0) {
            uint64_t *o = b.out_stats + 4 * (size_t)q;
            o[0] = st_scored;
            o[1] = SCORER == RV_SCORE_ADC ? 0 : st_scored;
            o[2] = st_expanded;
            o[3] = st_dedup;
        }
        __syncthreads();
    }
    if (gvis && !vq && lane == 0) vwords[-1] = (uint64_t)vstamp;
}

// ------------------------------------------------ several waves per query

// The best-first search of rv_search_kernel with WPQ wavefronts per query (a
// workgroup of 64 * WPQ lanes sharing one LDS pool), for the online
// correction on the 32-subspace x 4-dim shape with 8- or 4-bit residuals.
// Wave 0 owns the pool and the visited set: it picks the source, reads its
// row, tests the neighbors and writes the unvisited ones to a compact LDS
// list; then every wave scores that list, b.cand_tile candidates per pass
// with 64 * WPQ / cand_tile lanes each (the grouped sequential sum, bit-
// identical to one lane's), and writes the keys back to LDS; wave 0 merges
// them.  Expansion order, ids, distances and counters are those of
// rv_search_kernel.  Control flow between the workgroup barriers is decided
// by wave 0 and read by all waves from LDS, so every wave passes the same
// barriers.
template <int RB, int STOP, int WPQ>
__global__ __launch_bounds__(RV_WAVE * WPQ)
void rv_search_mw_kernel(RvIndex ix, RvBatch b, RvLayout lay) {
    extern __shared__ __align__(16) unsigned char smem[];
    float *qv = reinterpret_cast<float *>(smem + lay.off_q);
    uint64_t *pool_keys = reinterpret_cast<uint64_t *>(smem + lay.off_keys);
    uint8_t *pool_flags = smem + lay.off_flags;
    uint64_t *cand = reinterpret_cast<uint64_t *>(smem + lay.off_cand);
    uint32_t *hash = reinterpret_cast<uint32_t *>(smem + lay.off_hash);
    int32_t *cids = reinterpret_cast<int32_t *>(smem + lay.off_mw);
    uint64_t *ckeys = reinterpret_cast<uint64_t *>(smem + lay.off_mw + 4 * RV_WAVE);
    // control words: [0] query, [1] source position (-1: the search ends),
    // [2] unvisited neighbors to score, [3] a score was not finite
    int32_t *ctl = reinterpret_cast<int32_t *>(smem + lay.off_mw + 12 * RV_WAVE);
    constexpr int NT = RV_WAVE * WPQ;
    const int gl = threadIdx.x;
    const int lane = rv_lane();
    const int wave = gl / RV_WAVE;
    const int hash_n = 1 << b.hash_log2;
    const int expand = b.expand_count < ix.degree ? b.expand_count : ix.degree;
    const int tile = b.cand_tile;
    const int G = NT / tile;  // lanes per candidate: 1, 2, 4, 8 or 16

    int nq = b.nq;
    if (b.nq_dev) {
        const int live = *b.nq_dev;
        nq = live < nq ? (live > 0 ? live : 0) : nq;
    }
    nq = rv_uniform(nq);
    const bool gvis = b.visited != nullptr;
    uint64_t *vwords = gvis ? b.visited + (size_t)blockIdx.x * (b.visited_words + 1) : nullptr;
    uint32_t vstamp = (gvis && wave == 0) ? rv_uniform((int)(uint32_t)vwords[0]) : 0u;
    if (gvis) vwords += 1;
    auto visit = [&](uint32_t id) -> bool {
        return gvis ? rv_visit_exact(vwords, vstamp, id)
                    : rv_visit(hash, b.hash_log2, b.probe_limit, id);
    };
    for (;;) {
        if (gl == 0) ctl[0] = atomicAdd(b.queue, 1);
        __syncthreads();
        const int q = ctl[0];
        if (q >= nq) break;
        const int L = b.pool_sizes ? b.pool_sizes[q] : b.pool_size;
        const int K = b.num_outs ? b.num_outs[q] : b.num_out;
        if (L <= 0 || L > b.max_pool || K <= 0 || K > L || K > b.out_stride) {
            if (gl == 0) b.out_counts[q] = RV_QUERY_BAD_PARAM;
            __syncthreads();
            continue;
        }
        const float *qg = b.queries + (size_t)q * ix.dim;
        for (int d = gl; d < ix.dim; d += NT) qv[d] = qg[d];
        if (wave == 0) {
            if (gvis) {
                vstamp++;
                if (vstamp == 0u) {
                    for (int i = lane; i < b.visited_words; i += RV_WAVE) vwords[i] = 0ull;
                    __threadfence();
                    vstamp = 1u;
                }
            } else {
                for (int i = lane; i < hash_n; i += RV_WAVE) hash[i] = RV_EMPTY_SLOT;
            }
        }
        __syncthreads();

        // ---- wave 0: seeds, then the pool (held in wave 0's registers)
        int n = 0, status = 0;
        uint64_t st_scored = 0, st_expanded = 0, st_dedup = 0;
        RvStopState &ss = *reinterpret_cast<RvStopState *>(smem + lay.off_stop);
        auto merge = [&](uint64_t key, int track, int *first_new, int *tracked) {
            if (n == L && key != RV_KEY_MAX && key >= pool_keys[L - 1]) key = RV_KEY_MAX;
            *first_new = 0x7fffffff;
            *tracked = track;
            if (__ballot(key != RV_KEY_MAX) == 0ull) return;
            key = rv_sort_halves(key);
            if (__ballot(lane >= 32 && key != RV_KEY_MAX) != 0ull) {
                // keys came in the upper half (seeds): move them down
                const uint64_t up = rv_xor_lane64<32>(key);
                key = lane < 32 ? up : RV_KEY_MAX;
            }
            const uint64_t prev = rv_prev_lane64(key);
            bool valid = key != RV_KEY_MAX && !(lane > 0 && prev == key);
            int lb = 0;
            if (valid) {
                lb = rv_lower_bound(pool_keys, n, key);
                if (lb < n && pool_keys[lb] == key) valid = false;
            }
            const uint64_t vmask = __ballot(valid);
            const int cv = __popcll(vmask);
            if (cv == 0) return;
            const int rank = __popcll(vmask & ((1ull << lane) - 1ull));
            if (track >= 0 && track < n) {
                const uint64_t tk = pool_keys[track];
                *tracked = track + rv_uniform(__popcll(__ballot(valid && key < tk)));
            }
            // in-place move of the pool (see rv_search_kernel's merge)
            const int lb_min =
                __builtin_amdgcn_readlane(lb, __ffsll((unsigned long long)vmask) - 1);
            uint64_t rem = vmask;
            int above = 0;
            for (int j0 = ((n - 1) >> 6) << 6; j0 >= 0 && j0 + RV_WAVE > lb_min;
                 j0 -= RV_WAVE) {
                const int j = j0 + lane;
                int right = above;
                while (rem) {
                    const int top = 63 - __clzll((long long)rem);
                    const int lbc = __builtin_amdgcn_readlane(lb, top);
                    if (lbc < j0) break;
                    right += lbc > j ? 1 : 0;
                    above++;
                    rem &= ~(1ull << top);
                }
                const int shift = cv - right;
                if (j < n && shift > 0) {
                    const uint64_t pk = pool_keys[j];
                    const uint8_t pf = pool_flags[j];
                    const int np = j + shift;
                    if (np < L) {
                        pool_keys[np] = pk;
                        pool_flags[np] = pf;
                    }
                }
            }
            int pos = 0x7fffffff;
            if (valid) {
                pos = lb + rank;
                if (pos < L) {
                    pool_keys[pos] = key;
                    pool_flags[pos] = 0;
                } else {
                    pos = 0x7fffffff;
                }
            }
            *first_new = __builtin_amdgcn_readlane(pos, __ffsll((unsigned long long)vmask) - 1);
            n = (n + cv) < L ? (n + cv) : L;
            rv_wave_sync();
        };
        if (wave == 0) {
            if (STOP != 0) {
                ss.d_entry = 0.0f;
                ss.d_entry2 = 0.0f;
                ss.top1_0 = 0.0f;
                ss.topk_0 = 0.0f;
                ss.improve_ema = 1.0f;
                ss.last_change = 0;
                ss.last_top1 = 0;
            }
            // seeds (rpc_demo_pure_pq.c:192-208): 32 per round, lane slot
            // owns the entry, lane slot + 32 its key
            const int32_t *ents =
                b.entries + (b.entry_stride > 0 ? (size_t)q * b.entry_stride : 0);
            const int slot = lane & 31;
            for (int e0 = 0; e0 < b.n_entries && status == 0; e0 += 32) {
                const int e = e0 + slot;
                uint64_t key = RV_KEY_MAX;
                bool bad = false, nonfinite = false, dup = false;
                int id = -1, fresh = 0;
                if (lane < 32 && e < b.n_entries) {
                    id = ents[e];
                    if (id < 0 || id >= ix.n) bad = true;
                    else if (!visit((uint32_t)id)) dup = true;
                    else fresh = 1;
                }
                id = (int)rv_low_half((uint32_t)id);
                fresh = (int)rv_low_half((uint32_t)fresh);
                const bool active = fresh != 0;
                float d = 0.0f;
                const bool ok = rv_online_pair32<RB>(ix, qv, id, active, lane >= 32, &d);
                if (lane >= 32 && active) {
                    if (!ok || !isfinite(d) || d == FLT_MAX) nonfinite = true;
                    else key = rv_key(d + 0.0f, id);
                }
                if (__ballot(bad) != 0ull) status = RV_QUERY_BAD_ENTRY;
                else if (__ballot(nonfinite) != 0ull) status = RV_QUERY_NONFINITE;
                if (status) break;
                st_scored += __popcll(__ballot(key != RV_KEY_MAX));
                st_dedup += __popcll(__ballot(dup));
                if (STOP != 0 && e0 == 0 && lane == 32 && key != RV_KEY_MAX)
                    ss.d_entry = rv_key_dist(key);
                if (STOP != 0 && e0 == 0 && lane == 33 && key != RV_KEY_MAX)
                    ss.d_entry2 = rv_key_dist(key);
                int fn, tr;
                merge(key, -1, &fn, &tr);
            }
            if (status == 0 && n == 0) status = RV_QUERY_BAD_ENTRY;
            if (status == 0 && !gvis && b.n_entries * 4 > hash_n) {
                for (int i = lane; i < hash_n; i += RV_WAVE) hash[i] = RV_EMPTY_SLOT;
                rv_wave_sync();
                for (int i = lane; i < n; i += RV_WAVE)
                    rv_visit(hash, b.hash_log2, b.probe_limit, (uint32_t)rv_key_id(pool_keys[i]));
                rv_wave_sync();
            }
            if (STOP != 0 && status == 0) {
                ss.d_entry = rv_lane_value(ss.d_entry, 32);
                ss.d_entry2 = rv_lane_value(ss.d_entry2, 33);
                ss.top1_0 = rv_key_dist(pool_keys[0]);
                ss.topk_0 = rv_key_dist(pool_keys[(n < b.stop_label_k ? n : b.stop_label_k) - 1]);
            }
        }

        // ---- main loop
        int scan = 0, step = 0, next = -1;
        for (;;) {
            if (wave == 0) {
                next = -1;
                if (status == 0) {
                    while (scan < n) {
                        const int p = scan + lane;
                        const bool open = p < n && pool_flags[p] == 0;
                        const uint64_t m = __ballot(open);
                        if (m) {
                            next = scan + __ffsll((unsigned long long)m) - 1;
                            break;
                        }
                        scan += RV_WAVE;
                    }
                    next = rv_uniform(next);
                }
                if (STOP == 1 && next >= 0 && step > 0 && step % b.stop_every == 0 &&
                    step >= b.stop_min_steps) {
                    const float f =
                        rv_feature(ss, pool_keys, n, b.stop_label_k, step, next, lane);
                    const float logit =
                        rv_mlp_logit<false>(b.mlp, f, reinterpret_cast<float *>(cand));
                    const float p = 1.0f / (1.0f + __expf(-logit));
                    if (rv_uniform(p >= b.stop_threshold ? 1 : 0)) next = -1;
                }
                int nf = 0;
                if (next >= 0) {
                    const int src = rv_key_id(pool_keys[next]);
                    if (lane == 0) pool_flags[next] = 1;
                    int nb = -1;
                    if (lane < expand) nb = ix.graph[(size_t)src * ix.degree + lane];
                    const uint64_t pad = __ballot(lane < expand && nb == -1);
                    const int valid_n = pad ? (__ffsll((unsigned long long)pad) - 1) : expand;
                    const uint64_t after = __ballot(lane >= valid_n && lane < expand && nb != -1);
                    const uint64_t oob = __ballot(lane < valid_n && (nb < 0 || nb >= ix.n));
                    if (after | oob) {
                        status = RV_QUERY_BAD_ROW;
                        next = -1;
                    } else {
                        int fresh = 0;
                        bool dup = false;
                        if (lane < valid_n) {
                            fresh = visit((uint32_t)nb) ? 1 : 0;
                            dup = !fresh;
                        }
                        const uint64_t fmask = __ballot(fresh != 0);
                        nf = __popcll(fmask);
                        if (fresh) cids[__popcll(fmask & ((1ull << lane) - 1ull))] = nb;
                        st_expanded++;
                        st_scored += nf;
                        st_dedup += __popcll(__ballot(dup));
                        step++;
                    }
                }
                if (lane == 0) {
                    ctl[1] = next;
                    ctl[2] = nf;
                    ctl[3] = 0;
                }
            }
            __syncthreads();
            if (ctl[1] < 0) break;
            const int nf = ctl[2];
            // every wave scores the list: pass `base` covers candidates
            // base .. base + tile - 1, candidate r on lanes r*G .. r*G+G-1
            for (int base = 0; base < nf; base += tile) {
                const int r = gl / G, sub = gl % G;
                const int idx = base + r;
                const bool active = idx < nf;
                const int id = active ? cids[idx] : 0;
                float d = 0.0f;
                bool ok;
                switch (G) {
                case 1: ok = rv_online_lane<RB>(ix, qv, id, &d); break;
                case 2: ok = rv_online_group<2, RB>(ix, qv, id, active, sub, &d); break;
                case 4: ok = rv_online_group<4, RB>(ix, qv, id, active, sub, &d); break;
                case 8: ok = rv_online_group<8, RB>(ix, qv, id, active, sub, &d); break;
                default: ok = rv_online_group<16, RB>(ix, qv, id, active, sub, &d); break;
                }
                if (active && sub == G - 1) {
                    if (!ok || !isfinite(d) || d == FLT_MAX) ctl[3] = 1;
                    else ckeys[idx] = rv_key(d + 0.0f, id);
                }
            }
            __syncthreads();
            if (wave == 0) {
                if (ctl[3]) {
                    status = RV_QUERY_NONFINITE;
                } else {
                    const uint64_t key = lane < nf ? ckeys[lane] : RV_KEY_MAX;
                    int first_new, src_pos;
                    merge(key, next, &first_new, &src_pos);
                    if (STOP != 0) {
                        ss.improve_ema = 0.7f * ss.improve_ema +
                                         (first_new < b.stop_label_k ? 0.3f : 0.0f);
                        if (first_new < b.stop_label_k) ss.last_change = step;
                        if (first_new == 0) ss.last_top1 = step;
                    }
                    scan = first_new < src_pos + 1 ? first_new : src_pos + 1;
                }
            }
        }

        // ---- output (wave 0)
        if (wave == 0) {
            int32_t *oid = b.out_ids + (size_t)q * b.out_stride;
            float *odist = b.out_dists + (size_t)q * b.out_stride;
            const int refine = b.refine_candidates;
            if (status != 0) {
                if (lane == 0) b.out_counts[q] = status;
            } else if (refine == 0) {
                const int cnt = K < n ? K : n;
                for (int i = lane; i < cnt; i += RV_WAVE) {
                    oid[i] = rv_key_id(pool_keys[i]);
                    odist[i] = rv_key_dist(pool_keys[i]);
                }
                if (lane == 0) b.out_counts[q] = cnt;
            } else {
                // final refine of the first F pool members, as rv_search_kernel
                const int F = refine < n ? refine : n;
                const bool e4 = b.refine_error4 != 0;
                uint64_t *dbits = reinterpret_cast<uint64_t *>(smem + lay.off_refine);
                bool bad = false;
                for (int i = lane; i < F; i += RV_WAVE) {
                    const int id = rv_key_id(pool_keys[i]);
                    const uint8_t *pq = ix.codes + (size_t)id * ix.code_bytes;
                    const uint8_t *res = ix.res_codes + (size_t)id * ix.res_code_bytes;
                    const float sc = ix.res_qparams[2 * (size_t)id];
                    const float zc = ix.res_qparams[2 * (size_t)id + 1];
                    const uint8_t *c4 = nullptr;
                    float s4 = 0.0f, z4 = 0.0f;
                    if (e4) {
                        c4 = ix.e4_blocks + (size_t)(id / ix.e4_vpb) * RV_E4_BLOCK +
                             (size_t)(id % ix.e4_vpb) * ix.e4_code_bytes;
                        s4 = ix.e4_qparams[2 * (size_t)id];
                        z4 = ix.e4_qparams[2 * (size_t)id + 1];
                    }
                    double acc = 0.0;
                    if (!isfinite(sc) || !isfinite(zc) || !(sc > 0.0f)) bad = true;
                    int d = 0;
                    for (int sub = 0; sub < ix.m; sub++) {
                        const float *cent = ix.codebook +
                            ((size_t)sub * ix.ks + rv_code_get(pq, sub, ix.ks)) * ix.subdim;
                        for (int t = 0; t < ix.subdim; t++, d++) {
                            float v = rv_decode_value(cent[t], rv_res_code(res, ix.res_bits, d),
                                                      sc, zc);
                            if (!isfinite(v)) bad = true;
                            if (e4) {
                                const float level =
                                    (float)((c4[d >> 1] >> (4 * (d & 1))) & 15) * s4 + z4;
                                v = v + level;
                                if (!isfinite(v)) bad = true;
                            }
                            double delta = (double)qv[d] - (double)v;
                            acc += delta * delta;
                        }
                    }
                    if (!isfinite(acc) || acc >= (double)FLT_MAX) bad = true;
                    dbits[i] = (uint64_t)__double_as_longlong(acc);
                }
                if (__ballot(bad) != 0ull) {
                    if (lane == 0) b.out_counts[q] = RV_QUERY_NONFINITE;
                } else {
                    rv_wave_sync();
                    for (int i = lane; i < F; i += RV_WAVE) {
                        const uint64_t di = dbits[i];
                        const int gi = rv_key_id(pool_keys[i]);
                        const int oi = ix.graph_to_original ? ix.graph_to_original[gi] : gi;
                        int rank = 0;
                        for (int j = 0; j < F; j++) {
                            const uint64_t dj = dbits[j];
                            const int gj = rv_key_id(pool_keys[j]);
                            const int oj = ix.graph_to_original ? ix.graph_to_original[gj] : gj;
                            rank += (dj < di) || (dj == di && oj < oi);
                        }
                        if (rank < K) {
                            oid[rank] = gi;
                            odist[rank] = (float)__longlong_as_double((long long)di);
                        }
                    }
                    if (lane == 0) b.out_counts[q] = K < F ? K : F;
                }
            }
            if (b.out_stats && lane == 0) {
                uint64_t *o = b.out_stats + 4 * (size_t)q;
                o[0] = st_scored;
                o[1] = st_scored;
                o[2] = st_expanded;
                o[3] = st_dedup;
            }
        }
        __syncthreads();
    }
    if (gvis && gl == 0) vwords[-1] = (uint64_t)vstamp;
}

__global__ void rv_zero_kernel(uint32_t *p, size_t words) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < words;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

__global__ void rv_zero_bytes_kernel(uint8_t *p, size_t bytes) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < bytes;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// -------------------------------------------------------------- encoders

// retrieval_pq_core.c:226-268 (retrieval_pq_encode_batch): argmin over k of
// the sequential sub-L2, first minimum wins.  One thread per (vector, m).
__global__ void rv_pq_encode_kernel(const float *x, int n, int dim, int m,
                                    int ks, const float *codebook,
                                    uint8_t *codes, int code_bytes) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)n * m) return;
    const int v = (int)(t / m), s = (int)(t % m);
    const int sub = dim / m;
    const float *xv = x + (size_t)v * dim + (size_t)s * sub;
    const float *cb = codebook + (size_t)s * ks * sub;
    float best = FLT_MAX;
    int best_k = 0;
    for (int k = 0; k < ks; k++) {
        float acc = 0.0f;
        for (int d = 0; d < sub; d++) {
            float diff = xv[d] - cb[(size_t)k * sub + d];
            acc += diff * diff;
        }
        if (acc < best) {
            best = acc;
            best_k = k;
        }
    }
    if (ks <= 16) {
        // two subspaces share a byte: the even thread writes both nibbles
        if ((s & 1) == 0) {
            int hi = 0;
            if (s + 1 < m) {
                const float *xv2 = xv + sub;
                const float *cb2 = cb + (size_t)ks * sub;
                float best2 = FLT_MAX;
                for (int k = 0; k < ks; k++) {
                    float acc = 0.0f;
                    for (int d = 0; d < sub; d++) {
                        float diff = xv2[d] - cb2[(size_t)k * sub + d];
                        acc += diff * diff;
                    }
                    if (acc < best2) {
                        best2 = acc;
                        hi = k;
                    }
                }
            }
            codes[(size_t)v * code_bytes + (s >> 1)] = (uint8_t)(best_k | (hi << 4));
        }
    } else {
        codes[(size_t)v * code_bytes + s] = (uint8_t)best_k;
    }
}

// retrieval_pq_residual.c:314-385 (pq_residual_quantize_vector).  One
// wavefront per vector: lanes reduce min/max, then quantize.
__global__ __launch_bounds__(RV_WAVE) void rv_residual_encode_kernel(
    const float *x, const uint8_t *codes, int n, int dim, int m, int ks,
    const float *codebook, int code_bytes, int bits, uint8_t *res,
    int res_bytes, float *qparams) {
    const int v = blockIdx.x;
    const int lane = rv_lane();
    if (v >= n) return;
    const int sub = dim / m;
    const uint8_t *code = codes + (size_t)v * code_bytes;
    const float *xv = x + (size_t)v * dim;
    float lo = FLT_MAX, hi = -FLT_MAX;
    bool bad = false;
    for (int d = lane; d < dim; d += RV_WAVE) {
        const int s = d / sub;
        float c = codebook[((size_t)s * ks + rv_code_get(code, s, ks)) * sub + (d % sub)];
        float r = xv[d] - c;
        if (!isfinite(r)) bad = true;
        lo = r < lo ? r : lo;
        hi = r > hi ? r : hi;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        float a = __shfl_xor(lo, off), b2 = __shfl_xor(hi, off);
        lo = a < lo ? a : lo;
        hi = b2 > hi ? b2 : hi;
    }
    const int levels = bits == 4 ? 15 : 255;
    float scale = (hi - lo) / (float)levels;
    if (scale < 1.0e-12f) scale = 1.0e-12f;
    const float inv = 1.0f / scale;
    uint8_t *out = res + (size_t)v * res_bytes;
    if (bits == 4) {
        for (int byte = lane; byte < res_bytes; byte += RV_WAVE) {
            int packed = 0;
            for (int h = 0; h < 2; h++) {
                const int d = 2 * byte + h;
                if (d >= dim) break;
                const int s = d / sub;
                float c = codebook[((size_t)s * ks + rv_code_get(code, s, ks)) * sub + (d % sub)];
                float r = xv[d] - c;
                int qv = (int)roundf((r - lo) * inv);
                qv = qv < 0 ? 0 : (qv > levels ? levels : qv);
                packed |= qv << (4 * h);
            }
            out[byte] = (uint8_t)packed;
        }
    } else {
        for (int d = lane; d < dim; d += RV_WAVE) {
            const int s = d / sub;
            float c = codebook[((size_t)s * ks + rv_code_get(code, s, ks)) * sub + (d % sub)];
            float r = xv[d] - c;
            int qv = (int)roundf((r - lo) * inv);
            qv = qv < 0 ? 0 : (qv > levels ? levels : qv);
            out[d] = (uint8_t)qv;
        }
    }
    if (lane == 0) {
        qparams[2 * (size_t)v] = bad ? NAN : scale;
        qparams[2 * (size_t)v + 1] = lo;
    }
}

// retrieval_build_artifacts.c:25-66 (retrieval_build_error4), one vector
// per wavefront.  recon = the residual decode (retrieval_pq_residual.c:
// 954-981), e = x - recon; zero = the first minimum of e and hi the first
// maximum (the reference's strict scans keep the lowest dimension among
// equal values; the reduction carries the dimension to do the same),
// scale = max(hi - zero, 1e-12) / 15, code = clamp(floor((e - zero) *
// (1 / scale) + 0.5), 0, 15), two codes per byte, low nibble = even dim.
__device__ __forceinline__ float rv_e4_error(const float *xv, const uint8_t *code,
                                             const uint8_t *res, const float *codebook,
                                             int d, int sub, int ks, int res_bits,
                                             float s, float z) {
    const int sb = d / sub;
    const float c = codebook[((size_t)sb * ks + rv_code_get(code, sb, ks)) * sub + (d % sub)];
    const float rec = rv_decode_value(c, rv_res_code(res, res_bits, d), s, z);
    return xv[d] - rec;
}

__global__ __launch_bounds__(RV_WAVE) void rv_error4_encode_kernel(
    const float *x, const uint8_t *codes, const uint8_t *res, const float *res_qp,
    int n, int dim, int m, int ks, const float *codebook, int code_bytes,
    int res_bits, int res_bytes, uint8_t *e4, int e4_bytes, float *e4_qp) {
    const int v = blockIdx.x;
    const int lane = rv_lane();
    if (v >= n) return;
    const int sub = dim / m;
    const uint8_t *code = codes + (size_t)v * code_bytes;
    const uint8_t *rv = res + (size_t)v * res_bytes;
    const float s = res_qp[2 * (size_t)v], z = res_qp[2 * (size_t)v + 1];
    const float *xv = x + (size_t)v * dim;
    float lo = INFINITY, hi = -INFINITY;
    int lo_d = 0x7fffffff, hi_d = 0x7fffffff;
    bool bad = false;
    for (int d = lane; d < dim; d += RV_WAVE) {
        const float e = rv_e4_error(xv, code, rv, codebook, d, sub, ks, res_bits, s, z);
        if (!isfinite(e)) bad = true;
        if (e < lo) lo = e, lo_d = d;
        if (e > hi) hi = e, hi_d = d;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const float a = __shfl_xor(lo, off), b2 = __shfl_xor(hi, off);
        const int ad = __shfl_xor(lo_d, off), bd = __shfl_xor(hi_d, off);
        if (a < lo || (a == lo && ad < lo_d)) lo = a, lo_d = ad;
        if (b2 > hi || (b2 == hi && bd < hi_d)) hi = b2, hi_d = bd;
    }
    bad = __ballot(bad) != 0ull;
    const float scale = fmaxf(hi - lo, 1.0e-12f) / 15.0f;
    const float inv = 1.0f / scale;
    uint8_t *out = e4 + (size_t)v * e4_bytes;
    for (int byte = lane; byte < e4_bytes; byte += RV_WAVE) {
        int packed = 0;
        for (int h = 0; h < 2; h++) {
            const int d = 2 * byte + h;
            if (d >= dim) break;
            const float e = rv_e4_error(xv, code, rv, codebook, d, sub, ks, res_bits, s, z);
            const float t = (e - lo) * inv + 0.5f;
            int c = bad ? 0 : (int)floorf(t);
            c = c < 0 ? 0 : (c > 15 ? 15 : c);
            packed |= c << (4 * h);
        }
        out[byte] = (uint8_t)packed;
    }
    if (lane == 0) {
        e4_qp[2 * (size_t)v] = bad ? NAN : scale;
        e4_qp[2 * (size_t)v + 1] = lo;
    }
}

// ------------------------------------------------------------ host side

extern "C" int rv_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
extern "C" int rv_set_device(int device) {
    return rv_check(hipSetDevice(device), "hipSetDevice");
}
extern "C" int rv_malloc(void **p, size_t bytes) {
    return rv_check(hipMalloc(p, bytes ? bytes : 16), "hipMalloc");
}
extern "C" int rv_free(void *p) { return rv_check(hipFree(p), "hipFree"); }
extern "C" int rv_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    return rv_check(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
}
extern "C" int rv_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    return rv_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
}
extern "C" int rv_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    return rv_check(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice), "hipMemcpy D2D");
}
extern "C" int rv_synchronize(void) {
    return rv_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

extern "C" int rv_fill_zero(void *p, size_t bytes, void *stream) {
    hipStream_t s = (hipStream_t)stream;
    if (bytes % 4 == 0 && ((uintptr_t)p & 3) == 0) {
        size_t words = bytes / 4;
        int grid = (int)((words + 255) / 256);
        if (grid > 4096) grid = 4096;
        if (grid < 1) grid = 1;
        rv_zero_kernel<<<grid, 256, 0, s>>>((uint32_t *)p, words);
    } else {
        int grid = (int)((bytes + 255) / 256);
        if (grid > 4096) grid = 4096;
        if (grid < 1) grid = 1;
        rv_zero_bytes_kernel<<<grid, 256, 0, s>>>((uint8_t *)p, bytes);
    }
    return rv_check(hipGetLastError(), "zero kernel");
}

// ------------------------------------------------------ vector decoding

// One thread per (vector, dimension).  PQ: the residual decode
// (retrieval_pq_residual.c:954-981), ((centroid + code * scale) + zero).
// SQ: retrieval_final_refine_decode_sq (retrieval_final_refine.c:55-67),
// ((l1 * scale) + zero) + ((res * rscale) + rzero).  Each step a rounded
// f32.  out [count][dim].
__global__ void rv_decode_kernel(RvIndex ix, int sq, const int32_t *ids, int count,
                                 float *out) {
    const size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (t >= (size_t)count * ix.dim) return;
    const int i = (int)(t / ix.dim), d = (int)(t % ix.dim);
    const int id = ids[i];
    const float s = ix.res_qparams[2 * (size_t)id], z = ix.res_qparams[2 * (size_t)id + 1];
    const int rc = rv_res_code(ix.res_codes + (size_t)id * ix.res_code_bytes, ix.res_bits, d);
    float v;
    if (sq) {
        const uint8_t *l1 = ix.sq_codes + (size_t)id * ix.sq_code_bytes;
        const int lc = ix.sq_bits == 4 ? ((d & 1) ? (l1[d >> 1] >> 4) : (l1[d >> 1] & 15))
                                       : l1[d];
        const float lv = (float)lc * ix.sq_scale + ix.sq_zero;
        const float rv = (float)rc * s + z;
        v = lv + rv;
    } else {
        const int sub = d / ix.subdim;
        const int code = rv_code_get(ix.codes + (size_t)id * ix.code_bytes, sub, ix.ks);
        const float c = ix.codebook[((size_t)sub * ix.ks + code) * ix.subdim + d % ix.subdim];
        v = rv_decode_value(c, rc, s, z);
    }
    out[t] = v;
}

extern "C" int rv_decode_vectors(const RvIndex *ix, int sq, const int32_t *d_ids, int count,
                                 float *d_out, void *stream) {
    if (count <= 0) return 0;
    if (!ix || !d_ids || !d_out || !ix->res_codes || !ix->res_qparams ||
        (sq ? !ix->sq_codes : (!ix->codes || !ix->codebook))) {
        snprintf(g_rv_error, sizeof(g_rv_error), "rv_decode_vectors: bad arguments");
        return -1;
    }
    const size_t total = (size_t)count * ix->dim;
    rv_decode_kernel<<<(unsigned)((total + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        *ix, sq, d_ids, count, d_out);
    return rv_check(hipGetLastError(), "rv_decode_kernel launch");
}

// ------------------------------------------- fixed no-allocation workspace

// Workspace of the entry point that allocates nothing: the work counter in
// the first RV_NOALLOC_SEED_OFFSET bytes, then up to RV_NOALLOC_MAX_SEEDS
// seed ids.  Each device holds its own copy for the life of the process; the
// host reaches it only through the address hipGetSymbolAddress returns for
// that device (the symbol's name is a host-side stand-in, not a device
// address).
static __device__ __align__(256) uint8_t scratch[RV_NOALLOC_BYTES];

#define RV_MAX_DEVICES 64
static void *g_scratch_ptr[RV_MAX_DEVICES];
static pthread_mutex_t g_scratch_mu = PTHREAD_MUTEX_INITIALIZER;

extern "C" int rv_noalloc_scratch(int device, void **out) {
    int rc = 0;
    *out = nullptr;
    if (device < 0 || device >= RV_MAX_DEVICES) {
        snprintf(g_rv_error, sizeof(g_rv_error), "rv_noalloc_scratch: device %d", device);
        return -1;
    }
    pthread_mutex_lock(&g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int prev = 0;
        void *p = nullptr;
        rc = rv_check(hipGetDevice(&prev), "hipGetDevice");
        if (!rc && prev != device) rc = rv_check(hipSetDevice(device), "hipSetDevice");
        if (!rc) {
            rc = rv_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                          "hipGetSymbolAddress(scratch)");
            if (!rc && !p) {
                snprintf(g_rv_error, sizeof(g_rv_error), "hipGetSymbolAddress: null");
                rc = -1;
            }
            if (!rc) g_scratch_ptr[device] = p;
        }
        if (prev != device) hipSetDevice(prev);
    }
    *out = g_scratch_ptr[device];
    pthread_mutex_unlock(&g_scratch_mu);
    return rc;
}

// The reference's seed list (rpc_demo_pure_pq.c:192-208) written on the
// device: graph_entry first (when in range), then the LCG draws mapped by a
// multiply-shift.  One thread; count is at most a few thousand.
__global__ void rv_lcg_seeds_kernel(int32_t *out, int n, int count, int graph_entry,
                                    uint32_t seed) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int i = 0; i < count; i++) {
        if (i == 0 && graph_entry >= 0 && graph_entry < n) {
            out[i] = graph_entry;
        } else {
            seed = seed * 1103515245u + 12345u;
            out[i] = (int32_t)(((uint64_t)seed * (uint64_t)(uint32_t)n) >> 32);
        }
    }
}

extern "C" int rv_seed_list(int32_t *d_out, int n, int count, int graph_entry,
                            uint32_t seed, void *stream) {
    if (!d_out || n <= 0 || count <= 0 || count > n) {
        snprintf(g_rv_error, sizeof(g_rv_error), "rv_seed_list: bad parameters");
        return -1;
    }
    rv_lcg_seeds_kernel<<<1, 64, 0, (hipStream_t)stream>>>(d_out, n, count, graph_entry,
                                                          seed);
    return rv_check(hipGetLastError(), "rv_lcg_seeds_kernel launch");
}

static RvLayout rv_layout_for(const RvIndex *ix, const RvBatch *b) {
    return rv_layout(ix->dim, b->scorer, ix->m, ix->ks, b->max_pool, b->hash_log2,
                     b->stop_mode == 2, b->refine_candidates, b->waves_per_query > 0);
}


// One instantiation per (scorer, learned-stop mode, source window); traces
// are recorded with window 1 only.
template <int S, int T, int W>
static int rv_occupancy_t(int lds) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rv_search_kernel<S, T, W>,
                                                     RV_WAVE, lds) != hipSuccess)
        return 0;
    return blocks;
}

template <int S, int T, int W>
static int rv_configure_t() {
    static int configured = 0;
    if (!configured) {
        if (rv_check(hipFuncSetAttribute((const void *)rv_search_kernel<S, T, W>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         160 * 1024),
                     "hipFuncSetAttribute"))
            return -1;
        configured = 1;
    }
    return 0;
}

template <int S, int T, int W>
static int rv_launch_t(const RvIndex *ix, const RvBatch *b, int grid, RvLayout lay,
                       hipStream_t s) {
    rv_search_kernel<S, T, W><<<grid, RV_WAVE, lay.total, s>>>(*ix, *b, lay);
    return rv_check(hipGetLastError(), "rv_search_kernel launch");
}

// op 0: configure attributes, 1: occupancy (blocks per CU), 2: launch
template <int S, int T, int W>
static int rv_op_t(int op, const RvIndex *ix, const RvBatch *b, int grid, RvLayout lay,
                   hipStream_t s) {
    return op == 0 ? rv_configure_t<S, T, W>()
                   : op == 1 ? rv_occupancy_t<S, T, W>(lay.total)
                             : rv_launch_t<S, T, W>(ix, b, grid, lay, s);
}

template <int S>
static int rv_dispatch_stop(int op, const RvIndex *ix, const RvBatch *b, int grid,
                            RvLayout lay, hipStream_t s) {
    const int w = b->window == 2 ? 2 : 1;
    if (b->stop_mode == 0)
        return w == 2 ? rv_op_t<S, 0, 2>(op, ix, b, grid, lay, s)
                      : rv_op_t<S, 0, 1>(op, ix, b, grid, lay, s);
    if (b->stop_mode == 1 && b->suspend_step > 0 && w == 1)
        return rv_op_t<S, 3, 1>(op, ix, b, grid, lay, s);
    if (b->stop_mode == 1)
        return w == 2 ? rv_op_t<S, 1, 2>(op, ix, b, grid, lay, s)
                      : rv_op_t<S, 1, 1>(op, ix, b, grid, lay, s);
    if (b->stop_mode == 2 && w == 1) return rv_op_t<S, 2, 1>(op, ix, b, grid, lay, s);
    snprintf(g_rv_error, sizeof(g_rv_error), "rv_search: unknown stop mode / window");
    return -1;
}

// The multi-wave kernel: one instantiation per (residual bits, stop mode,
// waves per query).
template <int RB, int T, int W>
static int rv_mw_op_t(int op, const RvIndex *ix, const RvBatch *b, int grid, RvLayout lay,
                      hipStream_t s) {
    if (op == 0) {
        static int configured = 0;
        if (!configured) {
            if (rv_check(hipFuncSetAttribute((const void *)rv_search_mw_kernel<RB, T, W>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             160 * 1024),
                         "hipFuncSetAttribute"))
                return -1;
            configured = 1;
        }
        return 0;
    }
    if (op == 1) {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &blocks, rv_search_mw_kernel<RB, T, W>, RV_WAVE * W, lay.total) != hipSuccess)
            return 0;
        return blocks;
    }
    rv_search_mw_kernel<RB, T, W><<<grid, RV_WAVE * W, lay.total, s>>>(*ix, *b, lay);
    return rv_check(hipGetLastError(), "rv_search_mw_kernel launch");
}

template <int RB, int T>
static int rv_mw_dispatch_w(int op, const RvIndex *ix, const RvBatch *b, int grid,
                            RvLayout lay, hipStream_t s) {
    switch (b->waves_per_query) {
    case 1: return rv_mw_op_t<RB, T, 1>(op, ix, b, grid, lay, s);
    case 2: return rv_mw_op_t<RB, T, 2>(op, ix, b, grid, lay, s);
    case 4: return rv_mw_op_t<RB, T, 4>(op, ix, b, grid, lay, s);
    }
    snprintf(g_rv_error, sizeof(g_rv_error), "rv_search: waves per query 1, 2 or 4");
    return -1;
}

template <int RB>
static int rv_mw_dispatch(int op, const RvIndex *ix, const RvBatch *b, int grid,
                          RvLayout lay, hipStream_t s) {
    return b->stop_mode == 1 ? rv_mw_dispatch_w<RB, 1>(op, ix, b, grid, lay, s)
                             : rv_mw_dispatch_w<RB, 0>(op, ix, b, grid, lay, s);
}

static int rv_dispatch(int op, const RvIndex *ix, const RvBatch *b, int grid, RvLayout lay,
                       hipStream_t s) {
    if (b->waves_per_query > 0)
        return b->scorer == RV_SCORE_ONLINE_PAIR32_R4
                   ? rv_mw_dispatch<4>(op, ix, b, grid, lay, s)
                   : rv_mw_dispatch<8>(op, ix, b, grid, lay, s);
    switch (b->scorer) {
    case RV_SCORE_ADC: return rv_dispatch_stop<RV_SCORE_ADC>(op, ix, b, grid, lay, s);
    case RV_SCORE_ADC_ONLINE:
        return rv_dispatch_stop<RV_SCORE_ADC_ONLINE>(op, ix, b, grid, lay, s);
    case RV_SCORE_SQ: return rv_dispatch_stop<RV_SCORE_SQ>(op, ix, b, grid, lay, s);
    case RV_SCORE_ONLINE_PAIR32:
        return rv_dispatch_stop<RV_SCORE_ONLINE_PAIR32>(op, ix, b, grid, lay, s);
    case RV_SCORE_ONLINE_PAIR32_R4:
        return rv_dispatch_stop<RV_SCORE_ONLINE_PAIR32_R4>(op, ix, b, grid, lay, s);
    }
    snprintf(g_rv_error, sizeof(g_rv_error), "rv_search: unknown scorer");
    return -1;
}

// Validates a batch, picks the kernel instantiation and its LDS layout.
static int rv_resolve(const RvIndex *ix, const RvBatch *b_in, RvBatch *bb, RvLayout *lay) {
    if (!ix || !b_in || !b_in->queue || b_in->nq < 0 || b_in->hash_log2 < 4 ||
        b_in->hash_log2 > 16 || b_in->max_pool <= 0 || b_in->probe_limit <= 0 ||
        b_in->n_entries <= 0 || b_in->expand_count <= 0 || ix->degree <= 0 ||
        (b_in->expand_count < ix->degree ? b_in->expand_count : ix->degree) > RV_WAVE ||
        b_in->window < 0 || b_in->window > 2 || (b_in->window == 2 && b_in->stop_mode == 2) ||
        ix->dim <= 0 || ix->dim > 4096 || b_in->stop_mode < 0 || b_in->stop_mode > 2 ||
        (b_in->stop_mode && (b_in->stop_every <= 0 || b_in->stop_label_k <= 0)) ||
        (b_in->stop_mode == 2 &&
         (b_in->stop_label_k > RV_TRACE_TOPK || !b_in->trace_meta || !b_in->trace_feats ||
          !b_in->trace_steps || b_in->trace_cap <= 0))) {
        snprintf(g_rv_error, sizeof(g_rv_error), "rv_search_launch: bad parameters");
        return -1;
    }
    *bb = *b_in;
    // online correction on the 32 x 4-dim shape: two lanes per candidate
    if (bb->scorer == RV_SCORE_ADC_ONLINE &&
        rv_online_pairable(*ix, bb->expand_count < ix->degree ? bb->expand_count : ix->degree))
        bb->scorer = ix->res_bits == 4 ? RV_SCORE_ONLINE_PAIR32_R4 : RV_SCORE_ONLINE_PAIR32;
    if (bb->waves_per_query > 0 &&
        ((bb->waves_per_query != 1 && bb->waves_per_query != 2 && bb->waves_per_query != 4) ||
         (bb->cand_tile != 16 && bb->cand_tile != 32 && bb->cand_tile != 64) ||
         (bb->scorer != RV_SCORE_ONLINE_PAIR32 && bb->scorer != RV_SCORE_ONLINE_PAIR32_R4) ||
         bb->stop_mode > 1 || bb->window > 1 || bb->phase != 0)) {
        snprintf(g_rv_error, sizeof(g_rv_error),
                 "rv_search_launch: several waves per query need the online correction on "
                 "32 x 4-dim PQ (<= 32 neighbors), plain or learned stop, source window 1, "
                 "waves 1/2/4 and candidate tile 16/32/64");
        return -1;
    }
    if (bb->phase < 0 || bb->phase > 2 ||
        (bb->phase != 0 && (bb->window > 1 || bb->stop_mode == 2 || bb->suspend_step <= 0)) ||
        (bb->stop_probe_only && bb->stop_mode != 1)) {
        snprintf(g_rv_error, sizeof(g_rv_error),
                 "rv_search_launch: suspended searches need source window 1 and no traces");
        return -1;
    }
    *lay = rv_layout_for(ix, bb);
    if (lay->total > 160 * 1024) {
        snprintf(g_rv_error, sizeof(g_rv_error),
                 "rv_search_launch: %d bytes of LDS per query exceed 160 KiB", lay->total);
        return -1;
    }
    return 0;
}

extern "C" size_t rv_search_lds_bytes(const RvIndex *ix, const RvBatch *b) {
    RvBatch bb;
    RvLayout lay;
    if (rv_resolve(ix, b, &bb, &lay)) return 0;
    return (size_t)lay.total;
}

extern "C" int rv_search_slots_per_cu(const RvIndex *ix, const RvBatch *b) {
    RvBatch bb;
    RvLayout lay;
    if (rv_resolve(ix, b, &bb, &lay)) return 0;
    return rv_dispatch(1, ix, &bb, 0, lay, nullptr);
}

// Occupancy-derived grid: every CU filled with as many query slots as its
// registers and LDS allow, capped at the batch size.
static int rv_default_grid(const RvIndex *ix, const RvBatch *bb, RvLayout lay) {
    int dev = 0, cus = 0;
    if (rv_check(hipGetDevice(&dev), "hipGetDevice") ||
        rv_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                 "hipDeviceGetAttribute"))
        return -1;
    int per = rv_dispatch(1, ix, bb, 0, lay, nullptr);
    if (per < 1) per = 1;
    int grid = cus * per;
    if (grid > bb->nq) grid = bb->nq;
    return grid < 1 ? 1 : grid;
}

extern "C" int rv_search_prepare(const RvIndex *ix, const RvBatch *b) {
    RvBatch bb;
    RvLayout lay;
    if (rv_resolve(ix, b, &bb, &lay) || rv_dispatch(0, ix, &bb, 0, lay, nullptr)) return -1;
    return rv_default_grid(ix, &bb, lay);
}

extern "C" int rv_search_launch(const RvIndex *ix, const RvBatch *b_in, int grid_waves,
                                void *stream) {
    RvBatch bb;
    RvLayout lay;
    if (rv_resolve(ix, b_in, &bb, &lay)) return -1;
    if (bb.nq == 0) return 0;
    const bool reorder = bb.stop_mode == 1 && bb.suspend_step > 0 && bb.window == 1;
    if ((bb.phase != 0) != reorder ||
        (reorder &&
         (!bb.visited || !bb.susp ||
          (bb.phase == 2 &&
           bb.nq_dev != reinterpret_cast<const int32_t *>(
                            bb.susp + rv_susp_layout((size_t)bb.nq, (size_t)bb.max_pool).count))))) {
        snprintf(g_rv_error, sizeof(g_rv_error),
                 "rv_search_launch: suspended searches need per-query visited sets and state");
        return -1;
    }
    if (rv_dispatch(0, ix, &bb, 0, lay, nullptr)) return -1;
    if (grid_waves <= 0 && (grid_waves = rv_default_grid(ix, &bb, lay)) < 0) return -1;
    if (grid_waves > bb.nq) grid_waves = bb.nq;
    if (rv_fill_zero(bb.queue, sizeof(int32_t), stream)) return -1;
    return rv_dispatch(2, ix, &bb, grid_waves, lay, (hipStream_t)stream);
}

// Resumption order of the suspended queries: ascending stop probability in
// 4096 buckets (a probability below 0, none taken, counts as 0), so the
// searches the predictor expects to run longest resume first.  One
// workgroup: bucket counts, an exclusive scan, a scatter (the order inside
// a bucket is arbitrary; it changes when searches run, not their results).
#define RV_ORDER_BUCKETS 4096
__global__ __launch_bounds__(1024) void rv_susp_order_kernel(unsigned char *blk, int nq,
                                                              int max_pool) {
    __shared__ int cnt[RV_ORDER_BUCKETS];
    __shared__ int part[1024];
    const RvSuspLayout sl = rv_susp_layout((size_t)nq, (size_t)max_pool);
    const float *prob = reinterpret_cast<const float *>(blk + sl.prob);
    const int32_t *list = reinterpret_cast<const int32_t *>(blk + sl.list);
    int32_t *order = reinterpret_cast<int32_t *>(blk + sl.order);
    const int t = threadIdx.x;
    int total = *reinterpret_cast<const int32_t *>(blk + sl.count);
    total = total < nq ? total : nq;
    for (int i = t; i < RV_ORDER_BUCKETS; i += 1024) cnt[i] = 0;
    __syncthreads();
    auto bucket = [&](int i) {
        const float p = prob[list[i]];
        int k = (int)(p * (float)RV_ORDER_BUCKETS);
        return k < 0 ? 0 : (k >= RV_ORDER_BUCKETS ? RV_ORDER_BUCKETS - 1 : k);
    };
    for (int i = t; i < total; i += 1024) atomicAdd(&cnt[bucket(i)], 1);
    __syncthreads();
    // exclusive scan: four buckets per thread, then over the threads
    int local = 0;
    for (int j = 0; j < RV_ORDER_BUCKETS / 1024; j++) local += cnt[t * (RV_ORDER_BUCKETS / 1024) + j];
    part[t] = local;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - local;
    for (int j = 0; j < RV_ORDER_BUCKETS / 1024; j++) {
        const int k = t * (RV_ORDER_BUCKETS / 1024) + j;
        const int c = cnt[k];
        cnt[k] = run;
        run += c;
    }
    __syncthreads();
    for (int i = t; i < total; i += 1024) order[atomicAdd(&cnt[bucket(i)], 1)] = list[i];
}

extern "C" int rv_susp_order(unsigned char *susp, int nq, int max_pool, void *stream) {
    if (!susp || nq < 0 || max_pool <= 0) {
        snprintf(g_rv_error, sizeof(g_rv_error), "rv_susp_order: bad arguments");
        return -1;
    }
    if (nq == 0) return 0;
    rv_susp_order_kernel<<<1, 1024, 0, (hipStream_t)stream>>>(susp, nq, max_pool);
    return rv_check(hipGetLastError(), "rv_susp_order_kernel launch");
}

extern "C" int rv_graph_begin(void *stream) {
    if (!stream) {
        snprintf(g_rv_error, sizeof(g_rv_error), "graph capture needs a created stream");
        return -1;
    }
    return rv_check(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal),
                    "hipStreamBeginCapture");
}

extern "C" int rv_graph_end(void *stream, void **exec) {
    hipGraph_t g = nullptr;
    *exec = nullptr;
    if (rv_check(hipStreamEndCapture((hipStream_t)stream, &g), "hipStreamEndCapture"))
        return -1;
    hipGraphExec_t e = nullptr;
    const int rc = rv_check(hipGraphInstantiate(&e, g, nullptr, nullptr, 0),
                            "hipGraphInstantiate");
    (void)hipGraphDestroy(g);
    if (rc) return -1;
    *exec = (void *)e;
    return 0;
}

extern "C" int rv_graph_launch(void *exec, void *stream) {
    return rv_check(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)stream), "hipGraphLaunch");
}

extern "C" int rv_graph_destroy(void *exec) {
    return rv_check(hipGraphExecDestroy((hipGraphExec_t)exec), "hipGraphExecDestroy");
}

extern "C" int rv_pq_encode(const float *x, int n, int dim, int m, int ks,
                            const float *codebook, uint8_t *codes, void *stream) {
    if (n <= 0 || m <= 0 || dim % m || (ks != 16 && ks != 256)) return -1;
    const int code_bytes = ks <= 16 ? m / 2 + m % 2 : m;
    const size_t threads = (size_t)n * m;
    rv_pq_encode_kernel<<<(unsigned)((threads + 255) / 256), 256, 0,
                          (hipStream_t)stream>>>(x, n, dim, m, ks, codebook, codes,
                                                 code_bytes);
    return rv_check(hipGetLastError(), "rv_pq_encode");
}

extern "C" int rv_residual_encode(const float *x, const uint8_t *codes, int n,
                                  int dim, int m, int ks, const float *codebook,
                                  int bits, uint8_t *res, float *qparams,
                                  void *stream) {
    if (n <= 0 || m <= 0 || dim % m || (bits != 4 && bits != 8)) return -1;
    const int code_bytes = ks <= 16 ? m / 2 + m % 2 : m;
    const int res_bytes = bits == 4 ? dim / 2 + dim % 2 : dim;
    rv_residual_encode_kernel<<<n, RV_WAVE, 0, (hipStream_t)stream>>>(
        x, codes, n, dim, m, ks, codebook, code_bytes, bits, res, res_bytes, qparams);
    return rv_check(hipGetLastError(), "rv_residual_encode");
}

extern "C" int rv_error4_encode(const float *x, const uint8_t *codes,
                                const uint8_t *res, const float *res_qp, int n,
                                int dim, int m, int ks, const float *codebook,
                                int res_bits, uint8_t *e4, float *e4_qp,
                                void *stream) {
    if (n <= 0 || m <= 0 || dim <= 0 || dim % m || (res_bits != 4 && res_bits != 8))
        return -1;
    const int code_bytes = ks <= 16 ? m / 2 + m % 2 : m;
    const int res_bytes = res_bits == 4 ? dim / 2 + dim % 2 : dim;
    const int e4_bytes = (dim + 1) / 2;
    rv_error4_encode_kernel<<<n, RV_WAVE, 0, (hipStream_t)stream>>>(
        x, codes, res, res_qp, n, dim, m, ks, codebook, code_bytes, res_bits,
        res_bytes, e4, e4_bytes, e4_qp);
    return rv_check(hipGetLastError(), "rv_error4_encode");
}

// ----------------------------------------------------------- entry router
//
// Per-query entry selection: exact squared L2 from every query to every
// routing centroid (a tiled [queries x centroids] kernel, each distance a
// sequential f32 sum over the dimensions), then the S nearest centroids per
// query by (distance, centroid index), mapped to their representative node
// ids.  The ids become that query's entry list for the graph search.

#define RT_TQ 64  // queries per tile
#define RT_TC 64  // centroids per tile
#define RT_KD 32  // dimensions staged per step

__global__ __launch_bounds__(256) void rv_router_dist_kernel(
    const float *__restrict__ q, int nq, const float *__restrict__ cent, int nc,
    int dim, float *__restrict__ out) {
    __shared__ float Qs[RT_TQ][RT_KD + 1];
    __shared__ float Cs[RT_TC][RT_KD + 1];
    const int t = threadIdx.x;
    const int q0 = blockIdx.x * RT_TQ, c0 = blockIdx.y * RT_TC;
    const int tq = (t / 16) * 4, tc = (t % 16) * 4;  // 4x4 outputs per thread
    float acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0f;
    for (int k0 = 0; k0 < dim; k0 += RT_KD) {
        for (int i = t; i < RT_TQ * RT_KD; i += 256) {
            const int r = i / RT_KD, k = i % RT_KD;
            Qs[r][k] = (q0 + r < nq && k0 + k < dim) ? q[(size_t)(q0 + r) * dim + k0 + k] : 0.0f;
            Cs[r][k] = (c0 + r < nc && k0 + k < dim) ? cent[(size_t)(c0 + r) * dim + k0 + k] : 0.0f;
        }
        __syncthreads();
        const int kend = dim - k0 < RT_KD ? dim - k0 : RT_KD;
        for (int k = 0; k < kend; k++) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = Qs[tq + i][k];
#pragma unroll
            for (int j = 0; j < 4; j++) b[j] = Cs[tc + j][k];
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float d = a[i] - b[j];
                    acc[i][j] += d * d;
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int qq = q0 + tq + i, cc = c0 + tc + j;
            if (qq < nq && cc < nc) out[(size_t)qq * nc + cc] = acc[i][j];
        }
}

// One wavefront per query: the S smallest (distance, centroid) keys.
__global__ __launch_bounds__(RV_WAVE) void rv_router_select_kernel(
    const float *__restrict__ dist, int nq, int nc, const int32_t *__restrict__ rep_ids,
    int S, int32_t *__restrict__ out) {
    const int q = blockIdx.x;
    const int lane = rv_lane();
    if (q >= nq) return;
    uint64_t best[16];
#pragma unroll
    for (int i = 0; i < 16; i++) best[i] = RV_KEY_MAX;
    const float *row = dist + (size_t)q * nc;
    for (int c = lane; c < nc; c += RV_WAVE) {
        // insertion into the lane's sorted list: carry the larger key on
        uint64_t v = rv_key(row[c] + 0.0f, c);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            if (i < S) {
                const uint64_t lo = v < best[i] ? v : best[i];
                const uint64_t hi = v < best[i] ? best[i] : v;
                best[i] = lo;
                v = hi;
            }
        }
    }
    // S rounds of wave-wide minimum over the lanes' list heads
    int head = 0;
    for (int s = 0; s < S; s++) {
        uint64_t mine = RV_KEY_MAX;
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (i == head) mine = best[i];
        uint64_t m = mine;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            uint64_t o = __shfl_xor(m, off);
            m = o < m ? o : m;
        }
        if (mine == m && m != RV_KEY_MAX) head++;
        if (lane == 0)
            out[(size_t)q * S + s] = m == RV_KEY_MAX ? rep_ids[0] : rep_ids[rv_key_id(m)];
    }
}

extern "C" int rv_route_entries(const float *queries, int nq, int dim,
                                const float *centroids, int nc,
                                const int32_t *rep_ids, int S, float *scratch,
                                int32_t *out, void *stream) {
    if (nq <= 0) return 0;
    if (nc <= 0 || S <= 0 || S > 16 || S > nc || dim <= 0) {
        snprintf(g_rv_error, sizeof(g_rv_error), "rv_route_entries: bad parameters");
        return -1;
    }
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((nq + RT_TQ - 1) / RT_TQ, (nc + RT_TC - 1) / RT_TC);
    rv_router_dist_kernel<<<grid, 256, 0, s>>>(queries, nq, centroids, nc, dim, scratch);
    if (rv_check(hipGetLastError(), "rv_router_dist_kernel")) return -1;
    rv_router_select_kernel<<<nq, RV_WAVE, 0, s>>>(scratch, nq, nc, rep_ids, S, out);
    return rv_check(hipGetLastError(), "rv_router_select_kernel");
}

// ------------------------------------------------------------ copy kernel

// Plain 16-byte-vector copy; a kernel of known traffic for calibrating the
// HBM byte counters of a profiler.
__global__ void rv_copy_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                               size_t n16) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n16;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int rv_copy(void *dst, const void *src, size_t bytes, void *stream) {
    if ((bytes & 15) || ((uintptr_t)dst & 15) || ((uintptr_t)src & 15)) {
        snprintf(g_rv_error, sizeof(g_rv_error), "rv_copy: 16-byte alignment required");
        return -1;
    }
    rv_copy_kernel<<<4096, 256, 0, (hipStream_t)stream>>>((const uint4 *)src,
                                                          (uint4 *)dst, bytes / 16);
    return rv_check(hipGetLastError(), "rv_copy");
}

// ------------------------------------------------------ events and streams

extern "C" int rv_event_create(void **ev) {
    hipEvent_t e;
    if (rv_check(hipEventCreate(&e), "hipEventCreate")) return -1;
    *ev = (void *)e;
    return 0;
}
extern "C" int rv_event_record(void *ev, void *stream) {
    return rv_check(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream), "hipEventRecord");
}
extern "C" int rv_event_elapsed(void *start, void *stop, float *ms) {
    if (rv_check(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize")) return -1;
    return rv_check(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop),
                    "hipEventElapsedTime");
}
extern "C" int rv_event_destroy(void *ev) {
    return rv_check(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy");
}
// Streams are created at the default priority.
extern "C" int rv_memcpy_h2d_async(void *dst, const void *src, size_t bytes, void *stream) {
    return rv_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream),
                    "hipMemcpyAsync H2D");
}
extern "C" int rv_memcpy_d2h_async(void *dst, const void *src, size_t bytes, void *stream) {
    return rv_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream),
                    "hipMemcpyAsync D2H");
}
extern "C" int rv_host_alloc(void **p, size_t bytes) {
    return rv_check(hipHostMalloc(p, bytes ? bytes : 16, hipHostMallocDefault),
                    "hipHostMalloc");
}
extern "C" int rv_host_free(void *p) { return rv_check(hipHostFree(p), "hipHostFree"); }
extern "C" int rv_event_create_untimed(void **ev) {
    hipEvent_t e;
    if (rv_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate"))
        return -1;
    *ev = (void *)e;
    return 0;
}
extern "C" int rv_stream_wait_event(void *stream, void *ev) {
    return rv_check(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)ev, 0),
                    "hipStreamWaitEvent");
}
extern "C" int rv_stream_create(void **stream) {
    hipStream_t s;
    if (rv_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"))
        return -1;
    *stream = (void *)s;
    return 0;
}
extern "C" int rv_stream_destroy(void *stream) {
    return rv_check(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
}
extern "C" int rv_stream_synchronize(void *stream) {
    return rv_check(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

// ====================================================================== SQ
// traversal with distance-only heaps (retrieval_sqmd_common.c: the live SQ
// path's fake-DSP turn loop, :2052-2152; the DSP search op
// retrieval_async_dsp_search.c:368-509 is the same loop over its own seed
// list).  Ties between equal distances are resolved by the heaps' array
// mechanics, so the heaps are kept exactly: one wavefront per query, the
// heap updates done by lane 0 in the reference's order, the distances of a
// turn's work items computed by one lane each.
//
//   pool      max-heap on distance (:9-47); admits below capacity, else only
//             strictly below the worst (:51-79)
//   frontier  min-heap on distance with a bounded push (:81-147)
//   source    tops worse than a full pool's worst are discarded, then the
//             top is expanded (:478-523)
//   turn      neighbors in row order: out-of-range ids and the source are
//             skipped, duplicates within the turn and visited nodes are
//             skipped, the rest marked visited (:283-383); results applied
//             in that order (:638-708)
//   output    heap sort of the pool (:40-49), first k

__device__ __forceinline__ void rv_hp_up(RvSqmdCand *p, int i) {
    const RvSqmdCand t = p[i];
    while (i > 0) {
        const int parent = (i - 1) >> 1;
        const RvSqmdCand pp = p[parent];
        if (t.d <= pp.d) break;
        p[i] = pp;
        i = parent;
    }
    p[i] = t;
}

__device__ __forceinline__ void rv_hp_down(RvSqmdCand *p, int size, int i) {
    const RvSqmdCand t = p[i];
    for (;;) {
        int child = 2 * i + 1;
        if (child >= size) break;
        RvSqmdCand c = p[child];
        if (child + 1 < size) {
            const RvSqmdCand c2 = p[child + 1];
            if (c2.d > c.d) c = c2, child++;
        }
        if (t.d >= c.d) break;
        p[i] = c;
        i = child;
    }
    p[i] = t;
}

__device__ __forceinline__ void rv_hf_up(RvSqmdCand *f, int i) {
    const RvSqmdCand t = f[i];
    while (i > 0) {
        const int parent = (i - 1) >> 1;
        const RvSqmdCand pp = f[parent];
        if (t.d >= pp.d) break;
        f[i] = pp;
        i = parent;
    }
    f[i] = t;
}

__device__ __forceinline__ void rv_hf_down(RvSqmdCand *f, int size, int i) {
    const RvSqmdCand t = f[i];
    for (;;) {
        int child = 2 * i + 1;
        if (child >= size) break;
        RvSqmdCand c = f[child];
        if (child + 1 < size) {
            const RvSqmdCand c2 = f[child + 1];
            if (c2.d < c.d) c = c2, child++;
        }
        if (t.d <= c.d) break;
        f[i] = c;
        i = child;
    }
    f[i] = t;
}

// Applies one scored candidate: pool insert, frontier push when admitted.
__device__ __forceinline__ void rv_sqmd_apply(RvSqmdCand *pool, int &psize, int L,
                                              RvSqmdCand *front, int &fsize, int fcap,
                                              int id, float d) {
    bool admitted = false;
    if (psize < L) {
        pool[psize] = RvSqmdCand{id, d};
        rv_hp_up(pool, psize);
        psize++;
        admitted = true;
    } else if (d < pool[0].d) {
        pool[0] = RvSqmdCand{id, d};
        rv_hp_down(pool, psize, 0);
        admitted = true;
    }
    if (admitted && fsize < fcap) {
        front[fsize] = RvSqmdCand{id, d};
        rv_hf_up(front, fsize);
        fsize++;
    }
}

// Exact SQ + residual distance (retrieval_rpc_demo_support.c:1397-1443):
// value = ((code * scale) + zero) + ((res * rscale) + rzero), sequential
// f32 sum of (q - value)^2.
__device__ __forceinline__ float rv_sq_exact(const RvIndex &ix, const float *qv, int id) {
    const uint8_t *l1 = ix.sq_codes + (size_t)id * ix.sq_code_bytes;
    const uint8_t *res = ix.res_codes + (size_t)id * ix.res_code_bytes;
    const float rs = ix.res_qparams[2 * (size_t)id], rz = ix.res_qparams[2 * (size_t)id + 1];
    float dist = 0.0f;
    for (int d = 0; d < ix.dim; d++) {
        const int c = ix.sq_bits == 4 ? ((d & 1) ? (l1[d >> 1] >> 4) : (l1[d >> 1] & 15))
                                      : l1[d];
        float value = (float)c * ix.sq_scale;
        value = value + ix.sq_zero;
        float level = (float)rv_res_code(res, ix.res_bits, d) * rs;
        level = level + rz;
        value = value + level;
        const float diff = qv[d] - value;
        dist = dist + diff * diff;
    }
    return dist;
}

// The split distance (retrieval_rpc_demo_pipeline.c:936-1000 with the
// scalar kernels of retrieval_l1_int8_core.c:171-260): qu8 is the query on
// the SQ grid refined 17x (u8) and sum_q its sum.  coarse = the integer sum
// of (q - 17 c)^2 times (scale / 17)^2 in double, to float; the residual dot
// from the integer sums sum q v and sum v, combined in double; delta =
// (double)bias - 2 dot, to float; distance = coarse + delta in float, with
// (-1e-4, 0) clamped to 0.  NAN where the reference fails the candidate.
__device__ __forceinline__ float rv_sq_split(const RvIndex &ix, const uint8_t *qu8,
                                             uint64_t sum_q, int id) {
    const uint8_t *l1 = ix.sq_codes + (size_t)id * ix.sq_code_bytes;
    const uint8_t *res = ix.res_codes + (size_t)id * ix.res_code_bytes;
    const float rs = ix.res_qparams[2 * (size_t)id], rz = ix.res_qparams[2 * (size_t)id + 1];
    const double qs = (double)ix.sq_scale / 17.0;
    int64_t acc = 0;
    uint64_t sum_qv = 0, sum_v = 0;
    for (int d = 0; d < ix.dim; d++) {
        const int c = (d & 1) ? (l1[d >> 1] >> 4) : (l1[d >> 1] & 15);
        const int diff = (int)qu8[d] - 17 * c;
        const uint32_t v = (uint32_t)rv_res_code(res, ix.res_bits, d);
        acc += (int64_t)(diff * diff);
        sum_qv += (uint64_t)((uint32_t)qu8[d] * v);
        sum_v += v;
    }
    const float coarse = (float)((double)acc * qs * qs);
    if (!isfinite(rs) || !isfinite(rz) || rs < 0.0f) return __int_as_float(0x7fc00000);
    const double dot = qs * (double)rs * (double)sum_qv + qs * (double)rz * (double)sum_q +
                       (double)ix.sq_zero * (double)rs * (double)sum_v +
                       (double)ix.dim * (double)ix.sq_zero * (double)rz;
    const double delta = (double)ix.split_bias[id] - 2.0 * dot;
    if (!isfinite(delta)) return __int_as_float(0x7fc00000);
    float dist = coarse + (float)delta;
    if (dist < 0.0f && dist > -1.0e-4f) dist = 0.0f;
    if (!isfinite(dist) || dist < 0.0f) return __int_as_float(0x7fc00000);
    return dist;
}

struct RvSqmdLayout {
    int off_q, off_qsq, off_pool, off_wid, off_wd, off_refine, off_front, total;
};

static RvSqmdLayout rv_sqmd_layout(int dim, int max_pool, int correction, int refine,
                                   int front_lds) {
    RvSqmdLayout l;
    int off = 0;
    l.off_q = off;
    if (correction == 1 || refine > 0) off += rv_align(dim * 4, 16);
    l.off_qsq = off;
    off += rv_align(dim, 16);
    l.off_pool = off;
    off += rv_align(max_pool * 8, 16);
    l.off_wid = off;
    off += RV_WAVE * 4;
    l.off_wd = off;
    off += RV_WAVE * 4;
    l.off_refine = off;
    if (refine > 0) off += rv_align(8 * (refine < max_pool ? refine : max_pool), 16);
    l.off_front = off;
    off += front_lds;
    l.total = off;
    return l;
}

template <int CORR>
__global__ __launch_bounds__(RV_WAVE) void rv_sqmd_kernel(RvIndex ix, RvSqmdBatch b,
                                                          RvSqmdLayout lay) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = rv_lane();
    float *qv = reinterpret_cast<float *>(smem + lay.off_q);
    uint8_t *qsq = smem + lay.off_qsq;
    RvSqmdCand *pool = reinterpret_cast<RvSqmdCand *>(smem + lay.off_pool);
    int32_t *wid = reinterpret_cast<int32_t *>(smem + lay.off_wid);
    float *wd = reinterpret_cast<float *>(smem + lay.off_wd);
    RvSqmdCand *front = b.frontier
                            ? b.frontier + (size_t)blockIdx.x * b.frontier_phys
                            : reinterpret_cast<RvSqmdCand *>(smem + lay.off_front);
    uint32_t *vis = b.visited + (size_t)blockIdx.x * b.visited_words;
    const float scale_sq = ix.sq_scale * ix.sq_scale;
    const int ec = rv_uniform(b.expand_count < ix.degree ? b.expand_count : ix.degree);

    int nq = b.nq;
    if (b.nq_dev) {
        const int live = *b.nq_dev;
        nq = live < nq ? (live > 0 ? live : 0) : nq;
    }
    nq = rv_uniform(nq);
    for (;;) {
        int q = 0;
        if (lane == 0) q = atomicAdd(b.queue, 1);
        q = rv_uniform(__shfl(q, 0));
        if (q >= nq) break;
        const int L = b.pool_sizes ? b.pool_sizes[q] : b.pool_size;
        const int K = b.num_outs ? b.num_outs[q] : b.num_out;
        const long fcap_l = (long)b.fcap_entries + (long)L * b.fcap_expand;
        if (L <= 0 || L > b.max_pool || K <= 0 || K > b.out_stride ||
            fcap_l > b.frontier_phys || ec > b.max_work || ec > RV_WAVE) {
            if (lane == 0) b.out_counts[q] = RV_QUERY_BAD_PARAM;
            continue;
        }
        const int fcap = (int)fcap_l;

        // fresh visited bitmap; the clears land before this query's atomics
        for (int w = 4 * lane; w < b.visited_words; w += 4 * RV_WAVE)
            *reinterpret_cast<uint4 *>(vis + w) = make_uint4(0u, 0u, 0u, 0u);
        __threadfence();
        const float *qg = b.queries + (size_t)q * ix.dim;
        uint64_t sum_q = 0;
        bool query_bad = false;
        if (CORR == 2) {
            // retrieval_l1_int8_core.c:65-110: step scale / 17 in double,
            // round half away from zero, clip to [0, 255]
            const double qs = (double)ix.sq_scale / 17.0;
            uint32_t part = 0;
            bool bad = false;
            for (int d = lane; d < ix.dim; d += RV_WAVE) {
                const float x = qg[d];
                if (b.refine_candidates > 0) qv[d] = x;
                const double raw = ((double)x - (double)ix.sq_zero) / qs;
                if (!isfinite(x) || !isfinite(raw)) bad = true;
                long v = bad ? 0 : lround(raw);
                v = v < 0 ? 0 : (v > 255 ? 255 : v);
                qsq[d] = (uint8_t)v;
                part += (uint32_t)v;
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
            sum_q = (uint64_t)rv_uniform((int)part);
            query_bad = __ballot(bad) != 0ull;
        } else {
            const float inv = 1.0f / ix.sq_scale;
            const int hi = ix.sq_bits == 4 ? 15 : 255;
            for (int d = lane; d < ix.dim; d += RV_WAVE) {
                const float x = qg[d];
                if (CORR == 1 || b.refine_candidates > 0) qv[d] = x;
                int v = (int)roundf((x - ix.sq_zero) * inv);
                v = v < 0 ? 0 : (v > hi ? hi : v);
                qsq[d] = (uint8_t)v;
            }
        }
        __syncthreads();

        auto score = [&](int id) -> float {
            if (CORR == 2) return rv_sq_split(ix, qsq, sum_q, id);
            if (CORR == 1) return rv_sq_exact(ix, qv, id);
            return rv_sq(qsq, ix.sq_codes + (size_t)id * ix.sq_code_bytes, ix.dim,
                         ix.sq_bits, scale_sq);
        };

        int psize = 0, fsize = 0;  // meaningful in lane 0
        int status = query_bad ? RV_QUERY_NONFINITE : 0;
        uint64_t st_work = 0, st_invalid = 0, st_turns = 0, st_dedup = 0;

        // ---- seeds (:446-476), 64 at a time in list order
        const int32_t *seeds = b.seeds + (b.seed_stride > 0 ? (size_t)q * b.seed_stride : 0);
        const int ns = b.seed_stride > 0 ? b.seed_stride : b.n_seeds;
        for (int base = 0; base < ns && status == 0; base += RV_WAVE) {
            const int i = base + lane;
            const bool active = i < ns;
            const int id = active ? seeds[i] : -1;
            const bool bad = active && (id < 0 || id >= ix.n);
            // the reference stops at the first out-of-range seed
            const uint64_t badm = __ballot(bad);
            const int first_bad = badm ? (int)__builtin_ctzll(badm) : RV_WAVE;
            bool dup = false;
            for (int t = 0; t < RV_WAVE; t++) {
                const int other = __shfl(id, t);
                dup |= t < lane && other == id;
            }
            bool fresh = false;
            if (active && !bad && lane < first_bad && !dup) {
                const uint32_t bit = 1u << (id & 31);
                fresh = (atomicOr(vis + (id >> 5), bit) & bit) == 0u;
            }
            wid[lane] = id;
            const float sd = fresh ? score(id) : 0.0f;
            wd[lane] = sd;
            const uint64_t mask = __ballot(fresh);
            if (CORR == 2 && __ballot(fresh && isnan(sd)) != 0ull) {
                status = RV_QUERY_NONFINITE;
                break;
            }
            __syncthreads();
            if (lane == 0) {
                for (uint64_t m = mask; m; m &= m - 1) {
                    const int j = (int)__builtin_ctzll(m);
                    rv_sqmd_apply(pool, psize, L, front, fsize, fcap, wid[j], wd[j]);
                }
            }
            __syncthreads();
            if (badm) {
                status = RV_QUERY_BAD_ENTRY;
                break;
            }
        }

        // ---- turns
        while (status == 0) {
            int src = -1;
            if (lane == 0) {
                while (fsize > 0 && psize >= L && front[0].d > pool[0].d) {
                    fsize--;
                    if (fsize > 0) {
                        front[0] = front[fsize];
                        rv_hf_down(front, fsize, 0);
                    }
                }
                if (fsize > 0) {
                    src = front[0].id;
                    fsize--;
                    if (fsize > 0) {
                        front[0] = front[fsize];
                        rv_hf_down(front, fsize, 0);
                    }
                }
            }
            src = rv_uniform(__shfl(src, 0));
            if (src < 0) break;
            st_turns++;
            const int nb = lane < ec ? ix.graph[(size_t)src * ix.degree + lane] : -1;
            const bool in_row = lane < ec;
            const bool valid = in_row && nb >= 0 && nb < ix.n && nb != src;
            bool dup = false;
            for (int t = 0; t < ec; t++) {
                const int other = __shfl(nb, t);
                const bool ovalid = __shfl((int)valid, t) != 0;
                dup |= t < lane && ovalid && other == nb;
            }
            bool fresh = false;
            if (valid && !dup) {
                const uint32_t bit = 1u << (nb & 31);
                fresh = (atomicOr(vis + (nb >> 5), bit) & bit) == 0u;
            }
            const uint64_t mask = __ballot(fresh);
            st_invalid += (uint64_t)__popcll(__ballot(in_row && !valid));
            st_dedup += (uint64_t)__popcll(__ballot(valid && !fresh));
            st_work += (uint64_t)__popcll(mask);
            wid[lane] = nb;
            const float nd = fresh ? score(nb) : 0.0f;
            wd[lane] = nd;
            if (CORR == 2 && __ballot(fresh && isnan(nd)) != 0ull) {
                status = RV_QUERY_NONFINITE;
                break;
            }
            __syncthreads();
            if (lane == 0) {
                for (uint64_t m = mask; m; m &= m - 1) {
                    const int j = (int)__builtin_ctzll(m);
                    rv_sqmd_apply(pool, psize, L, front, fsize, fcap, wid[j], wd[j]);
                }
            }
            __syncthreads();
        }

        // ---- output: heap sort (:40-49), first k
        if (status != 0) {
            if (lane == 0) b.out_counts[q] = status;
        } else {
            if (lane == 0) {
                for (int i = psize - 1; i > 0; i--) {
                    const RvSqmdCand t = pool[0];
                    pool[0] = pool[i];
                    pool[i] = t;
                    rv_hp_down(pool, i, 0);
                }
            }
            psize = rv_uniform(__shfl(psize, 0));
            __syncthreads();
            int32_t *oid = b.out_ids + (size_t)q * b.out_stride;
            float *odist = b.out_dists + (size_t)q * b.out_stride;
            if (b.refine_candidates <= 0) {
                const int cnt = K < psize ? K : psize;
                for (int i = lane; i < cnt; i += RV_WAVE) {
                    oid[i] = pool[i].id;
                    odist[i] = pool[i].d;
                }
                if (lane == 0) b.out_counts[q] = cnt;
            } else {
                // final refine of the first F results (retrieval_final_refine.c:
                // 256-292) from the SQ decode (:55-67): per dimension
                // ((l1 * scale) + zero) + ((res * rscale) + rzero), plus the
                // error4 level ((code * s4) + z4) in ERROR4 mode, each step a
                // rounded f32; the distance a sequential f64 sum, ranked by
                // (distance, original id)
                const int F = b.refine_candidates < psize ? b.refine_candidates : psize;
                uint64_t *dbits = reinterpret_cast<uint64_t *>(smem + lay.off_refine);
                bool bad = false;
                for (int i = lane; i < F; i += RV_WAVE) {
                    const int id = pool[i].id;
                    const uint8_t *l1 = ix.sq_codes + (size_t)id * ix.sq_code_bytes;
                    const uint8_t *res = ix.res_codes + (size_t)id * ix.res_code_bytes;
                    const float rs = ix.res_qparams[2 * (size_t)id];
                    const float rz = ix.res_qparams[2 * (size_t)id + 1];
                    const uint8_t *c4 = nullptr;
                    float s4 = 0.0f, z4 = 0.0f;
                    if (b.refine_error4) {
                        c4 = ix.e4_blocks + (size_t)(id / ix.e4_vpb) * RV_E4_BLOCK +
                             (size_t)(id % ix.e4_vpb) * ix.e4_code_bytes;
                        s4 = ix.e4_qparams[2 * (size_t)id];
                        z4 = ix.e4_qparams[2 * (size_t)id + 1];
                    }
                    if (!isfinite(rs) || rs < 0.0f || !isfinite(rz)) bad = true;
                    double acc = 0.0;
                    for (int d = 0; d < ix.dim; d++) {
                        const int lc = ix.sq_bits == 4 ? ((d & 1) ? (l1[d >> 1] >> 4)
                                                                  : (l1[d >> 1] & 15))
                                                       : l1[d];
                        const float lv = (float)lc * ix.sq_scale + ix.sq_zero;
                        const float rv = (float)rv_res_code(res, ix.res_bits, d) * rs + rz;
                        float v = lv + rv;
                        if (!isfinite(v)) bad = true;
                        if (c4) {
                            const float level =
                                (float)((c4[d >> 1] >> (4 * (d & 1))) & 15) * s4 + z4;
                            v = v + level;
                            if (!isfinite(v)) bad = true;
                        }
                        const double delta = (double)qv[d] - (double)v;
                        acc += delta * delta;
                    }
                    if (!isfinite(acc) || acc >= (double)FLT_MAX) bad = true;
                    dbits[i] = (uint64_t)__double_as_longlong(acc);
                }
                if (__ballot(bad) != 0ull) {
                    if (lane == 0) b.out_counts[q] = RV_QUERY_NONFINITE;
                } else {
                    __syncthreads();
                    for (int i = lane; i < F; i += RV_WAVE) {
                        const uint64_t di = dbits[i];
                        const int gi = pool[i].id;
                        const int oi = ix.graph_to_original ? ix.graph_to_original[gi] : gi;
                        int rank = 0;
                        for (int j = 0; j < F; j++) {
                            const uint64_t dj = dbits[j];
                            const int gj = pool[j].id;
                            const int oj = ix.graph_to_original ? ix.graph_to_original[gj] : gj;
                            rank += (dj < di) || (dj == di && oj < oi);
                        }
                        if (rank < K) {
                            oid[rank] = gi;
                            odist[rank] = (float)__longlong_as_double((long long)di);
                        }
                    }
                    if (lane == 0) b.out_counts[q] = K < F ? K : F;
                }
            }
        }
        if (b.out_stats && lane == 0) {
            uint64_t *o = b.out_stats + 4 * (size_t)q;
            o[0] = st_work;
            o[1] = st_invalid;
            o[2] = st_turns;
            o[3] = st_dedup;
        }
        __syncthreads();
    }
}

static int rv_sqmd_cus(void) {
    int dev = 0, cus = 0;
    if (rv_check(hipGetDevice(&dev), "hipGetDevice") ||
        rv_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev),
                 "hipDeviceGetAttribute"))
        return -1;
    return cus;
}

// The frontier lives in LDS when a wave's whole footprint stays within
// this many bytes; otherwise it is an HBM slab per wave.
#define RV_SQMD_LDS_BUDGET (48 * 1024)

static RvSqmdLayout rv_sqmd_layout_for(const RvIndex *ix, const RvSqmdBatch *b,
                                       int *front_lds) {
    const long fmax = (long)b->fcap_entries + (long)b->max_pool * b->fcap_expand;
    RvSqmdLayout l =
        rv_sqmd_layout(ix->dim, b->max_pool, b->correction, b->refine_candidates, 0);
    const long with = l.total + fmax * 8;
    *front_lds = with <= RV_SQMD_LDS_BUDGET;
    if (*front_lds)
        l = rv_sqmd_layout(ix->dim, b->max_pool, b->correction, b->refine_candidates,
                           (int)(fmax * 8));
    return l;
}

template <int CORR>
static int rv_sqmd_occupancy(int lds) {
    static int configured = 0;
    int blocks = 0;
    if (!configured) {
        if (rv_check(hipFuncSetAttribute((const void *)rv_sqmd_kernel<CORR>,
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         160 * 1024),
                     "hipFuncSetAttribute"))
            return -1;
        configured = 1;
    }
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, rv_sqmd_kernel<CORR>, RV_WAVE,
                                                     lds) != hipSuccess)
        return -1;
    return blocks;
}

extern "C" int rv_sqmd_grid(const RvIndex *ix, const RvSqmdBatch *b, int *lds_frontier) {
    int front_lds = 0;
    const RvSqmdLayout lay = rv_sqmd_layout_for(ix, b, &front_lds);
    const int per = b->correction == 2   ? rv_sqmd_occupancy<2>(lay.total)
                    : b->correction == 1 ? rv_sqmd_occupancy<1>(lay.total)
                                         : rv_sqmd_occupancy<0>(lay.total);
    const int cus = rv_sqmd_cus();
    if (per < 0 || cus < 0) return -1;
    if (lds_frontier) *lds_frontier = front_lds;
    int grid = cus * (per > 0 ? per : 1);
    if (grid > b->nq) grid = b->nq;
    return grid < 1 ? 1 : grid;
}

extern "C" int rv_sqmd_launch(const RvIndex *ix, const RvSqmdBatch *b, int grid_waves,
                              void *stream) {
    int front_lds = 0;
    const RvSqmdLayout lay = rv_sqmd_layout_for(ix, b, &front_lds);
    if (b->nq == 0) return 0;
    if (!ix->sq_codes || !ix->graph || !b->visited || b->visited_words % 4 ||
        b->visited_words < (ix->n + 31) / 32 || (!front_lds && !b->frontier) ||
        ((b->correction || b->refine_candidates > 0) && (!ix->res_codes || !ix->res_qparams)) ||
        (b->refine_error4 && (!ix->e4_blocks || !ix->e4_qparams || ix->e4_vpb <= 0)) ||
        (b->correction == 2 && (!ix->split_bias || ix->sq_bits != 4))) {
        snprintf(g_rv_error, sizeof(g_rv_error), "rv_sqmd_launch: bad arguments");
        return -1;
    }
    RvSqmdBatch bb = *b;
    if (front_lds) {
        bb.frontier = nullptr;
        bb.frontier_phys = b->fcap_entries + b->max_pool * b->fcap_expand;
    }
    if (grid_waves <= 0 && (grid_waves = rv_sqmd_grid(ix, b, nullptr)) < 0) return -1;
    if (grid_waves > b->nq) grid_waves = b->nq;
    if (rv_fill_zero(b->queue, sizeof(int32_t), stream)) return -1;
    hipStream_t s = (hipStream_t)stream;
    if (b->correction == 2) {
        if (rv_sqmd_occupancy<2>(lay.total) < 0) return -1;
        rv_sqmd_kernel<2><<<grid_waves, RV_WAVE, lay.total, s>>>(*ix, bb, lay);
    } else if (b->correction) {
        if (rv_sqmd_occupancy<1>(lay.total) < 0) return -1;
        rv_sqmd_kernel<1><<<grid_waves, RV_WAVE, lay.total, s>>>(*ix, bb, lay);
    } else {
        if (rv_sqmd_occupancy<0>(lay.total) < 0) return -1;
        rv_sqmd_kernel<0><<<grid_waves, RV_WAVE, lay.total, s>>>(*ix, bb, lay);
    }
    return rv_check(hipGetLastError(), "rv_sqmd_kernel launch");
}
