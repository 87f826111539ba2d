// rover_learned.hip -- training kernels for the learned early-termination
// predictor (16 -> 64 -> 64 -> 1 MLP, ReLU, sigmoid output, weighted binary
// cross-entropy).  FP32 master weights; the forward pass uses the weights
// rounded to BF16 (round to nearest even), exactly what the search kernel
// runs, and gradients flow straight through the rounding to the masters.
//
// Parameter layout (flat FP32): w1[64][16] | b1[64] | w2[64][64] | b2[64] |
// w3[64] | b3[1].
//
// One workgroup of 256 threads processes tiles of 64 samples: the tile's
// activations and the weights sit in LDS, every thread owns a fixed slice of
// the weight gradients and accumulates it over its tiles in registers, then
// writes it to the workgroup's row of a partial-sum slab; a second kernel
// adds the rows in workgroup order, so a training run is reproducible bit
// for bit (no floating-point atomics).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "rover_hip.h"

#define MLP_IN RV_MLP_IN
#define MLP_H RV_MLP_HIDDEN
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + MLP_H * MLP_IN)
#define OFF_W2 (OFF_B1 + MLP_H)
#define OFF_B2 (OFF_W2 + MLP_H * MLP_H)
#define OFF_W3 (OFF_B2 + MLP_H)
#define OFF_B3 (OFF_W3 + MLP_H)
#define MLP_PARAMS (OFF_B3 + 1)
#define TILE 64
#define THREADS 256

extern "C" int rv_mlp_param_count(void) { return MLP_PARAMS; }

__host__ __device__ static inline uint16_t rv_f32_to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x007fffffu))
        return (uint16_t)((u >> 16) | 0x0040u);  // quiet NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

__device__ static inline float rv_round_bf16(float f) {
    return __uint_as_float((uint32_t)rv_f32_to_bf16(f) << 16);
}

// x: [n][16] normalized features, y: [n] labels in {0,1}, wt: [n] sample
// weights (0 drops a sample).  idx (optional): sample ids to gather.
// Writes this workgroup's share of the gradient of sum_i wt_i * BCE_i and of
// sum_i wt_i * BCE_i to part[blockIdx.x][0..MLP_PARAMS] (the loss last);
// writes the logits to out when out != NULL.
__global__ __launch_bounds__(THREADS) void rv_mlp_fwd_bwd_kernel(
    const float *__restrict__ params, const float *__restrict__ x,
    const float *__restrict__ y, const float *__restrict__ wt,
    const int32_t *__restrict__ idx, int n, float *part, float *out, int want_grad) {
    __shared__ float W1[MLP_H * MLP_IN];
    __shared__ float W2[MLP_H * MLP_H];
    __shared__ float W3[MLP_H];
    __shared__ float B1[MLP_H], B2[MLP_H];
    __shared__ float X[TILE * MLP_IN];
    __shared__ float H1[TILE * MLP_H];
    __shared__ float H2[TILE * MLP_H];   // then dH2
    __shared__ float D1[TILE * MLP_H];   // dH1
    __shared__ float DO[TILE];
    __shared__ float LS[TILE];
    const int t = threadIdx.x;
    for (int i = t; i < MLP_H * MLP_IN; i += THREADS) W1[i] = rv_round_bf16(params[OFF_W1 + i]);
    for (int i = t; i < MLP_H * MLP_H; i += THREADS) W2[i] = rv_round_bf16(params[OFF_W2 + i]);
    if (t < MLP_H) {
        W3[t] = rv_round_bf16(params[OFF_W3 + t]);
        B1[t] = params[OFF_B1 + t];
        B2[t] = params[OFF_B2 + t];
    }
    const float b3 = params[OFF_B3];
    // gradient slices owned by this thread
    float gw2[16], gw1[4];
    float gw3 = 0.f, gb2 = 0.f, gb1 = 0.f, gb3 = 0.f, lsum = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) gw2[i] = 0.f;
#pragma unroll
    for (int i = 0; i < 4; i++) gw1[i] = 0.f;
    __syncthreads();

    const int tiles = (n + TILE - 1) / TILE;
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int s0 = tile * TILE;
        for (int i = t; i < TILE * MLP_IN; i += THREADS) {
            const int s = s0 + i / MLP_IN;
            float v = 0.f;
            if (s < n) {
                const int src = idx ? idx[s] : s;
                v = x[(size_t)src * MLP_IN + (i % MLP_IN)];
            }
            X[i] = v;
        }
        __syncthreads();
        // H1 = relu(X W1^T + b1): thread -> 16 outputs (sample r, units c..c+15)
        for (int o = t; o < TILE * MLP_H; o += THREADS) {
            const int r = o / MLP_H, j = o % MLP_H;
            float a = B1[j];
#pragma unroll
            for (int f = 0; f < MLP_IN; f++) a += W1[j * MLP_IN + f] * X[r * MLP_IN + f];
            H1[o] = a > 0.f ? a : 0.f;
        }
        __syncthreads();
        for (int o = t; o < TILE * MLP_H; o += THREADS) {
            const int r = o / MLP_H, j = o % MLP_H;
            float a = B2[j];
#pragma unroll 16
            for (int i = 0; i < MLP_H; i++) a += W2[j * MLP_H + i] * H1[r * MLP_H + i];
            H2[o] = a > 0.f ? a : 0.f;
        }
        __syncthreads();
        if (t < TILE) {
            const int s = s0 + t;
            float a = b3;
            for (int j = 0; j < MLP_H; j++) a += W3[j] * H2[t * MLP_H + j];
            float dout = 0.f, l = 0.f;
            if (s < n) {
                const int src = idx ? idx[s] : s;
                const float yy = y[src], w = wt[src];
                // stable BCE with logits: max(a,0) - a*y + log(1+exp(-|a|));
                // a record of weight 0 contributes exactly nothing
                if (w != 0.f) {
                    l = w * (fmaxf(a, 0.f) - a * yy + log1pf(expf(-fabsf(a))));
                    const float p = 1.f / (1.f + expf(-a));
                    dout = w * (p - yy);
                }
                if (out) out[s] = a;
            }
            DO[t] = dout;
            LS[t] = l;
        }
        __syncthreads();
        if (t == 0) {
            float l = 0.f;
            for (int i = 0; i < TILE; i++) l += LS[i];
            lsum += l;
        }
        if (want_grad) {
            // dH2 = dout * w3 * [H2 > 0]; gw3 and gb3 before overwriting H2
            if (t < MLP_H) {
                float a = 0.f;
                for (int r = 0; r < TILE; r++) a += DO[r] * H2[r * MLP_H + t];
                gw3 += a;
            }
            if (t == 0) {
                float a = 0.f;
                for (int r = 0; r < TILE; r++) a += DO[r];
                gb3 += a;
            }
            __syncthreads();
            for (int o = t; o < TILE * MLP_H; o += THREADS) {
                const int r = o / MLP_H, j = o % MLP_H;
                H2[o] = H2[o] > 0.f ? DO[r] * W3[j] : 0.f;
            }
            __syncthreads();
            // dH1 = (dH2 W2) * [H1 > 0]
            for (int o = t; o < TILE * MLP_H; o += THREADS) {
                const int r = o / MLP_H, i = o % MLP_H;
                float a = 0.f;
                if (H1[o] > 0.f) {
#pragma unroll 16
                    for (int j = 0; j < MLP_H; j++) a += H2[r * MLP_H + j] * W2[j * MLP_H + i];
                }
                D1[o] = a;
            }
            __syncthreads();
            // gw2[j][i] = sum_r dH2[r][j] H1[r][i]: thread owns 16 (j, i)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int p = t * 16 + e, j = p / MLP_H, i = p % MLP_H;
                float a = 0.f;
#pragma unroll 4
                for (int r = 0; r < TILE; r++) a += H2[r * MLP_H + j] * H1[r * MLP_H + i];
                gw2[e] += a;
            }
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int p = t * 4 + e, j = p / MLP_IN, f = p % MLP_IN;
                float a = 0.f;
#pragma unroll 4
                for (int r = 0; r < TILE; r++) a += D1[r * MLP_H + j] * X[r * MLP_IN + f];
                gw1[e] += a;
            }
            if (t < MLP_H) {
                float a2 = 0.f, a1 = 0.f;
                for (int r = 0; r < TILE; r++) {
                    a2 += H2[r * MLP_H + t];
                    a1 += D1[r * MLP_H + t];
                }
                gb2 += a2;
                gb1 += a1;
            }
        }
        __syncthreads();
    }
    float *row = part + (size_t)blockIdx.x * (MLP_PARAMS + 1);
    if (t == 0) row[MLP_PARAMS] = lsum;
    if (!want_grad) return;
#pragma unroll
    for (int e = 0; e < 16; e++) row[OFF_W2 + t * 16 + e] = gw2[e];
#pragma unroll
    for (int e = 0; e < 4; e++) row[OFF_W1 + t * 4 + e] = gw1[e];
    if (t < MLP_H) {
        row[OFF_W3 + t] = gw3;
        row[OFF_B2 + t] = gb2;
        row[OFF_B1 + t] = gb1;
    }
    if (t == 0) row[OFF_B3] = gb3;
}

// grad[i] += sum over the workgroups' rows, in workgroup order; the last
// column is the loss.
__global__ void rv_mlp_reduce_kernel(const float *part, int rows, float *grad, float *loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > MLP_PARAMS) return;
    if (i < MLP_PARAMS && !grad) return;
    if (i == MLP_PARAMS && !loss) return;
    float s = 0.f;
    for (int b = 0; b < rows; b++) s += part[(size_t)b * (MLP_PARAMS + 1) + i];
    if (i < MLP_PARAMS) grad[i] += s;
    else *loss += s;
}

__global__ void rv_mlp_adam_kernel(float *p, const float *g, float *m1, float *m2,
                                   float lr_t, float b1, float b2, float eps,
                                   float gscale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= MLP_PARAMS) return;
    const float gi = g[i] * gscale;
    const float a = b1 * m1[i] + (1.f - b1) * gi;
    const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
    m1[i] = a;
    m2[i] = v;
    p[i] -= lr_t * a / (sqrtf(v) + eps);
}

__global__ void rv_mlp_bf16_kernel(const float *p, uint16_t *w) {
    // Inference copy as bf16, grouped so that hidden unit j (one lane) reads
    // four consecutive inputs' weights with one 8-byte load:
    // w1g[4][64][4] | w2g[16][64][4] | w3[64], element (input i, unit j) of
    // w1/w2 at ((i / 4) * 64 + j) * 4 + i % 4.
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n1 = MLP_H * MLP_IN, n2 = MLP_H * MLP_H, n3 = MLP_H;
    if (t < n1 + n2) {
        const bool first = t < n1;
        const int e = first ? t : t - n1;
        const int g = e / (MLP_H * 4), r = e % (MLP_H * 4);
        const int j = r / 4, in = g * 4 + r % 4;
        w[t] = first ? rv_f32_to_bf16(p[OFF_W1 + j * MLP_IN + in])
                     : rv_f32_to_bf16(p[OFF_W2 + j * MLP_H + in]);
    } else if (t < n1 + n2 + n3) {
        w[t] = rv_f32_to_bf16(p[OFF_W3 + (t - n1 - n2)]);
    }
}

// Turns search traces into samples: record r of query q is valid when
// r < records[q]; its label is 1 when the record's step is at or after the
// step at which all but `slack` of the final top-K members had been scored
// (a final top-K member stays in the top-K from the step it is scored, so
// with slack 0 this is the last change of the top-K: the top-K is final).
__global__ void rv_trace_labels_kernel(const int32_t *steps, const int32_t *meta,
                                       int nq, int cap, int need, float pos_w,
                                       float neg_w, float *y, float *weight) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= (size_t)nq * cap) return;
    const int q = (int)(i / cap), r = (int)(i % cap);
    const int32_t *mq = meta + (size_t)q * RV_TRACE_META;
    const int recs = mq[0], ready = mq[2 + need];
    const bool ok = r < recs;
    const bool pos = ok && steps[i] >= ready;
    y[i] = pos ? 1.f : 0.f;
    weight[i] = ok ? (pos ? pos_w : neg_w) : 0.f;
}

// SplitMix-style hash of (seed, step, i) -> sample id.
__global__ void rv_sample_indices_kernel(int32_t *idx, int n, uint64_t population,
                                         uint32_t seed, uint32_t step) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t z = ((uint64_t)seed << 32) ^ ((uint64_t)step << 20) ^ (uint64_t)i;
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    idx[i] = (int32_t)(z % population);
}

// In-place feature normalization (x - mean) * inv_std.
__global__ void rv_normalize_kernel(float *x, size_t n, const float *mean,
                                    const float *inv_std) {
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i >= n * MLP_IN) return;
    const int f = (int)(i % MLP_IN);
    x[i] = (x[i] - mean[f]) * inv_std[f];
}

static char g_err[256];
static int chk(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
        return -1;
    }
    return 0;
}
extern "C" const char *rv_learned_error(void) { return g_err; }

extern "C" size_t rv_mlp_slab_floats(void) {
    return (size_t)RV_MLP_SLAB_ROWS * (MLP_PARAMS + 1);
}

// slab: the caller's partial-sum workspace of rv_mlp_slab_floats() floats
// (one row per workgroup); concurrent calls need their own slabs.
extern "C" int rv_mlp_fwd_bwd_sampled(const float *params, const float *x,
                                      const float *y, const float *wt,
                                      const int32_t *idx, int n, float *grad,
                                      float *loss, float *out, float *slab,
                                      void *stream) {
    if (n <= 0) return 0;
    if (!slab) {
        snprintf(g_err, sizeof(g_err), "rv_mlp_fwd_bwd: no partial-sum slab");
        return -1;
    }
    int tiles = (n + TILE - 1) / TILE;
    int grid = tiles < RV_MLP_SLAB_ROWS ? tiles : RV_MLP_SLAB_ROWS;
    rv_mlp_fwd_bwd_kernel<<<grid, THREADS, 0, (hipStream_t)stream>>>(
        params, x, y, wt, idx, n, slab, out, grad != NULL);
    if (chk(hipGetLastError(), "rv_mlp_fwd_bwd")) return -1;
    if (!grad && !loss) return 0;
    rv_mlp_reduce_kernel<<<(MLP_PARAMS + 1 + 255) / 256, 256, 0, (hipStream_t)stream>>>(
        slab, grid, grad, loss);
    return chk(hipGetLastError(), "rv_mlp_reduce");
}

extern "C" int rv_mlp_adam(float *params, const float *grad, float *m1,
                           float *m2, int step, float lr, float beta1,
                           float beta2, float eps, float grad_scale,
                           void *stream) {
    const float c1 = 1.f - powf(beta1, (float)step);
    const float c2 = 1.f - powf(beta2, (float)step);
    const float lr_t = lr * sqrtf(c2) / c1;
    rv_mlp_adam_kernel<<<(MLP_PARAMS + 255) / 256, 256, 0, (hipStream_t)stream>>>(
        params, grad, m1, m2, lr_t, beta1, beta2, eps, grad_scale);
    return chk(hipGetLastError(), "rv_mlp_adam");
}

extern "C" int rv_mlp_to_bf16(const float *params, uint16_t *w, void *stream) {
    const int n = MLP_H * MLP_IN + MLP_H * MLP_H + MLP_H;
    rv_mlp_bf16_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(params, w);
    return chk(hipGetLastError(), "rv_mlp_to_bf16");
}

extern "C" int rv_trace_labels(const int32_t *steps, const int32_t *meta, int nq,
                               int cap, int label_k, int slack, float pos_w,
                               float neg_w, float *y, float *weight, void *stream) {
    if (label_k <= 0 || label_k > RV_TRACE_TOPK || slack < 0 || slack >= label_k) {
        snprintf(g_err, sizeof(g_err), "rv_trace_labels: bad label_k/slack");
        return -1;
    }
    const size_t n = (size_t)nq * cap;
    if (n == 0) return 0;
    rv_trace_labels_kernel<<<(unsigned)((n + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        steps, meta, nq, cap, label_k - 1 - slack, pos_w, neg_w, y, weight);
    return chk(hipGetLastError(), "rv_trace_labels");
}

extern "C" int rv_sample_indices(int32_t *idx, int n, size_t population,
                                 uint32_t seed, uint32_t step, void *stream) {
    rv_sample_indices_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(
        idx, n, (uint64_t)population, seed, step);
    return chk(hipGetLastError(), "rv_sample_indices");
}

extern "C" int rv_normalize(float *x, size_t n, const float *mean,
                            const float *inv_std, void *stream) {
    const size_t total = n * MLP_IN;
    rv_normalize_kernel<<<(unsigned)((total + 255) / 256), 256, 0, (hipStream_t)stream>>>(
        x, n, mean, inv_std);
    return chk(hipGetLastError(), "rv_normalize");
}

extern "C" uint16_t rv_bf16_round_host(float f) { return rv_f32_to_bf16(f); }
